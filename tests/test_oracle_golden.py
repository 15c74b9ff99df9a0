"""The CPU oracle against the known answers of the reference's own tests.

Fixtures: tests/golden/*.json|npz, written by tests/golden/make_golden.py from
the inputs and closed forms of KoningJasper/Offshore-SPH test/*.py.
Tolerances follow the reference tests (assertAlmostEqual: 7 places absolute;
Tait 3 places), tightened where the closed form allows.
"""
import json
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN


def _json(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def test_cubic_kat():
    """test/test_kernels_cubic.py:11-72"""
    d = _json('cubic_kat.json')
    L = oracle.load()
    for e in d['evaluate']:
        v = L.ref_cubic_evaluate(e['r'], e['h'])
        assert v == pytest.approx(e['w'], rel=1e-14, abs=1e-7)
    for e in d['gradient']:
        v = L.ref_cubic_gradient(e['x'], e['r'], e['h'])
        assert v == pytest.approx(e['g'], rel=1e-14, abs=1e-7)


def test_gaussian_vs_pysph():
    """test/test_kernels_Gaussian.py:28-101 against the PySPH forms (test/PySPH/Gaussian.py:46-96)."""
    z = np.load(os.path.join(GOLDEN, 'gaussian_pysph.npz'))
    a = z['data']
    xx, xy, r, h, w, dwdq, gx, gy = a.T
    np.testing.assert_allclose(oracle.kernel_evaluate('gaussian', r, h), w, rtol=1e-13, atol=1e-7)
    np.testing.assert_allclose(oracle.gaussian_derivative(r, h), dwdq, rtol=1e-13, atol=1e-7)
    np.testing.assert_allclose(oracle.kernel_gradient('gaussian', xx, r, h), gx, rtol=1e-13, atol=1e-7)
    np.testing.assert_allclose(oracle.kernel_gradient('gaussian', xy, r, h), gy, rtol=1e-13, atol=1e-7)


def test_momentum_kat():
    """test/test_numba_momentum.py:21-84: pressure term over 10,000 linspace records."""
    z = np.load(os.path.join(GOLDEN, 'momentum_kat.npz'))
    num = int(z['num'])
    lin = np.linspace(*z['lin'], num)
    comp = np.zeros(num, dtype=oracle.computed_dtype)
    comp['p'] = np.linspace(*z['p_lin'], num)
    comp['rho'] = float(z['rho'])
    comp['h'] = float(z['h'])
    for f in ('x', 'y', 'dw_x', 'dw_y', 'vx', 'vy'):
        comp[f] = lin
    comp['r'] = np.sqrt(2.0) * lin
    comp['m'] = 1.0
    p = np.zeros(1, dtype=oracle.particle_dtype)
    p['p'], p['rho'], p['h'], p['c'] = z['self']
    p['m'] = 1.0
    a = oracle.momentum(0.01, 0.0, p, comp)
    np.testing.assert_allclose(a, z['a'], rtol=1e-12)


def test_continuity_kat():
    """test/test_numba_continuity.py:12-52: 10^7 linspace records, closed-form sum."""
    d = _json('continuity_kat.json')
    n = d['n']
    lin = np.linspace(d['lo'], d['hi'], n)
    comp = np.zeros(n, dtype=oracle.computed_dtype)
    for f in ('vx', 'vy', 'dw_x', 'dw_y'):
        comp[f] = lin
    comp['m'] = d['m']
    arho = oracle.continuity(comp)
    assert arho / d['arho'] == pytest.approx(1.0, abs=10 ** -d['ratio_places'])
    del comp


def test_tait_kat():
    """test/test_numba_taiteos.py:11-46 (3 decimal places)."""
    z = np.load(os.path.join(GOLDEN, 'tait_kat.npz'))
    gamma, B, rho0, co = z['params']
    L = oracle.load()
    assert L.ref_tait_co(2.0) == pytest.approx(co, rel=1e-15)
    assert L.ref_tait_B(co, rho0, gamma) == pytest.approx(B, rel=1e-15)
    p = oracle.tait(gamma, B, rho0, z['rho'], np.zeros(len(z['rho']), np.int8))
    np.testing.assert_allclose(p, z['p'], rtol=1e-12, atol=1e-3)
    # non-fluid particles have zero Tait pressure (TaitEOS.py:27-28)
    assert oracle.tait(gamma, B, rho0, [1100.0], [1])[0] == 0.0


def test_boundary_force_kat():
    """test/test_eq_boundary.py:9-26; a_x = D (2^4 - 2^2) = 588.6."""
    d = _json('boundary_kat.json')
    comp = np.zeros(len(d['records']), dtype=oracle.computed_dtype)
    for k, rec in enumerate(d['records']):
        for f, v in rec.items():
            comp[k][f] = v
    a = oracle.boundary_force(d['r0'], d['D'], d['p1'], d['p2'], comp)
    assert a[0] > 0 and a[1] == 0
    np.testing.assert_allclose(a, d['a'], rtol=1e-15)


def _particle(d, label):
    p = np.zeros(1, dtype=oracle.particle_dtype)
    p['label'] = label
    for f, v in d['particle'].items():
        p[f] = v
    return p


def test_pec_kat():
    """test/test_integrators_pec.py:30-45"""
    d = _json('integrators_kat.json')
    p = _particle(d, 0)
    oracle.pec_predict(p, d['dt'], 0.0)
    for f, v in d['pec_predict'].items():
        assert p[f][0] == pytest.approx(v, abs=1e-7)
    oracle.pec_correct(p, d['dt'], 0.0)
    for f, v in d['pec_correct'].items():
        assert p[f][0] == pytest.approx(v, abs=1e-7)


def test_euler_kat_and_pec_equivalence():
    """test/test_integrators_euler.py:26-38; test/test_integrators_pec.py:47-65 (Euler == PEC
    when the forces are not recomputed)."""
    d = _json('integrators_kat.json')
    p = _particle(d, 0)
    oracle.euler_correct(p, d['dt'])
    for f, v in d['euler_correct'].items():
        assert p[f][0] == pytest.approx(v, abs=1e-7)
    a = _particle(d, 0)
    b = _particle(d, 0)
    oracle.pec_predict(a, d['dt'])
    oracle.pec_correct(a, d['dt'])
    oracle.euler_correct(b, d['dt'])
    for f in ('x', 'y', 'vx', 'vy', 'rho'):
        assert a[f][0] == pytest.approx(b[f][0], abs=1e-7)


def test_linked_list_kat():
    """test/test_linked_list.py:12-79: scale 3 on a 26x26 unit lattice gives the full q <= 3 ball."""
    d = _json('linked_list_kat.json')
    N = d['N']
    xv = np.linspace(0, 25, N + 1)
    X, Y = np.meshgrid(xv, xv, indexing='ij')
    pA = np.zeros(X.size, dtype=oracle.particle_dtype)
    pA['x'] = X.ravel()
    pA['y'] = Y.ravel()
    pA['h'] = d['h']
    nn = oracle.NN(d['scale'])
    nn.update(pA)
    dump = nn.dump(pA)
    near0 = dump['idx'][dump['off'][0]:dump['off'][1]]
    assert sorted(near0.tolist()) == d['near0']
    for k in d['must_contain']:
        assert k in near0


def test_timestep_kat():
    """src/Equations/TimeStep.py:40-56: dt = min(courant, force) over the fluid."""
    for case in _json('timestep_kat.json'):
        n = len(case['h'])
        pA = np.zeros(n + 1, dtype=oracle.particle_dtype)
        for f in ('h', 'c', 'ax', 'ay'):
            pA[f][:n] = case[f]
        pA['label'][n] = 1           # a wall particle with extreme values must be ignored
        pA['h'][n] = 1e-9
        pA['ax'][n] = 1e9
        out = oracle.timestep(pA)
        np.testing.assert_allclose(out, [case['dt'], case['dt_c'], case['dt_f']], rtol=1e-15)
