"""Settings of the reference's surface that the dam-break configs leave at
their defaults, each against the CPU oracle:
  * a non-zero background pressure Pb (src/Methods/WCSPH.py:149);
  * a non-zero quadratic viscosity beta (src/Equations/Momentum.py:48) on
    an approaching flow;
  * PEC's strict density clamp (src/Integrators/PEC.py:13,57-58,87-88) on a
    state that drives densities below zero;
  * the Gaussian and the float pair pass on the 200 x 200 stress lattice,
    where the Gaussian's 3h support is cut by the 3 x 3 cell stencil.
"""
import numpy as np
import pytest

import oracle
from gpu_util import FIELDS, diverging as _diverging, make_engine, oracle_loop, perturbed, rel_err, say
from hipsph import scenarios
from hipsph._lib import SPH_FLAG_ZERO_DENSITY

pytestmark = pytest.mark.gpu


def _steps_both(pA, m, steps, h, strict=False):
    e = make_engine(pA, m, strict=strict)
    e.step(steps)
    st = e.status()
    out = e.download()
    hist = e.dt_history(steps)
    e.close()
    s = oracle.Solver(pA.copy(), m, 'cubic', 'pec', h=h, strict=strict, nthreads=8)  # it steps in place
    rhist = s.run(steps)
    return out, st, hist, s, rhist


def test_background_pressure_fp64_bitexact():
    """Pb = 2e4 Pa on every pressure (Tait + Pb for fluid, Pb for the rest):
    200 steps bit-identical, and different from Pb = 0."""
    sc = scenarios.dam_break_c1()
    pA = perturbed(sc.pA, 11)
    m = sc.method()
    m.Pb = 2.0e4
    say('edges: Pb', m.Pb)
    out, st, hist, s, rhist = _steps_both(pA, m, 200, sc.h)
    assert st['steps'] == 200 and st['flags'] & 3 == 0
    np.testing.assert_array_equal(hist, rhist)
    assert st['t'] == s.t
    for k in FIELDS:
        np.testing.assert_array_equal(out[k], s.pA[k], err_msg=k)
    base, _, _, _, _ = _steps_both(pA, sc.method(), 200, sc.h)
    assert not np.array_equal(base['ax'], out['ax'])


def test_quadratic_viscosity_beta_fp64_bitexact():
    """beta = 1.5 with random velocities, so many pairs approach (v.x < 0)
    and the mu^2 term is live: 200 steps bit-identical, and different from
    beta = 0."""
    sc = scenarios.dam_break_c1()
    pA = perturbed(sc.pA, 12, vscale=1.0)
    m = sc.method()
    m.beta = 1.5
    say('edges: beta', m.beta)
    out, st, hist, s, rhist = _steps_both(pA, m, 200, sc.h)
    assert st['steps'] == 200 and st['flags'] & 3 == 0
    np.testing.assert_array_equal(hist, rhist)
    for k in FIELDS:
        np.testing.assert_array_equal(out[k], s.pA[k], err_msg=k)
    base, _, _, _, _ = _steps_both(pA, sc.method(), 200, sc.h)
    assert not np.array_equal(base['ax'], out['ax'])


def test_pec_strict_clamp_fp64_vs_oracle():
    """PEC(strict=True), the reference's default. Step 1: the corrector drives
    the centres' densities below zero and the clamp sets them to exactly 0;
    the whole state is bit-identical to the oracle's, and without the clamp
    (strict=False) the same densities come out negative. Step 2: the
    predictor clamps them again, p / rho^2 is -inf at the pair pass, and the
    reference's sums turn NaN (its zero-gradient self term is -inf x 0); the
    device raises SPH_FLAG_ZERO_DENSITY in that step and halts the next."""
    sc = scenarios.dam_break_c1()
    f = np.nonzero(sc.pA['label'] == 0)[0]
    centres = [int(f[len(f) // 3]), int(f[2 * len(f) // 3])]
    pA = _diverging(sc, centres)
    m = sc.method()
    say('edges: strict clamp, centres', centres)
    out, st, hist, s, rhist = _steps_both(pA, m, 1, sc.h, strict=True)
    assert st['flags'] & SPH_FLAG_ZERO_DENSITY == 0
    np.testing.assert_array_equal(hist, rhist)
    for k in FIELDS:
        np.testing.assert_array_equal(out[k], s.pA[k], err_msg=k)
    assert np.all(out['rho'][centres] == 0.0)
    loose, _, _, _, _ = _steps_both(pA, m, 1, sc.h, strict=False)
    assert np.all(loose['rho'][centres] < 0.0)
    # step 2: the predicted density is clamped to 0 before the pair pass
    e = make_engine(pA, m, strict=True)
    e.step(1)
    e.step(1)
    st2 = e.status()
    e.step(1)
    st3 = e.status()
    e.close()
    s.run(1)
    say('edges: strict step 2 flags', st2['flags'], 'oracle ax at centres', s.pA['ax'][centres])
    assert st2['steps'] == 2 and st2['flags'] & SPH_FLAG_ZERO_DENSITY
    assert st3['steps'] == 2  # halted
    assert np.all(np.isnan(s.pA['ax'][centres]))


@pytest.mark.parametrize('kernel,precision', [('gaussian', 'fp64'), ('cubic', 'fp32c'), ('gaussian', 'fp32c')])
def test_stress_lattice_200_vs_oracle(kernel, precision):
    """One pair pass on the jittered 200 x 200 lattice (40,000 fluid in a
    closed box, velocities N(0, (0.05 c0)^2)): the Gaussian in double within
    1e-12 (the device's exp), the float pass within 2e-5, both relative to
    the largest magnitude of each output."""
    sc = scenarios.stress_lattice(nside=200)
    m = sc.method()
    say('stress 200^2', kernel, precision)
    e = make_engine(sc.pA, m, kernel=kernel, precision=precision)
    e.nn_update()
    e.compute()
    out = e.download()
    e.close()
    ref, _ = oracle_loop(sc.pA, m, kernel, nthreads=16)
    f = sc.pA['label'] == 0
    tol = 1e-12 if precision == 'fp64' else 2e-5
    errs = {k: rel_err(out[k], ref[k], f) for k in ('drho', 'ax', 'ay')}
    say('stress 200^2 errors', kernel, precision, errs)
    for k, v in errs.items():
        assert v <= tol, (k, v)


def test_fp32c_out_of_grid_targets_vs_oracle():
    """One h and one mass for every particle (the float pass's uniform form),
    with targets whose own cell is outside the grid (x on the far edge: the
    key aliases, and the corner's key lies past the grid): each keeps its
    previous rates only when it has no neighbour at all, as in the
    reference; the rates of the others agree with the oracle to the float
    pass's 2e-5."""
    sc = scenarios.dam_break_c1()
    x = np.array([0.0, 1.5, 2.5, 4.0, 0.2, 4.0, 3.9, 3.7, 1.2, 1.7])
    y = np.array([0.0, 0.5, 1.5, 0.5, 1.6, 2.0, 1.9, 0.6, 0.4, 0.6])
    pA = np.zeros(len(x), dtype=sc.pA.dtype)
    pA['x'], pA['y'], pA['h'], pA['rho'], pA['m'] = x, y, 0.5, 1000.0, 1.0
    pA['ax'], pA['ay'], pA['drho'] = 7.0, -3.0, 11.0  # kept where a target finds no neighbour
    pA['vx'] = np.linspace(-1.0, 1.0, len(x))
    m = sc.method()
    e = make_engine(pA, m, precision='fp32c')
    e.nn_update()
    e.compute()
    out = e.download()
    e.close()
    ref, _ = oracle_loop(pA, m, 'cubic')
    say('fp32c out of grid: ax', out['ax'], 'oracle', ref['ax'])
    for k in ('drho', 'ax', 'ay'):
        kept = ref[k] == pA[k]
        np.testing.assert_array_equal(out[k][kept], pA[k][kept], err_msg=k)
        assert rel_err(out[k], ref[k], ~kept) < 2e-5, k
