"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle.

Bars (stated per test):
* neighbour search: grid, flat keys and every near(i) list bit-exact,
  including order, q and r;
* FP64 pair pass, cubic spline and Wendland C4: bit-exact (same operations,
  same order, no contraction on either side); Gaussian: within 1e-12 of the
  largest magnitude (its exp comes from a different math library);
* FP32C pair pass: within 2e-5 of the largest magnitude per field;
* multi-step trajectories: FP64 cubic bit-exact after 1,000 steps (C1),
  100 steps (C2, 500k) and 20 steps of the 1M headline workload (with FP32C
  pairs there within 1e-7 relative density), Wendland after 200.
"""
import json
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN
from gpu_util import FIELDS, make_engine, oracle_loop, perturbed, rel_err, say
from hipsph import scenarios
from hipsph.Kernels.CubicSpline import CubicSpline
from hipsph.Kernels.Gaussian import Gaussian
from hipsph.Kernels.Wendland import Wendland

pytestmark = pytest.mark.gpu


# --------------------------------------------------------------------------
# element-wise plugin kernels
# --------------------------------------------------------------------------
def test_kernels_match_oracle_and_golden():
    rng = np.random.default_rng(0)
    n = 20000
    h = rng.uniform(0.01, 0.05, n)
    r = rng.uniform(0, 3.2, n) * h
    r[:50] = 0.0
    x = rng.uniform(-1, 1, n) * r
    for K, name, exact in ((CubicSpline(), 'cubic', True), (Wendland(), 'wendland', True),
                           (Gaussian(), 'gaussian', False)):
        say('kernel', name)
        w = K.evaluate(r, h)
        g = K.gradient(x, r, h)
        wr = oracle.kernel_evaluate(name, r, h)
        gr = oracle.kernel_gradient(name, x, r, h)
        if exact:
            np.testing.assert_array_equal(w, wr)
            np.testing.assert_array_equal(g, gr)
        else:
            np.testing.assert_allclose(w, wr, rtol=1e-14, atol=0)
            np.testing.assert_allclose(g, gr, rtol=1e-14, atol=0)
    with open(os.path.join(GOLDEN, 'cubic_kat.json')) as f:
        kat = json.load(f)
    K = CubicSpline()
    for e in kat['evaluate']:
        assert K.evaluate([e['r']], [e['h']])[0] == pytest.approx(e['w'], rel=1e-14, abs=1e-7)
    for e in kat['gradient']:
        assert K.gradient([e['x']], [e['r']], [e['h']])[0] == pytest.approx(e['g'], rel=1e-14, abs=1e-7)
    z = np.load(os.path.join(GOLDEN, 'gaussian_pysph.npz'))['data']
    G = Gaussian()
    np.testing.assert_allclose(G.evaluate(z[:, 2], z[:, 3]), z[:, 4], rtol=1e-13, atol=1e-7)
    np.testing.assert_allclose(G.derivative(z[:, 2], z[:, 3]), z[:, 5], rtol=1e-13, atol=1e-7)
    np.testing.assert_allclose(G.gradient(z[:, 0], z[:, 2], z[:, 3]), z[:, 6], rtol=1e-13, atol=1e-7)


def test_tait_matches_oracle_and_golden():
    from hipsph.Equations.TaitEOS import TaitEOS
    z = np.load(os.path.join(GOLDEN, 'tait_kat.npz'))
    gamma, B, rho0, _ = z['params']
    lab = np.zeros(len(z['rho']), np.int8)
    lab[::7] = 1
    p = TaitEOS(gamma, B, rho0, z['rho'], lab)
    np.testing.assert_array_equal(p, oracle.tait(gamma, B, rho0, z['rho'], lab))
    f = lab == 0
    np.testing.assert_allclose(p[f], z['p'][f], rtol=1e-12, atol=1e-3)


# --------------------------------------------------------------------------
# neighbour search
# --------------------------------------------------------------------------
def _nn_parity(pA, method, scale=2.0):
    e = make_engine(pA, method, scale=scale)
    e.nn_update()
    g = e.grid()
    d = e.nn_dump()
    nn = oracle.NN(scale)
    a = oracle.as_particles(pA)
    nn.update(a)
    gr = nn.grid()
    for k in ('xmin', 'xmax', 'ymin', 'ymax', 'cell_size', 'ncx', 'ncy', 'n_cells'):
        assert g[k] == gr[k], k
    dr = nn.dump(a)
    np.testing.assert_array_equal(d['keys'], dr['keys'])
    np.testing.assert_array_equal(d['off'], dr['off'])
    np.testing.assert_array_equal(d['idx'], dr['idx'])
    np.testing.assert_array_equal(d['q'], dr['q'])
    np.testing.assert_array_equal(d['r'], dr['r'])
    e.close()
    return d


def test_nn_dam_break_c1():
    sc = scenarios.dam_break_c1()
    say('nn c1', sc.n)
    d = _nn_parity(sc.pA, sc.method())
    f = sc.pA['label'] == 0
    cnt = np.diff(d['off'])
    assert cnt[f].min() >= 1 and 35 < cnt[f].mean() < 50


@pytest.mark.parametrize('seed', [1, 2])
def test_nn_random_mixed_h(seed):
    rng = np.random.default_rng(seed)
    n = 3000
    sc = scenarios.dam_break_c1()
    pA = np.zeros(n, dtype=sc.pA.dtype)
    pA['x'] = rng.uniform(0, 1.0, n)
    pA['y'] = rng.uniform(0, 0.7, n)
    pA['h'] = rng.uniform(0.01, 0.02, n)
    pA['rho'] = 1000.0
    pA['m'] = 0.1
    pA['label'] = (rng.uniform(0, 1, n) < 0.1).astype(np.int8)
    say('nn random', seed)
    _nn_parity(pA, sc.method())


def test_nn_edge_alias_and_out_of_grid():
    """x on the far edge gives cx = ncx: the key aliases to cell (0, cy+1); the corner particle
    gets a key past the grid and is seen by no query (the reference writes it out of bounds)."""
    sc = scenarios.dam_break_c1()
    x = np.array([0.0, 1.5, 2.5, 4.0, 0.2, 4.0, 3.9])
    y = np.array([0.0, 0.5, 1.5, 0.5, 1.6, 2.0, 1.9])
    pA = np.zeros(len(x), dtype=sc.pA.dtype)
    pA['x'], pA['y'], pA['h'], pA['rho'], pA['m'] = x, y, 0.5, 1000.0, 1.0
    say('nn alias')
    d = _nn_parity(pA, sc.method())
    assert d['keys'][3] == 4          # (4, 0) -> cell (0, 1)
    assert d['keys'][5] == 4 + 4 * 2  # (4, 2) -> past the 4 x 2 grid


def test_nn_scale3_kat():
    """test/test_linked_list.py contract through the device search (scale 3)."""
    with open(os.path.join(GOLDEN, 'linked_list_kat.json')) as f:
        kat = json.load(f)
    sc = scenarios.dam_break_c1()
    xv = np.linspace(0, 25, kat['N'] + 1)
    X, Y = np.meshgrid(xv, xv, indexing='ij')
    pA = np.zeros(X.size, dtype=sc.pA.dtype)
    pA['x'], pA['y'], pA['h'], pA['rho'], pA['m'] = X.ravel(), Y.ravel(), kat['h'], 1000.0, 1.0
    d = _nn_parity(pA, sc.method(), scale=kat['scale'])
    assert sorted(d['idx'][d['off'][0]:d['off'][1]].tolist()) == kat['near0']


def test_nn_stress_lattice_small():
    sc = scenarios.stress_lattice(nside=300)
    say('nn stress 300^2', sc.n)
    _nn_parity(sc.pA, sc.method())


# --------------------------------------------------------------------------
# fused pair pass
# --------------------------------------------------------------------------
@pytest.mark.parametrize('kernel', ['cubic', 'gaussian', 'wendland'])
def test_compute_fp64_vs_oracle(kernel):
    sc = scenarios.dam_break_c1()
    pA = perturbed(sc.pA, 3)
    m = sc.method()
    say('compute fp64', kernel)
    e = make_engine(pA, m, kernel=kernel)
    e.nn_update()
    e.compute()
    out = e.download()
    ref, _ = oracle_loop(pA, m, kernel)
    f = pA['label'] == 0
    for k in ('drho', 'ax', 'ay'):
        if kernel != 'gaussian':  # the Gaussian's exp is the device's, not glibc's
            np.testing.assert_array_equal(out[k][f], ref[k][f], err_msg=k)
        else:
            assert rel_err(out[k], ref[k], f) < 1e-12, k
    # walls are not touched by the pass
    np.testing.assert_array_equal(out['ax'][~f], pA['ax'][~f])
    e.close()


@pytest.mark.parametrize('precision', ['fp64', 'fp32c'])
@pytest.mark.parametrize('kernel', ['cubic', 'gaussian', 'wendland'])
def test_uniform_pass_matches_general(kernel, precision):
    """One h and one fluid mass: the uniform pass (sph_compute) against the
    general pass on the same records (sph_compute_counts takes the general
    one): FP64 bit-identical (same expressions, same order; for the Gaussian
    also the exact q <= 3 cut at the cut-off q_max = 3), FP32C within 2e-6 of
    the largest value (the general pass sums in float with TwoSum, the
    uniform one in double)."""
    sc = scenarios.dam_break_c1()
    pA = perturbed(sc.pA, 11)
    m = sc.method()
    say('uniform vs general', kernel, precision)
    e = make_engine(pA, m, kernel=kernel, precision=precision)
    e.nn_update()
    assert e.pair_form() == precision + '-uniform'
    e.compute()
    uni = e.download()
    e.compute_counts()
    gen = e.download()
    e.close()
    f = pA['label'] == 0
    for k in ('drho', 'ax', 'ay'):
        if precision == 'fp64':
            np.testing.assert_array_equal(uni[k][f], gen[k][f], err_msg=k)
        else:
            err = rel_err(uni[k], gen[k], f)
            say(k, err)
            assert err < 2e-6, (k, err)


@pytest.mark.parametrize('kernel', ['cubic', 'gaussian', 'wendland'])
def test_compute_fp32c_vs_oracle(kernel):
    sc = scenarios.dam_break_c1()
    pA = perturbed(sc.pA, 4)
    m = sc.method()
    say('compute fp32c', kernel)
    e = make_engine(pA, m, kernel=kernel, precision='fp32c')
    e.nn_update()
    e.compute()
    out = e.download()
    ref, _ = oracle_loop(pA, m, kernel)
    f = pA['label'] == 0
    errs = {k: rel_err(out[k], ref[k], f) for k in ('drho', 'ax', 'ay')}
    say('fp32c rel errors', kernel, errs)
    for k, v in errs.items():
        assert v < 2e-5, (k, v)
    e.close()


@pytest.mark.parametrize('precision', ['fp64', 'fp32c'])
@pytest.mark.parametrize('kernel', ['cubic', 'gaussian', 'wendland'])
def test_split_passes_match_fused_and_oracle(kernel, precision):
    """sph_continuity writes drho only and sph_momentum ax, ay only; each
    equals the fused pass (same expressions, same order) and so the oracle.
    One h and one fluid mass: every kernel takes the uniform form (split and
    fused alike)."""
    sc = scenarios.dam_break_c1()
    pA = perturbed(sc.pA, 5)
    m = sc.method()
    say('split passes', kernel, precision)
    fused = make_engine(pA, m, kernel=kernel, precision=precision)
    fused.nn_update()
    fused.compute()
    want = fused.download()
    fused.close()
    e = make_engine(pA, m, kernel=kernel, precision=precision)
    e.nn_update()
    assert e.pair_form() == precision + '-uniform'
    e.continuity()
    c = e.download()
    e.momentum()
    mo = e.download()
    e.close()
    f = pA['label'] == 0
    np.testing.assert_array_equal(c['ax'], pA['ax'])   # continuity leaves ax, ay alone
    np.testing.assert_array_equal(c['ay'], pA['ay'])
    np.testing.assert_array_equal(mo['drho'], c['drho'])  # momentum leaves drho alone
    if precision == 'fp64':
        np.testing.assert_array_equal(c['drho'][f], want['drho'][f])
        np.testing.assert_array_equal(mo['ax'][f], want['ax'][f])
        np.testing.assert_array_equal(mo['ay'][f], want['ay'][f])
    else:
        for k, got in (('drho', c['drho']), ('ax', mo['ax']), ('ay', mo['ay'])):
            err = rel_err(got, want[k], f)
            say('split vs fused', k, err)
            assert err < 1e-6, k
    ref, _ = oracle_loop(pA, m, kernel)
    tol = 0 if (precision == 'fp64' and kernel != 'gaussian') else (1e-12 if precision == 'fp64' else 2e-5)
    assert rel_err(c['drho'], ref['drho'], f) <= tol
    assert rel_err(mo['ax'], ref['ax'], f) <= tol


def test_compute_stress_lattice_fp64():
    sc = scenarios.stress_lattice(nside=200)
    m = sc.method()
    say('compute stress 200^2')
    e = make_engine(sc.pA, m)
    e.nn_update()
    e.compute()
    out = e.download()
    ref, _ = oracle_loop(sc.pA, m, 'cubic')
    f = sc.pA['label'] == 0
    for k in ('drho', 'ax', 'ay'):
        np.testing.assert_array_equal(out[k][f], ref[k][f], err_msg=k)
    e.close()


# --------------------------------------------------------------------------
# whole steps
# --------------------------------------------------------------------------
def _run_both(sc, steps, kernel='cubic', integrator='pec', precision='fp64', graph=False):
    m = sc.method()
    e = make_engine(sc.pA, m, kernel=kernel, integrator=integrator, precision=precision)
    e.step(steps, graph=graph)
    st = e.status()
    out = e.download()
    hist = e.dt_history(steps)
    e.close()
    s = oracle.Solver(sc.pA, m, kernel, integrator, h=sc.h, nthreads=8)
    rhist = s.run(steps)
    return out, hist, st, s.pA, rhist, s


@pytest.mark.parametrize('case', ['dam_break_c2', 'lattice_h3'])
def test_staged_walk_matches_global_walk(case, precision='fp32c'):
    """The float passes with each workgroup's neighbour records staged in
    LDS (the default) against the same passes reading them from global
    memory: every field bit-identical after 30 PEC steps (graph replay), then
    one plain and one split pass. The dam break stages every workgroup; the
    lattice with h = 3 dx (about 36 particles per cell) puts some
    workgroups' ranges over the stage, which then walk global memory."""
    sc = (scenarios.dam_break_c2() if case == 'dam_break_c2'
          else scenarios.stress_lattice(nside=160, hfac=3.0))
    m = sc.method()
    out = []
    for staging in (True, False):
        say('staged walk', case, precision, staging)
        e = make_engine(sc.pA, m, precision=precision)
        assert e.pair_form() == precision + '-uniform'
        e.set_pair_staging(staging)
        e.step(30, graph=True)
        st = e.status()
        assert st['steps'] == 30 and st['flags'] & 3 == 0
        d = e.download()
        hist = e.dt_history(30)
        e.nn_update()
        e.compute()
        dc = e.download()
        e.momentum()
        dm = e.download()
        out.append((d, hist, dc, dm))
        e.close()
    np.testing.assert_array_equal(out[0][1], out[1][1])
    for k in FIELDS:
        np.testing.assert_array_equal(out[0][0][k], out[1][0][k], err_msg=k)
    for k in ('drho', 'ax', 'ay'):
        np.testing.assert_array_equal(out[0][2][k], out[1][2][k], err_msg='compute ' + k)
        np.testing.assert_array_equal(out[0][3][k], out[1][3][k], err_msg='momentum ' + k)


@pytest.mark.parametrize('precision,speed', [('fp64', 0.0), ('fp32c', 0.0), ('fp64', 5000.0)])
def test_border_bounds_match_full_scan(precision, speed):
    """The fused PEC step's neighbour update takes the grid's bounds from the
    border cells once the previous corrector's displacement bound allows:
    the same trajectory, dt series and grid as scanning every particle, over
    41 steps (the first after the upload scans everything, the other 40 take
    the border). Then a fluid particle is moved past the right wall by the
    host (sph_state_scatter): the next update scans everything again, and
    the two engines still agree after 10 more steps. With the particle sent
    off at 5000 m/s (dt v more than a cell) the bound refuses the border
    path on every one of those steps. A host predictor call (sph_predict)
    also sends the next update to the full scan."""
    sc = scenarios.dam_break_c2()
    m = sc.method()
    fluid = np.flatnonzero(sc.pA['label'] == 0)
    mover = int(fluid[len(fluid) // 2])
    out = []
    for border in (True, False):
        say('border bounds', precision, border)
        e = make_engine(sc.pA, m, precision=precision)
        assert e.pair_form() == precision + '-uniform'
        e.set_border_bounds(border)
        e.step(41, graph=True)
        st = e.status()
        assert st['steps'] == 41 and st['flags'] & 3 == 0
        n1 = e.border_count()
        d = e.download()
        g1 = e.grid()
        xmax = g1['xmax']
        e.scatter([mover], {'x': [xmax + 0.3], 'vx': [speed], 'vy': [0.0]})
        e.step(1, graph=True)
        n2 = e.border_count()
        e.step(10, graph=True)
        n3 = e.border_count()
        e.predict()  # moves every fluid particle: the next update scans everything
        e.step(3, graph=True)
        n4 = e.border_count()
        st = e.status()
        assert st['steps'] == 55 and st['flags'] & 3 == 0
        d2 = e.download()
        g2 = e.grid()
        hist = e.dt_history(55)
        out.append((d, g1, d2, g2, hist))
        e.close()
        if border:
            want = (40, 40, 50, 52) if speed == 0.0 else (40, 40, 40, 40)
            assert (n1, n2, n3, n4) == want, (n1, n2, n3, n4)
            assert g2['xmax'] >= xmax + 0.29
        else:
            assert (n1, n2, n3, n4) == (0, 0, 0, 0)
    assert out[0][1] == out[1][1] and out[0][3] == out[1][3]
    np.testing.assert_array_equal(out[0][4], out[1][4])
    for k in FIELDS:
        np.testing.assert_array_equal(out[0][0][k], out[1][0][k], err_msg=k)
        np.testing.assert_array_equal(out[0][2][k], out[1][2][k], err_msg='after the scatter ' + k)


def test_step_c1_fp64_bitexact_1000():
    """C1 (dx 0.02 dam break), cubic, PEC, 1,000 steps: every field bit-identical to the oracle."""
    sc = scenarios.dam_break_c1()
    say('c1 1000 steps fp64')
    out, hist, st, ref, rhist, s = _run_both(sc, 1000)
    assert st['steps'] == 1000 and st['flags'] & 3 == 0
    np.testing.assert_array_equal(hist, rhist)
    assert st['t'] == s.t
    for k in FIELDS:
        np.testing.assert_array_equal(out[k], ref[k], err_msg=k)


def test_step_c2_500k_fp64_bitexact_100():
    """C2 at 500k fluid (1.0 x 2.0 column, dx 0.002, 4 x 3 tank), cubic, PEC,
    100 steps with the graph replay: every field and the dt series bit-identical
    to the oracle (its pair loop on 16 host threads)."""
    sc = scenarios.dam_break_c2(million=False)
    m = sc.method()
    say('c2 500k: n', sc.n, '100 steps fp64')
    e = make_engine(sc.pA, m)
    e.step(100, graph=True)
    st = e.status()
    out = e.download()
    hist = e.dt_history(100)
    e.close()
    assert st['steps'] == 100 and st['flags'] & 3 == 0
    s = oracle.Solver(sc.pA, m, 'cubic', 'pec', h=sc.h, nthreads=16)
    rhist = s.run(100)
    say('c2 500k oracle done')
    np.testing.assert_array_equal(hist, rhist)
    assert st['t'] == s.t
    for k in FIELDS:
        np.testing.assert_array_equal(out[k], s.pA[k], err_msg=k)


def test_step_c2_1M_headline_vs_oracle():
    """The headline workload itself (1M dam break: 2.0 x 2.5 column at dx
    0.002, 1,265,021 particles), 20 steps replayed from the graph, both pair
    precisions against the oracle (its pair loop on 16 host threads):
    FP64 every field and the dt series bit-identical; FP32C pairs (FP64
    integrator) within 1e-7 relative density, 1e-6 dx in position and 1e-7
    c0 in velocity over the fluid, walls identical."""
    sc = scenarios.dam_break_c2(million=True)
    m = sc.method()
    steps = 20
    s = oracle.Solver(sc.pA.copy(), m, 'cubic', 'pec', h=sc.h, nthreads=16)  # it steps in place
    rhist = s.run(steps)
    say('c2 1M oracle done: n', sc.n)
    f = sc.pA['label'] == 0
    for prec in ('fp64', 'fp32c'):
        e = make_engine(sc.pA, m, precision=prec)
        assert e.pair_form() == prec + '-uniform'
        e.step(steps, graph=True)
        st = e.status()
        out = e.download()
        hist = e.dt_history(steps)
        e.close()
        assert st['steps'] == steps and st['flags'] & 3 == 0
        if prec == 'fp64':
            np.testing.assert_array_equal(hist, rhist)
            assert st['t'] == s.t
            for k in FIELDS:
                np.testing.assert_array_equal(out[k], s.pA[k], err_msg=k)
        else:
            e_rho = float(np.max(np.abs(out['rho'][f] - s.pA['rho'][f]) / s.pA['rho'][f]))
            e_x = float(np.max(np.hypot(out['x'][f] - s.pA['x'][f], out['y'][f] - s.pA['y'][f]))) / sc.dx
            e_v = float(np.max(np.hypot(out['vx'][f] - s.pA['vx'][f], out['vy'][f] - s.pA['vy'][f]))) / m.co
            say('c2 1M fp32c vs oracle: rho', e_rho, 'x/dx', e_x, 'v/c0', e_v)
            assert e_rho <= 1e-7 and e_x <= 1e-6 and e_v <= 1e-7
            for k in ('x', 'y', 'vx', 'vy', 'rho'):
                np.testing.assert_array_equal(out[k][~f], s.pA[k][~f], err_msg=k)


def test_step_wendland_fp64_bitexact():
    """Wendland C4 (the reference's DamBreak example kernel), PEC, 200 steps:
    every field and the dt series bit-identical to the oracle."""
    sc = scenarios.dam_break_c1()
    say('c1 wendland 200 steps fp64')
    out, hist, st, ref, rhist, s = _run_both(sc, 200, kernel='wendland')
    np.testing.assert_array_equal(hist, rhist)
    for k in FIELDS:
        np.testing.assert_array_equal(out[k], ref[k], err_msg=k)


def test_step_euler_and_gaussian():
    sc = scenarios.dam_break_c1()
    say('c1 euler 100')
    out, hist, st, ref, rhist, _ = _run_both(sc, 100, integrator='euler')
    for k in FIELDS:
        np.testing.assert_array_equal(out[k], ref[k], err_msg=k)
    say('c1 gaussian 100')
    out, hist, st, ref, rhist, _ = _run_both(sc, 100, kernel='gaussian')
    np.testing.assert_allclose(hist, rhist, rtol=1e-9)
    f = sc.pA['label'] == 0
    assert np.max(np.abs(out['x'] - ref['x'])) < 1e-9 * sc.dx
    assert rel_err(out['rho'], ref['rho'], f) < 1e-9


def test_graph_replay_matches_eager():
    """hipGraph replay of the step gives the eager result bit for bit, replay after replay, with
    eager work and synchronisation between replays."""
    sc = scenarios.dam_break_c1()
    m = sc.method()
    a = make_engine(sc.pA, m)
    b = make_engine(sc.pA, m)
    for chunk in range(4):
        say('graph chunk', chunk)
        a.step(5, graph=False)
        b.step(5, graph=True)
        oa, ob = a.download(), b.download()
        for k in FIELDS:
            np.testing.assert_array_equal(ob[k], oa[k], err_msg=f'{k} chunk {chunk}')
    assert a.status() == b.status()
    a.close()
    b.close()


def test_fp32c_density_error_after_1000_steps():
    """Precision form: FP32 pairs + FP64 integrator vs the FP64 engine, C1, 1,000 steps."""
    sc = scenarios.dam_break_c1()
    m = sc.method()
    a = make_engine(sc.pA, m, precision='fp64')
    b = make_engine(sc.pA, m, precision='fp32c')
    a.step(1000)
    b.step(1000)
    oa, ob = a.download(), b.download()
    f = sc.pA['label'] == 0
    drho = np.max(np.abs(ob['rho'][f] - oa['rho'][f]) / oa['rho'][f])
    dx = np.max(np.hypot(ob['x'][f] - oa['x'][f], ob['y'][f] - oa['y'][f])) / sc.dx
    say('fp32c after 1000 steps: max rel rho', drho, 'max |dx|/dx', dx)
    assert drho < 1e-5
    assert dx < 1e-3
    a.close()
    b.close()


# --------------------------------------------------------------------------
# host-forced failure handling (every kernel still exits normally)
# --------------------------------------------------------------------------
def test_capacity_flag_halts_step():
    from hipsph import _lib
    sc = scenarios.dam_break_c1()
    e = make_engine(sc.pA, sc.method(), cell_capacity=100)
    e.step(2)
    st = e.status()
    assert st['flags'] & _lib.SPH_FLAG_CAPACITY
    with pytest.raises(_lib.HipSphError):
        e.grid()
    out = e.download()
    f = sc.pA['label'] == 0
    assert np.all(np.isfinite(out['x'][f]))
    e.close()


def test_duration_stops_the_clock():
    sc = scenarios.dam_break_c1()
    e = make_engine(sc.pA, sc.method(), duration=3e-4)
    e.step(20)
    st = e.status()
    x1 = e.download()['x']
    e.step(5)
    x2 = e.download()['x']
    np.testing.assert_array_equal(x1, x2)
    assert st['t'] >= 3e-4 and st['steps'] < 20
    e.close()


def test_no_fluid_halts_without_touching_the_state():
    """Walls only: the reference's _minTimeStep has no fluid to reduce over
    (an empty min); the device raises SPH_FLAG_NO_FLUID and steps nothing."""
    from hipsph import _lib
    sc = scenarios.dam_break_c1()
    walls = sc.pA[sc.pA['label'] != 0].copy()
    e = make_engine(walls, sc.method())
    before = e.download()
    e.step(3)
    st = e.status()
    assert st['flags'] & _lib.SPH_FLAG_NO_FLUID and st['steps'] == 0 and st['t'] == 0.0
    after = e.download()
    for k in FIELDS:
        np.testing.assert_array_equal(after[k], before[k], err_msg=k)
    e.close()


def test_calls_before_upload_are_refused():
    """Before the first upload the device buffers hold no particle ids: every
    call that indexes by id (steps, download, gather, the neighbour dump, the
    kinetic energy) is refused with ESTATE and launches nothing; after the
    upload the same engine steps as usual. Likewise for a batch."""
    from hipsph import _lib
    from hipsph.batch import BatchEngine
    from hipsph.device import Engine, EngineParams, soa_from_particles
    sc = scenarios.dam_break_c1()
    m = sc.method()
    p = EngineParams.from_method(len(sc.pA), m)
    e = Engine(p)
    for fn in (lambda: e.step(1), lambda: e.step(2, graph=True), e.step_begin, e.nn_update, e.compute, e.dt,
               e.download, lambda: e.gather([0]), e.kinetic_energy, e.nn_dump):
        with pytest.raises(_lib.HipSphError) as ex:
            fn()
        assert _lib.ERRORS[ex.value.code] == 'ESTATE'
    e.upload(soa_from_particles(sc.pA))
    e.step(2)
    assert e.status()['steps'] == 2
    e.close()
    b = BatchEngine([p, p])
    for fn in (lambda: b.step(1), b.download):
        with pytest.raises(_lib.HipSphError) as ex:
            fn()
        assert _lib.ERRORS[ex.value.code] == 'ESTATE'
    b.close()


def test_kinetic_energy_of_the_fluid():
    """sph_kinetic_energy is KineticEnergy (src/Equations/KineticEnergy.py:5-12)
    over the fluid records, as the Solver passes them (src/Solver.py:202,416):
    vx = 1 on every fluid particle gives n_fluid * 0.5 * m (walls moving at 5
    add nothing); random velocities match the host sum. Both to the rounding
    of a different summation order."""
    from hipsph.Solver import KineticEnergy
    sc = scenarios.dam_break_c1()
    pA = sc.pA.copy()
    f = pA['label'] == 0
    pA['vx'][f] = 1.0
    pA['vx'][~f] = 5.0
    e = make_engine(pA, sc.method())
    assert e.kinetic_energy() == pytest.approx(f.sum() * 0.5 * pA['m'][f][0], rel=1e-13)
    e.close()
    q = perturbed(sc.pA, 4)
    e = make_engine(q, sc.method())
    want = KineticEnergy(int(f.sum()), q[f])
    assert e.kinetic_energy() == pytest.approx(want, rel=1e-13)
    e.close()


def test_single_particle_free_fall_fp64_bitexact():
    """One fluid particle, no walls: a one-cell grid, the particle its own
    only neighbour (no pair term), gravity alone; 50 steps bit-identical to
    the oracle and, to rounding, the closed-form free fall."""
    sc = scenarios.dam_break_c1()
    one = sc.pA[sc.pA['label'] == 0][:1].copy()
    one['vx'], one['vy'] = 0.3, 0.0
    m = sc.method()
    say('single particle, 50 steps')
    e = make_engine(one, m)
    e.step(50)
    st = e.status()
    out = e.download()
    e.close()
    s = oracle.Solver(one, m, 'cubic', 'pec', h=sc.h)
    hist = s.run(50)
    assert st['steps'] == 50 and st['t'] == s.t
    for k in FIELDS:
        np.testing.assert_array_equal(out[k], s.pA[k], err_msg=k)
    t = float(np.sum(hist[:, 0]))
    assert out['vy'][0] == pytest.approx(-9.81 * t, rel=1e-12)
    assert out['x'][0] == pytest.approx(sc.pA['x'][sc.pA['label'] == 0][0] + 0.3 * t, rel=1e-12)


def test_nonfinite_position_flags():
    from hipsph import _lib
    sc = scenarios.dam_break_c1()
    pA = sc.pA.copy()
    pA['x'][10] = np.nan
    e = make_engine(pA, sc.method())
    e.step(1)
    assert e.status()['flags'] & _lib.SPH_FLAG_NONFINITE
    e.close()


def test_boundary_pairs_dump_and_pair_pass_counts():
    """Pairs within a few ulp of q = 3: the device search and both pair passes reach the
    reference's verdict (a fused dx*dx + dy*dy would flip every one of these pairs)."""
    from test_oracle_invariants import boundary_system
    pA, z = boundary_system()
    sc = scenarios.dam_break_c1()
    full = np.zeros(len(pA), dtype=sc.pA.dtype)
    for f in ('x', 'y', 'h', 'm', 'rho', 'label'):
        full[f] = pA[f]
    d = _nn_parity(full, sc.method(), scale=4.0)
    for k in range(len(z['xi'])):
        assert (2 * k + 1 in d['idx'][d['off'][2 * k]:d['off'][2 * k + 1]]) == bool(z['ref_in'][k])
    want = np.diff(d['off']).astype(np.int32)
    for prec in ('fp64', 'fp32c'):
        e = make_engine(full, sc.method(), precision=prec, scale=4.0)
        e.nn_update()
        got = e.compute_counts()
        e.close()
        np.testing.assert_array_equal(got, want, err_msg=prec)


@pytest.mark.parametrize('prec', ['fp64', 'fp32c'])
def test_pair_pass_counts_match_near(prec):
    """The pair pass's own neighbour sets have the sizes of the reference's near(i)."""
    sc = scenarios.stress_lattice(nside=150)
    e = make_engine(sc.pA, sc.method(), precision=prec)
    e.nn_update()
    got = e.compute_counts()
    e.close()
    nn = oracle.NN(2.0)
    a = oracle.as_particles(sc.pA)
    nn.update(a)
    want = np.diff(nn.dump(a)['off'])
    f = sc.pA['label'] == 0
    np.testing.assert_array_equal(got[f], want[f])
    assert np.all(got[~f] == -1)
