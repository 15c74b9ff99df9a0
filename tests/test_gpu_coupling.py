"""The coupling hook on the device against the CPU oracle (FP64, bit-exact).

The reference's coupled step (src/Solver.py:374-398): predict the fluid and,
with Newmark-beta (src/Integrators/NewmarkBeta.py), the coupled particles;
_compute; the user's hook; correct both. Here, through the C ABI:
  * sph_step_begin + sph_step_end = sph_step, bit for bit, and both equal the
    oracle's step with Newmark-beta on the coupled particles;
  * sph_state_gather / sph_state_scatter by reference index, and their
    argument checks; a split step refuses other step calls until it ends;
  * sph_near_pos equals the oracle's nearPos (indices, q, r, order);
  * sph_probe_pressure equals the oracle's restatement of the example's
    pressure probe for the three kernels, edge cases included;
  * an ice-sheet hook in the example's form (nearPos, kernel, Shepard,
    summation density, Newmark-beta acceleration) run by the Solver mirror
    and by the oracle's split loop: same trajectory, dt series and hook
    pressures over 40 steps; the same hook with the device probe, driven
    through the engine, against the oracle's probe.
"""
import numpy as np
import pytest

import oracle
from gpu_util import FIELDS, make_engine, say
from hipsph import ParticleType, scenarios
from hipsph._lib import HipSphError
from hipsph.Equations.Shepard import Shepard
from hipsph.Equations.SummationDensity import SummationDensity
from hipsph.Helpers import Helpers
from hipsph.Integrators.NewmarkBeta import NewmarkBeta
from hipsph.Integrators.PEC import PEC
from hipsph.Kernels.CubicSpline import CubicSpline
from hipsph.Solver import Solver

pytestmark = pytest.mark.gpu

R0 = 0.02


def _beam(n, k=2.0e3, m=8.0):
    K = np.zeros((n, n))
    for i in range(n):
        K[i, i] = 2 * k if i < n - 1 else k
        if i > 0:
            K[i, i - 1] = K[i - 1, i] = -k
    return np.eye(n) * m, K, np.eye(n) * 0.5


def _scene():
    """A small dam break with an ice sheet (Coupled) resting on the water."""
    sc = scenarios.dam_break(0.4, 0.5, R0, tank_w=1.0, tank_h=0.8, name='ice')
    pA = sc.pA.copy()
    f = pA['label'] == 0
    top = pA['y'][f].max()
    sheet = Helpers.rect(xmin=0.06, xmax=0.34, ymin=top + R0, ymax=top + R0, r0=R0, mass=0.3, rho0=1000.0,
                         label=ParticleType.Coupled)
    sheet['h'] = 1.3 * R0
    sheet['rho'] = 1000.0
    return sc, np.concatenate([pA, sheet])


def _props(pA):
    c = pA['label'] == ParticleType.Coupled
    return {'y0': pA['y'][c].copy(), 'hh': R0, 'h': 1.3 * R0, 'b': R0, 'm': 8.0, 'g': 9.81, 'Pressure': []}


def ice_hook(pA, solver):
    """The example's hook (examples/IceBreak.py:110-178) reduced to its SPH
    part: the pressure under every node from nearPos + W + Shepard +
    summation density, the node force, the Newmark-beta acceleration."""
    P = solver.couplingProperties
    cInd = pA['label'] == ParticleType.Coupled
    y = pA[cInd]['y']
    w = y - P['y0']
    wdot = pA[cInd]['vy']
    x = pA[cInd]['x']
    P_n = np.zeros_like(w)
    rho_n = np.copy(pA['rho'][cInd])
    for i in range(len(w)):
        h_i, _, dist, near_arr = solver.nn.nearPos(x[i], y[i] - P['hh'], P['h'], pA)
        wk = solver.kernel.evaluate(dist, np.ones_like(h_i) * P['h'])
        w_tilde = Shepard(wk, pA[near_arr]['label'], pA[near_arr]['m'], pA[near_arr]['rho'])
        rho = SummationDensity(pA[near_arr]['label'], pA[near_arr]['m'], w_tilde)
        rho_n[i] = rho
        P_n[i] = ((rho / solver.method.rho0) ** solver.method.gamma - 1.0) * solver.method.B
    P['Pressure'].append(P_n)
    F = P_n * P['b'] - P['m'] * P['g']
    a = solver.couplingIntegrator.acceleration(solver.dt, F, w, wdot)
    pA['ay'][cInd] = a
    pA['p'][cInd] = P_n
    pA['rho'][cInd] = rho_n
    return pA


class _OracleNN:
    def __init__(self, nn):
        self.nn = nn

    def nearPos(self, x, y, h, pA):
        hh, q, r, idx = self.nn.near_pos(x, y, h, pA)
        return hh, q, r, idx.astype(np.uint64)


class _OracleKernel:
    def evaluate(self, r, h):
        return oracle.kernel_evaluate('cubic', r, h)


class _OracleSolver:
    """What the hook reads of a Solver, served by the oracle's split loop."""

    def __init__(self, s, method, nm, props):
        self.nn = _OracleNN(s.nn)
        self.kernel = _OracleKernel()
        self.method = method
        self.couplingIntegrator = nm
        self.couplingProperties = props
        self.dt = 0.0


def _oracle_coupled(pA, method, nm, props, steps, hook, damping=0.0):
    s = oracle.Solver(pA, method, 'cubic', 'pec', h=1.3 * R0, newmark=(nm.beta, nm.gamma), damping=damping)
    view = _OracleSolver(s, method, nm, props)
    hist = np.zeros((steps, 3))
    for k in range(steps):
        hist[k] = s.begin()
        view.dt = hist[k, 0]
        out = hook(s.pA, view)
        s.pA[:] = out
        s.end()
    return s, hist


@pytest.mark.parametrize('integrator', ['pec', 'euler'])
def test_split_step_equals_step_and_oracle(integrator):
    sc, pA = _scene()
    c = pA['label'] == ParticleType.Coupled
    pA['ay'][c] = np.linspace(-2.0, 2.0, c.sum())
    pA['vy'][c] = 0.05
    y0 = pA['y'][c].copy()  # the oracle below steps pA in place
    kw = dict(coupled_integrator='newmark', newmark_beta=0.25, newmark_gamma=0.5)
    a = make_engine(pA, sc.method(), integrator=integrator, **kw)
    b = make_engine(pA, sc.method(), integrator=integrator, **kw)
    say('split step', integrator, ': n', len(pA), 'coupled', int(c.sum()))
    a.step(30)
    for _ in range(30):
        b.step_begin()
        b.step_end()
    da, db = a.download(), b.download()
    s = oracle.Solver(pA, sc.method(), 'cubic', integrator, h=sc.h, newmark=(0.25, 0.5))
    s.run(30)
    for k in FIELDS + ('h',):
        np.testing.assert_array_equal(da[k], db[k], err_msg=k)
        np.testing.assert_array_equal(da[k], s.pA[k], err_msg=k)
    assert np.ptp(da['y'][c] - y0) > 0
    a.close()
    b.close()


def test_split_step_state_machine():
    sc, pA = _scene()
    e = make_engine(pA, sc.method())
    with pytest.raises(HipSphError):
        e.step_end()
    e.step_begin()
    for fn in (lambda: e.step(1), e.step_begin, e.compute, e.nn_update):
        with pytest.raises(HipSphError):
            fn()
    e.step_end()
    e.step(2)
    e.close()


def test_gather_scatter():
    sc, pA = _scene()
    e = make_engine(pA, sc.method())
    e.step(3)
    d = e.download()
    rng = np.random.default_rng(3)
    ids = rng.choice(len(pA), 50, replace=False)
    g = e.gather(ids, ('x', 'y', 'vx', 'vy', 'rho', 'ax', 'ay', 'drho', 'm', 'h', 'p', 'label'))
    for k, v in g.items():
        np.testing.assert_array_equal(v, d[k][ids], err_msg=k)
    new = {'ay': rng.normal(size=50), 'rho': 1000.0 + rng.normal(size=50), 'x': d['x'][ids] + 1e-4}
    e.scatter(ids, new)
    d2 = e.download()
    for k, v in new.items():
        np.testing.assert_array_equal(d2[k][ids], v, err_msg=k)
        rest = np.setdiff1d(np.arange(len(pA)), ids)
        np.testing.assert_array_equal(d2[k][rest], d[k][rest], err_msg=k)
    for bad in ([0, 0], [-1], [len(pA)]):
        with pytest.raises(HipSphError):
            e.scatter(bad, {'ay': np.zeros(len(bad))})
    with pytest.raises(KeyError):
        e.scatter([1], {'label': [0]})
    e.close()


def test_near_pos_vs_oracle():
    sc, pA = _scene()
    nm = (0.25, 0.5)
    e = make_engine(pA, sc.method(), coupled_integrator='newmark')
    e.step(4)
    e.step_begin()
    s = oracle.Solver(pA, sc.method(), 'cubic', 'pec', h=sc.h, newmark=nm)
    s.run(4)
    s.begin()
    rng = np.random.default_rng(4)
    px = rng.uniform(-0.1, 1.1, 40)
    py = rng.uniform(-0.1, 0.7, 40)
    ph = rng.uniform(0.01, 0.04, 40)
    off, idx, q, r = e.near_pos(px, py, ph)
    for i in range(40):
        _, qo, ro, io = s.nn.near_pos(px[i], py[i], ph[i], s.pA)
        a, b = off[i], off[i + 1]
        np.testing.assert_array_equal(idx[a:b], io)
        np.testing.assert_array_equal(q[a:b], qo)
        np.testing.assert_array_equal(r[a:b], ro)
    assert off[-1] > 0
    e.step_end()
    e.close()


@pytest.mark.parametrize('kernel', ['cubic', 'gaussian', 'wendland'])
def test_probe_vs_oracle(kernel):
    """On a binned, perturbed state: cubic and Wendland bit-exact; Gaussian
    (exp) to 1e-12 relative, as its element-wise kernel."""
    sc, pA = _scene()
    pA = oracle.as_particles(pA)
    f = pA['label'] == 0
    rng = np.random.default_rng(5)
    pA['rho'][f] *= 1 + rng.uniform(-0.01, 0.01, f.sum())
    pA['x'][f] += rng.uniform(-0.003, 0.003, f.sum())
    e = make_engine(pA, sc.method(), kernel=kernel)
    e.nn_update()
    nn = oracle.NN()
    nn.update(pA)
    px = np.concatenate([rng.uniform(0.0, 0.4, 30), [0.2, 0.5, 0.2]])
    py = np.concatenate([rng.uniform(0.0, 0.5, 30), [-0.3, 0.45, -0.04]])
    h = np.concatenate([np.full(30, 1.3 * R0), [0.026, 0.026, 0.005]])
    say('probe', kernel, 'points', len(px))
    rho, p = e.probe_pressure(px, py, h)
    ro, po = oracle.probe_pressure(nn, pA, kernel, oracle.wcsph(sc.method()), px, py, h)
    if kernel != 'gaussian':  # the Gaussian's exp is the device's
        np.testing.assert_array_equal(rho, ro)
        np.testing.assert_array_equal(p, po)
    else:
        np.testing.assert_allclose(rho, ro, rtol=1e-12)
        np.testing.assert_allclose(p, po, rtol=1e-9, atol=1e-6)
    assert rho[30] == 0.0 and p[30] == -sc.method().B  # outside every cell: no neighbour
    assert np.all(np.isfinite(rho[:30])) and np.all(rho[:30] > 900.0)
    e.close()


def test_solver_mirror_coupling_hook_fp64_bitexact():
    sc, pA = _scene()
    m = sc.method()
    c = pA['label'] == ParticleType.Coupled
    M, K, Cd = _beam(int(c.sum()))
    nm = NewmarkBeta(0.25, 0.5, M, K, Cd)
    props = _props(pA)
    sv = Solver(m, PEC(useXSPH=False, strict=False), CubicSpline(), duration=1.0, incrementalWriteout=False,
                exportProperties=[], h=1.3 * R0, coupling=ice_hook, couplingIntegrator=nm,
                couplingProperties=props, damping=0.0, settle=False, chunk=8)
    sv.addParticles(pA.copy())
    sv.setup()
    ref = oracle.as_particles(sv.particleArray[sv.indexes].copy())
    K_ = 40
    props_o = _props(pA)
    s, hist = _oracle_coupled(ref, m, nm, props_o, K_, ice_hook)
    t = np.cumsum(hist[:, 0])
    sv.duration = 0.5 * (t[K_ - 2] + t[K_ - 1])
    say('coupled solver: n', sv.num_particles, 'coupled', int(c.sum()), 'steps', K_)
    sv.run()
    sv.close()
    assert sv.t_step == K_
    np.testing.assert_array_equal(np.asarray(sv.dt_a), hist[:, 0])
    got = sv.particleArray[sv.indexes]
    for f in FIELDS:
        np.testing.assert_array_equal(got[f], s.pA[f], err_msg=f)
    np.testing.assert_array_equal(got['p'][got['label'] == 3], s.pA['p'][s.pA['label'] == 3])
    assert len(props['Pressure']) == K_
    for a, b in zip(props['Pressure'], props_o['Pressure']):
        np.testing.assert_array_equal(a, b)
    assert np.ptp(got['y'][got['label'] == 3] - props['y0']) > 0  # the sheet moved


def test_solver_mirror_coupling_with_settling_fp64_bitexact():
    """The hook also runs while settling (src/Solver.py:374-398 with damping,
    the clock held): maxSettle + 1 damped coupled steps, then undamped ones."""
    sc, pA = _scene()
    m = sc.method()
    c = pA['label'] == ParticleType.Coupled
    M, K, Cd = _beam(int(c.sum()))
    nm = NewmarkBeta(0.25, 0.5, M, K, Cd)
    props = _props(pA)
    max_settle = 3
    sv = Solver(m, PEC(useXSPH=False, strict=False), CubicSpline(), duration=1.0, incrementalWriteout=False,
                exportProperties=[], h=1.3 * R0, coupling=ice_hook, couplingIntegrator=nm,
                couplingProperties=props, damping=0.05, maxSettle=max_settle, chunk=8)
    sv.addParticles(pA.copy())
    sv.setup()
    ref = oracle.as_particles(sv.particleArray[sv.indexes].copy())
    props_o = _props(pA)
    n_settle = max_settle + 1  # KE0 = 0: the check fires at t_step > maxSettle
    s1, h1 = _oracle_coupled(ref, m, nm, props_o, n_settle, ice_hook, damping=0.05)
    K_ = 12
    s2, h2 = _oracle_coupled(oracle.as_particles(s1.pA.copy()), m, nm, props_o, K_, ice_hook)
    t = np.cumsum(h2[:, 0])
    sv.duration = 0.5 * (t[K_ - 2] + t[K_ - 1])
    say('coupled solver with settling: settle', n_settle, 'steps', K_)
    sv.run()
    sv.close()
    assert sv.settled and sv.t_step == n_settle + K_
    np.testing.assert_array_equal(np.asarray(sv.dt_a), np.concatenate((h1[:, 0], h2[:, 0])))
    got = sv.particleArray[sv.indexes]
    for f in FIELDS:
        np.testing.assert_array_equal(got[f], s2.pA[f], err_msg=f)
    assert len(props['Pressure']) == n_settle + K_
    for a, b in zip(props['Pressure'], props_o['Pressure']):
        np.testing.assert_array_equal(a, b)


def test_engine_device_probe_hook_fp64_bitexact():
    """The same coupling driven through the engine with the device probe and
    records moved by index only (no full download per step)."""
    sc, pA = _scene()
    m = sc.method()
    c = np.nonzero(pA['label'] == ParticleType.Coupled)[0]
    M, K, Cd = _beam(len(c))
    nm = NewmarkBeta(0.25, 0.5, M, K, Cd)
    P = _props(pA)
    e = make_engine(pA, m, coupled_integrator='newmark', newmark_beta=0.25, newmark_gamma=0.5)
    steps = 40
    dts = []
    for _ in range(steps):
        e.step_begin()
        dt = e.status()['dt']
        dts.append(dt)
        g = e.gather(c, ('x', 'y', 'vy'))
        rho, p = e.probe_pressure(g['x'], g['y'] - P['hh'], P['h'])
        a = nm.acceleration(dt, p * P['b'] - P['m'] * P['g'], g['y'] - P['y0'], g['vy'])
        e.scatter(c, {'ay': a, 'rho': rho})
        e.step_end()
    d = e.download()
    e.close()
    s = oracle.Solver(pA, m, 'cubic', 'pec', h=sc.h, newmark=(0.25, 0.5))
    for k in range(steps):
        assert s.begin()[0] == dts[k]
        rho, p = s.probe(s.pA['x'][c], s.pA['y'][c] - P['hh'], P['h'])
        a = nm.acceleration(dts[k], p * P['b'] - P['m'] * P['g'], s.pA['y'][c] - P['y0'], s.pA['vy'][c])
        s.pA['ay'][c] = a
        s.pA['rho'][c] = rho
        s.end()
    for f in FIELDS:
        np.testing.assert_array_equal(d[f], s.pA[f], err_msg=f)


def test_near_pos_and_probe_outside_the_grid():
    """Points far outside the grid, or NaN, have no cells to walk: nearPos
    is empty and the probe gives the empty sum (rho = 0, p = -B), as for a
    point whose 3 x 3 cells hold no fluid (examples/IceBreak.py:253-285)."""
    sc, pA = _scene()
    m = sc.method()
    e = make_engine(pA, m)
    e.step(2)
    off, idx, q, r = e.near_pos([1e12, np.nan, -5.0], [0.1, 0.1, np.nan], 0.026)
    assert len(idx) == 0 and list(off) == [0, 0, 0, 0]
    rho, p = e.probe_pressure([1e12, np.nan, 0.1], [0.1, 0.1, -1e300], 0.026)
    assert np.all(rho == 0.0) and np.all(p == -m.B)
    e.close()
