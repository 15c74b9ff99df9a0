"""The oracle's coupling pieces pinned by mathematics (CPU).

* Newmark-beta (src/Integrators/NewmarkBeta.py:41-92): predict + correct give
  the textbook update x1 = x0 + dt v0 + dt^2 ((1/2 - beta) a0 + beta a1),
  v1 = v0 + dt ((1 - gamma) a0 + gamma a1); a constant acceleration is
  integrated exactly; with average acceleration (beta 1/4, gamma 1/2) and the
  host ``acceleration`` solve, an undamped chain of springs keeps its discrete
  energy 1/2 v'Mv + 1/2 w'Kw, which gamma > 1/2 damps.
* The solver body split at the coupling hook: begin + end = step, bit for
  bit; without a coupling integrator coupled particles do not move.
* The pressure probe of the coupled example (examples/IceBreak.py:253-285):
  with Shepard weights the summation density is the weighted harmonic mean
  of the neighbours' densities with weights m W, so a uniform field returns
  its density and p = 0 at rho0; a brute-force 3 x 3-cell neighbour set with
  that closed form; no fluid neighbour gives rho = 0 and p = -B, fluid
  neighbours all outside the probe's support give NaN.
"""
import math

import numpy as np
import pytest

import oracle
from hipsph.Integrators.NewmarkBeta import NewmarkBeta
from test_oracle_invariants import _M, _block_with_walls, _particles


def _coupled(n, seed):
    rng = np.random.default_rng(seed)
    pA = _particles(rng.uniform(0, 1, n), rng.uniform(0, 1, n), 0.026, label=np.full(n, 3, np.int8),
                    vx=rng.normal(0, 1, n), vy=rng.normal(0, 1, n))
    pA['ax'] = rng.normal(0, 5, n)
    pA['ay'] = rng.normal(0, 5, n)
    return pA


@pytest.mark.parametrize('beta,gamma', [(0.25, 0.5), (1.0 / 6.0, 0.5), (0.3025, 0.6)])
def test_newmark_textbook_update(beta, gamma):
    pA = _coupled(50, 1)
    dt = 1.7e-3
    x0, y0, vx0, vy0, ax0, ay0 = (pA[k].copy() for k in ('x', 'y', 'vx', 'vy', 'ax', 'ay'))
    oracle.newmark_predict(pA, dt, beta, gamma)
    rng = np.random.default_rng(2)
    ax1, ay1 = rng.normal(0, 5, len(pA)), rng.normal(0, 5, len(pA))
    pA['ax'], pA['ay'] = ax1, ay1
    oracle.newmark_correct(pA, dt, beta, gamma)
    for got, p0, v0, a0, a1 in (('x', x0, vx0, ax0, ax1), ('y', y0, vy0, ay0, ay1)):
        want = p0 + dt * v0 + dt * dt * ((0.5 - beta) * a0 + beta * a1)
        np.testing.assert_allclose(pA[got], want, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(pA['vx'], vx0 + dt * ((1 - gamma) * ax0 + gamma * ax1), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(pA['vy'], vy0 + dt * ((1 - gamma) * ay0 + gamma * ay1), rtol=1e-13, atol=1e-15)


def test_newmark_host_mirror_matches_oracle():
    """hipsph's NewmarkBeta.predict/correct (host) and the oracle agree bit for bit."""
    nm = NewmarkBeta(0.25, 0.5, np.eye(1), np.eye(1), np.zeros((1, 1)))
    a, b = _coupled(40, 3), _coupled(40, 3)
    oracle.newmark_predict(a, 2.3e-3, 0.25, 0.5)
    nm.predict(2.3e-3, b, 0.0)
    oracle.newmark_correct(a, 2.3e-3, 0.25, 0.5)
    nm.correct(2.3e-3, b, 0.0)
    for k in ('x', 'y', 'vx', 'vy'):
        np.testing.assert_array_equal(a[k], b[k])


def test_newmark_constant_acceleration_exact():
    pA = _coupled(10, 4)
    x0, vx0, ax = pA['x'].copy(), pA['vx'].copy(), pA['ax'].copy()
    dt, n = 2.0 ** -9, 256
    for _ in range(n):
        oracle.newmark_predict(pA, dt, 0.25, 0.5)
        oracle.newmark_correct(pA, dt, 0.25, 0.5)
    t = n * dt
    np.testing.assert_allclose(pA['x'], x0 + vx0 * t + 0.5 * ax * t * t, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(pA['vx'], vx0 + ax * t, rtol=1e-12, atol=1e-12)


def _chain(n, k=4.0e3, m=2.0):
    """A fixed-free chain of n masses and springs: M = m I, K tridiagonal SPD."""
    K = np.zeros((n, n))
    for i in range(n):
        K[i, i] = 2 * k if i < n - 1 else k
        if i > 0:
            K[i, i - 1] = K[i - 1, i] = -k
    return np.eye(n) * m, K


@pytest.mark.parametrize('beta,gamma,conserved', [(0.25, 0.5, True), (0.3025, 0.6, False)])
def test_newmark_average_acceleration_conserves_energy(beta, gamma, conserved):
    """The trapezoidal member (1/4, 1/2) keeps 1/2 v'Mv + 1/2 w'Kw of an
    undamped linear system; gamma > 1/2 (0.3025, 0.6) damps it."""
    n = 8
    M, K = _chain(n)
    nm = NewmarkBeta(beta, gamma, M, K, np.zeros((n, n)))
    pA = _particles(np.arange(n) * 0.1, np.zeros(n), 0.026, label=np.full(n, 3, np.int8))
    w0 = np.sin(np.linspace(0.3, 2.0, n)) * 1e-2
    pA['y'] = w0
    pA['ay'] = np.linalg.solve(M, -K @ w0)
    dt = 0.4 * 2.0 / math.sqrt(np.max(np.linalg.eigvalsh(np.linalg.solve(M, K))))

    def energy():
        v, w = pA['vy'], pA['y']
        return 0.5 * v @ M @ v + 0.5 * w @ K @ w

    e0 = energy()
    drift = 0.0
    for _ in range(2000):
        oracle.newmark_predict(pA, dt, beta, gamma)
        pA['ay'] = nm.acceleration(dt, np.zeros(n), pA['y'].copy(), pA['vy'].copy())
        oracle.newmark_correct(pA, dt, beta, gamma)
        drift = max(drift, abs(energy() - e0) / e0)
    if conserved:
        assert drift < 1e-10
    else:
        assert drift > 1e-4


def _scene(seed):
    pA = _block_with_walls(seed)
    f = pA['label'] == 0
    pA['c'][f] = _M().co
    # a sheet of coupled particles on the free surface
    top = pA['y'][f].max()
    xs = np.arange(2, 14) * 0.02
    sheet = _particles(xs, np.full(xs.size, top + 0.02), 0.026, label=np.full(xs.size, 3, np.int8),
                       rho=1000.0, m=0.4)
    sheet['ay'] = np.random.default_rng(seed).normal(0, 1, xs.size)
    return np.concatenate([pA, sheet])


@pytest.mark.parametrize('integrator', ['pec', 'euler'])
def test_solver_begin_end_equals_step(integrator):
    a = oracle.Solver(_scene(5), _M(), 'cubic', integrator, h=0.026, newmark=(0.25, 0.5))
    b = oracle.Solver(_scene(5), _M(), 'cubic', integrator, h=0.026, newmark=(0.25, 0.5))
    ha = a.run(5)
    for k in range(5):
        np.testing.assert_array_equal(b.begin(), ha[k])
        b.end()
    assert a.t == b.t and a.s.t_step == b.s.t_step == 5
    np.testing.assert_array_equal(a.pA, b.pA)
    c = b.pA['label'] == 3
    assert np.ptp(b.pA['y'][c] - _scene(5)['y'][c]) > 0  # the coupled particles moved


def test_coupled_particles_fixed_without_coupling_integrator():
    """src/Solver.py:381,397: the coupling integrator runs only with a coupling."""
    pA = _scene(6)
    s = oracle.Solver(pA, _M(), 'cubic', 'pec', h=0.026)
    s.run(5)
    c = pA['label'] == 3
    for k in ('x', 'y', 'vx', 'vy'):
        np.testing.assert_array_equal(s.pA[k][c], pA[k][c])


def _brute_probe(pA, grid, x, y, h, kernel, m):
    """nearPos by brute force over the 3 x 3-cell rule, then the closed form
    rho = sum a_j / sum (a_j / rho_j), a_j = m_j W_j over the fluid (the
    Shepard-normalised summation density)."""
    cs, xmin, ymin, ncx, ncy = grid['cell_size'], grid['xmin'], grid['ymin'], grid['ncx'], grid['ncy']
    cx = np.floor((pA['x'] - xmin) / cs)
    cy = np.floor((pA['y'] - ymin) / cs)
    inside = (cx >= 0) & (cx < ncx) & (cy >= 0) & (cy < ncy)
    px, py = math.floor((x - xmin) / cs), math.floor((y - ymin) / cs)
    r = np.hypot(x - pA['x'], y - pA['y'])
    q = r / (0.5 * (h + pA['h']))
    near = inside & (np.abs(cx - px) <= 1) & (np.abs(cy - py) <= 1) & (q <= 3.0)
    fl = near & (pA['label'] == 0)
    w = oracle.kernel_evaluate(kernel, r[fl], np.full(fl.sum(), h))
    a = pA['m'][fl] * w
    if a.size == 0:
        return 0.0
    rho = a.sum() / np.sum(a / pA['rho'][fl])
    return rho


@pytest.mark.parametrize('kernel', ['cubic', 'gaussian', 'wendland'])
def test_probe_uniform_field_and_brute_force(kernel):
    pA = _block_with_walls(7)
    m = _M()
    nn = oracle.NN()
    nn.update(pA)
    g = nn.grid()
    rng = np.random.default_rng(8)
    px = rng.uniform(0.04, 0.26, 12)
    py = rng.uniform(0.0, 0.2, 12)
    # uniform density: the probe returns it, p = 0 at rho0
    u = pA.copy()
    u['rho'][u['label'] == 0] = 1000.0
    rho, p = oracle.probe_pressure(nn, u, kernel, m, px, py, 0.026)
    np.testing.assert_allclose(rho, 1000.0, rtol=1e-13)
    np.testing.assert_allclose(p, 0.0, atol=1e-7)
    # random densities: the weighted harmonic mean over the brute-force set
    rho, p = oracle.probe_pressure(nn, pA, kernel, m, px, py, 0.026)
    want = np.array([_brute_probe(pA, g, x, y, 0.026, kernel, m) for x, y in zip(px, py)])
    np.testing.assert_allclose(rho, want, rtol=1e-13)
    np.testing.assert_allclose(p, ((want / m.rho0) ** m.gamma - 1.0) * m.B, rtol=1e-9, atol=1e-7)


def test_probe_without_fluid_weight():
    """No fluid neighbour: the empty sum, rho = 0 and p = -B. Fluid neighbours
    (q <= 3 with the mixed h) that all lie outside the probe's own support
    W(r, h) = 0: a zero Shepard normaliser and 0 / 0 weights, NaN as numpy's
    w / w_tilde gives (src/Equations/Shepard.py:12)."""
    pA = _block_with_walls(9)
    m = _M()
    nn = oracle.NN()
    nn.update(pA)
    rho, p = oracle.probe_pressure(nn, pA, 'cubic', m, [0.1, 0.1], [-0.2, -0.05], 0.01)
    assert rho[0] == 0.0 and p[0] == -m.B
    _, _, r, idx = nn.near_pos(0.1, -0.05, 0.01, pA)
    fl = pA['label'][idx] == 0
    assert fl.any() and np.all(r[fl] > 2 * 0.01)
    assert np.isnan(rho[1]) and np.isnan(p[1])
