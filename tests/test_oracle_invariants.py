"""Pinning the composed oracle step where the reference's tests are silent.

The reference has no working test of NNLinkedList (current API), _loop,
TimeStep or the Solver step. These checks pin the oracle by mathematics and
by an independent formulation:

* neighbour sets against an O(N^2) brute force of the stated rule
  (3x3 block of cells of size 2 min h around the query's cell, q <= 3),
  including the reference's order (x-shift outer, y-shift inner, descending
  index inside a cell), and completeness of the cubic support (r <= 2h);
* the fused per-particle sums (continuity, momentum + artificial viscosity +
  gravity, Lennard-Jones wall force) against a dense all-pairs numpy model;
* closed forms (two particles), exact zeros (uniform velocity), momentum
  conservation (antisymmetric pair forces), mirror and rotation symmetry.
A dropped term, a wrong sign, a swapped x/y or a wrong neighbour rule fails
at least one of them.
"""
import math

import numpy as np
import pytest

import oracle

G = 9.81


def _particles(x, y, h, label=None, vx=None, vy=None, rho=None, m=None):
    n = len(x)
    pA = np.zeros(n, dtype=oracle.particle_dtype)
    pA['x'], pA['y'], pA['h'] = x, y, h
    pA['label'] = 0 if label is None else label
    pA['vx'] = 0 if vx is None else vx
    pA['vy'] = 0 if vy is None else vy
    pA['rho'] = 1000.0 if rho is None else rho
    pA['m'] = 0.4 if m is None else m
    return pA


def _random_set(seed, n=400, mixed_h=False):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1.0, n)
    y = rng.uniform(0, 0.7, n)
    h = np.full(n, 0.03) if not mixed_h else rng.uniform(0.02, 0.04, n)
    return _particles(x, y, h)


def brute_near(pA, scale=2.0):
    """near(i) for all i by brute force: returns list of index arrays in reference order."""
    x, y, h = pA['x'], pA['y'], pA['h']
    xmin, ymin = x.min(), y.min()
    cs = h.min() * scale
    if cs < 1e-6:
        cs = 1.0
    ncx = max(1, math.ceil((1.0 / cs) * (x.max() - xmin)))
    ncy = max(1, math.ceil((1.0 / cs) * (y.max() - ymin)))
    cx = np.floor((x - xmin) / cs).astype(np.int64)
    cy = np.floor((y - ymin) / cs).astype(np.int64)
    key = cx + ncx * cy
    kx, ky = key % ncx, key // ncx
    inside = (key >= 0) & (key < ncx * ncy)
    out = []
    for i in range(len(x)):
        dx = x[i] - x
        dy = y[i] - y
        r = np.sqrt(dx * dx + dy * dy)
        q = r / (0.5 * (h[i] + h))
        a = kx - cx[i]
        b = ky - cy[i]
        ok = inside & (np.abs(a) <= 1) & (np.abs(b) <= 1) & (q <= 3.0)
        j = np.nonzero(ok)[0]
        order = np.lexsort((-j, b[j], a[j]))  # x-shift outer, y-shift inner, index descending
        out.append(j[order])
    return out


@pytest.mark.parametrize('seed,mixed', [(1, False), (2, True), (3, False)])
def test_nn_bruteforce(seed, mixed):
    pA = _random_set(seed, mixed_h=mixed)
    nn = oracle.NN(2.0)
    nn.update(pA)
    d = nn.dump(pA)
    ref = brute_near(pA)
    for i in range(len(pA)):
        got = d['idx'][d['off'][i]:d['off'][i + 1]]
        np.testing.assert_array_equal(got, ref[i])


def test_nn_cubic_support_complete():
    """With uniform h the 3x3 block (cells 2h wide) holds every pair with r <= 2h."""
    pA = _random_set(7, n=600)
    nn = oracle.NN(2.0)
    nn.update(pA)
    d = nn.dump(pA)
    x, y, h = pA['x'], pA['y'], pA['h']
    for i in range(len(pA)):
        r = np.hypot(x[i] - x, y[i] - y)
        must = set(np.nonzero(r <= 2 * h[i])[0].tolist())
        got = set(d['idx'][d['off'][i]:d['off'][i + 1]].tolist())
        assert must <= got


def test_nn_edge_alias():
    """A particle whose x lands exactly on xmin + ncx*cs gets cx = ncx: its flat key aliases to
    cell (0, cy + 1) (src/Tools/NNLinkedList.py:129-141 has no validity check)."""
    # h = 0.5 -> cell size 1.0; x spans exactly 0..4 -> ncx = 4; the particle at x = 4 has cx = 4
    x = np.array([0.0, 1.5, 2.5, 4.0, 0.2])
    y = np.array([0.0, 0.5, 1.5, 0.5, 1.6])
    pA = _particles(x, y, 0.5)
    nn = oracle.NN(2.0)
    nn.update(pA)
    g = nn.grid()
    assert (g['ncx'], g['ncy'], g['cell_size']) == (4, 2, 1.0)
    d = nn.dump(pA)
    assert d['keys'][3] == 4 + 4 * 0  # == cell (0, 1)
    ref = brute_near(pA)
    for i in range(len(pA)):
        np.testing.assert_array_equal(d['idx'][d['off'][i]:d['off'][i + 1]], ref[i])
    # particle 4 sits in cell (0, 1): it sees the aliased particle 3 as a candidate but q > 3
    assert 3 not in d['idx'][d['off'][4]:d['off'][5]]


def dense_rates(pA, method, kernel='cubic', scale=2.0):
    """Continuity, momentum + AV + gravity, LJ wall force for every fluid particle, all pairs at once."""
    x, y, h, lab = pA['x'], pA['y'], pA['h'], pA['label']
    rho, vx, vy, m = pA['rho'], pA['vx'], pA['vy'], pA['m']
    n = len(x)
    xmin, ymin = x.min(), y.min()
    cs = h.min() * scale
    ncx = max(1, math.ceil((1.0 / cs) * (x.max() - xmin)))
    ncy = max(1, math.ceil((1.0 / cs) * (y.max() - ymin)))
    cx = np.floor((x - xmin) / cs).astype(np.int64)
    cy = np.floor((y - ymin) / cs).astype(np.int64)
    key = cx + ncx * cy
    kx, ky = key % ncx, key // ncx
    inside = (key >= 0) & (key < ncx * ncy)
    block = inside[None, :] & (np.abs(kx[None, :] - cx[:, None]) <= 1) & (np.abs(ky[None, :] - cy[:, None]) <= 1)
    dx = x[:, None] - x[None, :]
    dy = y[:, None] - y[None, :]
    r = np.sqrt(dx * dx + dy * dy)
    hij = 0.5 * (h[:, None] + h[None, :])
    q = r / hij
    nb = block & (q <= 3.0)
    p = np.where(lab == 0, ((rho / method.rho0) ** method.gamma - 1.0) * method.B, 0.0) + method.Pb
    prs = p / (rho * rho)
    with np.errstate(divide='ignore', invalid='ignore'):
        if kernel == 'cubic':
            alpha = 10.0 / (7.0 * np.pi) / (hij * hij)
            g = np.where(q <= 1.0, -3.0 * q * (1.0 - 0.75 * q), -0.75 * (2.0 - q) ** 2)
            norm = np.where((q > 2.0) | (r < 1e-10), 0.0, alpha * g / (hij * r))
        elif kernel == 'gaussian':
            dwdq = -2.0 * q / (np.pi * hij * hij) * np.exp(-q * q)
            norm = np.where((r * hij > 1e-12) & (q <= 3.0), dwdq / (r * hij), 0.0)
        else:
            alpha = 9.0 / (4.0 * np.pi) / (hij * hij)
            inner = 1.0 - 0.5 * q
            g = inner ** 5 * (-14.0 / 3.0) * q * (1 + 2.5 * q)
            norm = np.where((q >= 2.0) | (r < 1e-10), 0.0, alpha * g / (hij * r))
        dwx, dwy = norm * dx, norm * dy
        vxij = vx[:, None] - vx[None, :]
        vyij = vy[:, None] - vy[None, :]
        W = nb & (lab[None, :] == 0)
        drho = np.sum(np.where(W, m[None, :] * (vxij * dwx + vyij * dwy), 0.0), axis=1)
        dv = vxij * dx + vyij * dy
        hm = 0.5 * (h[:, None] + hij)
        mu = hm * dv / (r * r + 0.01 * hm * hm)
        cij = 0.5 * method.co
        PI = np.where(dv < 0, mu * (method.beta * mu - method.alpha * cij) / (0.5 * (rho[:, None] + rho[None, :])),
                      0.0)
        fac = -m[None, :] * (prs[:, None] + prs[None, :] + PI)
        ax = np.sum(np.where(W, fac * dwx, 0.0), axis=1)
        ay = np.sum(np.where(W, fac * dwy, 0.0), axis=1)
        Bm = nb & (lab[None, :] != 0) & (r <= method.r0) & (r > 1e-12)
        fr = method.r0 / r
        lj = method.D * (fr ** method.p1 - fr ** method.p2) / (r * r)
        fx = np.sum(np.where(Bm, lj * dx, 0.0), axis=1)
        fy = np.sum(np.where(Bm, lj * dy, 0.0), axis=1)
    return drho, ax + fx, ay - G + fy


class _M:
    """WCSPH constants (src/Methods/WCSPH.py:35-79) for a column height H."""

    def __init__(self, H=1.0, r0=0.02, rho0=1000.0):
        self.height, self.r0, self.rho0 = H, r0, rho0
        self.useXSPH = False
        self.useSummationDensity = False
        self.gamma = 7.0
        self.co = 10.0 * math.sqrt(2 * 9.81 * H)
        self.B = self.co * self.co * rho0 / self.gamma
        self.Pb = 0.0
        self.alpha, self.beta = 0.01, 0.0
        self.epsilon = 0.5
        self.D = 5 * 9.81 * H
        self.p1, self.p2 = 4, 2


def _block_with_walls(seed, nx=16, ny=12, dx=0.02):
    """A jittered fluid block resting on a wall row, random velocities and densities."""
    rng = np.random.default_rng(seed)
    gx, gy = np.meshgrid(np.arange(nx) * dx, np.arange(ny) * dx, indexing='ij')
    xf = gx.ravel() + rng.uniform(-0.15 * dx, 0.15 * dx, gx.size)
    yf = gy.ravel() + rng.uniform(-0.15 * dx, 0.15 * dx, gx.size)
    xw = np.arange(-3, nx + 3) * dx
    yw = np.full(xw.size, -0.9 * dx)
    x = np.concatenate([xf, xw])
    y = np.concatenate([yf, yw])
    lab = np.concatenate([np.zeros(xf.size, np.int8), np.ones(xw.size, np.int8)])
    vx = np.concatenate([rng.normal(0, 0.5, xf.size), np.zeros(xw.size)])
    vy = np.concatenate([rng.normal(0, 0.5, xf.size), np.zeros(xw.size)])
    rho = 1000.0 * (1 + rng.uniform(-0.01, 0.01, x.size))
    return _particles(x, y, 1.3 * dx, lab, vx, vy, rho, 1000.0 * dx * dx)


@pytest.mark.parametrize('kernel', ['cubic', 'gaussian', 'wendland'])
@pytest.mark.parametrize('seed', [11, 12])
def test_loop_vs_dense_pairs(kernel, seed):
    pA = _block_with_walls(seed)
    m = _M()
    ref = dense_rates(pA, m, kernel)
    oracle.loop(pA, m, kernel)
    f = pA['label'] == 0
    for got, want in ((pA['drho'], ref[0]), (pA['ax'], ref[1]), (pA['ay'], ref[2])):
        scale = np.max(np.abs(want[f]))
        assert scale > 0
        np.testing.assert_allclose(got[f], want[f], rtol=0, atol=1e-12 * scale)
    # the AV branch and the wall force were both exercised
    assert np.any(np.abs(ref[1][f]) > 0)


def test_two_particles_closed_form():
    """Two fluid particles at rest, d = 0.6 h apart along x (q = 0.6, cubic inner branch)."""
    h, d, m0 = 0.026, 0.6 * 0.026, 0.4
    pA = _particles(np.array([0.0, d]), np.array([0.0, 0.0]), h, rho=np.array([1002.0, 1001.0]), m=m0)
    m = _M()
    oracle.loop(pA, m, 'cubic')
    p = ((pA['rho'] / m.rho0) ** 7 - 1) * m.B
    q = d / h
    dwdx = 10 / (7 * math.pi * h * h) * (-3 * q * (1 - 0.75 * q)) / (h * d) * (0.0 - d)  # x_0 - x_1 = -d
    a0 = -m0 * (p[0] / pA['rho'][0] ** 2 + p[1] / pA['rho'][1] ** 2) * dwdx
    assert pA['ax'][0] == pytest.approx(a0, rel=1e-13)
    assert pA['ax'][1] == pytest.approx(-a0, rel=1e-13)
    assert pA['ay'][0] == pytest.approx(-G, rel=1e-15)
    assert pA['drho'][0] == 0.0


def test_uniform_velocity_gives_zero_density_rate():
    pA = _block_with_walls(5)
    f = pA['label'] == 0
    pA['vx'][f] = 0.7
    pA['vy'][f] = -0.2
    pA['vx'][~f] = 0.7
    pA['vy'][~f] = -0.2
    oracle.loop(pA, _M(), 'cubic')
    assert np.all(pA['drho'][f] == 0.0)


def test_momentum_conservation_fluid_only():
    """Pair forces are antisymmetric (uniform h, m): sum m (a + g) over a fluid-only set is ~0."""
    pA = _block_with_walls(9)
    pA = pA[pA['label'] == 0].copy()
    oracle.loop(pA, _M(), 'cubic')
    tot_x = np.sum(pA['m'] * pA['ax'])
    tot_y = np.sum(pA['m'] * (pA['ay'] + G))
    scale = np.sum(pA['m'] * np.abs(pA['ax']))
    assert abs(tot_x) < 1e-12 * scale
    assert abs(tot_y) < 1e-12 * scale


def test_mirror_and_rotation_symmetry():
    pA = _block_with_walls(4)
    m = _M()
    a = pA.copy()
    oracle.loop(a, m, 'cubic')
    # mirror x -> -x
    b = pA.copy()
    b['x'] = -b['x']
    b['vx'] = -b['vx']
    oracle.loop(b, m, 'cubic')
    f = pA['label'] == 0
    sc = np.max(np.abs(a['ax'][f]))
    np.testing.assert_allclose(b['ax'][f], -a['ax'][f], atol=1e-10 * sc)
    np.testing.assert_allclose(b['ay'][f], a['ay'][f], atol=1e-10 * sc)
    np.testing.assert_allclose(b['drho'][f], a['drho'][f], atol=1e-10 * np.max(np.abs(a['drho'][f])))
    # rotate by +90 degrees: (x, y) -> (-y, x); gravity stays along -y, so compare a + g
    c = pA.copy()
    c['x'], c['y'] = -pA['y'], pA['x']
    c['vx'], c['vy'] = -pA['vy'], pA['vx']
    oracle.loop(c, m, 'cubic')
    np.testing.assert_allclose(c['ax'][f], -(a['ay'][f] + G), atol=1e-10 * sc)
    np.testing.assert_allclose(c['ay'][f] + G, a['ax'][f], atol=1e-10 * sc)


def test_solver_step_masks_and_walls_fixed():
    """The loop body moves fluid only and the first dt is the Courant step (a = 0)."""
    pA = _block_with_walls(3)
    m = _M()
    pA['c'][pA['label'] == 0] = m.co  # Solver.setup (src/Solver.py:196)
    s = oracle.Solver(pA, m, 'cubic', 'pec', h=0.026)
    w = s.pA['label'] != 0
    xw = s.pA['x'][w].copy()
    hist = s.run(3)
    assert hist[0][0] == pytest.approx(0.25 * 0.026 / m.co, rel=1e-15)
    assert hist[0][2] == 1e10
    np.testing.assert_array_equal(s.pA['x'][w], xw)
    assert s.t == pytest.approx(hist[:, 0].sum(), rel=1e-15)


def boundary_system():
    """The fixture's boundary pairs as one particle set: (i_k, j_k) at indices (2k, 2k+1), h = 1."""
    import os
    from conftest import GOLDEN
    z = np.load(os.path.join(GOLDEN, 'boundary_pairs.npz'))
    n = 2 * len(z['xi'])
    x = np.empty(n)
    y = np.empty(n)
    x[0::2], y[0::2] = z['xi'], z['yi']
    x[1::2], y[1::2] = z['xj'], z['yj']
    return _particles(x, y, float(z['h']), m=1.0), z


def test_nn_boundary_pairs_follow_ieee_expression():
    """Near q = 3 the verdict is the reference's unfused expression, not a fused one."""
    pA, z = boundary_system()
    nn = oracle.NN(4.0)  # cell 4 h wide: every pair at r ~ 3h is a candidate
    nn.update(pA)
    d = nn.dump(pA)
    for k in range(len(z['xi'])):
        near_i = d['idx'][d['off'][2 * k]:d['off'][2 * k + 1]]
        near_j = d['idx'][d['off'][2 * k + 1]:d['off'][2 * k + 2]]
        assert (2 * k + 1 in near_i) == bool(z['ref_in'][k])
        assert (2 * k in near_j) == bool(z['ref_in'][k])  # x_j - x_i = -(x_i - x_j) exactly
    assert z['ref_in'].any() and (~z['ref_in']).any() and np.any(z['ref_in'] != z['fused_in'])


@pytest.mark.parametrize('kernel,support,mass', [('cubic', 2.0, 1.0), ('wendland', 2.0, 1.0),
                                                 ('gaussian', 3.0, 1.0 - math.exp(-9.0))])
def test_kernel_normalisation_and_gradient(kernel, support, mass):
    """Pins W and grad W of each oracle kernel by what the mathematics fixes,
    independently of how the formula is written: the 2D normalisation
    (integral of W over the support = 1; the Gaussian truncated at 3h keeps
    1 - e^-9), the value at the origin (the normalisation constant:
    cubic 10/(7 pi h^2), Wendland C4 9/(4 pi h^2), Gaussian 1/(pi h^2)), and
    grad W = dW/dr * x / r against a central difference of W. A dropped term,
    a wrong constant or a wrong power in either function fails one of these."""
    h = 0.37
    n = 200001
    r = np.linspace(0.0, support * h, n)
    hv = np.full_like(r, h)
    w = oracle.kernel_evaluate(kernel, r, hv)
    integral = np.trapezoid(w * 2 * math.pi * r, r)
    assert integral == pytest.approx(mass, rel=2e-9)
    w0 = {'cubic': 10 / (7 * math.pi), 'wendland': 9 / (4 * math.pi), 'gaussian': 1 / math.pi}[kernel] / h ** 2
    assert w[0] == pytest.approx(w0, rel=1e-15)
    rr = np.linspace(0.05, support * h * 0.995, 997)
    eps = 1e-6 * h
    dwdr = (oracle.kernel_evaluate(kernel, rr + eps, np.full_like(rr, h)) -
            oracle.kernel_evaluate(kernel, rr - eps, np.full_like(rr, h))) / (2 * eps)
    x = 0.6 * rr  # a component of the separation
    g = oracle.kernel_gradient(kernel, x, rr, np.full_like(rr, h))
    np.testing.assert_allclose(g, dwdr * x / rr, rtol=2e-6, atol=1e-9 * np.max(np.abs(dwdr)))


def _av_pair(u, alpha, beta=0.0):
    h, d, m0 = 0.026, 0.6 * 0.026, 0.4
    pA = _particles(np.array([0.0, d]), np.array([0.0, 0.0]), h, vx=np.array([u, -u]),
                    rho=np.array([1002.0, 1001.0]), m=m0)
    m = _M()
    m.alpha, m.beta = alpha, beta
    pA['c'] = m.co  # WCSPH.compute_speed_of_sound (src/Methods/WCSPH.py:110-129)
    oracle.loop(pA, m, 'cubic')
    return pA, m, h, d, m0


def test_two_particles_artificial_viscosity_closed_form():
    """Monaghan (1992) artificial viscosity on two approaching fluid particles,
    with the reference's choices (src/Equations/Momentum.py:38-53): the mean
    sound speed is 0.5 c_i (the neighbour's c is never filled, Q5) and the
    smoothing length is re-averaged, h_bar = 0.5 (h_i + h_ij) = h here.
    Also: receding particles get no viscous term, the term removes relative
    kinetic energy, and at beta = 0 it is linear in alpha."""
    u, alpha = 0.8, 0.3
    pA, m, h, d, m0 = _av_pair(u, alpha)
    p = ((pA['rho'] / m.rho0) ** 7 - 1) * m.B
    q = d / h
    xij = -d                      # x_0 - x_1
    vij = 2 * u                   # v_0 - v_1 (approaching: v.x < 0)
    dwdx = 10 / (7 * math.pi * h * h) * (-3 * q * (1 - 0.75 * q)) / (h * d) * xij
    mu = h * (vij * xij) / (d * d + 0.01 * h * h)
    rho_bar = 0.5 * (pA['rho'][0] + pA['rho'][1])
    pi_ij = mu * (0.0 * mu - alpha * 0.5 * m.co) / rho_bar
    a0 = -m0 * (p[0] / pA['rho'][0] ** 2 + p[1] / pA['rho'][1] ** 2 + pi_ij) * dwdx
    assert pA['ax'][0] == pytest.approx(a0, rel=1e-12)
    assert pA['drho'][0] == pytest.approx(m0 * vij * dwdx, rel=1e-12)
    # branch: receding pair -> pressure term only
    rec, *_ = _av_pair(-u, alpha)
    a_rec = -m0 * (p[0] / pA['rho'][0] ** 2 + p[1] / pA['rho'][1] ** 2) * dwdx
    assert rec['ax'][0] == pytest.approx(a_rec, rel=1e-12)
    # dissipation: the viscous part decelerates the approach (a_0 < pressure-only value for u > 0)
    base, *_ = _av_pair(u, 0.0)
    assert pA['ax'][0] - base['ax'][0] < 0 and pA['ax'][1] - base['ax'][1] > 0
    # linear in alpha at beta = 0
    dbl, *_ = _av_pair(u, 2 * alpha)
    assert (dbl['ax'][0] - base['ax'][0]) == pytest.approx(2 * (pA['ax'][0] - base['ax'][0]), rel=1e-10)


def test_torch_dense_model_matches_oracle_trajectory():
    """The FP64 dense-pair torch model (the reference gradient of the adjoint
    tests) follows the oracle's trajectory: PEC, dynamic dt, walls, AV."""
    import torch
    from torch_ref import DenseWCSPH
    from hipsph import scenarios
    sc = scenarios.dam_break(0.2, 0.2, 0.02, tank_w=0.5, tank_h=0.4, name='tiny')
    pA = oracle.as_particles(sc.pA.copy())
    rng = np.random.default_rng(11)
    f = pA['label'] == 0
    pA['vx'][f] = rng.normal(0, 0.3, f.sum())
    pA['vy'][f] = rng.normal(0, 0.3, f.sum())
    m = sc.method()
    model = DenseWCSPH(pA, m, sc.h)
    K = 12
    x, y, vx, vy, rho = model.run(torch.tensor(m.alpha, dtype=torch.float64),
                                  torch.tensor(m.co, dtype=torch.float64),
                                  torch.as_tensor(pA['vx'][f].copy()), torch.as_tensor(pA['vy'][f].copy()), K)
    o = oracle.Solver(pA, m, 'cubic', 'pec', h=sc.h)
    o.run(K)
    for got, k in ((x, 'x'), (y, 'y'), (vx, 'vx'), (vy, 'vy'), (rho, 'rho')):
        want = o.pA[k][f]
        assert np.max(np.abs(got.numpy() - want)) <= 1e-11 * max(1.0, np.max(np.abs(want))), k


# ---------------------------------------------------------------------------
# XSPH and summation density (src/Equations/XSPH.py, SummationDensity.py;
# src/Tools/SolverTools.py:129-140,173). The reference has no test of either
# on the loop; its momentum test keeps an older all-pairs form of the XSPH
# term, eta * sum m_j / rho_ij * W_ij * (v_j - v_i) with eta = 0.5
# (test/test_numba_momentum.py:117-123), which the dense model below follows.
# ---------------------------------------------------------------------------
def dense_w(kernel, r, hij):
    q = r / hij
    if kernel == 'cubic':
        a = 10.0 / (7.0 * np.pi) / (hij * hij)
        return np.where(q > 2.0, 0.0, np.where(q > 1.0, a * 0.25 * (2.0 - q) ** 3,
                                               a * (1.0 - 1.5 * q * q * (1.0 - 0.5 * q))))
    if kernel == 'gaussian':
        return np.where(q <= 3.0, np.exp(-q * q) / (np.pi * hij * hij), 0.0)
    a = 9.0 / (4.0 * np.pi) / (hij * hij)
    inner = 1.0 - 0.5 * q
    return np.where(q >= 2.0, 0.0, a * inner ** 6 * (35.0 / 12.0 * q * q + 3.0 * q + 1.0))


def dense_neighbours(pA, scale=2.0):
    """Neighbour matrix of the 3x3-cell rule (q <= 3), distances, h_ij."""
    x, y, h = pA['x'], pA['y'], pA['h']
    xmin, ymin = x.min(), y.min()
    cs = h.min() * scale
    ncx = max(1, math.ceil((1.0 / cs) * (x.max() - xmin)))
    ncy = max(1, math.ceil((1.0 / cs) * (y.max() - ymin)))
    cx = np.floor((x - xmin) / cs).astype(np.int64)
    cy = np.floor((y - ymin) / cs).astype(np.int64)
    key = cx + ncx * cy
    kx, ky = key % ncx, key // ncx
    inside = (key >= 0) & (key < ncx * ncy)
    block = inside[None, :] & (np.abs(kx[None, :] - cx[:, None]) <= 1) & (np.abs(ky[None, :] - cy[:, None]) <= 1)
    r = np.hypot(x[:, None] - x[None, :], y[:, None] - y[None, :])
    hij = 0.5 * (h[:, None] + h[None, :])
    return block & (r / hij <= 3.0), r, hij


def dense_xsph(pA, kernel, eps=0.5):
    nb, r, hij = dense_neighbours(pA)
    W = np.where(nb, dense_w(kernel, r, hij), 0.0)
    rho, m = pA['rho'], pA['m']
    fac = eps * m[None, :] / (0.5 * (rho[:, None] + rho[None, :])) * W
    return (np.sum(fac * (pA['vx'][None, :] - pA['vx'][:, None]), axis=1),
            np.sum(fac * (pA['vy'][None, :] - pA['vy'][:, None]), axis=1))


def dense_summation_density(pA, kernel):
    nb, r, hij = dense_neighbours(pA)
    W = np.where(nb & (pA['label'][None, :] == 0), dense_w(kernel, r, hij), 0.0)
    return np.sum(pA['m'][None, :] * W, axis=1)


@pytest.mark.parametrize('kernel', ['cubic', 'gaussian', 'wendland'])
def test_xsph_vs_dense_pairs(kernel):
    """XSPH over all near neighbours, walls included; xsph = v + correction."""
    pA = _block_with_walls(21)
    pA['m'] *= 1 + np.random.default_rng(3).uniform(-0.2, 0.2, len(pA))
    want = dense_xsph(pA, kernel)
    m = _M()
    m.useXSPH = True
    oracle.loop(pA, m, kernel)
    f = pA['label'] == 0
    for got, v, w in ((pA['xsphx'], pA['vx'], want[0]), (pA['xsphy'], pA['vy'], want[1])):
        scale = np.max(np.abs(w[f]))
        assert scale > 0
        np.testing.assert_allclose(got[f] - v[f], w[f], rtol=0, atol=1e-12 * scale + 1e-15)
    assert np.all(pA['xsphx'][~f] == 0.0)  # walls are not updated
    m.useXSPH = False
    oracle.loop(pA, m, kernel)
    assert np.all(pA['xsphx'][f] == 0.0) and np.all(pA['xsphy'][f] == 0.0)  # WCSPH.py:188-189


def test_xsph_galilean_invariance_and_momentum():
    """The correction depends on v_j - v_i only, and sum_i m_i corr_i = 0 over
    a fluid-only set with uniform h (the pair terms are antisymmetric)."""
    pA = _block_with_walls(22)
    pA = pA[pA['label'] == 0].copy()
    pA['m'] *= 1 + np.random.default_rng(4).uniform(-0.3, 0.3, len(pA))
    m = _M()
    m.useXSPH = True
    a = pA.copy()
    oracle.loop(a, m, 'cubic')
    cx, cy = a['xsphx'] - a['vx'], a['xsphy'] - a['vy']
    scale = np.max(np.abs(cx))
    assert scale > 0
    assert abs(np.sum(a['m'] * cx)) < 1e-12 * np.sum(a['m'] * np.abs(cx))
    assert abs(np.sum(a['m'] * cy)) < 1e-12 * np.sum(a['m'] * np.abs(cy))
    b = pA.copy()
    b['vx'] += 3.0
    b['vy'] -= 1.5
    oracle.loop(b, m, 'cubic')
    np.testing.assert_allclose(b['xsphx'] - b['vx'], cx, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(b['xsphy'] - b['vy'], cy, rtol=0, atol=1e-12 * scale)
    m.epsilon = 0.25  # linear in epsilon
    c = pA.copy()
    oracle.loop(c, m, 'cubic')
    np.testing.assert_allclose(c['xsphx'] - c['vx'], 0.5 * cx, rtol=0, atol=1e-12 * scale)


@pytest.mark.parametrize('kernel', ['cubic', 'gaussian', 'wendland'])
def test_summation_density_vs_dense_pairs(kernel):
    """rho_i = sum over fluid neighbours (self included) of m_j W_ij; the
    pressure of the same loop is the one of the old density (SolverTools.py:
    122-127 run before :129-140), continuity is 0 (WCSPH.py:200-201), and the
    momentum equation sees the new densities."""
    pA = _block_with_walls(23)
    rho_old = pA['rho'].copy()
    want_rho = dense_summation_density(pA, kernel)
    m = _M()
    m.useSummationDensity = True
    b = pA.copy()
    b['rho'] = np.where(b['label'] == 0, want_rho, b['rho'])
    p_old = np.where(pA['label'] == 0, ((rho_old / m.rho0) ** m.gamma - 1.0) * m.B, 0.0) + m.Pb
    oracle.loop(pA, m, kernel)
    f = pA['label'] == 0
    np.testing.assert_allclose(pA['rho'][f], want_rho[f], rtol=1e-13)
    np.testing.assert_array_equal(pA['rho'][~f], rho_old[~f])
    # numpy's pow against the oracle's left-to-right products: (rho/rho0)^7 - 1 cancels
    np.testing.assert_allclose(pA['p'], p_old, rtol=1e-11)
    assert np.all(pA['drho'][f] == 0.0)
    # momentum with p from rho_old and rho_new in p / rho^2 and the AV mean
    _, ax, ay = dense_rates_with_p(b, m, kernel, p_old)
    scale = np.max(np.abs(ax[f]))
    np.testing.assert_allclose(pA['ax'][f], ax[f], rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(pA['ay'][f], ay[f], rtol=0, atol=1e-12 * np.max(np.abs(ay[f])))


def dense_rates_with_p(pA, method, kernel, p):
    """dense_rates with a given pressure field instead of Tait(rho)."""
    rho = pA['rho']
    lab = pA['label']
    x, y, h = pA['x'], pA['y'], pA['h']
    nb, r, hij = dense_neighbours(pA)
    dx = x[:, None] - x[None, :]
    dy = y[:, None] - y[None, :]
    q = r / hij
    with np.errstate(divide='ignore', invalid='ignore'):
        if kernel == 'cubic':
            a = 10.0 / (7.0 * np.pi) / (hij * hij)
            g = np.where(q <= 1.0, -3.0 * q * (1.0 - 0.75 * q), -0.75 * (2.0 - q) ** 2)
            norm = np.where((q > 2.0) | (r < 1e-10), 0.0, a * g / (hij * r))
        elif kernel == 'gaussian':
            dwdq = -2.0 * q / (np.pi * hij * hij) * np.exp(-q * q)
            norm = np.where((r * hij > 1e-12) & (q <= 3.0), dwdq / (r * hij), 0.0)
        else:
            a = 9.0 / (4.0 * np.pi) / (hij * hij)
            inner = 1.0 - 0.5 * q
            g = inner ** 5 * (-14.0 / 3.0) * q * (1 + 2.5 * q)
            norm = np.where((q >= 2.0) | (r < 1e-10), 0.0, a * g / (hij * r))
        dwx, dwy = norm * dx, norm * dy
        vxij = pA['vx'][:, None] - pA['vx'][None, :]
        vyij = pA['vy'][:, None] - pA['vy'][None, :]
        W = nb & (lab[None, :] == 0)
        prs = p / (rho * rho)
        dv = vxij * dx + vyij * dy
        hm = 0.5 * (h[:, None] + hij)
        mu = hm * dv / (r * r + 0.01 * hm * hm)
        PI = np.where(dv < 0, mu * (method.beta * mu - method.alpha * 0.5 * method.co) /
                      (0.5 * (rho[:, None] + rho[None, :])), 0.0)
        fac = -pA['m'][None, :] * (prs[:, None] + prs[None, :] + PI)
        ax = np.sum(np.where(W, fac * dwx, 0.0), axis=1)
        ay = np.sum(np.where(W, fac * dwy, 0.0), axis=1)
        Bm = nb & (lab[None, :] != 0) & (r <= method.r0) & (r > 1e-12)
        fr = method.r0 / r
        lj = method.D * (fr ** method.p1 - fr ** method.p2) / (r * r)
        fx = np.sum(np.where(Bm, lj * dx, 0.0), axis=1)
        fy = np.sum(np.where(Bm, lj * dy, 0.0), axis=1)
    return None, ax + fx, ay - G + fy


def test_summation_density_lattice_normalisation():
    """On a square lattice of spacing dx with h = 1.3 dx the interior density
    is m / dx^2 times the lattice sum of W dx^2: within 1e-4 of 1 for the
    cubic spline, 0.44 % for the Gaussian (its 3h support reaches past the
    3x3 cells of size 2h, so the sum depends on the position in the cell) and
    0.85 % for Wendland C4 at this h / dx. A wrong normalisation constant or a
    missing self term (W(0) dx^2 / m ~ 0.27) is off by far more."""
    dx, n = 0.02, 24
    gx, gy = np.meshgrid(np.arange(n) * dx, np.arange(n) * dx, indexing='ij')
    pA = _particles(gx.ravel(), gy.ravel(), 1.3 * dx, m=1000.0 * dx * dx)
    m = _M()
    m.useSummationDensity = True
    for kernel, tol in (('cubic', 1e-4), ('gaussian', 5e-3), ('wendland', 1e-2)):
        a = pA.copy()
        oracle.loop(a, m, kernel)
        ii = (np.abs(a['x'] - 0.5 * (n - 1) * dx) < 4 * dx) & (np.abs(a['y'] - 0.5 * (n - 1) * dx) < 4 * dx)
        assert np.all(np.abs(a['rho'][ii] / 1000.0 - 1.0) < tol), kernel


def test_pec_xsph_branch_reduces_to_plain_branch():
    """PEC with XSPH moves x by the XSPH velocity (src/Integrators/PEC.py:
    43-48,66-71): with xsph = v (the velocity the plain branch uses at that
    point: the start velocity in the predictor, the half-step one in the
    corrector) it gives the plain integrator bit for bit, and walls ignore
    xsph."""
    pA = _block_with_walls(24)
    rng = np.random.default_rng(5)
    pA['ax'] = rng.normal(0, 3, len(pA))
    pA['ay'] = rng.normal(0, 3, len(pA))
    pA['drho'] = rng.normal(0, 10, len(pA))
    a = pA.copy()
    a['xsphx'], a['xsphy'] = a['vx'], a['vy']
    b = pA.copy()
    b['xsphx'], b['xsphy'] = 7.0, -9.0  # ignored without XSPH
    dt = 1e-4
    oracle.pec_predict(a, dt, 0.02, useXSPH=True)
    oracle.pec_predict(b, dt, 0.02, useXSPH=False)
    for f in ('x', 'y', 'vx', 'vy', 'rho', 'x0', 'y0'):
        np.testing.assert_array_equal(a[f], b[f], err_msg=f)
    # a zero correction in the loop leaves xsph = the half-step velocity
    a['xsphx'], a['xsphy'] = a['vx'], a['vy']
    oracle.pec_correct(a, dt, 0.02, useXSPH=True)
    oracle.pec_correct(b, dt, 0.02, useXSPH=False)
    for f in ('x', 'y', 'vx', 'vy', 'rho'):
        np.testing.assert_array_equal(a[f], b[f], err_msg=f)
    c = pA.copy()
    c['xsphx'] = c['vx'] + 1.0  # an XSPH velocity differing by 1 m/s moves fluid x by dt/2 more
    d = pA.copy()
    oracle.pec_predict(c, dt, 0.0, useXSPH=True)
    oracle.pec_predict(d, dt, 0.0, useXSPH=False)
    w = c['label'] != 0
    np.testing.assert_array_equal(c['x'][w], d['x'][w])
    np.testing.assert_allclose(c['x'][~w] - d['x'][~w], 0.5 * dt, rtol=1e-9)


def test_dynamic_h_from_predicted_density():
    """Dynamic h (src/Tools/SolverTools.py:106-118, Solver._compute :242-246):
    after a step every fluid h is 1.3 (m / rho)^(1/2) of the density the loop
    saw, the predicted one rho + dt/2 drho (PEC.py:57), and walls keep theirs.
    On a lattice at rho = m / dx^2 that is h = 1.3 dx (sigma and the 2D
    exponent), and a zero density gives h = 0 (:115-116)."""
    pA = _block_with_walls(25)
    f = pA['label'] == 0
    rng = np.random.default_rng(6)
    pA['drho'][f] = rng.normal(0, 50, f.sum())
    pA['c'][f] = _M().co  # the Courant dt reads c (WCSPH.compute_speed_of_sound)
    h_wall = pA['h'][~f].copy()
    s = oracle.Solver(pA, _M(), 'cubic', 'pec', h=None)
    rho0, drho0 = s.pA['rho'].copy(), s.pA['drho'].copy()
    hist = s.run(1)
    dt = hist[0, 0]
    rho_half = rho0 + 0.5 * dt * drho0
    np.testing.assert_allclose(s.pA['h'][f], 1.3 * np.sqrt(s.pA['m'][f] / rho_half[f]), rtol=1e-15)
    np.testing.assert_array_equal(s.pA['h'][~f], h_wall)
    dx = 0.02
    gx, gy = np.meshgrid(np.arange(12) * dx, np.arange(12) * dx, indexing='ij')
    lat = _particles(gx.ravel(), gy.ravel(), 1.3 * dx, rho=1000.0, m=1000.0 * dx * dx)
    lat['rho'][5] = 0.0
    lat['c'] = _M().co
    s = oracle.Solver(lat, _M(), 'cubic', 'euler', h=None)
    s.run(1)
    g = np.arange(len(lat)) != 5
    np.testing.assert_allclose(s.pA['h'][g], 1.3 * dx, rtol=1e-15)
    assert s.pA['h'][5] == 0.0
