"""Host-side logic: data model, lattice generator, scenarios, parameter plumbing."""
import math

import numpy as np
import pytest

import oracle
from hipsph import Common, scenarios
from hipsph.Helpers import Helpers
from hipsph.Methods.WCSPH import WCSPH
from hipsph.device import EngineParams


def test_dtypes_match_reference_layout():
    assert Common.particle_dtype.itemsize == 154
    assert Common.computed_dtype.itemsize == 113
    assert Common.particle_dtype == oracle.particle_dtype
    assert Common.computed_dtype == oracle.computed_dtype
    assert Common.get_label_code('boundary') == Common.ParticleType.Boundary
    with pytest.raises(Exception):
        Common.get_label_code('nope')


def test_rect_counts_and_pack():
    """src/Helpers.py:45-51: N = ceil(range / r0) points per axis, inclusive ends."""
    a = Helpers.rect(0, 1.0, 0, 2.0, 0.02)
    assert len(a) == 50 * 100
    assert a['x'].min() == 0 and a['x'].max() == 1.0 and a['h'].max() == 0
    b = Helpers.rect(0, 1.0, 0, 1.0, 0.25, pack=True)
    assert len(b) == 16
    assert np.allclose(b['y'][0:4] - a['y'][0:4] * 0, np.linspace(0, 1, 4) + 0.125)
    line = Helpers.rect(-0.06, 4.06, -0.02, -0.02, 0.02)
    assert len(line) == math.ceil(4.12 / 0.02)


def test_scenario_sizes():
    c1 = scenarios.dam_break_c1()
    assert c1.n_fluid == 5000
    assert 1200 < c1.n - c1.n_fluid < 1700
    assert np.all(c1.pA['h'] == 1.3 * 0.02)
    m = c1.method()
    f = c1.pA['label'] == 0
    # hydrostatic density decreases with height, equals rho0 at the free surface y = H
    top = np.argmax(c1.pA['y'][f])
    assert c1.pA['rho'][f][top] == pytest.approx(1000.0)
    assert m.co == pytest.approx(10 * math.sqrt(2 * 9.81 * 2.0))


def test_wcsph_constants_match_oracle():
    m = WCSPH(height=2.0, r0=0.02, rho0=1000.0, useXSPH=False, Pb=0, useSummationDensity=False)
    r = oracle.wcsph(height=2.0, r0=0.02, rho0=1000.0)
    for f in ('co', 'B', 'D', 'p1', 'p2', 'gamma', 'alpha', 'beta', 'epsilon'):
        assert getattr(m, f) == getattr(r, f), f


def test_engine_params_to_c():
    m = WCSPH(2.0, 0.02, 1000.0)
    p = EngineParams.from_method(123, m, kernel='gaussian', precision='fp32c', integrator='euler').to_c()
    assert (p.n, p.kernel, p.precision, p.integrator) == (123, 1, 1, 1)
    assert p.co == m.co and p.B == m.B and p.D == m.D and p.qmax == 3.0 and p.scale == 2.0
    assert (p.use_xsph, p.use_summation_density, p.epsilon) == (0, 0, 0.5)
    q = EngineParams.from_method(1, WCSPH(2.0, 0.02, 1000.0, useXSPH=True, useSummationDensity=True)).to_c()
    assert (q.use_xsph, q.use_summation_density, q.epsilon) == (1, 1, 0.5)


def test_solver_mirror_signature_and_host_policy():
    """Solver keeps the reference constructor (src/Solver.py:46) and fails loudly
    for what the device path does not do; the host-stop policy is pure logic."""
    import inspect
    from hipsph.Solver import Solver
    ref_params = ['method', 'integrator', 'kernel', 'duration', 'quick', 'incrementalWriteout', 'incrementalFile',
                  'incrementalFreq', 'exportProperties', 'kE', 'maxSettle', 'timeStep', 'h', 'coupling',
                  'couplingIntegrator', 'couplingProperties', 'customSettle', 'damping']
    sig = inspect.signature(Solver.__init__)
    names = [p for p in sig.parameters if p != 'self']
    assert names[:len(ref_params)] == ref_params
    assert all(sig.parameters[p].kind == inspect.Parameter.KEYWORD_ONLY for p in names[len(ref_params):])
    d = sig.parameters
    assert (d['duration'].default, d['incrementalFreq'].default, d['kE'].default, d['maxSettle'].default,
            d['damping'].default) == (1.0, 1000, 0.8, 500, 0.05)
    sv = Solver(None, None, None, exportProperties=[], incrementalWriteout=True, incrementalFreq=10, chunk=64)
    # write-outs happen at t_step % incrementalFreq == 0 (src/Solver.py:485-486)
    assert sv._next_stop(0) == 1 and sv._next_stop(1) == 10 and sv._next_stop(5) == 6
    sv.exportProperties = ['x']
    assert sv._next_stop(3) == 1
    sv2 = Solver(None, None, None, exportProperties=[], incrementalWriteout=False, chunk=64)
    assert sv2._next_stop(7) == 64
    # a coupling hook (src/Solver.py:381-398) needs the coupled particles'
    # integrator on the device: Newmark-beta; without one, run() refuses
    from hipsph import scenarios
    from hipsph.Integrators.PEC import PEC
    from hipsph.Kernels.CubicSpline import CubicSpline
    sc = scenarios.dam_break(0.4, 0.5, 0.02, tank_w=1.0, tank_h=0.8)
    sv3 = Solver(sc.method(), PEC(useXSPH=False, strict=False), CubicSpline(), duration=1.0,
                 incrementalWriteout=False, exportProperties=[], h=sc.h, coupling=lambda *a: None)
    sv3.addParticles(sc.pA.copy())
    sv3.num_particles = len(sv3.particleArray)  # as setup would leave it (setup needs the device)
    with pytest.raises(NotImplementedError):
        sv3.run()


def test_solver_rejects_mismatched_xsph_flags():
    """src/Methods/WCSPH.py:56: XSPH has to be on in both the integrator and
    the method; the mirror refuses a PEC/WCSPH pair that disagree."""
    from hipsph import scenarios
    from hipsph.Integrators.PEC import PEC
    from hipsph.Kernels.CubicSpline import CubicSpline
    from hipsph.Solver import Solver
    sc = scenarios.dam_break(0.4, 0.5, 0.02, tank_w=1.0, tank_h=0.8)
    for pec_xsph, m_xsph in ((True, False), (False, True)):
        m = sc.method()
        m.useXSPH = m_xsph
        sv = Solver(m, PEC(useXSPH=pec_xsph, strict=False), CubicSpline(), duration=1.0, incrementalWriteout=False,
                    exportProperties=[], h=sc.h)
        sv.addParticles(sc.pA.copy())
        with pytest.raises(ValueError):
            sv.setup()


def test_kinetic_energy_and_find_active_known_answers():
    """The reference's own known answers: test/test_kinetic_energy.py:11-23
    (zero records give 0; m = 1, vx = 1 on 100 records give 100 * 0.5) and
    test/test_tools.py:7-21 (ten records, one deleted: count 9, the mask
    false at the deleted one)."""
    from hipsph.Solver import KineticEnergy, findActive
    pA = np.zeros(100, dtype=Common.particle_dtype)
    assert KineticEnergy(len(pA), pA) == 0
    pA['m'] = 1
    pA['vx'] = 1
    assert KineticEnergy(len(pA), pA) == 100 * 0.5 * 1 * 1
    no = np.zeros(10, dtype=Common.particle_dtype)
    count, indexes = findActive(len(no), no)
    assert count == 10 and len(indexes) == 10 and indexes.all()
    no['deleted'][5] = True
    count, indexes = findActive(len(no), no)
    assert count == 9 and len(indexes) == 10 and not indexes[5] and indexes.sum() == 9


def test_courant_known_answers():
    """test/test_eq_courant.py:8-11 (alpha h / c for h, c = 0, 1; 1, 1; 2, 4)
    on TimeStep.courant (src/Equations/TimeStep.py:47-49), the form the
    solver's time step uses."""
    from hipsph.Equations.TimeStep import TimeStep
    ts = TimeStep()
    assert ts.courant(0.4, 0.0, 1.0) == 0.0
    assert ts.courant(0.4, 1.0, 1.0) == 0.4
    assert ts.courant(0.4, 2.0, 4.0) == 0.5 * 0.4
    assert ts.force(0.25, 1.0, 0.0) == 1e10 and ts.force(0.25, 4.0, 1.0) == 0.5


# --------------------------------------------------------------------------
# the border-cell bounds of the fused PEC step's neighbour update: the two
# facts k_bounds and k_dt_final rely on, checked in numpy
# --------------------------------------------------------------------------
def _bin(x, y, xmin, ymin, cs, ncx, ncy):
    """The neighbour update's cells: floor((x - xmin) / cs) per axis, flat key
    cx + ncx cy, anything outside [0, ncx ncy) in the out-of-grid bin."""
    cx = np.floor((x - xmin) / cs)
    cy = np.floor((y - ymin) / cs)
    key = cx + ncx * cy
    return np.where((key >= 0) & (key < ncx * ncy), key, ncx * ncy).astype(np.int64)


def _border_keys(ncx, ncy):
    """Rows 0, ncy - 2, ncy - 1, the out-of-grid bin, and for each row r <=
    ncy - 3 the keys of columns ncx - 2, ncx - 1 of row r and column 0 of
    row r + 1 (one run of consecutive keys)."""
    keys = set(range(0, ncx)) | set(range((ncy - 2) * ncx, ncy * ncx + 1))
    for r in range(ncy - 2):
        keys |= set(range(r * ncx + ncx - 2, (r + 1) * ncx + 1))
    return keys


@pytest.mark.parametrize('seed', range(6))
def test_border_cells_hold_the_extremes(seed):
    """Bin particles on the grid of their bounds, then move each by at most D
    per axis with 2 D just under the cell size: the new extremes all lie among
    the particles of the border cells. Seeds 0-2 put the old maximum on an
    exact multiple of the cell size (so a flat key carries it past the last
    column or row), seeds 3-5 crowd particles against cell edges."""
    rng = np.random.default_rng(seed)
    cs = 0.01
    n = 20000
    ncx0, ncy0 = int(rng.integers(3, 40)), int(rng.integers(3, 40))
    x = rng.uniform(0.0, ncx0 * cs, n)
    y = rng.uniform(0.0, ncy0 * cs, n)
    if seed < 3:
        x[0], y[0] = 0.0, 0.0
        x[1], y[1] = ncx0 * cs, ncy0 * cs  # the maximum on a cell edge
    else:
        x = np.round(x / cs) * cs + rng.choice([-1e-13, 0.0, 1e-13], n)
        y = np.round(y / cs) * cs + rng.choice([-1e-13, 0.0, 1e-13], n)
    xmin, xmax, ymin, ymax = x.min(), x.max(), y.min(), y.max()
    ncx = max(1, int(math.ceil((1.0 / cs) * (xmax - xmin))))
    ncy = max(1, int(math.ceil((1.0 / cs) * (ymax - ymin))))
    assert ncx >= 3 and ncy >= 3
    key = _bin(x, y, xmin, ymin, cs, ncx, ncy)
    border = np.isin(key, np.fromiter(_border_keys(ncx, ncy), np.int64))
    D = 0.5 * cs * (1.0 - 2e-6)
    for trial in range(20):
        # the worst case: every interior particle moves D toward the outside
        sx = rng.choice([-1.0, 1.0], n) if trial else np.sign(x - 0.5 * (xmin + xmax))
        sy = rng.choice([-1.0, 1.0], n) if trial else np.sign(y - 0.5 * (ymin + ymax))
        nx = x + D * sx * (rng.uniform(0.0, 1.0, n) if trial % 2 else 1.0)
        ny = y + D * sy * (rng.uniform(0.0, 1.0, n) if trial % 2 else 1.0)
        assert nx[border].min() == nx.min() and nx[border].max() == nx.max()
        assert ny[border].min() == ny.min() and ny[border].max() == ny.max()


def test_border_cells_need_the_last_two_columns():
    """Plants the mistake of walking only the last column: with the old
    maximum just past a cell edge, a particle of column ncx - 2 overtakes it,
    so the last two columns (and rows) are needed."""
    cs = 0.01
    x = np.array([0.0, 3.0 * cs - 1e-9, 3.0 * cs + 1e-6, 0.5 * cs, 0.5 * cs])
    y = np.array([0.0, 1.5 * cs, 1.5 * cs, 3.5 * cs, 0.0])
    ncx = int(math.ceil((1.0 / cs) * (x.max() - x.min())))
    ncy = int(math.ceil((1.0 / cs) * (y.max() - y.min())))
    key = _bin(x, y, x.min(), y.min(), cs, ncx, ncy)
    assert key[1] % ncx == ncx - 2 and key[2] % ncx == ncx - 1
    nx = x + np.array([0.0, 0.4 * cs, -0.4 * cs, 0.0, 0.0])
    assert nx.argmax() == 1  # column ncx - 2 holds the new maximum


@pytest.mark.parametrize('seed', range(4))
def test_pec_displacement_bound(seed):
    """k_dt_final's bound: a fluid particle binned at x + dt0 v / 2, corrected
    by PEC (src/Integrators/PEC.py:61-89, with damping) and predicted again
    with dt1 lies within dt0 (V + dt0 A0 / 2) + dt0 V / 2 + dt1 (V + dt0 A1) / 2
    of where it was binned, per axis, and |v'| <= V + dt0 A1, for |v| <= V,
    the predictor's rates |a0| <= A0 and the new rates |a1| <= A1."""
    rng = np.random.default_rng(seed)
    n = 200000
    x0 = rng.uniform(-5.0, 5.0, n)
    v0 = rng.uniform(-8.0, 8.0, n) * rng.choice([0.0, 1.0], n, p=[0.05, 0.95])
    a0 = rng.uniform(-1e4, 1e4, n) * rng.uniform(0.0, 1.0, n) ** 3
    a1 = rng.uniform(-1e4, 1e4, n) * rng.uniform(0.0, 1.0, n) ** 3
    dt0 = rng.uniform(1e-6, 1e-3, n)
    dt1 = dt0 * rng.uniform(0.5, 2.0, n)
    damp = rng.choice([0.0, 0.0, 0.3, 5.0], n)
    binned = x0 + 0.5 * dt0 * v0
    vh = (v0 + 0.5 * dt0 * a0) / (1 + 0.5 * damp)
    x1 = 2 * (x0 + 0.5 * dt0 * vh) - x0
    v1 = 2 * ((v0 + 0.5 * dt0 * a1) / (1 + 0.5 * damp)) - v0
    pred = x1 + 0.5 * dt1 * v1
    V, A0, A1 = np.abs(v0), np.abs(a0), np.abs(a1)
    V1 = V + dt0 * A1
    D = dt0 * (V + 0.5 * dt0 * A0) + 0.5 * dt0 * V + 0.5 * dt1 * V1
    slack = 1e-12 * (np.abs(x0) + 1.0)
    assert np.all(np.abs(v1) <= V1 * (1 + 1e-12) + 1e-300)
    assert np.all(np.abs(pred - binned) <= D * (1 + 1e-6) + slack)
    # the bound is not loose by more than the damping and the signs allow
    tight = (damp == 0) & (np.sign(v0) == np.sign(a0)) & (np.sign(v0) == np.sign(a1))
    assert np.max((np.abs(pred - binned) / np.maximum(D, 1e-300))[tight]) > 0.5
