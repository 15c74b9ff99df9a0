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
