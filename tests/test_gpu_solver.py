"""Solver, NNLinkedList and TimeStep host mirrors driven on the device.

The Solver mirror runs the reference's time loop (src/Solver.py:346-475) on
the device engine; these tests pin it against the CPU oracle's loop body
(oracle/refcpu.c ref_solver_run) step for step, bit-exactly in FP64:
  * the plain loop (settle=False): same final state and dt series;
  * the settling phase: maxSettle + 1 damped steps that do not advance time,
    then TempBoundary deletion and undamped steps (src/Solver.py:404-457),
    compared with the oracle run in the same two phases.
"""
import numpy as np
import pytest

import oracle
from gpu_util import say
from hipsph import ParticleType
from hipsph.Helpers import Helpers
from hipsph import scenarios
from hipsph.Integrators.PEC import PEC
from hipsph.Kernels.CubicSpline import CubicSpline
from hipsph.Solver import Solver

pytestmark = pytest.mark.gpu

FIELDS = ('x', 'y', 'vx', 'vy', 'rho', 'ax', 'ay', 'drho')


def _small():
    return scenarios.dam_break(0.4, 0.5, 0.02, tank_w=1.0, tank_h=0.8, name='solver_small')


def _duration_for(hist, k):
    """A duration that makes the loop take exactly k steps of the dt series."""
    t = np.cumsum(hist[:, 0])
    return 0.5 * (t[k - 2] + t[k - 1]) if k > 1 else 0.5 * t[0]


def _oracle_after_setup(sv):
    return oracle.as_particles(sv.particleArray[sv.indexes].copy())


def test_solver_loop_fp64_bitexact():
    sc = _small()
    m = sc.method()
    sv = Solver(m, PEC(useXSPH=False, strict=False), CubicSpline(), duration=1.0, incrementalWriteout=False,
                exportProperties=[], h=sc.h, damping=0.0, settle=False, chunk=16)
    sv.addParticles(sc.pA.copy())
    sv.setup()
    ref = _oracle_after_setup(sv)
    K = 40
    o = oracle.Solver(ref, m, 'cubic', 'pec', h=sc.h, damping=0.0)
    hist = o.run(K)
    sv.duration = _duration_for(hist, K)
    say('solver loop: n', sv.num_particles, 'steps', K)
    sv.run()
    sv.close()
    assert sv.t_step == K
    np.testing.assert_array_equal(np.asarray(sv.dt_a), hist[:, 0])
    np.testing.assert_array_equal(np.asarray(sv.dt_c), hist[:, 1])
    np.testing.assert_array_equal(np.asarray(sv.dt_f), hist[:, 2])
    got = sv.particleArray[sv.indexes]
    for f in FIELDS:
        np.testing.assert_array_equal(got[f], o.pA[f], err_msg=f)


def test_solver_loop_xsph_summation_density_fp64_bitexact():
    """The same loop with WCSPH(useXSPH=True, useSummationDensity=True) and
    PEC(useXSPH=True): XSPH velocities and summed densities on the device."""
    sc = _small()
    m = sc.method()
    m.useXSPH = True
    m.useSummationDensity = True
    sv = Solver(m, PEC(useXSPH=True, strict=False), CubicSpline(), duration=1.0, incrementalWriteout=False,
                exportProperties=[], h=sc.h, damping=0.0, settle=False, chunk=16)
    pA = sc.pA.copy()
    f = pA['label'] == 0
    pA['vx'][f] = np.random.default_rng(8).normal(0, 0.1, f.sum())
    sv.addParticles(pA)
    sv.setup()
    ref = _oracle_after_setup(sv)
    K = 40
    o = oracle.Solver(ref, m, 'cubic', 'pec', h=sc.h, damping=0.0)
    hist = o.run(K)
    sv.duration = _duration_for(hist, K)
    say('solver loop xsph+sd: n', sv.num_particles, 'steps', K)
    sv.run()
    sv.close()
    assert sv.t_step == K
    np.testing.assert_array_equal(np.asarray(sv.dt_a), hist[:, 0])
    got = sv.particleArray[sv.indexes]
    for fld in FIELDS + ('xsphx', 'xsphy'):
        np.testing.assert_array_equal(got[fld], o.pA[fld], err_msg=fld)


def test_solver_dynamic_h_fp64_bitexact():
    """Solver(h=None): h from the input density at setup (src/Solver.py:185-189),
    then every step on the device (:242-246)."""
    sc = _small()
    m = sc.method()
    sv = Solver(m, PEC(useXSPH=False, strict=False), CubicSpline(), duration=1.0, incrementalWriteout=False,
                exportProperties=[], h=None, damping=0.0, settle=False, chunk=16)
    pA = sc.pA.copy()
    f = pA['label'] == 0
    pA['rho'][f] *= 1 + np.random.default_rng(9).uniform(-0.01, 0.01, f.sum())
    h0 = 1.3 * np.sqrt(pA['m'][f] / pA['rho'][f])  # before initialize resets rho in place
    sv.addParticles(pA)
    sv.setup()
    ref = _oracle_after_setup(sv)
    np.testing.assert_allclose(ref['h'][ref['label'] == 0], h0, rtol=1e-15)
    assert np.ptp(h0) > 0
    K = 40
    o = oracle.Solver(ref, m, 'cubic', 'pec', h=None, damping=0.0)
    hist = o.run(K)
    sv.duration = _duration_for(hist, K)
    say('solver loop dynamic h: n', sv.num_particles, 'steps', K)
    sv.run()
    sv.close()
    assert sv.t_step == K
    np.testing.assert_array_equal(np.asarray(sv.dt_a), hist[:, 0])
    got = sv.particleArray[sv.indexes]
    for fld in FIELDS + ('h',):
        np.testing.assert_array_equal(got[fld], o.pA[fld], err_msg=fld)


def test_solver_settling_and_temp_boundary_fp64_bitexact():
    sc = _small()
    dx = sc.dx
    # a temporary gate right of the column, removed once settled
    gate = Helpers.rect(0.4 + 2 * dx, 0.4 + 2 * dx, 0.0, 0.5, dx, mass=1000.0 * dx * dx,
                        label=ParticleType.TempBoundary)
    gate['h'] = sc.h
    pA = np.concatenate((sc.pA.copy(), gate))
    m = sc.method()
    max_settle = 3
    sv = Solver(m, PEC(useXSPH=False, strict=False), CubicSpline(), duration=1.0, incrementalWriteout=False,
                exportProperties=[], h=sc.h, damping=0.05, maxSettle=max_settle, chunk=8)
    sv.addParticles(pA)
    sv.setup()
    ref = _oracle_after_setup(sv)
    n_settle = max_settle + 1  # KE0 = 0: the settle check fires at t_step > maxSettle
    o1 = oracle.Solver(ref, m, 'cubic', 'pec', h=sc.h, damping=0.05)
    h1 = o1.run(n_settle)
    keep = o1.pA['label'] != ParticleType.TempBoundary
    ref2 = oracle.as_particles(o1.pA[keep].copy())
    K = 12
    o2 = oracle.Solver(ref2, m, 'cubic', 'pec', h=sc.h, damping=0.0)
    h2 = o2.run(K)
    sv.duration = _duration_for(h2, K)
    say('solver settle: n', sv.num_particles, 'gate', len(gate), 'settle', n_settle, 'steps', K)
    sv.run()
    sv.close()
    assert sv.settled
    assert sv.t_step == n_settle + K
    assert sv.settleTime == pytest.approx(float(np.sum(h1[:, 0])), rel=1e-15)
    np.testing.assert_array_equal(np.asarray(sv.dt_a), np.concatenate((h1[:, 0], h2[:, 0])))
    deleted = sv.particleArray['deleted']
    assert np.all(deleted == (sv.particleArray['label'] == ParticleType.TempBoundary))
    assert np.all(sv.particleArray['p'][deleted] == -1e15)
    got = sv.particleArray[~deleted]
    for f in FIELDS:
        np.testing.assert_array_equal(got[f], o2.pA[f], err_msg=f)


def test_solver_export_and_store():
    sc = _small()
    m = sc.method()
    sv = Solver(m, PEC(useXSPH=False, strict=False), CubicSpline(), duration=1.0, incrementalWriteout=False,
                exportProperties=['x', 'p'], h=sc.h, damping=0.0, settle=False)
    sv.addParticles(sc.pA.copy())
    sv.setup()
    ref = _oracle_after_setup(sv)
    K = 5
    o = oracle.Solver(ref, m, 'cubic', 'pec', h=sc.h, damping=0.0)
    hist = o.run(K)
    sv.duration = _duration_for(hist, K)
    sv.run()
    sv.close()
    assert len(sv.export['x']) == K and len(sv.export['p']) == K
    np.testing.assert_array_equal(sv.export['x'][-1][sv.indexes], o.pA['x'])
    assert len(sv.data) == K + 1


def test_nnlinkedlist_mirror_matches_oracle():
    from hipsph.Tools.NNLinkedList import NNLinkedList
    sc = _small()
    pA = sc.pA.copy()
    nn = NNLinkedList(2.0)
    nn.update(pA)
    ref = oracle.NN(2.0)
    ref.update(oracle.as_particles(pA))
    d = ref.dump(oracle.as_particles(pA))
    for i in (0, 17, len(pA) // 2, len(pA) - 1):
        h_ij, q, r, idx = nn.near(i, pA)
        s, e = d['off'][i], d['off'][i + 1]
        np.testing.assert_array_equal(idx.astype(np.int64), d['idx'][s:e])
        np.testing.assert_array_equal(q, d['q'][s:e])
        np.testing.assert_array_equal(r, d['r'][s:e])
        np.testing.assert_array_equal(h_ij, 0.5 * (pA['h'][i] + pA['h'][d['idx'][s:e]]))
    nn.close()


def test_timestep_mirror_matches_oracle():
    from hipsph.Equations.TimeStep import TimeStep
    rng = np.random.default_rng(7)
    sc = _small()
    f = sc.pA[sc.pA['label'] == 0].copy()
    f['c'] = sc.method().co
    f['ax'] = rng.normal(0, 50, len(f))
    f['ay'] = rng.normal(0, 50, len(f))
    got = TimeStep().compute(len(f), f)
    want = oracle.timestep(f)
    assert got == tuple(want)
