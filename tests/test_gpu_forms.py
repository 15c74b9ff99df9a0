"""The loop's other forms on the device against the CPU oracle: XSPH,
summation density and dynamic h.

The oracle's XSPH (src/Equations/XSPH.py via WCSPH.compute_velocity), its
summation density (src/Tools/SolverTools.py:129-140), PEC's XSPH branch
(src/Integrators/PEC.py:43-48,66-71) and dynamic h (SolverTools.py:106-118,
Solver._compute) are pinned by a dense-pair model and invariants in
tests/test_oracle_invariants.py. Here, through the C ABI:
  * one loop on a perturbed dam break, every combination of the two: FP64
    cubic and Wendland bit-identical (XSPH velocity, summed density, rates),
    Gaussian FP64 to 1e-12, FP32C pairs to 2e-5 of the largest value;
  * whole steps: FP64 cubic trajectories and dt series bit-identical over
    200 steps for PEC with XSPH, PEC with summation density, both, and Euler
    with summation density;
  * dynamic h: FP64 cubic trajectories, dt series and h bit-identical over
    200 steps (PEC, Euler) and with XSPH and summation density on as well;
  * hipGraph replay equals the eager steps with both on;
  * FP32 pairs against FP64 with each form on: relative density within 1e-5
    and positions within 1e-3 dx after 1,000 steps;
  * without XSPH the fluid's XSPH velocity reads 0 after a loop, the walls'
    keeps its value;
  * what the device path does not cover fails loudly: the split passes and
    the adjoint tape with either form on.
"""
import numpy as np
import pytest

import oracle
from gpu_util import make_engine, oracle_loop, perturbed, rel_err, say
from hipsph import scenarios
from hipsph._lib import HipSphError

pytestmark = pytest.mark.gpu

OUT = ('xsphx', 'xsphy', 'rho', 'ax', 'ay', 'drho')
STATE = ('x', 'y', 'vx', 'vy', 'rho', 'ax', 'ay', 'drho', 'xsphx', 'xsphy')


def _method(sc, xsph, sd):
    m = sc.method()
    m.useXSPH = xsph
    m.useSummationDensity = sd
    return m


@pytest.mark.parametrize('kernel', ['cubic', 'gaussian', 'wendland'])
@pytest.mark.parametrize('flags', [(True, False), (False, True), (True, True)])
def test_loop_vs_oracle(kernel, flags):
    sc = scenarios.dam_break_c1()
    pA = perturbed(sc.pA, 7)
    m = _method(sc, *flags)
    ref, _ = oracle_loop(pA, m, kernel)
    f = pA['label'] == 0
    for prec in ('fp64', 'fp32c'):
        say('loop', kernel, prec, 'xsph', flags[0], 'sd', flags[1])
        e = make_engine(pA, m, kernel=kernel, precision=prec)
        e.nn_update()
        e.compute()
        d = e.download()
        e.close()
        for k in OUT:
            if prec == 'fp64' and kernel != 'gaussian':
                np.testing.assert_array_equal(d[k], ref[k], err_msg=k)
            else:
                tol = 1e-12 if prec == 'fp64' else 2e-5
                assert rel_err(d[k], ref[k], f) <= tol, (k, rel_err(d[k], ref[k], f))
        if flags[0]:
            assert np.max(np.abs(ref['xsphx'][f] - ref['vx'][f])) > 0  # the correction is exercised
        if flags[1]:
            assert np.all(d['drho'][f] == 0.0)


def _run_both(pA, m, h, steps, integrator='pec', graph=False, dynamic_h=False):
    e = make_engine(pA, m, integrator=integrator, dynamic_h=dynamic_h)
    e.step(steps, graph=graph)
    st = e.status()
    out = e.download()
    hist = e.dt_history(steps)
    e.close()
    s = oracle.Solver(pA, m, 'cubic', integrator, h=None if dynamic_h else h, nthreads=8)
    rhist = s.run(steps)
    return out, hist, st, s.pA, rhist, s


@pytest.mark.parametrize('case', ['pec_xsph', 'pec_sd', 'pec_xsph_sd', 'euler_sd'])
def test_steps_fp64_bitexact(case):
    sc = scenarios.dam_break_c1()
    pA = perturbed(sc.pA, 3, vscale=0.1, rscale=0.001)
    m = _method(sc, 'xsph' in case, 'sd' in case)
    integrator = case.split('_')[0]
    say('steps', case)
    out, hist, st, ref, rhist, s = _run_both(pA, m, sc.h, 200, integrator)
    assert st['steps'] == 200 and st['flags'] & 3 == 0
    np.testing.assert_array_equal(hist, rhist)
    assert st['t'] == s.t
    for k in STATE:
        np.testing.assert_array_equal(out[k], ref[k], err_msg=k)


@pytest.mark.parametrize('case', ['pec', 'euler', 'pec_xsph_sd'])
def test_steps_dynamic_h_fp64_bitexact(case):
    """Solver(h=None): the grid of each step from the previous h, then
    h = 1.3 (m / rho)^(1/2) of the density the loop sees."""
    sc = scenarios.dam_break_c1()
    pA = perturbed(sc.pA, 5, vscale=0.1, rscale=0.01)
    m = _method(sc, 'xsph' in case, 'sd' in case)
    integrator = case.split('_')[0]
    say('dynamic h steps', case)
    out, hist, st, ref, rhist, s = _run_both(pA, m, sc.h, 200, integrator, dynamic_h=True)
    assert st['steps'] == 200 and st['flags'] & 3 == 0
    np.testing.assert_array_equal(hist, rhist)
    f = pA['label'] == 0
    assert np.ptp(out['h'][f]) > 0  # h does vary
    for k in STATE + ('h',):
        np.testing.assert_array_equal(out[k], ref[k], err_msg=k)


@pytest.mark.parametrize('case', ['xsph', 'sd', 'xsph_sd_dynh'])
def test_fp32c_forms_vs_fp64_after_1000_steps(case):
    """FP32 pairs with FP64 integration against the FP64 engine with the
    loop's other forms on: relative density within 1e-5 and positions within
    1e-3 dx after 1,000 steps, the bar of the plain step."""
    sc = scenarios.dam_break_c1()
    pA = perturbed(sc.pA, 11, vscale=0.1, rscale=0.002)
    m = _method(sc, 'xsph' in case, 'sd' in case)
    dyn = 'dynh' in case
    a = make_engine(pA, m, precision='fp64', dynamic_h=dyn)
    b = make_engine(pA, m, precision='fp32c', dynamic_h=dyn)
    a.step(1000)
    b.step(1000)
    oa, ob = a.download(), b.download()
    a.close()
    b.close()
    f = pA['label'] == 0
    drho = np.max(np.abs(ob['rho'][f] - oa['rho'][f]) / oa['rho'][f])
    dx = np.max(np.hypot(ob['x'][f] - oa['x'][f], ob['y'][f] - oa['y'][f])) / sc.dx
    say('fp32c', case, 'after 1000 steps: max rel rho', drho, 'max |dx|/dx', dx)
    assert drho < 1e-5
    assert dx < 1e-3


def test_graph_replay_with_xsph_and_summation_density():
    sc = scenarios.dam_break_c1()
    pA = perturbed(sc.pA, 4, vscale=0.1, rscale=0.001)
    m = _method(sc, True, True)
    outs = []
    for graph in (False, True):
        e = make_engine(pA, m)
        e.step(60, graph=graph)
        outs.append(e.download())
        e.close()
    for k in STATE:
        np.testing.assert_array_equal(outs[1][k], outs[0][k], err_msg=k)


def test_xsph_velocity_without_xsph():
    """Without XSPH the loop sets the fluid's XSPH velocity to 0 and leaves
    the walls' (WCSPH.py:187-189): after steps the download says so, however
    the buffer memory was used before; use_xsph is fixed for a context."""
    sc = scenarios.dam_break_c1()
    pA = sc.pA.copy()
    rng = np.random.default_rng(12)
    pA['xsphx'] = rng.normal(size=len(pA))
    pA['xsphy'] = rng.normal(size=len(pA))
    warm = make_engine(pA, _method(sc, True, False))  # leaves XSPH velocities in device memory
    warm.step(3)
    warm.close()
    e = make_engine(pA, sc.method())
    d0 = e.download()
    np.testing.assert_array_equal(d0['xsphx'], pA['xsphx'])  # before any loop: as uploaded
    e.step(5)
    d = e.download()
    f = pA['label'] == 0
    assert np.all(d['xsphx'][f] == 0.0) and np.all(d['xsphy'][f] == 0.0)
    np.testing.assert_array_equal(d['xsphx'][~f], pA['xsphx'][~f])
    np.testing.assert_array_equal(d['xsphy'][~f], pA['xsphy'][~f])
    ids = np.nonzero(f)[0][:5]
    with pytest.raises(HipSphError):
        e.scatter(ids, {'xsphx': np.ones(5)})
    with pytest.raises(HipSphError):
        e.set_params(use_xsph=True)
    e.close()


def test_unsupported_combinations_fail_loudly():
    from hipsph.adjoint import Tape
    sc = scenarios.dam_break_c1()
    e = make_engine(sc.pA, sc.method(), dynamic_h=True)
    with pytest.raises(HipSphError):
        Tape(e, 2)
    e.close()
    e = make_engine(sc.pA, _method(sc, True, False))
    e.nn_update()
    with pytest.raises(HipSphError):
        e.continuity()
    with pytest.raises(HipSphError):
        e.momentum()
    with pytest.raises(HipSphError):
        Tape(e, 2)
    e.close()
