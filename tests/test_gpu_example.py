"""The reference's own dam-break example, run through the drop-in unchanged.

examples/DamBreak.py:57-82 builds its solver from the plugin classes; here
the same calls run with hipsph's classes (only the imports differ) on the
example's particles (scenarios.dam_break_example, the geometry of
examples/DamBreak.py:16-53):
  * 2,500 hex-packed fluid particles in a 25 m column (r0 = 0.5 m);
  * zero-mass, zero-density walls with h = 0, so the cell size falls back to
    1 m (src/Tools/NNLinkedList.py:103-110), and a TempBoundary gate that is
    deleted after the settling phase (501 damped steps: KE0 = 0);
  * Wendland, XSPH in the method and in PEC(strict=False), h = 1.6 r0.

Checks:
  * settling plus 300 steps: every field and the dt series bit-identical to
    the CPU oracle run through the same two phases;
  * the whole 5 s run: bit-identical to the oracle at the end, and the surge
    front Z and the column height H at t* = 0.71, 1.39, 2.10, 3.10, computed
    as examples/post-2d.ipynb cell 4 computes them from the saved file,
    set beside the notebook's own output of the reference ("JJK" column),
    which the code as it stands does not reproduce (see that test).
"""
import math
import os

import numpy as np
import pytest

import oracle
from gpu_util import FIELDS, say
from hipsph import ParticleType, scenarios
from hipsph.Integrators.PEC import PEC
from hipsph.Kernels.Wendland import Wendland
from hipsph.Methods.WCSPH import WCSPH
from hipsph.Solver import Solver

pytestmark = pytest.mark.gpu

N = 50
MAX_SETTLE = 500
# examples/post-2d.ipynb cell 5 output: the reference's run of this example
T_STAR = (0.71, 1.39, 2.10, 3.10)
H_JJK = (0.99, 0.89, 0.77, 0.58)
Z_JJK = (1.63, 2.69, 4.05, 6.12)
# the notebook's experiment columns (hard-coded there, source not named)
H_EXP = (0.90, 0.76, 0.57, 0.32)
Z_EXP = (1.33, 2.25, 3.22, 4.80)


def _example(duration):
    """examples/DamBreak.py:57-78 with hipsph's classes."""
    rho0 = 1000.0
    XSPH = True
    height = 25.0
    r0, pA = scenarios.dam_break_example(N)
    kernel = Wendland()
    method = WCSPH(height=height, r0=r0, rho0=rho0, useXSPH=XSPH, Pb=0, useSummationDensity=False)
    integrator = PEC(useXSPH=XSPH, strict=False)
    solver = Solver(method, integrator, kernel, duration, incrementalWriteout=False, h=1.6 * r0)
    solver.addParticles(pA)
    solver.setup()
    return solver, method, r0


def _initial(sv):
    """The active records after setup, before the run (the oracle's input)."""
    return oracle.as_particles(sv.particleArray[sv.indexes].copy())


def _oracle_phases(ref, method, r0, steps_after):
    """The oracle through the example's two phases: maxSettle + 1 damped
    steps (time held), the gate deleted, then `steps_after` undamped steps."""
    o1 = oracle.Solver(ref, method, 'wendland', 'pec', h=1.6 * r0, damping=0.05, nthreads=16)
    h1 = o1.run(MAX_SETTLE + 1)
    keep = o1.pA['label'] != ParticleType.TempBoundary
    o2 = oracle.Solver(oracle.as_particles(o1.pA[keep].copy()), method, 'wendland', 'pec', h=1.6 * r0,
                       damping=0.0, nthreads=16)
    h2 = o2.run(steps_after) if steps_after else np.zeros((0, 3))
    return h1, h2, o2


def _check_bitexact(sv, h1, h2, o2):
    np.testing.assert_array_equal(np.asarray(sv.dt_a), np.concatenate((h1[:, 0], h2[:, 0])))
    np.testing.assert_array_equal(np.asarray(sv.dt_c), np.concatenate((h1[:, 1], h2[:, 1])))
    np.testing.assert_array_equal(np.asarray(sv.dt_f), np.concatenate((h1[:, 2], h2[:, 2])))
    act = sv.particleArray[~sv.particleArray['deleted']]
    for f in FIELDS + ('xsphx', 'xsphy'):
        np.testing.assert_array_equal(act[f], o2.pA[f], err_msg=f)


def test_example_grid_falls_back_to_unit_cells():
    """Zero-h walls: min h over the active set is 0, so the cell size is 1 m."""
    sv, method, r0 = _example(1.0)
    sv.t = 0.0
    e = sv._make_engine(0.05, advance_time=False)
    e.nn_update()
    g = e.grid()
    sv.close()
    say('example grid', g)
    assert g['cell_size'] == 1.0
    assert (g['ncx'], g['ncy']) == (181, 31)


def test_example_settle_and_300_steps_bitexact_vs_oracle():
    sv, method, r0 = _example(1.0)
    K = 300
    h1, h2, o2 = _oracle_phases(_initial(sv), method, r0, K)
    t = np.cumsum(h2[:, 0])
    sv.duration = 0.5 * (t[K - 2] + t[K - 1])
    say('example: n', sv.num_particles, 'settle', MAX_SETTLE + 1, 'then', K, 'steps')
    sv.run()
    sv.close()
    assert sv.settled and sv.t_step == MAX_SETTLE + 1 + K
    assert int(np.sum(sv.particleArray['deleted'])) == 61
    _check_bitexact(sv, h1, h2, o2)


def _notebook_h_z(path):
    """examples/post-2d.ipynb cells 2 and 4 on a saved run."""
    with np.load(path, allow_pickle=False) as z:
        pA, x, y, dta, stl = z['particleArray'], z['x'], z['y'], z['dt_a'], float(z['settleTime'])
    times = np.array(T_STAR) * math.sqrt(25 / 9.81)
    ind = pA['label'] == 0
    t = -stl
    t_i = 0
    t_a, h, zz = [], [], []
    for i in range(len(x)):
        if t >= times[t_i]:
            t_a.append(t)
            h.append(np.max(y[i, ind]))
            zz.append(np.max(x[i, ind]))
            t_i += 1
            if t_i == 4:
                break
        t += dta[i]
    return np.array(t_a), np.array(h) / 25, np.array(zz) / 25


def test_example_full_run_vs_oracle_and_notebook(tmp_path):
    """The example's 5 s (34,430 steps): bit-identical to the oracle at the
    end, dt series included; then H and Z as the notebook computes them.

    The notebook's "JJK" column cannot be reproduced by the reference's code
    as it stands: the oracle, which restates that code and which the device
    matches bit for bit, gives H = 0.954, 0.81, 0.657, 0.405 and Z = 1.495,
    2.474, 3.689, 5.571, below JJK by 0.04-0.18 in H and 8-9 % in Z, and
    neither the cubic spline, XSPH off nor alpha = 0 moves Z by more than
    3 % (oracle runs of those variants), so the notebook's run came from
    another state of the code, whose settings it does not record. The check
    against it is therefore a bracket, not parity: at all four times H and Z
    lie between the notebook's experiment column and its JJK column, and
    within 0.2 (H) and 10 % (Z) of JJK."""
    sv, method, r0 = _example(5.0)
    ref = _initial(sv)
    say('example full 5 s run: n', sv.num_particles)
    sv.run()
    sv.close()
    n_after = len(sv.dt_a) - (MAX_SETTLE + 1)
    say('example: steps', len(sv.dt_a), 'after settling', n_after, 't', sv.t)
    path = os.path.join(str(tmp_path), 'dam-break-2d.hdf5')
    sv.save(path)
    stored = path[:-5] + '.npz'  # no h5py here: the same datasets in a .npz
    t_a, H, Z = _notebook_h_z(stored)
    say('example t*', T_STAR, 't', t_a.tolist(), 'H', np.round(H, 3).tolist(), 'JJK', H_JJK,
        'Z', np.round(Z, 3).tolist(), 'JJK', Z_JJK)
    h1, h2, o2 = _oracle_phases(ref, method, r0, n_after)
    _check_bitexact(sv, h1, h2, o2)
    assert len(H) == 4
    assert np.all((np.array(H_EXP) < H) & (H < np.array(H_JJK))), H
    assert np.all((np.array(Z_EXP) < Z) & (Z < np.array(Z_JJK))), Z
    assert np.all(np.abs(H - np.array(H_JJK)) <= 0.2), H
    assert np.all(np.abs(Z / np.array(Z_JJK) - 1) <= 0.10), Z
