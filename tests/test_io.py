"""The Solver mirror's file output on the host (no GPU): save / load and the
incremental write-out (src/Solver.py:143-158,477-487,497-531).

h5py is not importable here, so the mirror writes the reference's datasets
into a .npz of the same base name; these tests pin the dataset names, the
round trip and the write-out cadence on that path.
"""
import os

import numpy as np
import pytest

from hipsph import ParticleType, scenarios
from hipsph.Common import particle_dtype
from hipsph.Integrators.PEC import PEC
from hipsph.Kernels.CubicSpline import CubicSpline
from hipsph.Solver import Solver


def _solver(tmp, **kw):
    sc = scenarios.dam_break(0.2, 0.2, 0.02, tank_w=0.4, tank_h=0.4, name='io')
    sv = Solver(sc.method(), PEC(useXSPH=False, strict=False), CubicSpline(), duration=1.0, h=sc.h,
                incrementalFile=os.path.join(tmp, 'export'), **kw)
    sv.addParticles(sc.pA.copy())
    # what setup() leaves for the file output (setup itself evaluates the
    # pressure on the device)
    for key in sv.exportProperties:
        sv.export[key] = []
    return sv


def _advance(sv, t_step, rng):
    """What a step leaves on the host: a changed state and one more dt."""
    f = sv.particleArray['label'] == ParticleType.Fluid
    sv.particleArray['x'][f] += rng.normal(0, 1e-3, f.sum())
    sv.particleArray['p'][f] = rng.normal(0, 1e3, f.sum())
    sv.dt_a.append(1e-4 * (1 + t_step))
    sv.dt_c.append(2e-4)
    sv.dt_f.append(1e-4 * (1 + t_step))
    sv._store(t_step)


def test_save_load_round_trip(tmp_path):
    sv = _solver(str(tmp_path), incrementalWriteout=False, exportProperties=['x', 'y', 'p'])
    rng = np.random.default_rng(1)
    for k in range(5):
        _advance(sv, k, rng)
    sv.settleTime = 0.0125
    path = os.path.join(str(tmp_path), 'run.hdf5')
    sv.save(path, extraProperties={'probe': np.arange(3.0)})
    stored = os.path.join(str(tmp_path), 'run.npz')
    assert os.path.exists(stored)
    with np.load(stored, allow_pickle=False) as z:
        names = set(z.files)
        assert names == {'particleArray', 'dt_a', 'dt_c', 'dt_f', 'settleTime', 'x', 'y', 'p', 'probe'}
        assert z['particleArray'].dtype == particle_dtype
        assert z['x'].shape == (5, len(sv.particleArray))  # np.stack of the export list
        np.testing.assert_array_equal(z['x'][-1], sv.particleArray['x'])
        np.testing.assert_array_equal(z['p'][2], sv.export['p'][2])
        assert float(z['settleTime']) == 0.0125
        np.testing.assert_array_equal(z['probe'], np.arange(3.0))
    other = _solver(str(tmp_path), incrementalWriteout=False)
    other.load(path)
    np.testing.assert_array_equal(other.particleArray, sv.particleArray)
    assert other.particleArray.dtype == particle_dtype
    assert other.dt_a == sv.dt_a and other.dt_c == sv.dt_c and other.dt_f == sv.dt_f


def test_incremental_writeout_cadence(tmp_path):
    """A file export-<t_step> at every t_step % incrementalFreq == 0, holding
    the state of that step; the in-memory history keeps incrementalFreq
    entries."""
    freq = 3
    sv = _solver(str(tmp_path), incrementalWriteout=True, incrementalFreq=freq, exportProperties=['x'])
    rng = np.random.default_rng(2)
    states = {}
    for k in range(8):
        _advance(sv, k, rng)
        states[k] = sv.particleArray['x'].copy()
    files = sorted(f for f in os.listdir(str(tmp_path)) if f.startswith('export-'))
    assert files == ['export-0.npz', 'export-3.npz', 'export-6.npz']
    for k in (0, 3, 6):
        with np.load(os.path.join(str(tmp_path), f'export-{k}.npz'), allow_pickle=False) as z:
            np.testing.assert_array_equal(z['particleArray']['x'], states[k])
            assert len(z['dt_a']) == k + 1
            assert z['x'].shape[0] == k + 1
    assert len(sv.data) == freq


def test_next_stop_honours_the_writeout_cadence(tmp_path):
    """The device runs in chunks between host points; with write-out on, a
    chunk never crosses a multiple of incrementalFreq."""
    sv = _solver(str(tmp_path), incrementalWriteout=True, incrementalFreq=10, exportProperties=[], chunk=64)
    assert sv._next_stop(0) == 1  # t_step 0 writes at once
    assert sv._next_stop(1) == 10  # steps 1 .. 10: the next write at t_step 10 ...
    assert sv._next_stop(7) == 4
    sv2 = _solver(str(tmp_path), incrementalWriteout=False, exportProperties=[], chunk=64)
    assert sv2._next_stop(5) == 64
    sv3 = _solver(str(tmp_path), incrementalWriteout=False, exportProperties=['x'], chunk=64)
    assert sv3._next_stop(5) == 1
    with pytest.raises(TypeError):
        sv3.addParticles(np.zeros(3))
