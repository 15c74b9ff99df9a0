"""Solver: the time loop of KoningJasper/Offshore-SPH on the device engine.

Mirrors ``src/Solver.py:26-531``: the same constructor, ``addParticles``,
``setup``, ``run``, ``timing``, ``save`` and ``load``, the same settling policy
(damped steps that do not advance time until the kinetic energy drops below
``kE`` x its initial value, ``maxSettle`` steps pass or ``customSettle`` says
so; then TempBoundary particles are deleted, their pressure set to -1e15 and
damping switched off) and the same dt series (``dt_a``, ``dt_c``, ``dt_f``).

What runs where: the loop body (dt, predict, neighbour update, pair pass,
correct, ``src/Solver.py:366-410``) runs in libhipsph as one device step per
call; the host only decides how many steps to launch between the points where
the reference's loop needs the host (a settle check, a store of export
properties, an incremental write-out, the end of ``duration``). Host arrays
are the reference's ``particle_dtype`` records; the device keeps its own
structure-of-arrays state and they are synchronised at those points.

Differences from the reference, each deliberate:
  * State lives on the instance. The reference keeps ``data``, ``export``,
    ``dt_a`` ... as class attributes shared by every Solver (Q13).
  * ``addParticles`` concatenates (the reference passes the second array as
    numpy's ``axis`` argument).
  * ``settle=False`` (keyword only) skips the settling phase and starts the
    time-advancing loop at once, with the constructor's ``damping`` (pass
    0.0 for the reference's post-settling loop): the parity harness drives
    the loop body this way.
  * ``p`` written back after ``run`` is the Tait pressure of the final
    density; the reference leaves the pressure of the last predicted density.
  * Dynamic h (``h=None``), XSPH and summation density run on the device;
    PEC's ``useXSPH`` must equal the method's (the reference's WCSPH
    docstring requires both on), otherwise ``ValueError``.
  * A coupling hook (``coupling``, ``couplingIntegrator=NewmarkBeta(...)``)
    runs on the host between the device's loop and correctors, every step:
    the active records come down, the hook edits the array, and the changed
    x, y, vx, vy, rho, ax, ay, drho, m, h (and xsphx, xsphy with XSPH) go
    back; its p (and xsphx, xsphy without XSPH) stay on the host; a change to
    any other field raises ``NotImplementedError``.
    Inside the hook ``solver.nn`` answers ``nearPos`` from the step's device
    grid with the active records' positions (the reference reads the array it
    is passed at the active indices, which differ from the full array's once
    TempBoundary particles are deleted); ``hipsph.coupling.probe_pressure``
    is the example's pressure probe on the device.
  * ``save``/``load`` use h5py with the reference's dataset names when h5py
    is importable, else a ``.npz`` with the same names.
"""
import os
from time import perf_counter

import numpy as np

from . import coupling as _coupling
from .Common import ParticleType, particle_dtype

# device flags the host loop stops on: cell capacity, non-finite position,
# a key outside the grid (the reference writes out of bounds there), zero
# fluid density (the reference's sums turn NaN there)
_HALT_FLAGS = 1 | 2 | 4 | 32
_TIMING_KEYS = ('total', 'storage', 'integrate_correction', 'integrate_prediction', 'compute', 'time_step',
                'neighbour_hood', 'coupling')


def println(text):
    print(text, flush=True)


def KineticEnergy(J, pA):
    """src/Equations/KineticEnergy.py:5-12 on host records (setup only)."""
    pA = pA[:J]
    return float(np.sum(0.5 * pA['m'] * (pA['vx'] * pA['vx'] + pA['vy'] * pA['vy'])))


def findActive(J, pA):
    """src/Tools/SolverTools.py:46-72: active = not deleted."""
    a_i = np.invert(np.array(pA['deleted'], copy=True, dtype=np.bool_))
    return int(np.sum(a_i)), a_i


class Solver:
    def __init__(self, method, integrator, kernel, duration: float = 1.0, quick: bool = True,
                 incrementalWriteout: bool = True, incrementalFile: str = 'export', incrementalFreq: int = 1000,
                 exportProperties=('x', 'y', 'p'), kE: float = 0.8, maxSettle: int = 500, timeStep: float = None,
                 h: float = None, coupling=None, couplingIntegrator=None, couplingProperties=None,
                 customSettle=None, damping: float = 0.05, *, precision: str = 'fp64', device: str = 'cuda:0',
                 chunk: int = 256, graph: bool = True, settle: bool = True):
        self.method = method
        self.integrator = integrator
        self.kernel = kernel
        self.duration = duration
        self.quick = quick
        self.kEF = kE
        self.maxSettle = maxSettle
        self.timeStep = timeStep
        self.h = h
        self.ts_error = 0
        self.damping = damping
        self.coupling = coupling
        self.couplingIntegrator = couplingIntegrator
        self.couplingProperties = couplingProperties
        self.customSettle = customSettle
        self.incrementalWriteout = incrementalWriteout
        self.incrementalFreq = incrementalFreq
        self.incrementalFile = incrementalFile
        self.exportProperties = list(exportProperties)
        # device options (no counterpart in the reference)
        self.precision = precision
        self.device = device
        self.chunk = max(1, int(chunk))
        self.graph = graph
        self.settle = settle
        # per-instance state (class attributes in the reference)
        self.particleArray = None
        self.num_particles = 0
        self.timing_data = {k: 0.0 for k in _TIMING_KEYS}
        self.data = []
        self.export = {}
        self.settleTime = 0.0
        self.dt_a, self.dt_c, self.dt_f = [], [], []
        self.dt = 0.0
        self.nn = None
        self._engine = None
        self._host_hook = {}

    # -- particles ---------------------------------------------------------
    def load(self, file: str):
        """src/Solver.py:143-158"""
        d = _read_store(file)
        self.particleArray = np.asarray(d['particleArray']).astype(particle_dtype)
        self.dt_a = list(d['dt_a'])
        self.dt_c = list(d['dt_c'])
        self.dt_f = list(d['dt_f'])

    def addParticles(self, particles):
        """src/Solver.py:160-171"""
        particles = np.asarray(particles)
        if particles.dtype != particle_dtype:
            raise TypeError('particles must have particle_dtype')
        if self.particleArray is None:
            self.particleArray = particles
        else:
            self.particleArray = np.concatenate((self.particleArray, particles))

    # -- setup ---------------------------------------------------------------
    def setup(self):
        """src/Solver.py:173-208"""
        println('Starting setup.')
        if (self.integrator.name == 'pec' and
                bool(getattr(self.integrator, 'useXSPH', False)) != bool(getattr(self.method, 'useXSPH', False))):
            # src/Methods/WCSPH.py:56: "useXSPH has to be enabled in both the integrator and the method!"
            raise ValueError('PEC(useXSPH=...) and WCSPH(useXSPH=...) must agree')
        pA = self.particleArray
        self.num_particles, self.indexes = findActive(self.num_particles, pA)
        self._masks()
        println(f'{self.num_particles} total particles, {self.fluid_count} fluid particles.')
        if self.h is None:
            # SolverTools.computeH (src/Tools/SolverTools.py:106-118) from the input density,
            # before initialize (src/Solver.py:185-189); the device repeats it every step
            f = pA[self.f_indexes]
            pA['h'][self.f_indexes] = np.where(f['rho'] > 1e-12, 1.3 * np.sqrt(f['m'] / np.where(
                f['rho'] > 1e-12, f['rho'], 1.0)), 0.0)
        else:
            pA['h'][self.f_indexes] = self.h
        pA[self.f_indexes] = self.method.initialize(pA[self.f_indexes])
        pA['p'][self.f_indexes] = self.method.compute_pressure(pA[self.f_indexes])
        pA['c'][self.f_indexes] = self.method.compute_speed_of_sound(pA[self.f_indexes])
        self.data.append(pA[:])
        self.kE = KineticEnergy(self.fluid_count, pA[self.f_indexes]) * self.kEF
        for key in self.exportProperties:
            self.export[key] = []
        println('Setup complete.')

    def _masks(self):
        pA = self.particleArray
        self.f_indexes = (pA['label'] == ParticleType.Fluid) & (pA['deleted'] == False)  # noqa: E712
        self.c_indexes = (pA['label'] == ParticleType.Coupled) & (pA['deleted'] == False)  # noqa: E712
        self.fluid_count = int(np.sum(self.f_indexes))

    # -- device plumbing ------------------------------------------------------
    def _check_coupling(self):
        ci = self.couplingIntegrator
        if ci is None or not hasattr(ci, 'beta') or not hasattr(ci, 'gamma'):
            raise NotImplementedError('a coupling hook needs couplingIntegrator=NewmarkBeta(...): the '
                                      'device integrates the coupled particles with Newmark-beta')

    def _make_engine(self, damping, advance_time):
        from .device import Engine, EngineParams, soa_from_particles
        if self._engine is not None:
            self._engine.close()
        act = self.particleArray[self.indexes]
        p = EngineParams.from_method(len(act), self.method, kernel=self.kernel.name,
                                     integrator=self.integrator.name, precision=self.precision,
                                     strict=bool(getattr(self.integrator, 'strict', False)))
        p.damping = damping
        p.dynamic_h = self.h is None
        if self.coupling is not None:
            self._check_coupling()
            p.coupled_integrator = 'newmark'
            p.newmark_beta = float(self.couplingIntegrator.beta)
            p.newmark_gamma = float(self.couplingIntegrator.gamma)
        p.fixed_dt = float(self.timeStep) if self.timeStep is not None else 0.0
        p.duration = float(self.duration)
        e = Engine(p, device=self.device)
        e.upload(soa_from_particles(act))
        e.set_clock(self.t, advance_time)
        self._engine = e
        self._hist_seen = 0
        return e

    def _pull(self, e):
        """Device state -> active records of particleArray."""
        d = e.download()
        idx = np.nonzero(self.indexes)[0]
        for k in ('x', 'y', 'vx', 'vy', 'rho', 'ax', 'ay', 'drho', 'p', 'xsphx', 'xsphy', 'h', 'm'):
            self.particleArray[k][idx] = d[k]
        for f, (rows, vals) in self._host_hook.items():  # the last coupling hook's values stand
            self.particleArray[f][rows] = vals

    def _steps(self, e, n):
        """Launch n device steps; returns the steps actually taken."""
        if self.coupling is not None:
            return sum(self._coupled_step(e) for _ in range(n))
        before = e.status()['steps']
        e.step(n, graph=self.graph and n > 1)
        st = e.status()
        bad = st['flags'] & _HALT_FLAGS
        if bad:
            raise RuntimeError(f'device step halted (flags {st["flags"]:#x}: 1 cell capacity, 2 non-finite, '
                               f'4 key outside the grid, 32 zero fluid density)')
        if st['flags'] & 8 and self.ts_error == 0:
            self.ts_error = len(self.dt_a)  # src/Solver.py:220-224
        taken = st['steps'] - before
        self._collect_dt(e, st['steps'])
        self.t = st['t']
        return taken

    _HOOK_WRITABLE = ('x', 'y', 'vx', 'vy', 'rho', 'ax', 'ay', 'drho', 'm', 'h', 'xsphx', 'xsphy')
    _HOOK_HOST_ONLY = ('p', 'c')

    def _coupled_step(self, e):
        """One step split at the coupling hook (src/Solver.py:374-398): the
        device runs dt, the predictors (fluid and Newmark-beta) and the loop;
        the hook sees and edits the particle array; what it changed in the
        active records goes back to the device before the correctors. The
        hook's pressures (and, without XSPH, XSPH velocities) are kept on the
        host: the device derives p from rho and keeps no XSPH velocity then."""
        before = e.status()['steps']
        e.step_begin()
        st = e.status()
        taken = st['steps'] - before
        if st['flags'] & _HALT_FLAGS:
            e.step_end()
            raise RuntimeError(f'device step halted (flags {st["flags"]:#x})')
        if taken == 0:  # past the duration: nothing ran
            e.step_end()
            return 0
        self._collect_dt(e, st['steps'])
        self.dt = st['dt']
        t0 = perf_counter()
        self._host_hook = {}
        self._pull(e)
        idx = np.nonzero(self.indexes)[0]
        snap = self.particleArray.copy()
        self.nn = _coupling.EngineNN(e, self)
        out = self.coupling(self.particleArray, self)
        if out is not None:
            self.particleArray = out
        pA = self.particleArray
        host_only = self._HOOK_HOST_ONLY
        writable = self._HOOK_WRITABLE
        if not e.params.use_xsph:  # the device keeps no XSPH velocity then
            host_only = host_only + ('xsphx', 'xsphy')
            writable = tuple(f for f in writable if f not in ('xsphx', 'xsphy'))
        for f in pA.dtype.names:
            if f in writable or f in host_only:
                continue
            if not _same(pA[f], snap[f]).all():
                raise NotImplementedError(f'the coupling hook changed {f!r}; the device takes back only '
                                          + ', '.join(writable))
        rows = np.zeros(len(idx), dtype=bool)
        for f in writable:
            rows |= ~_same(pA[f][idx], snap[f][idx])
        if rows.any():
            sel = np.nonzero(rows)[0]
            e.scatter(sel, {f: pA[f][idx[sel]] for f in writable})
        for f in host_only:
            ch = ~_same(pA[f][idx], snap[f][idx])
            if ch.any():
                self._host_hook[f] = (idx[ch], pA[f][idx[ch]].copy())
        self.timing_data['coupling'] += perf_counter() - t0
        e.step_end()
        self.t = e.status()['t']
        return taken

    def _collect_dt(self, e, total):
        new = total - self._hist_seen
        if new <= 0:
            return
        from .device import _lib
        cap = _lib.SPH_DT_HISTORY
        if new > cap:
            raise RuntimeError(f'more than {cap} steps between two host points; lower chunk')
        h = e.dt_history(new)
        self.dt_a.extend(h[:, 0].tolist())
        if self.timeStep is None:  # the fixed-step branch records dt_a only (src/Solver.py:370-372)
            self.dt_c.extend(h[:, 1].tolist())
            self.dt_f.extend(h[:, 2].tolist())
        self._hist_seen = total

    def _store(self, t_step):
        """src/Solver.py:477-487"""
        self.data.append(np.copy(self.particleArray))
        for key in self.exportProperties:
            self.export[key].append(np.copy(self.particleArray[key]))
        if self.incrementalWriteout and t_step % self.incrementalFreq == 0:
            self.save('{0}-{1}.hdf5'.format(self.incrementalFile, t_step), printLocation=False)
        if len(self.data) > self.incrementalFreq:
            self.data.pop(0)

    def _needs_host_every_step(self):
        return bool(self.exportProperties)

    def _next_stop(self, t_step):
        """Steps until the next point where the host is needed."""
        if self._needs_host_every_step():
            return 1
        n = self.chunk
        if self.incrementalWriteout:
            k = t_step % self.incrementalFreq
            n = min(n, 1 if k == 0 else self.incrementalFreq - k + 1)
        return max(1, n)

    # -- the loop ----------------------------------------------------------------
    def run(self):
        """src/Solver.py:346-475"""
        start_all = perf_counter()
        if self.particleArray is None or len(self.particleArray) == 0 or \
                int(np.sum(~self.particleArray['deleted'])) != self.num_particles:
            raise Exception('No or invalid particles set!')
        if self.coupling is not None:
            self._check_coupling()
        println('Started solving...')
        t_step = 0
        self.t = 0.0
        self.settled = not self.settle

        if not self.settled:
            println('Settling particles...')
            e = self._make_engine(self.damping, advance_time=False)
            # With no custom criterion and a threshold KE can never go below
            # (kE <= 0, e.g. a fluid at rest: KE0 = 0), settling lasts exactly
            # maxSettle + 1 steps and needs no host check in between.
            bulk = self.customSettle is None and self.kE <= 0.0
            while True:
                t0 = perf_counter()
                n = self._next_stop(t_step)
                if bulk:
                    n = min(n, self.maxSettle + 1 - t_step)
                else:
                    n = 1
                taken = self._steps(e, n)
                self.timing_data['compute'] += perf_counter() - t0
                if taken == 0:
                    break
                t_step += taken
                t0 = perf_counter()
                if not bulk or self._needs_host_every_step() or self.incrementalWriteout:
                    self._pull(e)
                    self._store(t_step - 1)
                self.timing_data['storage'] += perf_counter() - t0
                if t_step > 1:
                    ke, cs = 1e12, False
                    if self.customSettle is None:
                        ke = e.kinetic_energy()
                    else:
                        self._pull(e)
                        cs = self.customSettle(self.particleArray, self)
                    if ke < self.kE or t_step > self.maxSettle or cs:
                        if t_step > self.maxSettle:
                            println('WARNING! Maximum settle steps reached, check configuration, maybe increase '
                                    'spacing between wall and particles.')
                        break
            self._pull(e)
            pA = self.particleArray
            inds = np.nonzero((pA['label'] == ParticleType.TempBoundary) & (pA['deleted'] == False))[0]  # noqa
            pA['deleted'][inds] = True
            self.num_particles, self.indexes = findActive(self.num_particles, pA)
            self._masks()
            pA['p'][inds] = -1e15
            self.settleTime = sum(self.dt_a)
            self.damping = 0.0
            self.settled = True
            println('Settling Complete.')

        e = self._make_engine(0.0 if self.settle else self.damping, advance_time=True)
        while self.t < self.duration:
            t0 = perf_counter()
            taken = self._steps(e, self._next_stop(t_step))
            self.timing_data['compute'] += perf_counter() - t0
            if taken == 0:
                break
            t_step += taken
            t0 = perf_counter()
            if self._needs_host_every_step() or self.incrementalWriteout:
                self._pull(e)
                self._store(t_step - 1)
            self.timing_data['storage'] += perf_counter() - t0
        self._pull(e)
        self.timing_data['total'] = perf_counter() - start_all
        self.t_step = t_step
        println('Solved!')
        println(f'Solved {self.num_particles} particles for {self.duration:f} [s].')
        println(f'Completed solve in {self.timing_data["total"]:f} [s] and {t_step} steps')

    def timing(self):
        """src/Solver.py:489-495"""
        total = self.timing_data['total'] or 1.0
        println('Detailed timing statistics:')
        println(f'{"Name":24s} {"Time [s]":>10s} {"Percentage [%]":>15s}')
        for k, v in self.timing_data.items():
            println(f'{k:24s} {round(v, 3):10.3f} {round(v / total * 100, 2):15.2f}')

    def save(self, location: str, printLocation: bool = True, extraProperties: dict = None):
        """src/Solver.py:497-531 (h5py when importable, else .npz with the same names)."""
        d = {'particleArray': self.particleArray, 'dt_a': np.asarray(self.dt_a), 'dt_c': np.asarray(self.dt_c),
             'dt_f': np.asarray(self.dt_f), 'settleTime': np.asarray(self.settleTime)}
        for key in self.exportProperties:
            if self.export.get(key):
                d[key] = np.stack(self.export[key])
        if extraProperties:
            d.update({k: np.asarray(v) for k, v in extraProperties.items()})
        path = _write_store(location, d)
        if printLocation:
            println(f'Exported arrays to: "{path}".')

    def close(self):
        if self._engine is not None:
            self._engine.close()
            self._engine = None


def _same(a, b):
    """Element-wise bit-level equality that treats NaN as equal to NaN."""
    return (a == b) | ((a != a) & (b != b))


def _write_store(location, d):
    try:
        import h5py
    except ImportError:
        h5py = None
    if h5py is not None:
        with h5py.File(location, 'w') as h5f:
            for k, v in d.items():
                if np.ndim(v) == 0:
                    h5f.create_dataset(k, data=v)
                else:
                    h5f.create_dataset(k, data=v, shuffle=True, compression='gzip')
        return location
    root, _ = os.path.splitext(location)
    path = root + '.npz'
    np.savez_compressed(path, **d)
    return path


def _read_store(file):
    try:
        import h5py
    except ImportError:
        h5py = None
    if h5py is not None and os.path.exists(file) and not file.endswith('.npz'):
        with h5py.File(file, 'r') as h5f:
            return {k: h5f[k][()] for k in h5f.keys()}
    root, _ = os.path.splitext(file)
    path = file if file.endswith('.npz') else root + '.npz'
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


__all__ = ['Solver', 'KineticEnergy', 'findActive']
