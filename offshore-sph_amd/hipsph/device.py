"""Device engine: one libhipsph context on one GPU.

PyTorch-ROCm supplies the device buffers the state is uploaded from and
downloaded to, and the stream every call is ordered on
(``torch.cuda.current_stream()``). The WCSPH step itself runs in the HIP
kernels of libhipsph; nothing here computes physics.
"""
import ctypes as C
from dataclasses import dataclass, asdict
import math

import numpy as np

from . import _lib
from ._lib import call, sph_params_t, sph_state_t

IO_F64 = ('x', 'y', 'vx', 'vy', 'rho', 'ax', 'ay', 'drho', 'm', 'h', 'p', 'xsphx', 'xsphy')

KERNELS = {'cubic': _lib.SPH_KERNEL_CUBIC, 'gaussian': _lib.SPH_KERNEL_GAUSSIAN, 'wendland': _lib.SPH_KERNEL_WENDLAND}
PRECISIONS = {'fp64': _lib.SPH_PREC_FP64, 'fp32c': _lib.SPH_PREC_FP32C}
INTEGRATORS = {'pec': _lib.SPH_INTEGRATOR_PEC, 'euler': _lib.SPH_INTEGRATOR_EULER}
COUPLED = {'none': _lib.SPH_COUPLED_NONE, 'newmark': _lib.SPH_COUPLED_NEWMARK}
GATHER_FIELDS = ('x', 'y', 'vx', 'vy', 'rho', 'ax', 'ay', 'drho', 'm', 'h', 'p', 'label', 'xsphx', 'xsphy')
SCATTER_FIELDS = ('x', 'y', 'vx', 'vy', 'rho', 'ax', 'ay', 'drho', 'm', 'h', 'xsphx', 'xsphy')


def host_array(a, dtype=None):
    """A fresh C-ordered copy with the element's own stride. A field of a
    particle_dtype record array is strided by the 154-byte record, and numpy
    calls a one-element such view contiguous, which torch rejects."""
    return np.array(a, dtype=dtype, order='C', copy=True)


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise _lib.HipSphUnavailable('no GPU visible to torch: the WCSPH engine runs on the device only')
    return torch


@dataclass
class EngineParams:
    """Constants of one run (mirrors sph_params_t; defaults are the reference's)."""
    n: int
    kernel: str = 'cubic'
    precision: str = 'fp64'
    integrator: str = 'pec'
    strict: bool = False
    rho0: float = 1000.0
    gamma: float = 7.0
    co: float = 0.0
    B: float = 0.0
    Pb: float = 0.0
    alpha: float = 0.01
    beta: float = 0.0
    r0: float = 0.0
    D: float = 0.0
    p1: float = 4.0
    p2: float = 2.0
    gravity: float = 9.81
    scale: float = 2.0
    qmax: float = 3.0
    gamma_c: float = 0.25
    gamma_f: float = 0.25
    damping: float = 0.0
    fixed_dt: float = 0.0
    duration: float = math.inf
    cell_capacity: int = 0
    use_xsph: bool = False
    use_summation_density: bool = False
    epsilon: float = 0.5
    dynamic_h: bool = False
    coupled_integrator: str = 'none'
    newmark_beta: float = 0.25
    newmark_gamma: float = 0.5

    def to_c(self):
        p = sph_params_t()
        p.n = int(self.n)
        p.kernel = KERNELS[self.kernel] if isinstance(self.kernel, str) else int(self.kernel)
        p.precision = PRECISIONS[self.precision] if isinstance(self.precision, str) else int(self.precision)
        p.integrator = INTEGRATORS[self.integrator] if isinstance(self.integrator, str) else int(self.integrator)
        p.strict = int(bool(self.strict))
        p.use_xsph = int(bool(self.use_xsph))
        p.use_summation_density = int(bool(self.use_summation_density))
        p.dynamic_h = int(bool(self.dynamic_h))
        ci = self.coupled_integrator
        p.coupled_integrator = COUPLED[ci] if isinstance(ci, str) else int(ci)
        p.newmark_beta = float(self.newmark_beta)
        p.newmark_gamma = float(self.newmark_gamma)
        for f in ('rho0', 'gamma', 'co', 'B', 'Pb', 'alpha', 'beta', 'r0', 'D', 'p1', 'p2', 'gravity', 'scale',
                  'qmax', 'gamma_c', 'gamma_f', 'damping', 'fixed_dt', 'duration', 'epsilon'):
            setattr(p, f, float(getattr(self, f)))
        p.cell_capacity = int(self.cell_capacity)
        return p

    @classmethod
    def from_method(cls, n, method, kernel='cubic', integrator='pec', precision='fp64', **kw):
        """Take the WCSPH constants from a method object (src/Methods/WCSPH.py:35-79),
        XSPH and summation density included (with PEC the integrator's useXSPH
        must match the method's: the reference requires both on, WCSPH.py:56)."""
        kw.setdefault('use_xsph', bool(getattr(method, 'useXSPH', False)))
        kw.setdefault('use_summation_density', bool(getattr(method, 'useSummationDensity', False)))
        kw.setdefault('epsilon', float(getattr(method, 'epsilon', 0.5)))
        return cls(n=n, kernel=kernel, integrator=integrator, precision=precision, rho0=method.rho0,
                   gamma=method.gamma, co=method.co, B=method.B, Pb=method.Pb, alpha=method.alpha,
                   beta=method.beta, r0=method.r0, D=method.D, p1=method.p1, p2=method.p2, **kw)


class Engine:
    """A libhipsph context plus the torch buffers its state moves through."""

    def __init__(self, params: EngineParams, device='cuda:0'):
        torch = _torch()
        _lib.load()
        self.params = params
        self.device = torch.device(device)
        self.n = int(params.n)
        self._ctx = C.c_void_p()
        with torch.cuda.device(self.device):
            call('sph_ctx_create', C.byref(params.to_c()), C.byref(self._ctx))
            self.io = {k: torch.zeros(self.n, dtype=torch.float64, device=self.device) for k in IO_F64}
            self.io['label'] = torch.zeros(self.n, dtype=torch.int8, device=self.device)

    # -- plumbing --------------------------------------------------------
    def _stream(self):
        import torch
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _state(self, fields=None):
        st = sph_state_t()
        for f, _ in sph_state_t._fields_:
            if fields is None or f in fields:
                setattr(st, f, self.io[f].data_ptr())
        return st

    def close(self):
        if self._ctx:
            _lib.load().sph_ctx_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, **changes):
        for k, v in changes.items():
            setattr(self.params, k, v)
        call('sph_ctx_set_params', self._ctx, C.byref(self.params.to_c()))

    # -- state -----------------------------------------------------------
    def upload(self, arrays, reset_clock=True):
        """Copy x, y, vx, vy, rho, m, h, label (and ax, ay, drho, xsphx, xsphy if given) in."""
        import torch
        need = ('x', 'y', 'vx', 'vy', 'rho', 'm', 'h', 'label')
        for k in need:
            if k not in arrays:
                raise KeyError(f'upload needs {k}')
        fields = set(need)
        for k, v in arrays.items():
            if k not in self.io:
                continue
            t = torch.as_tensor(host_array(v) if isinstance(v, np.ndarray) else v)
            if t.numel() != self.n:
                raise ValueError(f'{k} has {t.numel()} entries, engine has n={self.n}')
            self.io[k].copy_(t.to(self.io[k].dtype).reshape(-1))
            fields.add(k)
        # The pair records carry the label in the sign of the mass slot.
        bad = (self.io['label'] == 0) & ~(self.io['m'] >= 0)
        if bool(bad.any()):
            raise ValueError('fluid particles need a finite, non-negative mass')
        call('sph_state_upload', self._ctx, C.byref(self._state(fields)), int(bool(reset_clock)), self._stream())

    def restore(self, arrays, t):
        """A saved state (device tensors or arrays, as from download) back in,
        clock at t, without a host round trip."""
        fields = set()
        for k, v in arrays.items():
            if k in self.io and k != 'p':
                self.io[k].copy_(v.reshape(-1) if hasattr(v, 'reshape') else v)
                fields.add(k)
        call('sph_state_restore', self._ctx, C.byref(self._state(fields)), float(t), self._stream())

    def download(self, to_numpy=True):
        """State in reference index order (plus p = Tait + Pb)."""
        import torch
        call('sph_state_download', self._ctx, C.byref(self._state()), self._stream())
        if not to_numpy:
            return {k: v.clone() for k, v in self.io.items()}
        torch.cuda.current_stream(self.device).synchronize()
        return {k: v.cpu().numpy() for k, v in self.io.items()}

    # -- the step ----------------------------------------------------------
    def dt(self):
        call('sph_dt', self._ctx, self._stream())

    def predict(self):
        call('sph_predict', self._ctx, self._stream())

    def nn_update(self):
        call('sph_nn_update', self._ctx, self._stream())

    def compute(self):
        call('sph_compute', self._ctx, self._stream())

    def continuity(self):
        """Continuity only (drho); ax, ay untouched."""
        call('sph_continuity', self._ctx, self._stream())

    def momentum(self):
        """Momentum + AV + gravity + wall force only (ax, ay); drho untouched."""
        call('sph_momentum', self._ctx, self._stream())

    def compute_counts(self):
        """One pair pass; returns the neighbour count it found per particle (-1 for walls)."""
        import torch
        out = torch.empty(self.n, dtype=torch.int32, device=self.device)
        call('sph_compute_counts', self._ctx, out.data_ptr(), self._stream())
        return out.cpu().numpy()

    def correct(self):
        call('sph_correct', self._ctx, self._stream())

    def step(self, nsteps=1, graph=False):
        call('sph_step', self._ctx, int(nsteps), int(bool(graph)), self._stream())

    def step_begin(self):
        """First half of a step, up to the coupling hook (sph_step_begin)."""
        call('sph_step_begin', self._ctx, self._stream())

    def step_end(self):
        """Second half: the correctors and the clock (sph_step_end)."""
        call('sph_step_end', self._ctx, self._stream())

    # -- records by reference index (the coupling hook's view) ---------------
    @staticmethod
    def _ids(ids):
        return np.ascontiguousarray(np.asarray(ids, dtype=np.int64).reshape(-1))

    def gather(self, ids, fields=('x', 'y', 'vx', 'vy', 'rho', 'ax', 'ay')):
        """{field: numpy array} of the records with reference indices ids."""
        import torch
        ids = self._ids(ids)
        k = len(ids)
        out = {f: torch.empty(k, dtype=torch.int8 if f == 'label' else torch.float64, device=self.device)
               for f in fields}
        st = sph_state_t()
        for f in fields:
            if f not in GATHER_FIELDS:
                raise KeyError(f)
            setattr(st, f, out[f].data_ptr() if k else None)
        call('sph_state_gather', self._ctx, ids.ctypes.data_as(C.c_void_p), k, C.byref(st), self._stream())
        return {f: v.cpu().numpy() for f, v in out.items()}

    def scatter(self, ids, values):
        """Write {field: array in the order of ids} into the records ids."""
        import torch
        ids = self._ids(ids)
        k = len(ids)
        st = sph_state_t()
        keep = []
        for f, v in values.items():
            if f not in SCATTER_FIELDS:
                raise KeyError(f'{f} cannot be written (writable: {", ".join(SCATTER_FIELDS)})')
            t = torch.as_tensor(host_array(v, np.float64) if not isinstance(v, torch.Tensor) else v,
                                dtype=torch.float64).to(self.device).reshape(-1).contiguous()
            if t.numel() != k:
                raise ValueError(f'{f} has {t.numel()} entries for {k} ids')
            keep.append(t)
            setattr(st, f, t.data_ptr() if k else None)
        call('sph_state_scatter', self._ctx, ids.ctypes.data_as(C.c_void_p), k, C.byref(st), self._stream())

    def near_pos(self, x, y, h):
        """NNLinkedList.nearPos for points against the last update: CSR (off, idx, q, r), numpy."""
        import torch
        dev = self.device
        pts = [torch.as_tensor(np.atleast_1d(host_array(a, np.float64))).to(dev).contiguous() for a in (x, y)]
        if pts[0].shape != pts[1].shape or pts[0].dim() != 1:
            raise ValueError('x and y must be 1-D and of one length')
        hh = torch.as_tensor(np.broadcast_to(np.asarray(h, dtype=np.float64), pts[0].shape).copy()).to(dev)
        npts = pts[0].numel()
        off = torch.zeros(npts + 1, dtype=torch.int64, device=dev)
        total = C.c_int64(0)
        args = (pts[0].data_ptr(), pts[1].data_ptr(), hh.data_ptr(), off.data_ptr())
        rc = _lib.load().sph_near_pos(self._ctx, npts, *args, None, None, None, 0, C.byref(total), self._stream())
        if rc not in (0, -4):
            _lib.check(rc)
        cap = max(int(total.value), 1)
        idx = torch.zeros(cap, dtype=torch.int32, device=dev)
        q = torch.zeros(cap, dtype=torch.float64, device=dev)
        r = torch.zeros(cap, dtype=torch.float64, device=dev)
        call('sph_near_pos', self._ctx, npts, *args, idx.data_ptr(), q.data_ptr(), r.data_ptr(), cap,
             C.byref(total), self._stream())
        t = int(total.value)
        return off.cpu().numpy(), idx[:t].cpu().numpy(), q[:t].cpu().numpy(), r[:t].cpu().numpy()

    def probe_pressure(self, x, y, h):
        """pressure_SPH of the coupled example at points (x, y) with length h: (rho, p), numpy."""
        import torch
        dev = self.device
        px, py = [torch.as_tensor(np.atleast_1d(host_array(a, np.float64))).to(dev).contiguous()
                  for a in (x, y)]
        if px.shape != py.shape or px.dim() != 1:
            raise ValueError('x and y must be 1-D and of one length')
        ph = torch.as_tensor(np.broadcast_to(np.asarray(h, dtype=np.float64), px.shape).copy()).to(dev)
        n = px.numel()
        rho = torch.empty(n, dtype=torch.float64, device=dev)
        p = torch.empty(n, dtype=torch.float64, device=dev)
        call('sph_probe_pressure', self._ctx, n, px.data_ptr(), py.data_ptr(), ph.data_ptr(), rho.data_ptr(),
             p.data_ptr(), self._stream())
        return rho.cpu().numpy(), p.cpu().numpy()

    def status(self):
        d = (C.c_double * 4)()
        i = (C.c_int64 * 2)()
        call('sph_status', self._ctx, d, i, self._stream())
        return {'t': d[0], 'dt': d[1], 'dt_c': d[2], 'dt_f': d[3], 'steps': int(i[0]), 'flags': int(i[1])}

    def set_clock(self, t, advance_time=True):
        """Device time and whether steps advance it (False while settling)."""
        call('sph_set_clock', self._ctx, float(t), int(bool(advance_time)), self._stream())

    def kinetic_energy(self):
        """KineticEnergy (src/Equations/KineticEnergy.py:5-12) of the fluid, on the device."""
        out = C.c_double()
        call('sph_kinetic_energy', self._ctx, C.byref(out), self._stream())
        return out.value

    def dt_history(self, count):
        out = np.zeros((int(count), 3), dtype=np.float64)
        call('sph_dt_history', self._ctx, out.ctypes.data_as(C.POINTER(C.c_double)), int(count), self._stream())
        return out

    # -- neighbour search --------------------------------------------------
    def grid(self):
        d = (C.c_double * 5)()
        i = (C.c_int64 * 3)()
        call('sph_nn_grid', self._ctx, d, i, self._stream())
        return {'xmin': d[0], 'xmax': d[1], 'ymin': d[2], 'ymax': d[3], 'cell_size': d[4],
                'ncx': int(i[0]), 'ncy': int(i[1]), 'n_cells': int(i[2])}

    def nn_dump(self, with_qr=True):
        """CSR of NNLinkedList.near(i) for every i: keys, off, idx, q, r (numpy)."""
        import torch
        dev = self.device
        keys = torch.zeros(self.n, dtype=torch.int64, device=dev)
        off = torch.zeros(self.n + 1, dtype=torch.int64, device=dev)
        total = C.c_int64(0)
        rc = _lib.load().sph_nn_dump(self._ctx, keys.data_ptr(), off.data_ptr(), None, None, None, 0,
                                     C.byref(total), self._stream())
        if rc not in (0, -4):
            _lib.check(rc)
        cap = max(int(total.value), 1)
        idx = torch.zeros(cap, dtype=torch.int32, device=dev)
        q = torch.zeros(cap, dtype=torch.float64, device=dev) if with_qr else None
        r = torch.zeros(cap, dtype=torch.float64, device=dev) if with_qr else None
        call('sph_nn_dump', self._ctx, keys.data_ptr(), off.data_ptr(), idx.data_ptr(),
             q.data_ptr() if with_qr else None, r.data_ptr() if with_qr else None, cap, C.byref(total),
             self._stream())
        out = {'keys': keys.cpu().numpy(), 'off': off.cpu().numpy(), 'idx': idx[:total.value].cpu().numpy()}
        if with_qr:
            out['q'] = q[:total.value].cpu().numpy()
            out['r'] = r[:total.value].cpu().numpy()
        return out

    def set_pair_staging(self, on):
        """Whether the uniform float pair passes stage each workgroup's
        neighbour records in LDS (the default) or walk global memory; the
        results are bit-identical either way."""
        call('sph_ctx_set_pair_staging', self._ctx, int(bool(on)))

    def set_border_bounds(self, on):
        """Whether the fused PEC step's neighbour update takes the grid's
        bounds from the border cells when the previous corrector's
        displacement bound allows (the default) or scans every particle; the
        grid is the same either way."""
        call('sph_ctx_set_border_bounds', self._ctx, int(bool(on)))

    def border_count(self):
        """Neighbour updates that took their bounds from the border cells."""
        v = C.c_int64()
        call('sph_nn_border_count', self._ctx, C.byref(v), self._stream())
        return int(v.value)

    def pair_form(self):
        """Which form of the fused pair pass the context runs: 'fp64', 'fp32c'
        (general) or 'fp64-uniform', 'fp32c-uniform' (one h for every
        particle and one fluid mass)."""
        f = C.c_int32()
        call('sph_pair_form', self._ctx, C.byref(f))
        return {0: 'fp64', 1: 'fp32c', 2: 'fp32c-uniform', 3: 'fp64-uniform'}[f.value]

    # -- timing --------------------------------------------------------------
    def timing_enable(self, on=True):
        call('sph_timing_enable', self._ctx, int(bool(on)))

    def timing_read(self, reset=True):
        p = C.c_double()
        q = C.c_double()
        k = C.c_int64()
        call('sph_timing_read', self._ctx, C.byref(p), C.byref(q), C.byref(k), int(bool(reset)))
        return {'pair_ms': p.value, 'nn_ms': q.value, 'count': int(k.value)}


def soa_from_particles(pA, fields=('x', 'y', 'vx', 'vy', 'rho', 'ax', 'ay', 'drho', 'm', 'h', 'label', 'xsphx',
                                   'xsphy')):
    """Structure-of-arrays view of a particle_dtype array (active records only)."""
    return {f: host_array(pA[f]) for f in fields}
