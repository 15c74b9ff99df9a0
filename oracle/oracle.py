"""Python access to the CPU oracle (librefcpu.so, built from refcpu.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline, never as part of the product.

The C structs mirror the reference's numpy dtypes byte for byte, so
particle arrays are passed without conversion.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'librefcpu.so')

PARTICLE_FIELDS = ['deleted', 'label', 'm', 'rho', 'p', 'c', 'drho', 'h', 'x', 'y', 'vx', 'vy',
                   'ax', 'ay', 'xsphx', 'xsphy', 'x0', 'y0', 'vx0', 'vy0', 'rho0']
particle_dtype = np.dtype({'names': PARTICLE_FIELDS, 'formats': [np.bool_, np.int8] + [np.double] * 19})
COMPUTED_FIELDS = ['label', 'm', 'p', 'rho', 'h', 'q', 'c', 'r', 'w', 'dw_x', 'dw_y', 'x', 'y', 'vx', 'vy']
computed_dtype = np.dtype({'names': COMPUTED_FIELDS, 'formats': [np.int8] + [np.double] * 14})

KERNELS = {'cubic': 0, 'gaussian': 1, 'wendland': 2}
INTEGRATORS = {'pec': 0, 'euler': 1}


class ref_wcsph_t(C.Structure):
    _fields_ = [('useXSPH', C.c_int32), ('useSummationDensity', C.c_int32)] + [
        (f, C.c_double) for f in ('height', 'rho0', 'r0', 'D', 'p1', 'p2', 'epsilon', 'gamma', 'co', 'B', 'Pb',
                                  'alpha', 'beta')]


class ref_nn_t(C.Structure):
    _fields_ = [('scale', C.c_double), ('xmin', C.c_double), ('xmax', C.c_double), ('ymin', C.c_double),
                ('ymax', C.c_double), ('cell_size', C.c_double), ('ncx', C.c_int64), ('ncy', C.c_int64),
                ('n_cells', C.c_int64), ('heads', C.c_void_p), ('nexts', C.c_void_p), ('cap_cells', C.c_int64),
                ('cap_n', C.c_int64), ('oob_binned', C.c_int64)]


class ref_solver_t(C.Structure):
    _fields_ = [('kernel', C.c_int32), ('integrator', C.c_int32), ('pec_useXSPH', C.c_int32),
                ('pec_strict', C.c_int32), ('h_fixed', C.c_double), ('damping', C.c_double),
                ('gamma_c', C.c_double), ('gamma_f', C.c_double), ('t', C.c_double), ('t_step', C.c_int64),
                ('nthreads', C.c_int32), ('coupled_integrator', C.c_int32), ('nm_beta', C.c_double),
                ('nm_gamma', C.c_double), ('dt', C.c_double)]


_lib = None


def build():
    subprocess.run(['make', '-s', '-C', HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    d, i64, i32, p = C.c_double, C.c_int64, C.c_int32, C.c_void_p
    sig = {
        'ref_cubic_evaluate': (d, [d, d]), 'ref_cubic_gradient': (d, [d, d, d]),
        'ref_gaussian_evaluate': (d, [d, d]), 'ref_gaussian_derivative': (d, [d, d]),
        'ref_gaussian_gradient': (d, [d, d, d]),
        'ref_kernel_evaluate_n': (None, [i32, i64, p, p, p]), 'ref_kernel_gradient_n': (None, [i32, i64, p, p, p, p]),
        'ref_tait_eos': (d, [d, d, d, d, C.c_int8]), 'ref_tait_co': (d, [d]), 'ref_tait_B': (d, [d, d, d]),
        'ref_tait_height': (d, [d, d, d, d, d]),
        'ref_continuity': (d, [p, i64]), 'ref_momentum': (None, [d, d, p, p, i64, p]),
        'ref_boundary_force': (None, [d, d, d, d, p, i64, p]), 'ref_timestep': (None, [i64, p, d, d, p]),
        'ref_wcsph_init': (None, [C.POINTER(ref_wcsph_t), d, d, d, i32, d, i32]),
        'ref_pec_predict': (None, [i32, i32, d, p, i64, d]), 'ref_pec_correct': (None, [i32, i32, d, p, i64, d]),
        'ref_euler_correct': (None, [d, p, i64, d]),
        'ref_nn_init': (None, [C.POINTER(ref_nn_t), d]), 'ref_nn_free': (None, [C.POINTER(ref_nn_t)]),
        'ref_nn_update': (C.c_int, [C.POINTER(ref_nn_t), p, i64]),
        'ref_nn_dump': (i64, [C.POINTER(ref_nn_t), p, i64, p, p, p, p, p, i64]),
        'ref_loop': (C.c_int, [p, i64, i32, C.POINTER(ref_wcsph_t), C.POINTER(ref_nn_t), i32]),
        'ref_solver_step': (C.c_int, [p, i64, C.POINTER(ref_solver_t), C.POINTER(ref_wcsph_t),
                                      C.POINTER(ref_nn_t), p]),
        'ref_solver_run': (C.c_int, [p, i64, C.POINTER(ref_solver_t), C.POINTER(ref_wcsph_t),
                                     C.POINTER(ref_nn_t), i64, p]),
        'ref_solver_begin': (C.c_int, [p, i64, C.POINTER(ref_solver_t), C.POINTER(ref_wcsph_t),
                                       C.POINTER(ref_nn_t), p]),
        'ref_solver_end': (C.c_int, [p, i64, C.POINTER(ref_solver_t)]),
        'ref_nn_near_pos': (i64, [C.POINTER(ref_nn_t), d, d, d, p, p, p, p, p, i64]),
        'ref_newmark_predict': (None, [d, d, d, p, i64]), 'ref_newmark_correct': (None, [d, d, d, p, i64]),
        'ref_probe_pressure': (i64, [C.POINTER(ref_nn_t), p, i64, i32, C.POINTER(ref_wcsph_t), d, d, d, p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def as_particles(pA):
    a = np.ascontiguousarray(pA)
    if a.dtype != particle_dtype:
        b = np.zeros(len(a), dtype=particle_dtype)
        for f in PARTICLE_FIELDS:
            b[f] = a[f]
        a = b
    return a


def wcsph(method=None, height=1.0, r0=0.02, rho0=1000.0, **overrides):
    """ref_wcsph_t from a method object (reads its attributes) or from arguments."""
    m = ref_wcsph_t()
    if method is not None:
        height, r0, rho0 = method.height, method.r0, method.rho0
    load().ref_wcsph_init(C.byref(m), height, r0, rho0, 0, 0.0, 0)
    if method is not None:
        for f in ('useXSPH', 'useSummationDensity'):
            setattr(m, f, int(bool(getattr(method, f))))
        for f in ('D', 'p1', 'p2', 'epsilon', 'gamma', 'co', 'B', 'Pb', 'alpha', 'beta'):
            setattr(m, f, float(getattr(method, f)))
    for k, v in overrides.items():
        setattr(m, k, v)
    return m


class NN:
    """The reference's NNLinkedList over a particle array (CPU)."""

    def __init__(self, scale=2.0):
        self.s = ref_nn_t()
        load().ref_nn_init(C.byref(self.s), scale)

    def update(self, pA):
        rc = load().ref_nn_update(C.byref(self.s), _ptr(pA), len(pA))
        if rc:
            raise RuntimeError(f'ref_nn_update failed: {rc}')

    def grid(self):
        s = self.s
        return {'xmin': s.xmin, 'xmax': s.xmax, 'ymin': s.ymin, 'ymax': s.ymax, 'cell_size': s.cell_size,
                'ncx': s.ncx, 'ncy': s.ncy, 'n_cells': s.n_cells, 'oob_binned': s.oob_binned}

    def dump(self, pA):
        n = len(pA)
        keys = np.zeros(n, np.int64)
        off = np.zeros(n + 1, np.int64)
        tot = load().ref_nn_dump(C.byref(self.s), _ptr(pA), n, _ptr(keys), _ptr(off), None, None, None, 0)
        cap = max(-tot, 1)
        idx = np.zeros(cap, np.int64)
        q = np.zeros(cap)
        r = np.zeros(cap)
        tot = load().ref_nn_dump(C.byref(self.s), _ptr(pA), n, _ptr(keys), _ptr(off), _ptr(idx), _ptr(q), _ptr(r),
                                 cap)
        assert tot >= 0
        return {'keys': keys, 'off': off, 'idx': idx[:tot], 'q': q[:tot], 'r': r[:tot]}

    def near_pos(self, x, y, h, pA):
        """nearPos (src/Tools/NNLinkedList.py:41-80): (h_ij, q, r, idx) against the last update."""
        lib = load()
        k = lib.ref_nn_near_pos(C.byref(self.s), x, y, h, _ptr(pA), None, None, None, None, 0)
        hh, q, r, idx = np.zeros(k), np.zeros(k), np.zeros(k), np.zeros(k, np.int64)
        lib.ref_nn_near_pos(C.byref(self.s), x, y, h, _ptr(pA), _ptr(hh), _ptr(q), _ptr(r), _ptr(idx), k)
        return hh, q, r, idx

    def __del__(self):
        try:
            load().ref_nn_free(C.byref(self.s))
        except Exception:
            pass


def loop(pA, method, kernel='cubic', scale=2.0, nthreads=1):
    """nn.update + _loop on an active array (in place); returns the NN."""
    nn = NN(scale)
    nn.update(pA)
    rc = load().ref_loop(_ptr(pA), len(pA), KERNELS[kernel], C.byref(wcsph(method)), C.byref(nn.s), nthreads)
    if rc:
        raise RuntimeError(f'ref_loop failed: {rc}')
    return nn


class Solver:
    """The reference's Solver.run loop body on a particle array (CPU, in place)."""

    def __init__(self, pA, method, kernel='cubic', integrator='pec', h=None, damping=0.0, strict=False,
                 nthreads=1, scale=2.0, gamma_c=0.25, gamma_f=0.25, pec_xsph=None, newmark=None):
        """pec_xsph: PEC(useXSPH=...); default: the method's useXSPH (the
        reference requires the two to agree). newmark: (beta, gamma) of the
        coupling integrator for the coupled particles, None for none."""
        self.pA = as_particles(pA)
        self.m = wcsph(method)
        self.nn = NN(scale)
        s = ref_solver_t()
        s.kernel = KERNELS[kernel]
        s.integrator = INTEGRATORS[integrator]
        s.pec_useXSPH = int(bool(getattr(method, 'useXSPH', False) if pec_xsph is None else pec_xsph))
        s.pec_strict = int(bool(strict))
        s.h_fixed = float(h) if h else 0.0
        s.damping = damping
        s.gamma_c = gamma_c
        s.gamma_f = gamma_f
        s.nthreads = nthreads
        if newmark is not None:
            s.coupled_integrator = 1
            s.nm_beta, s.nm_gamma = float(newmark[0]), float(newmark[1])
        self.s = s
        self.kernel = kernel

    def run(self, nsteps):
        hist = np.zeros((nsteps, 3))
        rc = load().ref_solver_run(_ptr(self.pA), len(self.pA), C.byref(self.s), C.byref(self.m),
                                   C.byref(self.nn.s), nsteps, _ptr(hist))
        if rc:
            raise RuntimeError(f'ref_solver_run failed: {rc}')
        return hist

    def begin(self):
        """The loop body up to the coupling hook (src/Solver.py:366-392); returns (dt, dt_c, dt_f)."""
        out = np.zeros(3)
        rc = load().ref_solver_begin(_ptr(self.pA), len(self.pA), C.byref(self.s), C.byref(self.m),
                                     C.byref(self.nn.s), _ptr(out))
        if rc:
            raise RuntimeError(f'ref_solver_begin failed: {rc}')
        return out

    def end(self):
        """The loop body after the coupling hook (src/Solver.py:394-410)."""
        rc = load().ref_solver_end(_ptr(self.pA), len(self.pA), C.byref(self.s))
        if rc:
            raise RuntimeError(f'ref_solver_end failed: {rc}')

    def probe(self, x, y, h):
        """IceBreak's pressure_SPH at points (x, y) with smoothing length h,
        against this solver's last neighbour update; returns (rho, p)."""
        return probe_pressure(self.nn, self.pA, self.kernel, self.m, x, y, h)

    @property
    def t(self):
        return self.s.t


# -- element-wise and per-record functions ---------------------------------
def kernel_evaluate(kernel, r, h):
    r = np.ascontiguousarray(r, dtype=np.float64)
    h = np.ascontiguousarray(h, dtype=np.float64)
    out = np.zeros_like(r)
    load().ref_kernel_evaluate_n(KERNELS[kernel], r.size, _ptr(r), _ptr(h), _ptr(out))
    return out


def kernel_gradient(kernel, x, r, h):
    x = np.ascontiguousarray(x, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    h = np.ascontiguousarray(h, dtype=np.float64)
    out = np.zeros_like(r)
    load().ref_kernel_gradient_n(KERNELS[kernel], r.size, _ptr(x), _ptr(r), _ptr(h), _ptr(out))
    return out


def gaussian_derivative(r, h):
    f = load().ref_gaussian_derivative
    return np.array([f(a, b) for a, b in zip(np.ravel(r), np.ravel(h))])


def tait(gamma, B, rho0, rho, label):
    f = load().ref_tait_eos
    return np.array([f(gamma, B, rho0, a, int(b)) for a, b in zip(np.ravel(rho), np.ravel(label))])


def continuity(comp):
    comp = np.ascontiguousarray(comp, dtype=computed_dtype)
    return load().ref_continuity(_ptr(comp), len(comp))


def momentum(alpha, beta, p, comp):
    comp = np.ascontiguousarray(comp, dtype=computed_dtype)
    p = np.ascontiguousarray(np.atleast_1d(p), dtype=particle_dtype)
    out = np.zeros(2)
    load().ref_momentum(alpha, beta, _ptr(p), _ptr(comp), len(comp), _ptr(out))
    return out


def boundary_force(r0, D, p1, p2, comp):
    comp = np.ascontiguousarray(comp, dtype=computed_dtype)
    out = np.zeros(2)
    load().ref_boundary_force(r0, D, p1, p2, _ptr(comp), len(comp), _ptr(out))
    return out


def timestep(pA, gamma_c=0.25, gamma_f=0.25):
    pA = as_particles(pA)
    out = np.zeros(3)
    load().ref_timestep(len(pA), _ptr(pA), gamma_c, gamma_f, _ptr(out))
    return out


def pec_predict(pA, dt, damping=0.0, useXSPH=False, strict=False):
    load().ref_pec_predict(int(useXSPH), int(strict), dt, _ptr(pA), len(pA), damping)
    return pA


def pec_correct(pA, dt, damping=0.0, useXSPH=False, strict=False):
    load().ref_pec_correct(int(useXSPH), int(strict), dt, _ptr(pA), len(pA), damping)
    return pA


def euler_correct(pA, dt, damping=0.0):
    load().ref_euler_correct(dt, _ptr(pA), len(pA), damping)
    return pA


def newmark_predict(pA, dt, beta, gamma):
    load().ref_newmark_predict(beta, gamma, dt, _ptr(pA), len(pA))
    return pA


def newmark_correct(pA, dt, beta, gamma):
    load().ref_newmark_correct(beta, gamma, dt, _ptr(pA), len(pA))
    return pA


def probe_pressure(nn, pA, kernel, method, x, y, h):
    """examples/IceBreak.py:253-285 at every point (x[k], y[k]): (rho, p)."""
    m = method if isinstance(method, ref_wcsph_t) else wcsph(method)
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    y = np.atleast_1d(np.asarray(y, dtype=np.float64))
    h = np.broadcast_to(np.asarray(h, dtype=np.float64), x.shape)
    rho, p = np.zeros(len(x)), np.zeros(len(x))
    out = np.zeros(2)
    for k in range(len(x)):
        if load().ref_probe_pressure(C.byref(nn.s), _ptr(pA), len(pA), KERNELS[kernel], C.byref(m), x[k], y[k],
                                     h[k], _ptr(out)) < 0:
            raise MemoryError('ref_probe_pressure')
        rho[k], p[k] = out
    return rho, p
