"""ctypes binding of libhipsph.so (include/hipsph.h).

The library is loaded from this package directory only. There is no
fallback: if it is missing or does not load, every entry point raises
``HipSphUnavailable`` naming the build command.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libhipsph.so')

SPH_KERNEL_CUBIC, SPH_KERNEL_GAUSSIAN, SPH_KERNEL_WENDLAND = 0, 1, 2
SPH_PREC_FP64, SPH_PREC_FP32C = 0, 1
SPH_INTEGRATOR_PEC, SPH_INTEGRATOR_EULER = 0, 1
SPH_COUPLED_NONE, SPH_COUPLED_NEWMARK = 0, 1

SPH_FLAG_CAPACITY = 1
SPH_FLAG_NONFINITE = 2
SPH_FLAG_OOB_KEY = 4
SPH_FLAG_DT_SMALL = 8
SPH_FLAG_NO_FLUID = 16
SPH_FLAG_ZERO_DENSITY = 32
SPH_DT_HISTORY = 65536
SPH_NOALLOC_MAX_PARTICLES = 32768
SPH_NOALLOC_MAX_CELLS = 131072

ERRORS = {-1: 'EINVAL', -2: 'EHIP', -3: 'ENOMEM', -4: 'ECAPACITY', -5: 'ESTATE', -6: 'ENUMERIC',
          -7: 'EUNSUPPORTED'}


class HipSphError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f'hipsph {ERRORS.get(code, code)}: {msg}')
        self.code = code


class HipSphUnavailable(RuntimeError):
    pass


class sph_params_t(C.Structure):
    _fields_ = [
        ('n', C.c_int64),
        ('kernel', C.c_int32), ('precision', C.c_int32), ('integrator', C.c_int32), ('strict', C.c_int32),
        ('use_xsph', C.c_int32), ('use_summation_density', C.c_int32),
        ('rho0', C.c_double), ('gamma', C.c_double), ('co', C.c_double), ('B', C.c_double), ('Pb', C.c_double),
        ('alpha', C.c_double), ('beta', C.c_double),
        ('r0', C.c_double), ('D', C.c_double), ('p1', C.c_double), ('p2', C.c_double),
        ('gravity', C.c_double), ('scale', C.c_double), ('qmax', C.c_double),
        ('gamma_c', C.c_double), ('gamma_f', C.c_double), ('damping', C.c_double),
        ('fixed_dt', C.c_double), ('duration', C.c_double),
        ('cell_capacity', C.c_int64),
        ('epsilon', C.c_double),
        ('dynamic_h', C.c_int32),
        ('coupled_integrator', C.c_int32),
        ('newmark_beta', C.c_double), ('newmark_gamma', C.c_double),
    ]


class sph_state_t(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in ('x', 'y', 'vx', 'vy', 'rho', 'ax', 'ay', 'drho', 'm', 'h', 'p', 'label',
                                           'xsphx', 'xsphy')]


class sph_adjoint_t(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in ('x', 'y', 'vx', 'vy', 'rho', 'ax', 'ay', 'drho', 'g_alpha', 'g_co')]


_P = C.c_void_p
_I64 = C.c_int64
_I32 = C.c_int32
_D = C.c_double

_PROTOS = {
    'sph_ctx_create': (C.c_int, [C.POINTER(sph_params_t), C.POINTER(C.c_void_p)]),
    'sph_ctx_destroy': (C.c_int, [_P]),
    'sph_ctx_set_params': (C.c_int, [_P, C.POINTER(sph_params_t)]),
    'sph_state_upload': (C.c_int, [_P, C.POINTER(sph_state_t), _I32, _P]),
    'sph_state_download': (C.c_int, [_P, C.POINTER(sph_state_t), _P]),
    'sph_nn_update': (C.c_int, [_P, _P]),
    'sph_nn_grid': (C.c_int, [_P, C.POINTER(_D), C.POINTER(_I64), _P]),
    'sph_nn_dump': (C.c_int, [_P, _P, _P, _P, _P, _P, _I64, C.POINTER(_I64), _P]),
    'sph_compute': (C.c_int, [_P, _P]),
    'sph_compute_counts': (C.c_int, [_P, _P, _P]),
    'sph_dt': (C.c_int, [_P, _P]),
    'sph_predict': (C.c_int, [_P, _P]),
    'sph_correct': (C.c_int, [_P, _P]),
    'sph_step': (C.c_int, [_P, _I64, _I32, _P]),
    'sph_step_begin': (C.c_int, [_P, _P]),
    'sph_step_end': (C.c_int, [_P, _P]),
    'sph_state_gather': (C.c_int, [_P, _P, _I64, C.POINTER(sph_state_t), _P]),
    'sph_state_scatter': (C.c_int, [_P, _P, _I64, C.POINTER(sph_state_t), _P]),
    'sph_near_pos': (C.c_int, [_P, _I64, _P, _P, _P, _P, _P, _P, _P, _I64, C.POINTER(_I64), _P]),
    'sph_probe_pressure': (C.c_int, [_P, _I64, _P, _P, _P, _P, _P, _P]),
    'sph_status': (C.c_int, [_P, C.POINTER(_D), C.POINTER(_I64), _P]),
    'sph_dt_history': (C.c_int, [_P, C.POINTER(_D), _I64, _P]),
    'sph_set_clock': (C.c_int, [_P, _D, C.c_int32, _P]),
    'sph_continuity': (C.c_int, [_P, _P]),
    'sph_tape_create': (C.c_int, [_P, _I32, C.POINTER(C.c_void_p)]),
    'sph_tape_destroy': (C.c_int, [_P]),
    'sph_tape_force_steps': (C.c_int, [_P, C.POINTER(_I64)]),
    'sph_step_record': (C.c_int, [_P, _P, _I32, _P]),
    'sph_adjoint_step': (C.c_int, [_P, _P, _I32, C.POINTER(sph_adjoint_t), _P]),
    'sph_adjoint_segment': (C.c_int, [_P, _P, _I32, C.POINTER(sph_adjoint_t), _P]),
    'sph_state_restore': (C.c_int, [_P, C.POINTER(sph_state_t), _D, _P]),
    'sph_batch_create': (C.c_int, [C.POINTER(sph_params_t), _I32, C.POINTER(C.c_void_p)]),
    'sph_batch_destroy': (C.c_int, [_P]),
    'sph_batch_upload': (C.c_int, [_P, C.POINTER(sph_state_t), _P]),
    'sph_batch_download': (C.c_int, [_P, C.POINTER(sph_state_t), _P]),
    'sph_batch_step': (C.c_int, [_P, _I64, _I32, _P]),
    'sph_batch_status': (C.c_int, [_P, C.POINTER(_D), C.POINTER(_I64), _P]),
    'sph_batch_set_tuning': (C.c_int, [_P, _I32]),
    'sph_batch_set_list_len': (C.c_int, [_P, _I32]),
    'sph_batch_set_persistent': (C.c_int, [_P, _I32]),
    'sph_batch_set_border_bounds': (C.c_int, [_P, _I32]),
    'sph_batch_border_count': (C.c_int, [_P, C.POINTER(_I64), _P]),
    'sph_batch_timing_enable': (C.c_int, [_P, _I32]),
    'sph_batch_timing_read': (C.c_int, [_P, C.POINTER(_D), C.POINTER(_I64), _I32]),
    'sph_momentum': (C.c_int, [_P, _P]),
    'sph_kinetic_energy': (C.c_int, [_P, C.POINTER(_D), _P]),
    'sph_kernel_evaluate': (C.c_int, [_I32, _I64, _P, _P, _P, _P]),
    'sph_kernel_gradient': (C.c_int, [_I32, _I64, _P, _P, _P, _P, _P]),
    'sph_kernel_derivative': (C.c_int, [_I32, _I64, _P, _P, _P, _P]),
    'sph_tait_eos': (C.c_int, [_I64, _D, _D, _D, _D, _P, _P, _P, _P]),
    'sph_last_error': (C.c_char_p, []),
    'sph_version': (C.c_char_p, []),
    'sph_ctx_set_tuning': (C.c_int, [_P, _I32]),
    'sph_ctx_get_tuning': (C.c_int, [_P, C.POINTER(_I32)]),
    'sph_ctx_set_pair_staging': (C.c_int, [_P, _I32]),
    'sph_ctx_set_border_bounds': (C.c_int, [_P, _I32]),
    'sph_nn_border_count': (C.c_int, [_P, C.POINTER(_I64), _P]),
    'sph_timing_enable': (C.c_int, [_P, _I32]),
    'sph_timing_read': (C.c_int, [_P, C.POINTER(_D), C.POINTER(_D), C.POINTER(_I64), _I32]),
    'sph_pair_form': (C.c_int, [_P, C.POINTER(_I32)]),
    'sph_batch_pair_form': (C.c_int, [_P, C.POINTER(_I32)]),
    'sph_loop_noalloc': (C.c_int, [C.POINTER(sph_params_t), C.POINTER(sph_state_t), _P, _P, _P, _P, _P]),
}

EXPORTED = tuple(_PROTOS)

_lib = None


def load():
    """Load libhipsph.so (once). Raises HipSphUnavailable if it cannot."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HipSphUnavailable(f'{LIB_PATH} is missing: build it with `python offshore-sph_amd/build.py` '
                                '(or __graft_entry__.build())')
    try:
        lib = C.CDLL(LIB_PATH)
    except OSError as e:
        raise HipSphUnavailable(f'cannot load {LIB_PATH}: {e}') from e
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        raise HipSphError(rc, load().sph_last_error().decode())
    return rc


def call(name, *args):
    return check(getattr(load(), name)(*args))
