"""sph_loop_noalloc: nn.update + _loop with its workspace in a static device
array, against sph_nn_update + sph_compute on a context (the entry points it
is a variant of), through the C ABI.

Two calls in a row, the second with a smaller input (the static array keeps
the first call's contents), each equal to the context path bit for bit in
FP64, and in FP32C within the float pair pass's 2e-5 (the context takes the
pass's uniform form, this entry point its general one); then the limits: more particles than the workspace holds is
refused on the host, and a grid of more cells than it holds comes back as
SPH_FLAG_CAPACITY in the status word with the outputs untouched.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from gpu_util import make_engine, perturbed, say
from hipsph import _lib, scenarios
from hipsph._lib import sph_state_t
from hipsph.device import EngineParams, soa_from_particles

pytestmark = pytest.mark.gpu


def _device(arrs):
    out = {}
    for k, v in arrs.items():
        t = torch.as_tensor(np.ascontiguousarray(v))
        out[k] = t.to('cuda')
    return out


def _noalloc(pA, method, precision):
    n = len(pA)
    p = EngineParams.from_method(n, method, kernel='cubic', precision=precision).to_c()
    d = _device(soa_from_particles(pA))
    st = sph_state_t()
    for f, _ in sph_state_t._fields_:
        if f in d:
            setattr(st, f, d[f].data_ptr())
    out = {k: torch.full((n,), -7.0, dtype=torch.float64, device='cuda') for k in ('drho', 'ax', 'ay')}
    status = torch.full((1,), -1, dtype=torch.int32, device='cuda')
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.call('sph_loop_noalloc', C.byref(p), C.byref(st), out['drho'].data_ptr(), out['ax'].data_ptr(),
              out['ay'].data_ptr(), status.data_ptr(), stream)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}, int(status.item())


def _context(pA, method, precision):
    e = make_engine(pA, method, kernel='cubic', precision=precision)
    e.nn_update()
    e.compute()
    d = e.download()
    e.close()
    return d


@pytest.mark.parametrize('precision', ['fp64', 'fp32c'])
def test_noalloc_loop_matches_context_twice(precision):
    big = scenarios.dam_break_c1()
    small = scenarios.dam_break(0.4, 0.5, 0.02, tank_w=1.0, tank_h=0.8, name='noalloc_small')
    assert small.n < big.n <= _lib.SPH_NOALLOC_MAX_PARTICLES
    for k, sc in enumerate((big, small)):
        pA = perturbed(sc.pA, 20 + k)
        f = pA['label'] == 0
        pA['ax'][~f] = 3.0  # walls keep their rates: the outputs carry them through
        m = sc.method()
        say('noalloc', precision, 'call', k, 'n', sc.n)
        got, status = _noalloc(pA, m, precision)
        want = _context(pA, m, precision)
        assert status & 3 == 0, status
        for fld in ('drho', 'ax', 'ay'):
            if precision == 'fp64':
                np.testing.assert_array_equal(got[fld], want[fld], err_msg=f'{precision} call {k} {fld}')
            else:  # the context takes the float pass's uniform form, this entry point the general one
                scale = float(np.max(np.abs(want[fld][f])))
                assert np.max(np.abs(got[fld] - want[fld])) <= 2e-5 * scale, (k, fld)
                np.testing.assert_array_equal(got[fld][~f], want[fld][~f])


def test_noalloc_limits():
    sc = scenarios.dam_break_c1()
    m = sc.method()
    n = _lib.SPH_NOALLOC_MAX_PARTICLES + 1
    p = EngineParams.from_method(n, m).to_c()
    st = sph_state_t()
    with pytest.raises(_lib.HipSphError, match='workspace'):
        _lib.call('sph_loop_noalloc', C.byref(p), C.byref(st), 1, 1, 1, 1, None)
    # two particles 400 m apart with h = 0.026: 7,700 x 1 ... cells per axis
    # beyond the workspace's cell count
    pA = sc.pA[:2].copy()
    pA['x'] = [0.0, 400.0]
    pA['y'] = [0.0, 400.0]
    got, status = _noalloc(pA, m, 'fp64')
    say('noalloc capacity status', status)
    assert status & _lib.SPH_FLAG_CAPACITY
    assert np.all(got['ax'] == -7.0) and np.all(got['drho'] == -7.0)
