"""Kernel plugin interface (src/Kernels/Kernel.py:5-12).

A kernel names the device kernel id the pair pass specialises on
(``kernel_id``) and evaluates W and grad W element-wise on the device for
host or torch arrays.
"""
import abc

import numpy as np

from .. import _lib
from ..device import host_array


def _dev_call(name, kid, arrays):
    import torch
    if not torch.cuda.is_available():
        raise _lib.HipSphUnavailable('no GPU visible to torch: kernels evaluate on the device only')
    dev = torch.device('cuda:0')
    host = not isinstance(arrays[0], torch.Tensor)
    ts = [torch.as_tensor(host_array(a, np.float64) if host else a, dtype=torch.float64).to(dev).contiguous()
          for a in arrays]
    n = ts[0].numel()
    out = torch.empty(n, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    _lib.call(name, kid, n, *[t.data_ptr() for t in ts], out.data_ptr(), stream)
    if host:
        return out.cpu().numpy()
    return out


class Kernel(abc.ABC):
    kernel_id = None
    name = None

    def evaluate(self, r, h):
        return _dev_call('sph_kernel_evaluate', self.kernel_id, [r, h])

    def gradient(self, x, r, h):
        return _dev_call('sph_kernel_gradient', self.kernel_id, [x, r, h])
