"""Tait equation of state on the device (src/Equations/TaitEOS.py:6-31)."""
import numpy as np

from .. import _lib
from ..device import host_array


def TaitEOS(gamma, B, rho0, rho, label, Pb=0.0):
    import torch
    if not torch.cuda.is_available():
        raise _lib.HipSphUnavailable('no GPU visible to torch: TaitEOS evaluates on the device only')
    dev = torch.device('cuda:0')
    r = torch.as_tensor(host_array(rho, np.float64)).to(dev)
    lab = torch.as_tensor(host_array(label, np.int8)).to(dev)
    out = torch.empty_like(r)
    _lib.call('sph_tait_eos', r.numel(), float(gamma), float(B), float(rho0), float(Pb), r.data_ptr(),
              lab.data_ptr(), out.data_ptr(), torch.cuda.current_stream(dev).cuda_stream)
    return out.cpu().numpy()
