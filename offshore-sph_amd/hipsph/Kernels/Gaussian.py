"""Gaussian kernel 1 / (pi h^2) exp(-q^2), truncated at q <= 3
(src/Kernels/Gaussian.py:6-59). Device implementation: libhipsph."""
from .. import _lib
from .Kernel import Kernel, _dev_call


class Gaussian(Kernel):
    kernel_id = _lib.SPH_KERNEL_GAUSSIAN
    name = 'gaussian'

    def derivative(self, r, h):
        return _dev_call('sph_kernel_derivative', self.kernel_id, [r, h])
