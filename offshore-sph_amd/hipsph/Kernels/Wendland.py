"""Wendland C4 kernel 9 / (4 pi h^2), support 2h (src/Kernels/Wendland.py:6-64).
Device implementation: libhipsph."""
from .. import _lib
from .Kernel import Kernel


class Wendland(Kernel):
    kernel_id = _lib.SPH_KERNEL_WENDLAND
    name = 'wendland'
