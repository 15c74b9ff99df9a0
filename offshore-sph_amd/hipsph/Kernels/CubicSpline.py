"""Cubic spline kernel, 2-D normalisation 10 / (7 pi h^2), support 2h
(src/Kernels/CubicSpline.py:6-70). Device implementation: libhipsph."""
from math import pi

from .. import _lib
from .Kernel import Kernel


class CubicSpline(Kernel):
    kernel_id = _lib.SPH_KERNEL_CUBIC
    name = 'cubic'

    def __init__(self):
        self.alpha = 10 / (7 * pi)
