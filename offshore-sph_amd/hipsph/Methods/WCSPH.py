"""WCSPH method constants and setup hooks.

Mirrors src/Methods/WCSPH.py:8-203 of KoningJasper/Offshore-SPH: the same
constructor, the same constants (gamma = 7, c0 = 10 sqrt(2 g H) via
TaitEOS_co, B = c0^2 rho0 / gamma via TaitEOS_B, alpha = 0.01, beta = 0,
epsilon = 0.5, Lennard-Jones D = 5 g H, p1 = 4, p2 = 2) held as plain
attributes. alpha, beta and co may be reassigned before Solver.setup(), as on
the reference jitclass.

The per-particle hooks of the reference (compute_density_change,
compute_acceleration, compute_velocity) have no per-particle host form here:
the device pair pass (libhipsph sph_compute) evaluates them for every fluid
particle at once. initialize (hydrostatic density, src/Equations/TaitEOS.py:46-65)
runs once at setup on the host arrays it is given; compute_pressure runs the
device Tait kernel.
"""
import math

import numpy as np

from .Method import Method

G = 9.81


def TaitEOS_co(H):
    """src/Equations/TaitEOS.py:40-44"""
    return 10.0 * math.sqrt(2 * 9.81 * H)


def TaitEOS_B(co, rho0, gamma):
    """src/Equations/TaitEOS.py:33-38"""
    return co * co * rho0 / gamma


def TaitEOS_height(rho0, H, B, gamma, y):
    """src/Equations/TaitEOS.py:46-65 (element-wise over y)."""
    frac = rho0 * 9.81 * (H - np.asarray(y, dtype=np.float64)) / B
    return rho0 * (1 + frac) ** (1 / gamma)


class WCSPH(Method):
    def __init__(self, height: float, r0: float, rho0: float, useXSPH: bool = False, Pb: float = 0.0,
                 useSummationDensity: bool = False):
        self.height = height
        self.rho0 = rho0
        self.useXSPH = useXSPH
        self.useSummationDensity = useSummationDensity
        self.epsilon = 0.5
        self.gamma = 7.0
        self.co = TaitEOS_co(self.height)
        self.B = TaitEOS_B(self.co, self.rho0, self.gamma)
        self.Pb = Pb
        self.alpha = 0.01
        self.beta = 0.0
        self.r0 = r0
        self.D = 5 * 9.81 * self.height
        self.p1 = 4
        self.p2 = 2

    def initialize(self, pA):
        """Hydrostatic density from height (src/Methods/WCSPH.py:82-108)."""
        pA['rho'] = TaitEOS_height(self.rho0, self.height, self.B, self.gamma, pA['y'])
        return pA

    def compute_speed_of_sound(self, pA):
        """c = co for every particle (src/Methods/WCSPH.py:110-129)."""
        return np.full(len(pA), self.co, dtype=np.float64)

    def compute_pressure(self, pA):
        """Tait pressure + Pb on the device (src/Methods/WCSPH.py:131-149)."""
        from ..Equations.TaitEOS import TaitEOS
        return TaitEOS(self.gamma, self.B, self.rho0, pA['rho'], pA['label']) + self.Pb
