"""Lattice generator used to build particle arrays.

Restates ``Helpers.rect`` of KoningJasper/Offshore-SPH (src/Helpers.py:7-87):
N = max(1, ceil(range / r0)) points per axis from an inclusive
``np.linspace``, ``meshgrid(indexing='ij')`` ravelled (x outer, y inner), the
label/mass/density fields filled and ``h`` left at 0, with optional hex
packing. Host-side geometry only: it builds inputs, it computes nothing on
the step path.
"""
from math import ceil, floor

import numpy as np

from .Common import ParticleType, particle_dtype


class Helpers:
    @staticmethod
    def rect(xmin: float, xmax: float, ymin: float, ymax: float, r0: float, mass: float = 1.0,
             rho0: float = 1000.0, pack: bool = False, label: ParticleType = ParticleType.Fluid,
             strict: bool = False, packDirection: str = 'y') -> np.ndarray:
        xrange = xmax - xmin
        yrange = ymax - ymin
        Nx = max(1, ceil(xrange / r0))
        Ny = max(1, ceil(yrange / r0))
        xv = np.linspace(xmin, xmax, Nx)
        yv = np.linspace(ymin, ymax, Ny)
        x, y = np.meshgrid(xv, yv, indexing='ij')
        x = x.ravel()
        y = y.ravel()
        pA = np.zeros(len(x), dtype=particle_dtype)
        pA['label'] = label
        pA['x'] = x
        pA['y'] = y
        pA['m'] = mass
        pA['rho'] = rho0
        if pack and xrange > 0:
            if packDirection == 'y':
                for i in range(floor(Nx / 2)):
                    pA['y'][2 * i * Ny: 2 * i * Ny + Ny] += r0 / 2
            elif packDirection == 'x':
                for i in range(floor(Ny / 2)):
                    pA['x'][2 * i::Nx] += r0 / 2
            else:
                raise Exception('Invalid packing direction.')
            if strict:
                keep = (pA['x'] <= xmax) & (pA['x'] >= xmin) & (pA['y'] <= ymax) & (pA['y'] >= ymin)
                pA = pA[keep]
        return pA
