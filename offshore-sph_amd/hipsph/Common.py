"""Particle data model of the WCSPH step.

Mirrors ``src/Common.py`` of KoningJasper/Offshore-SPH: the ``ParticleType``
enum (src/Common.py:8-12), ``get_label_code`` (:14-24), the 154-byte
``particle_dtype`` record (:26-57) and the 113-byte ``computed_dtype``
neighbour record (:60-98), field for field and in the same order, so host
arrays built for the reference can be handed to this package unchanged.
``np.bool`` (removed from numpy) is spelled ``np.bool_``.

On the device the engine keeps the state as structure-of-arrays, not as these
records; ``hipsph.device`` converts between the two.
"""
import enum

import numpy as np


class ParticleType(enum.IntEnum):
    Fluid = 0
    Boundary = 1
    TempBoundary = 2
    Coupled = 3


def get_label_code(label: str):
    if label == 'fluid':
        return ParticleType.Fluid
    elif label == 'boundary':
        return ParticleType.Boundary
    elif label == 'temp-boundary':
        return ParticleType.TempBoundary
    elif label == 'coupled':
        return ParticleType.Coupled
    else:
        raise Exception('Argument out of range')


PARTICLE_FIELDS = ['deleted', 'label', 'm', 'rho', 'p', 'c', 'drho', 'h', 'x', 'y', 'vx', 'vy',
                   'ax', 'ay', 'xsphx', 'xsphy', 'x0', 'y0', 'vx0', 'vy0', 'rho0']

particle_dtype = np.dtype({
    'names': PARTICLE_FIELDS,
    'formats': [np.bool_, np.int8] + [np.double] * 19,
})

COMPUTED_FIELDS = ['label', 'm', 'p', 'rho', 'h', 'q', 'c', 'r', 'w', 'dw_x', 'dw_y', 'x', 'y', 'vx', 'vy']

computed_dtype = np.dtype({
    'names': COMPUTED_FIELDS,
    'formats': [np.int8] + [np.double] * 14,
})

assert particle_dtype.itemsize == 154
assert computed_dtype.itemsize == 113
