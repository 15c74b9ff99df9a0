"""Summation density (src/Equations/SummationDensity.py:5-13): the sum of
m w over the fluid neighbours, in neighbour order. Host function for
coupling hooks; the loop's form runs on the device (``k_sumden``)."""
import numpy as np

from ..Common import ParticleType


def SummationDensity(labels: np.ndarray, m: np.ndarray, w: np.ndarray) -> float:
    rho = 0.0
    for j in range(len(labels)):
        if labels[j] != ParticleType.Fluid:
            continue
        rho += m[j] * w[j]
    return rho


__all__ = ['SummationDensity']
