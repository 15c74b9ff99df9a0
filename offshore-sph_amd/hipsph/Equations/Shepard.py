"""Shepard filter (src/Equations/Shepard.py:5-12): kernel weights normalised
by the sum of w m / rho over the fluid neighbours with rho >= 1e-3.
Host function for coupling hooks; the device form is ``sph_probe_pressure``."""
import numpy as np

from ..Common import ParticleType


def Shepard(w: np.ndarray, labels: np.ndarray, m_b: np.ndarray, rho_b: np.ndarray) -> np.ndarray:
    w_tilde = 0.0
    for i in range(len(w)):
        if rho_b[i] < 1e-3 or labels[i] != ParticleType.Fluid:
            continue
        w_tilde += w[i] * m_b[i] / rho_b[i]
    with np.errstate(divide='ignore', invalid='ignore'):
        return np.asarray(w, dtype=np.float64) / w_tilde


__all__ = ['Shepard']
