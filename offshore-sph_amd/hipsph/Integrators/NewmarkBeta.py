"""Newmark-beta for the coupled particles (src/Integrators/NewmarkBeta.py:12-95).

The coupling integrator of ``Solver(coupling=..., couplingIntegrator=...)``.
On the device the predictor and the corrector run inside the step for every
particle labelled Coupled (``sph_params_t.coupled_integrator``, ``beta``,
``gamma``); ``predict`` and ``correct`` here are the same updates on a host
record array. ``acceleration`` solves the structure's linear system on the
host, as the reference does: the coupling hook calls it once per step with
the node forces.
"""
import numpy as np

from ..Common import ParticleType
from .Integrator import Integrator


class NewmarkBeta(Integrator):
    name = 'newmark'

    def __init__(self, beta: float, gamma: float, M: np.ndarray, K: np.ndarray, C: np.ndarray):
        """beta, gamma: the Newmark parameters (commonly 0.25, 0.5); M, K, C:
        N x N mass, stiffness and damping matrices of the coupled nodes."""
        assert beta * 2 >= gamma
        self.beta = float(beta)
        self.gamma = float(gamma)
        self.M = np.asarray(M, dtype=np.float64)
        self.K = np.asarray(K, dtype=np.float64)
        self.C = np.asarray(C, dtype=np.float64)

    def isMultiStage(self) -> bool:
        return False

    def predict(self, dt: float, pA: np.ndarray, damping: float) -> np.ndarray:
        """:41-52 (damping unused). dt ** 2 is the product dt * dt, as numba
        lowers a power with the integer literal 2."""
        self._check(pA)
        dt2 = dt * dt
        pA['x'] = pA['x'] + pA['vx'] * dt + pA['ax'] * (0.5 - self.beta) * dt2
        pA['y'] = pA['y'] + pA['vy'] * dt + pA['ay'] * (0.5 - self.beta) * dt2
        pA['vx'] = pA['vx'] + pA['ax'] * (1 - self.gamma) * dt
        pA['vy'] = pA['vy'] + pA['ay'] * (1 - self.gamma) * dt
        return pA

    def correct(self, dt: float, pA: np.ndarray, damping: float) -> np.ndarray:
        """:54-66"""
        self._check(pA)
        dt2 = dt * dt
        pA['x'] = pA['x'] + pA['ax'] * self.beta * dt2
        pA['y'] = pA['y'] + pA['ay'] * self.beta * dt2
        pA['vx'] = pA['vx'] + pA['ax'] * self.gamma * dt
        pA['vy'] = pA['vy'] + pA['ay'] * self.gamma * dt
        return pA

    def acceleration(self, dt: float, F: np.ndarray, r: np.ndarray, v: np.ndarray) -> np.ndarray:
        """:68-92: a = (M + gamma dt C + beta dt^2 K)^-1 (F - K r - C v) with
        the predicted displacement r and velocity v."""
        rhs = F - np.dot(self.K, r) - np.dot(self.C, v)
        lhs = self.M + self.C * self.gamma * dt + self.K * self.beta * dt ** 2
        return np.dot(np.linalg.inv(lhs), rhs)

    @staticmethod
    def _check(pA):
        # checkCoupling :94-96 warns, and integrates the record anyway
        if np.any(pA['label'] != ParticleType.Coupled):
            print('\nNewmark-Beta integration method is only suited for solid coupling particles.')


__all__ = ['NewmarkBeta']
