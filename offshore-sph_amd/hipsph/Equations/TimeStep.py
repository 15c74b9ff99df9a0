"""Courant / force time step (src/Equations/TimeStep.py:7-90).

``TimeStep().compute(J, pA, gamma_c, gamma_f) -> (dt, dt_c, dt_f)`` over the
first J records (the fluid, as ``Solver._minTimeStep`` passes them):
dt_c = gamma_c h_min / c_max, dt_f = gamma_f sqrt(h_min / max(ax^2 + ay^2))
(1e10 when max a^2 < 1e-12), dt = min(dt_c, dt_f). The reduction runs on the
device (the engine's dt kernels, which use c = c0 for every fluid particle,
as WCSPH.compute_speed_of_sound sets it); the records' own ``c`` is checked to
be that single value.
"""
import numpy as np


class TimeStep:
    def compute(self, J, pA, gamma_c: float = 0.25, gamma_f: float = 0.25):
        from ..device import Engine, EngineParams
        pA = pA[:J]
        c = np.asarray(pA['c'], dtype=np.float64)
        if len(c) == 0:
            raise ValueError('no particles')
        if not np.all(c == c[0]):
            raise NotImplementedError('the device dt reduction assumes one speed of sound for the fluid')
        e = Engine(EngineParams(n=len(pA), co=float(c[0]), gamma_c=gamma_c, gamma_f=gamma_f))
        try:
            soa = {k: np.ascontiguousarray(pA[k]) for k in ('x', 'y', 'vx', 'vy', 'rho', 'm', 'h', 'ax', 'ay',
                                                           'drho')}
            soa['label'] = np.zeros(len(pA), dtype=np.int8)  # every record counts, as in computeVars
            soa['m'] = np.abs(soa['m'])
            e.upload(soa)
            e.dt()
            st = e.status()
            return st['dt'], st['dt_c'], st['dt_f']
        finally:
            e.close()

    def courant(self, cfl, h_min, c_max) -> float:
        """:47-49"""
        return cfl * h_min / c_max

    def force(self, cfl, min_h, max_a) -> float:
        """:51-56"""
        if max_a < 1e-12:
            return 1e10
        return cfl * np.sqrt(min_h / max_a)


__all__ = ['TimeStep']
