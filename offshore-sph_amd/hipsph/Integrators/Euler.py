"""Explicit Euler (src/Integrators/Euler.py:6-26): predict is a no-op,
correct advances x, v and rho with the forces of the step."""
from .. import _lib
from .Integrator import Integrator


class Euler(Integrator):
    integrator_id = _lib.SPH_INTEGRATOR_EULER
    name = 'euler'
    useXSPH = False
    strict = False

    def isMultiStage(self) -> bool:
        return False
