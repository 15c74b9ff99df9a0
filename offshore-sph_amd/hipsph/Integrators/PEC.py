"""Monaghan (1992) predictor-corrector (src/Integrators/PEC.py:7-89).

Selects the device PEC update (libhipsph sph_predict / sph_correct) for the
solver's fluid particles.
"""
from .. import _lib
from .Integrator import Integrator


class PEC(Integrator):
    integrator_id = _lib.SPH_INTEGRATOR_PEC
    name = 'pec'

    def __init__(self, useXSPH: bool = True, strict: bool = True):
        self.useXSPH = useXSPH
        self.strict = strict

    def isMultiStage(self) -> bool:
        return False
