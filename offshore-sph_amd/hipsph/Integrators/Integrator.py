"""Integrator plugin interface (src/Integrators/Integrator.py:4-36)."""
import abc


class Integrator(abc.ABC):
    integrator_id = None

    @abc.abstractmethod
    def isMultiStage(self) -> bool:
        ...
