"""Method plugin interface (src/Methods/Method.py:5-36)."""
import abc


class Method(abc.ABC):
    """Generic method: constants plus the per-array and per-particle hooks the solver calls."""

    @abc.abstractmethod
    def initialize(self, pA):
        ...

    @abc.abstractmethod
    def compute_speed_of_sound(self, pA):
        ...

    @abc.abstractmethod
    def compute_pressure(self, pA):
        ...
