"""The coupling hook's view of the device state (src/Solver.py:381-398).

A coupling hook ``coupling(pA, solver)`` runs between the loop and the
correctors of every step (the reference's IceBreak example computes the
pressure on an ice sheet there and returns the sheet's accelerations).
The ``Solver`` mirror splits each device step at that point
(``sph_step_begin`` / ``sph_step_end``), hands the hook the particle array
and carries what the hook changed back to the device.

Inside a hook, ``solver.nn`` is an ``EngineNN``: ``nearPos`` / ``near``
answered from the step's device grid, as the reference's ``NNLinkedList``
after ``_compute``. ``probe_pressure`` is the device form of the example's
``pressure_SPH`` (``sph_probe_pressure``): one lane per point instead of a
neighbour list per point on the host.
"""
import numpy as np


class EngineNN:
    """NNLinkedList.nearPos / near (src/Tools/NNLinkedList.py:41-84) against
    the grid of the engine's last neighbour update and its current positions.
    Indices are those of the active records the engine holds."""

    def __init__(self, engine, solver):
        self._engine = engine
        self._solver = solver

    def _h(self):
        s = self._solver
        return s.particleArray['h'][np.nonzero(s.indexes)[0]]

    def nearPos(self, x: float, y: float, h: float, pA: np.ndarray):
        off, idx, q, r = self._engine.near_pos([x], [y], [h])
        idx = idx.astype(np.int64)
        h_ij = 0.5 * (h + self._h()[idx])
        return h_ij, q, r, idx.astype(np.uint64)

    def near(self, i: int, pA: np.ndarray):
        a = self._solver.particleArray[np.nonzero(self._solver.indexes)[0][i]]
        return self.nearPos(float(a['x']), float(a['y']), float(a['h']), pA)

    def update(self, pA):
        raise NotImplementedError('the device updates the grid inside the step')


def probe_pressure(solver, x, y, h):
    """pressure_SPH (examples/IceBreak.py:253-285) at the points (x, y) with
    smoothing length h, on the device against the current step's grid:
    (rho, p) as numpy arrays. Call it from a coupling hook."""
    e = solver._engine
    if e is None:
        raise RuntimeError('no device step is running')
    return e.probe_pressure(x, y, h)


__all__ = ['EngineNN', 'probe_pressure']
