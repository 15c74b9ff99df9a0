"""Cell-linked-list neighbour search (src/Tools/NNLinkedList.py:21-208).

Same interface as the reference's ``NNLinkedList``: ``update(pA)`` bins the
active particle records, ``near(i, pA)`` / ``nearPos(x, y, h, pA)`` return
``(h_ij, q, r, idx)`` for every particle of the 3 x 3 cells around the query
whose q = r / (0.5 (h_i + h_j)) <= 3, in the reference's order (x shift
outer, each cell in descending index order). The binning and the queries run
on the device (libhipsph ``sph_nn_update`` / ``sph_nn_dump``); ``update``
fetches the neighbour lists of every particle at once, so ``near`` is a
lookup. ``nearPos`` for an arbitrary point is answered from the same device
grid through a one-particle query context.
"""
import numpy as np

from .. import _lib


class NNLinkedList:
    def __init__(self, scale: float = 2.0, device: str = 'cuda:0'):
        self.scale = float(scale)
        self.device = device
        self.shifts = np.array([-1, 0, 1], dtype=np.int64)
        self._engine = None
        self._n = -1

    def _make(self, n):
        from ..device import Engine, EngineParams
        if self._engine is not None and self._n == n:
            return self._engine
        if self._engine is not None:
            self._engine.close()
        # the neighbour search does not read the WCSPH constants
        self._engine = Engine(EngineParams(n=n, scale=self.scale), device=self.device)
        self._n = n
        return self._engine

    def update(self, pA: np.ndarray):
        """_init + _bin over the active records pA (:37-39)."""
        n = len(pA)
        if n == 0:
            raise ValueError('no particles')
        e = self._make(n)
        soa = {k: np.ascontiguousarray(pA[k]) for k in ('x', 'y', 'vx', 'vy', 'rho', 'm', 'h', 'label')}
        soa['m'] = np.abs(soa['m'])  # the search ignores mass; the engine only checks it is finite
        e.upload(soa)
        e.nn_update()
        g = e.grid()
        self.xmin, self.xmax, self.ymin, self.ymax = g['xmin'], g['xmax'], g['ymin'], g['ymax']
        self.cell_size = g['cell_size']
        self.ncells_per_dim = np.array([g['ncx'], g['ncy']], dtype=np.int64)
        self.n_cells = g['n_cells']
        self._dump = e.nn_dump(with_qr=True)
        self._h = np.ascontiguousarray(pA['h'], dtype=np.float64)

    def near(self, i: int, pA: np.ndarray):
        """(h_ij, q, r, idx) of particle i (:82-84)."""
        d = self._dump
        s, e = int(d['off'][i]), int(d['off'][i + 1])
        idx = d['idx'][s:e].astype(np.uint64)
        h_ij = 0.5 * (self._h[i] + self._h[idx.astype(np.int64)])
        return h_ij, d['q'][s:e].copy(), d['r'][s:e].copy(), idx

    def nearPos(self, x: float, y: float, h: float, pA: np.ndarray):
        """Neighbours of an arbitrary point (:41-80) against the last update."""
        # a point query = the query of an extra particle appended to the set,
        # with the grid of the last update (the extra particle must not move
        # the bounds, so it is answered only inside them)
        if not (self.xmin <= x <= self.xmax and self.ymin <= y <= self.ymax):
            raise _lib.HipSphError(-1, 'nearPos is answered inside the bounds of the last update only')
        ext = np.concatenate([pA, pA[:1]])
        ext[-1]['x'], ext[-1]['y'], ext[-1]['h'] = x, y, h
        ext[-1]['label'] = 1
        sub = NNLinkedList(self.scale, self.device)
        sub.update(ext)
        hh, q, r, idx = sub.near(len(ext) - 1, ext)
        keep = idx != len(ext) - 1
        sub.close()
        return hh[keep], q[keep], r[keep], idx[keep]

    def close(self):
        if self._engine is not None:
            self._engine.close()
            self._engine = None


__all__ = ['NNLinkedList']
