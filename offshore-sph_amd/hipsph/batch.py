"""Batched ensembles: many independent WCSPH simulations on one GPU.

``BatchEngine`` holds K instances of uneven particle counts (C ABI
``sph_batch_*``). Every stage of a step is one launch over all instances
(segmented cell grid, per-instance clock, dt, grid and constants) and a step
replays as one hipGraph; each instance evolves exactly as an ``Engine`` with
the same parameters would advance it alone.

``autotune(batch)`` picks the pair pass's launch shape for the batch's own
workload, the workgroup size, the per-thread contributor-list length (its
LDS tile) and the schedule (one thread per slot, or a persistent kernel
draining a work queue), by timing each candidate on the device.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import call, sph_state_t
from .device import IO_F64, EngineParams, _torch

FIELDS_IN = ('x', 'y', 'vx', 'vy', 'rho', 'm', 'h', 'label')


class BatchEngine:
    def __init__(self, params, device='cuda:0'):
        torch = _torch()
        _lib.load()
        self.params = list(params)
        self.K = len(self.params)
        if self.K == 0:
            raise ValueError('empty batch')
        self.sizes = np.array([int(p.n) for p in self.params], dtype=np.int64)
        self.offsets = np.concatenate(([0], np.cumsum(self.sizes)))
        self.n = int(self.offsets[-1])
        self.device = torch.device(device)
        arr = (_lib.sph_params_t * self.K)(*[p.to_c() for p in self.params])
        self._b = C.c_void_p()
        with torch.cuda.device(self.device):
            call('sph_batch_create', arr, self.K, C.byref(self._b))
            self.io = {k: torch.zeros(self.n, dtype=torch.float64, device=self.device) for k in IO_F64}
            self.io['label'] = torch.zeros(self.n, dtype=torch.int8, device=self.device)

    def _stream(self):
        import torch
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _state(self, fields=None):
        st = sph_state_t()
        for f, _ in sph_state_t._fields_:
            if fields is None or f in fields:
                setattr(st, f, self.io[f].data_ptr())
        return st

    def upload(self, states):
        """states[k]: dict of instance k's arrays (reference order), as for Engine.upload."""
        import torch
        if len(states) != self.K:
            raise ValueError(f'{len(states)} states for {self.K} instances')
        fields = set(FIELDS_IN)
        for f in IO_F64 + ('label',):
            if all(f in st for st in states):
                parts = [np.ascontiguousarray(st[f]) for st in states]
                for k, a in enumerate(parts):
                    if a.size != self.sizes[k]:
                        raise ValueError(f'instance {k}: {f} has {a.size} entries, expected {self.sizes[k]}')
                cat = np.concatenate(parts)
                self.io[f].copy_(torch.as_tensor(cat).to(self.io[f].dtype))
                fields.add(f)
            elif f in FIELDS_IN:
                raise KeyError(f'upload needs {f} for every instance')
        fields &= set(self.io)
        call('sph_batch_upload', self._b, C.byref(self._state(fields)), self._stream())

    def download(self):
        """Per-instance dicts of numpy arrays (reference order, p = Tait + Pb)."""
        import torch
        call('sph_batch_download', self._b, C.byref(self._state()), self._stream())
        torch.cuda.current_stream(self.device).synchronize()
        host = {k: v.cpu().numpy() for k, v in self.io.items()}
        return [{k: v[self.offsets[i]:self.offsets[i + 1]].copy() for k, v in host.items()} for i in range(self.K)]

    def step(self, nsteps=1, graph=True):
        call('sph_batch_step', self._b, int(nsteps), int(bool(graph)), self._stream())

    def status(self):
        d = (C.c_double * (4 * self.K))()
        i = (C.c_int64 * (2 * self.K))()
        call('sph_batch_status', self._b, d, i, self._stream())
        return [{'t': d[4 * k], 'dt': d[4 * k + 1], 'dt_c': d[4 * k + 2], 'dt_f': d[4 * k + 3],
                 'steps': int(i[2 * k]), 'flags': int(i[2 * k + 1])} for k in range(self.K)]

    def pair_form(self):
        """The pair pass the batch runs, chosen at upload: 'fp64', 'fp32c', or
        'fp64-uniform', 'fp32c-uniform' when every instance has one h and one
        fluid mass (the uniform pass has its own workgroup size and list, so
        only the schedule of set_tuning applies to it)."""
        f = C.c_int32()
        call('sph_batch_pair_form', self._b, C.byref(f))
        return {0: 'fp64', 1: 'fp32c', 2: 'fp32c-uniform', 3: 'fp64-uniform'}[f.value]

    def set_tuning(self, pair_threads, list_len=None, persistent=None):
        call('sph_batch_set_tuning', self._b, int(pair_threads))
        if list_len is not None:
            call('sph_batch_set_list_len', self._b, int(list_len))
        if persistent is not None:
            call('sph_batch_set_persistent', self._b, int(bool(persistent)))

    def set_border_bounds(self, on):
        """Whether each instance's fused PEC step takes the grid's bounds
        from its border cells when its carried displacement bound allows (the
        default) or scans all its particles; the grids are the same."""
        call('sph_batch_set_border_bounds', self._b, int(bool(on)))

    def border_count(self):
        """Per instance: neighbour updates that took the border-cell bounds."""
        v = (C.c_int64 * self.K)()
        call('sph_batch_border_count', self._b, v, self._stream())
        return [int(x) for x in v]

    def timing_enable(self, on=True):
        call('sph_batch_timing_enable', self._b, int(bool(on)))

    def timing_read(self, reset=True):
        """{'pair_ms': mean ms of the pair pass + corrector over the timed
        eager steps, 'count': how many}."""
        p = C.c_double()
        k = C.c_int64()
        call('sph_batch_timing_read', self._b, C.byref(p), C.byref(k), int(bool(reset)))
        return {'pair_ms': p.value, 'count': int(k.value)}

    def close(self):
        if self._b:
            _lib.load().sph_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(batch, candidates=(64, 128, 256), lists=(16, 24, 32), schedules=(False, True), steps=4):
    """Pick the pair pass's workgroup size, contributor-list length and
    schedule (persistent work-queue kernel or not) by timing each combination
    over a few eager steps of the batch itself (advances the batch by
    len(candidates) * len(lists) * len(schedules) * (steps + 1) steps; no
    choice changes a result); returns ((threads, list_len, persistent), table
    of (threads, list_len, persistent, ms))."""
    table = []
    if batch.pair_form().endswith('uniform'):  # one launch shape: time the schedules only
        candidates, lists = candidates[-1:], lists[-1:]
    for threads in candidates:
        for ll in lists:
            for per in schedules:
                batch.set_tuning(threads, ll, per)
                batch.step(1, graph=False)
                batch.timing_enable(True)
                batch.step(steps, graph=False)
                ms = batch.timing_read()['pair_ms']
                batch.timing_enable(False)
                table.append((threads, ll, bool(per), ms))
    best = min(table, key=lambda r: r[3])
    batch.set_tuning(best[0], best[1], best[2])
    return (best[0], best[1], best[2]), table


def params_for(scenario, alpha=None, **kw):
    """EngineParams of one ensemble member (WCSPH constants of its column height)."""
    m = scenario.method()
    if alpha is not None:
        m.alpha = alpha
    return EngineParams.from_method(scenario.n, m, **kw)


__all__ = ['BatchEngine', 'autotune', 'params_for']
