"""Autotuner for the launch shape of the fused pair pass.

The pair pass runs one thread per particle with a per-thread contributor list
in LDS; the workgroup size (64, 128 or 256 threads) trades LDS per workgroup
against scheduling granularity. For a given engine (state uploaded, so the
cell grid is the real one) the tuner times each candidate on the device with
HIP events and keeps the fastest. The result is cached per (GPU name,
precision, kernel, particle-count bucket) in a JSON file, so one workload pays
for the search once per process tree.

Timing: one neighbour update, then `reps` pair passes per shape after one
warm-up pass. The pair pass is a pure function of the cell-ordered state, so
repeating it leaves the state unchanged.
"""
import ctypes as C
import json
import math
import os

from . import _lib

CACHE = os.environ.get('HIPSPH_TUNE_CACHE', os.path.join(os.path.expanduser('~'), '.cache', 'hipsph', 'tune.json'))

CANDIDATES = (64, 128, 256)


def get_tuning(engine):
    out = C.c_int32()
    _lib.call('sph_ctx_get_tuning', engine._ctx, C.byref(out))
    return out.value


def set_tuning(engine, pair_threads):
    _lib.call('sph_ctx_set_tuning', engine._ctx, int(pair_threads))


def time_pair(engine, reps=10):
    """Mean device time (ms) of the pair pass at the current tuning."""
    engine.compute()
    engine.timing_enable(True)
    for _ in range(reps):
        engine.compute()
    t = engine.timing_read()
    engine.timing_enable(False)
    return t['pair_ms']


def _key(engine):
    import torch
    name = torch.cuda.get_device_name(engine.device)
    bucket = int(round(math.log2(max(engine.n, 1)) * 2))
    p = engine.params
    return f'{name}|{p.precision}|{p.kernel}|{bucket}'


def _load_cache():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except Exception:
        return {}


def _save_cache(d):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        tmp = CACHE + '.tmp'
        with open(tmp, 'w') as f:
            json.dump(d, f, indent=1)
        os.replace(tmp, CACHE)
    except OSError:
        pass


def autotune(engine, candidates=None, reps=10, use_cache=True, verbose=False):
    """Pick the fastest pair-pass shape for this engine's workload; returns (best, table)."""
    key = _key(engine)
    cache = _load_cache() if use_cache else {}
    if use_cache and key in cache:
        best = int(cache[key]['best'])
        set_tuning(engine, best)
        return best, cache[key].get('table', [])
    engine.nn_update()
    table = []
    for threads in (candidates or CANDIDATES):
        set_tuning(engine, threads)
        ms = time_pair(engine, reps)
        table.append((threads, ms))
        if verbose:
            print(f'# tune {engine.params.precision} pair_threads={threads}: {ms * 1e3:.1f} us', flush=True)
    best = min(table, key=lambda r: r[1])
    set_tuning(engine, best[0])
    if use_cache:
        cache[key] = {'best': best[0], 'ms': best[1], 'table': table}
        _save_cache(cache)
    return best[0], table
