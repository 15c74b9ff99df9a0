"""Helpers shared by the GPU tests: engines built from particle arrays, oracle runs."""
import sys

import numpy as np

import oracle
from hipsph.device import Engine, EngineParams, soa_from_particles

FIELDS = ('x', 'y', 'vx', 'vy', 'rho', 'ax', 'ay', 'drho')


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_engine(pA, method, kernel='cubic', precision='fp64', integrator='pec', **kw):
    p = EngineParams.from_method(len(pA), method, kernel=kernel, precision=precision, integrator=integrator, **kw)
    e = Engine(p)
    e.upload(soa_from_particles(pA))
    return e


def perturbed(pA, seed, vscale=0.3, rscale=0.004):
    """Random velocities and densities on the fluid so that every branch (AV, walls) is taken."""
    rng = np.random.default_rng(seed)
    a = pA.copy()
    f = a['label'] == 0
    a['vx'][f] = rng.normal(0, vscale, f.sum())
    a['vy'][f] = rng.normal(0, vscale, f.sum())
    a['rho'][f] = a['rho'][f] * (1 + rng.uniform(-rscale, rscale, f.sum()))
    return a


def oracle_loop(pA, method, kernel='cubic', nthreads=8):
    a = oracle.as_particles(pA.copy())
    nn = oracle.loop(a, method, kernel, nthreads=nthreads)
    return a, nn


def rel_err(got, want, mask=None):
    if mask is not None:
        got, want = got[mask], want[mask]
    scale = np.max(np.abs(want))
    return float(np.max(np.abs(got - want)) / scale) if scale > 0 else float(np.max(np.abs(got)))
