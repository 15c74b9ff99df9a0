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


def diverging(sc, centres, speed=5.0, rho_low=0.2):
    """Fluid around each centre moves radially outwards at `speed` and the
    centre particle has a density of `rho_low`: continuity then gives it
    dt * drho/dt far below -rho_low in the first step."""
    pA = sc.pA.copy()
    f = np.nonzero(pA['label'] == 0)[0]
    for c in centres:
        d = np.hypot(pA['x'][f] - pA['x'][c], pA['y'][f] - pA['y'][c])
        near = f[(d > 0) & (d < 3 * sc.h)]
        rx = pA['x'][near] - pA['x'][c]
        ry = pA['y'][near] - pA['y'][c]
        r = np.hypot(rx, ry)
        pA['vx'][near] = speed * rx / r
        pA['vy'][near] = speed * ry / r
        pA['rho'][c] = rho_low
    return pA
