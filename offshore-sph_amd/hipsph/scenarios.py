"""Synthetic inputs: dam breaks and the pair-kernel stress lattice.

Geometry follows the reference's own builder, ``Helpers.rect``
(src/Helpers.py:7-87), and its dam-break example (examples/DamBreak.py:16-72):

* fluid column: ``Helpers.rect(0, W, 0, H, dx)`` with mass rho0 * dx^2 and a
  hydrostatic initial density (WCSPH.initialize, src/Methods/WCSPH.py:82-108);
* tank walls: ``layers`` rows of fixed Boundary particles at spacing dx
  outside the tank [0, tank_w] x [0, tank_h]: bottom rows at y = -k dx
  spanning x in [-layers dx, tank_w + layers dx], side columns at
  x = -k dx and x = tank_w + k dx spanning y in [0, tank_h];
* h = 1.3 dx for every particle, walls included (Helpers.rect leaves h = 0,
  which would make the reference's cell size fall back to 1 m;
  src/Tools/NNLinkedList.py:103-110).

``Scenario`` carries the particle array (reference particle_dtype) and the
WCSPH height H the method is built with.
"""
from dataclasses import dataclass, field

import numpy as np

from .Common import ParticleType, particle_dtype
from .Helpers import Helpers
from .Methods.WCSPH import WCSPH

SEED = 20261202


@dataclass
class Scenario:
    name: str
    pA: np.ndarray
    height: float
    dx: float
    rho0: float = 1000.0
    h: float = 0.0
    meta: dict = field(default_factory=dict)

    @property
    def n(self):
        return len(self.pA)

    @property
    def n_fluid(self):
        return int(np.sum(self.pA['label'] == ParticleType.Fluid))

    def method(self, **kw):
        m = WCSPH(height=self.height, r0=self.dx, rho0=self.rho0, useXSPH=False, Pb=0,
                  useSummationDensity=False)
        for k, v in kw.items():
            setattr(m, k, v)
        if 'co' in kw and 'B' not in kw:
            m.B = m.co * m.co * m.rho0 / m.gamma
        return m


def tank_walls(tank_w, tank_h, dx, layers=3, rho0=1000.0, top=False):
    parts = []
    mass = rho0 * dx * dx
    for k in range(1, layers + 1):
        parts.append(Helpers.rect(-layers * dx, tank_w + layers * dx, -k * dx, -k * dx, dx, mass=mass, rho0=rho0,
                                  label=ParticleType.Boundary))
        if top:
            parts.append(Helpers.rect(-layers * dx, tank_w + layers * dx, tank_h + k * dx, tank_h + k * dx, dx,
                                      mass=mass, rho0=rho0, label=ParticleType.Boundary))
        parts.append(Helpers.rect(-k * dx, -k * dx, 0.0, tank_h, dx, mass=mass, rho0=rho0,
                                  label=ParticleType.Boundary))
        parts.append(Helpers.rect(tank_w + k * dx, tank_w + k * dx, 0.0, tank_h, dx, mass=mass, rho0=rho0,
                                  label=ParticleType.Boundary))
    return np.concatenate(parts)


def dam_break(width=1.0, height=2.0, dx=0.02, tank_w=4.0, tank_h=3.0, layers=3, rho0=1000.0, hfac=1.3,
              name=None):
    """Dam-break column of width x height in a tank_w x tank_h tank."""
    fluid = Helpers.rect(0.0, width, 0.0, height, dx, mass=rho0 * dx * dx, rho0=rho0)
    walls = tank_walls(tank_w, tank_h, dx, layers, rho0)
    pA = np.concatenate((fluid, walls))
    h = hfac * dx
    pA['h'] = h
    sc = Scenario(name or f'dam_break_{width}x{height}_dx{dx}', pA, height, dx, rho0, h,
                  meta={'width': width, 'height': height, 'tank_w': tank_w, 'tank_h': tank_h, 'layers': layers})
    m = sc.method()
    f = pA['label'] == ParticleType.Fluid
    pA['rho'][f] = m.initialize(pA[f])['rho']
    pA['c'][f] = m.co
    return sc


def dam_break_c1():
    """Plain 2-D dam break, dx = 0.02: 1.0 x 2.0 column, 4.0 x 3.0 tank (~5,000 fluid)."""
    return dam_break(1.0, 2.0, 0.02, name='dam_break_dx0.02')


def dam_break_c2(million=True):
    """Scaled dam break, dx = 0.002: 2.0 x 2.5 column (1.25M fluid) or 1.0 x 2.0 (500k)."""
    if million:
        return dam_break(2.0, 2.5, 0.002, name='dam_break_1M_dx0.002')
    return dam_break(1.0, 2.0, 0.002, name='dam_break_500k_dx0.002')


def dam_break_c5(small=False):
    """Differentiable dam break: 0.5 x 1.0 column at dx = 0.005 (20,000 fluid) in a
    2.0 x 1.5 tank; small: 0.4 x 0.5 at dx = 0.01 (2,000 fluid), same tank."""
    if small:
        return dam_break(0.4, 0.5, 0.01, tank_w=2.0, tank_h=1.5, name='dam_break_2k_dx0.01')
    return dam_break(0.5, 1.0, 0.005, tank_w=2.0, tank_h=1.5, name='dam_break_20k_dx0.005')


def ensemble_c4(count=256, seed=SEED, dxs=(0.005, 0.01, 0.02), members=None):
    """Batched ensemble: dam breaks with column width U(0.5, 1.5) m, height
    U(0.5, 2.5) m and dx drawn from dxs in a 4.0 x 3.0 tank; instance k uses
    PCG64(seed + k), which also draws its alpha ~ U(0.005, 0.05). `members`
    (instance numbers) builds only those instances of the ensemble, each as
    it is in the whole one. Returns (scenarios, alphas)."""
    out, alphas = [], []
    for k in (range(count) if members is None else members):
        rng = np.random.Generator(np.random.PCG64(seed + k))
        w = float(rng.uniform(0.5, 1.5))
        hgt = float(rng.uniform(0.5, 2.5))
        dx = float(dxs[int(rng.integers(len(dxs)))])
        alphas.append(float(rng.uniform(0.005, 0.05)))
        out.append(dam_break(w, hgt, dx, name=f'ensemble_{k}'))
    return out, alphas


def dam_break_example(N=50, obstacle=False, rho0=1000.0):
    """The reference's own 2-D dam-break example (examples/DamBreak.py:16-53,
    57-64): a 25 m x 25 m column of N x N hex-packed fluid particles
    (r0 = 25 / N, mass 25^2 / N^2 rho0 each), a bottom row from x = -r0 to 180 m, a left column at x = -r0 and a
    TempBoundary gate at x = 25 + r0, both up to y = 30 m, and optionally
    the example's 2.5 m triangular obstacle at x = 50 m. The walls keep the
    example's zero mass, zero density and zero h (Helpers.rect leaves h at 0),
    so the cell size falls back to 1 m (src/Tools/NNLinkedList.py:103-110).
    The column height is 25 m; the example runs Wendland, XSPH and
    h = 1.6 r0. Returns (r0, particle array) in the example's order: fluid,
    bottom, left, gate[, obstacle]."""
    r0 = 25 / N
    mass = 25 * 25 / N ** 2 * rho0
    lo = -1 * r0
    fluid = Helpers.rect(xmin=0, xmax=25, ymin=0, ymax=25, r0=r0, mass=mass, rho0=rho0, pack=True)
    walls = [Helpers.rect(xmin=lo, xmax=180, ymin=lo, ymax=lo, r0=r0, mass=0.0, rho0=0.0,
                          label=ParticleType.Boundary),
             Helpers.rect(xmin=lo, xmax=lo, ymin=lo, ymax=30, r0=r0, mass=0.0, rho0=0.0,
                          label=ParticleType.Boundary),
             Helpers.rect(xmin=25 - lo, xmax=25 - lo, ymin=lo, ymax=30, r0=r0, mass=0.0, rho0=0.0,
                          label=ParticleType.TempBoundary)]
    if obstacle:
        side = 2.5
        k = int(np.ceil(side / r0))
        run = np.linspace(0.0, side, k)
        ob = np.zeros(3 * k, dtype=particle_dtype)
        ob['x'] = np.concatenate((50 + run, np.full(k, 50 + side), 50 + run))  # diagonal, right side, base
        ob['y'] = np.concatenate((run, run, np.zeros(k)))
        ob['label'] = ParticleType.Boundary
        walls.append(ob)
    return r0, np.concatenate([fluid] + walls)


def stress_lattice(nside=2000, dx=1e-3, H=2.0, layers=3, rho0=1000.0, hfac=1.3, seed=SEED):
    """Pair-kernel stress: nside^2 jittered fluid particles in a closed box.

    Positions i*dx + U(-0.1 dx, 0.1 dx); velocities N(0, (0.05 c0)^2) per
    component with c0 = 10 sqrt(2 g H); densities rho0 (1 + U(-0.01, 0.01)).
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    g = np.arange(nside, dtype=np.float64) * dx
    x, y = np.meshgrid(g, g, indexing='ij')
    x = x.ravel()
    y = y.ravel()
    n = x.size
    fluid = np.zeros(n, dtype=particle_dtype)
    fluid['label'] = ParticleType.Fluid
    fluid['x'] = x + rng.uniform(-0.1 * dx, 0.1 * dx, n)
    fluid['y'] = y + rng.uniform(-0.1 * dx, 0.1 * dx, n)
    co = 10.0 * np.sqrt(2 * 9.81 * H)
    fluid['vx'] = rng.normal(0.0, 0.05 * co, n)
    fluid['vy'] = rng.normal(0.0, 0.05 * co, n)
    fluid['rho'] = rho0 * (1 + rng.uniform(-0.01, 0.01, n))
    fluid['m'] = rho0 * dx * dx
    side = (nside - 1) * dx
    walls = tank_walls(side, side, dx, layers, rho0, top=True)
    pA = np.concatenate((fluid, walls))
    pA['h'] = hfac * dx
    pA['c'][pA['label'] == ParticleType.Fluid] = co
    return Scenario(f'stress_{nside}x{nside}_dx{dx}', pA, H, dx, rho0, hfac * dx, meta={'nside': nside})
