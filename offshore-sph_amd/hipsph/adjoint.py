"""Differentiable WCSPH runs: checkpointed forward, hand-written adjoint.

``SPHRun`` advances a fixed set of particles ``nsteps`` steps on the device
engine and is differentiable in the artificial-viscosity ``alpha``, the sound
speed ``c0`` (with B = rho0 c0^2 / gamma kept consistent) and the initial
fluid velocity field. The gradient comes from the adjoint kernels of
libhipsph (``sph_step_record`` / ``sph_adjoint_step``), not from autograd:

* forward: the state is saved every ``checkpoint_every`` steps (device
  tensors), the steps themselves run as captured device steps;
* backward: for each segment, last first, the saved state is restored, the
  segment is recomputed while recording a tape (bit-identical to the forward:
  the engine is deterministic), and the recorded steps are reversed.

``simulate(run, alpha, c0, v0)`` wraps this as a ``torch.autograd.Function``
so that a loss written in torch on the final positions back-propagates into
``alpha``, ``c0`` and ``v0``. Torch does no SPH arithmetic here.

Everything not a parameter is held fixed: positions, densities, masses and h
of the initial state, the walls, gamma, rho0, beta and the integrator. The
dynamic dt = min(gamma_c h_min / c0, gamma_f sqrt(h_min / max |a|^2)) is
differentiated on both branches; the force branch routes its sensitivity to
the particle attaining the maximum (smallest index on ties), one valid
subgradient where several particles tie (``SPHRun.force_steps`` counts the
force-limited steps). With ``fixed_dt`` the step size is a constant.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import call
from .device import Engine, EngineParams, soa_from_particles

ADJ_FIELDS = ('x', 'y', 'vx', 'vy', 'rho', 'ax', 'ay', 'drho')


class Tape:
    """``nslots`` recorded steps of one engine (C ABI sph_tape_t)."""

    def __init__(self, engine: Engine, nslots: int):
        self.engine = engine
        self.nslots = int(nslots)
        self._t = C.c_void_p()
        call('sph_tape_create', engine._ctx, self.nslots, C.byref(self._t))

    def record(self, slot):
        call('sph_step_record', self.engine._ctx, self._t, int(slot), self.engine._stream())

    @staticmethod
    def _adj(adj, g_alpha, g_co):
        a = _lib.sph_adjoint_t()
        for f in ADJ_FIELDS:
            setattr(a, f, adj[f].data_ptr())
        a.g_alpha = g_alpha.data_ptr()
        a.g_co = g_co.data_ptr()
        return a

    def adjoint(self, slot, adj, g_alpha, g_co):
        a = self._adj(adj, g_alpha, g_co)
        call('sph_adjoint_step', self.engine._ctx, self._t, int(slot), C.byref(a), self.engine._stream())

    def segment(self, nsteps, adj, g_alpha, g_co):
        """Record nsteps from the current state and reverse them (hipGraph replays)."""
        a = self._adj(adj, g_alpha, g_co)
        call('sph_adjoint_segment', self.engine._ctx, self._t, int(nsteps), C.byref(a), self.engine._stream())

    def force_steps(self):
        out = C.c_int64()
        call('sph_tape_force_steps', self._t, C.byref(out))
        return int(out.value)

    def close(self):
        if self._t:
            _lib.load().sph_tape_destroy(self._t)
            self._t = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SPHRun:
    """nsteps device steps from the particle records ``pA`` (active set)."""

    def __init__(self, pA, method, nsteps, checkpoint_every=20, kernel='cubic', precision='fp64',
                 integrator='pec', device='cuda:0', fixed_dt=None, strict=False):
        import torch
        self.pA = np.ascontiguousarray(pA).copy()
        self.n = len(self.pA)
        self.method = method
        self.nsteps = int(nsteps)
        self.k = max(1, int(checkpoint_every))
        self.params = EngineParams.from_method(self.n, method, kernel=kernel, precision=precision,
                                               integrator=integrator, fixed_dt=float(fixed_dt or 0.0),
                                               strict=bool(strict))
        self.engine = Engine(self.params, device=device)
        self.device = self.engine.device
        fl = self.pA['label'] == 0
        self.fluid = torch.as_tensor(np.nonzero(fl)[0], device=self.device)
        self.n_fluid = int(fl.sum())
        self._base = {k: torch.as_tensor(v, device=self.device) for k, v in soa_from_particles(self.pA).items()}
        self._tape = None
        self._ckpt = None
        self.force_steps = 0
        # adjoint buffers kept for the run's lifetime (the reverse graphs bind them)
        self._adjbuf = {f: torch.zeros(self.n, dtype=torch.float64, device=self.device) for f in ADJ_FIELDS}
        self._gbuf = torch.zeros(2, dtype=torch.float64, device=self.device)

    # -- plumbing ------------------------------------------------------------
    def _set(self, alpha, co):
        g, rho0 = self.params.gamma, self.params.rho0
        self.engine.set_params(alpha=float(alpha), co=float(co), B=float(co) * float(co) * rho0 / g)

    def _upload_initial(self, v0):
        import torch
        st = {k: v.clone() for k, v in self._base.items()}
        v0 = v0.to(device=self.device, dtype=torch.float64)
        st['vx'][self.fluid] = v0[:, 0]
        st['vy'][self.fluid] = v0[:, 1]
        self.engine.restore(st, 0.0)

    def _save(self):
        d = self.engine.download(to_numpy=False)
        d['_t'] = self.engine.status()['t']
        return d

    def _restore(self, ck):
        self.engine.restore({k: v for k, v in ck.items() if not k.startswith('_')}, ck['_t'])

    # -- forward / backward ----------------------------------------------------------
    def forward(self, alpha, co, v0, keep=True):
        """Final fluid positions [n_fluid, 2] (device float64)."""
        import torch
        self._set(alpha, co)
        self._upload_initial(v0)
        ck = []
        done = 0
        while done < self.nsteps:
            if keep:
                ck.append(self._save())
            m = min(self.k, self.nsteps - done)
            self.engine.step(m, graph=True)
            done += m
        self._ckpt = ck if keep else None
        fin = self.engine.download(to_numpy=False)
        return torch.stack((fin['x'][self.fluid], fin['y'][self.fluid]), 1)

    def backward(self, g_xy, g_state=None, graphs=True):
        """Gradients (dL/dalpha, dL/dc0, dL/dv0 [n_fluid, 2]) for dL/d(final xy) = g_xy.

        g_state optionally adds dL/d(final field) for any of x, y, vx, vy, rho
        (full-length device arrays); the returned dict ``self.last_adjoint``
        then holds dL/d(initial field) for every field."""
        import torch
        if self._ckpt is None:
            raise RuntimeError('backward needs a forward with keep=True')
        dev = self.device
        adj = self._adjbuf
        for v in adj.values():
            v.zero_()
        g_xy = g_xy.to(device=dev, dtype=torch.float64)
        adj['x'][self.fluid] += g_xy[:, 0]
        adj['y'][self.fluid] += g_xy[:, 1]
        for f, v in (g_state or {}).items():
            adj[f] += v.to(device=dev, dtype=torch.float64)
        self._gbuf.zero_()
        ga, gc = self._gbuf[0:1], self._gbuf[1:2]
        if self._tape is None:
            self._tape = Tape(self.engine, self.k)
        before = self._tape.force_steps()
        for si in reversed(range(len(self._ckpt))):
            self._restore(self._ckpt[si])
            m = min(self.k, self.nsteps - si * self.k)
            if graphs:
                self._tape.segment(m, adj, ga, gc)
            else:  # one launch per kernel (same kernels, same order)
                for s in range(m):
                    self._tape.record(s)
                for s in reversed(range(m)):
                    self._tape.adjoint(s, adj, ga, gc)
        self.force_steps += self._tape.force_steps() - before
        self.last_adjoint = {k: v.clone() for k, v in adj.items()}
        gv = torch.stack((adj['vx'][self.fluid], adj['vy'][self.fluid]), 1)
        return ga[0].clone(), gc[0].clone(), gv

    def close(self):
        if self._tape is not None:
            self._tape.close()
            self._tape = None
        self.engine.close()


def _fn():
    import torch

    class _Run(torch.autograd.Function):
        @staticmethod
        def forward(ctx, alpha, co, v0, run):
            ctx.run = run
            ctx.devs = (alpha.device, co.device, v0.device)
            return run.forward(float(alpha), float(co), v0.detach())

        @staticmethod
        def backward(ctx, g):
            ga, gc, gv = ctx.run.backward(g.contiguous())
            da, dc, dv = ctx.devs
            return ga.to(da).reshape(()), gc.to(dc).reshape(()), gv.to(dv), None

    return _Run


_RUN = None


def simulate(run: SPHRun, alpha, co, v0):
    """Final fluid positions as a torch tensor differentiable in alpha, co, v0."""
    global _RUN
    if _RUN is None:
        _RUN = _fn()
    return _RUN.apply(alpha, co, v0, run)


__all__ = ['Tape', 'SPHRun', 'simulate']
