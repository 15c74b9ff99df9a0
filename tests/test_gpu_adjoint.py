"""Reverse mode of the device step (sph_step_record / sph_adjoint_step).

Checks, all FP64 unless stated:
  * directional derivatives of a linear functional of the final state
    (positions, velocities, densities) with respect to alpha, c0, the initial
    velocities, positions and densities, against central finite differences
    of the device forward, across checkpoint segments;
  * the checkpoint interval does not change the gradient (bit-identical);
  * PEC's strict clamp: bit-identical to strict=False where it never acts,
    and against central differences on a step where it clamps densities;
  * the 2,000-particle differentiable dam break (200 steps) against the
    gradient of the FP64 dense-pair torch model (tests/torch_ref.py, itself
    tied to the CPU oracle), whose final state also matches the device's;
  * the 20,000-particle differentiable dam break (200 steps, checkpoints every
    20) against central finite differences with eps = 1e-6 relative; its
    FP32-pair forward gives the same gradient to the FP32 pair error.
Loss of the dam-break cases: mean over fluid particles of the squared
distance between final positions and those of a target run at alpha = 0.03
and c0 x 1.1.
"""
import numpy as np
import pytest
import torch

from gpu_util import say
from hipsph import scenarios
from hipsph.adjoint import SPHRun, simulate

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def _linear_loss_setup(run, seed):
    rng = np.random.default_rng(seed)
    dev = run.device
    f = run.fluid
    n = run.n
    w = {}
    for k in ('x', 'y', 'vx', 'vy', 'rho'):
        a = torch.zeros(n, dtype=torch.float64, device=dev)
        a[f] = torch.as_tensor(rng.normal(0, 1, run.n_fluid), device=dev)
        w[k] = a
    if 'rho' in w:
        w['rho'] *= 1e-3  # densities are ~1000: keep the terms comparable
    return w


def _linear_loss(run, w, alpha, co, v0):
    run.forward(alpha, co, v0, keep=False)
    d = run.engine.download(to_numpy=False)
    return float(sum(torch.sum(w[k] * d[k]) for k in w))


def test_adjoint_directional_derivatives_vs_fd():
    sc = scenarios.dam_break(0.4, 0.5, 0.02, tank_w=1.0, tank_h=0.8, name='adj_small')
    m = sc.method()
    rng = np.random.default_rng(3)
    f = sc.pA['label'] == 0
    v0 = torch.as_tensor(rng.normal(0, 0.3, (int(f.sum()), 2)), device='cuda')  # exercises the viscosity branch
    run = SPHRun(sc.pA, m, nsteps=7, checkpoint_every=3)
    w = _linear_loss_setup(run, 5)
    alpha, co = m.alpha, m.co
    run.forward(alpha, co, v0)
    g_xy = torch.stack((w['x'][run.fluid], w['y'][run.fluid]), 1)
    ga, gc, gv = run.backward(g_xy, {'vx': w['vx'], 'vy': w['vy'], 'rho': w['rho']})
    adj = {k: v.clone() for k, v in run.last_adjoint.items()}
    say('adjoint small: dL/dalpha', float(ga), 'dL/dc0', float(gc), 'force-criterion steps', run.force_steps)

    def L(a=alpha, c=co, v=v0):
        return _linear_loss(run, w, a, c, v)

    eps = 1e-6
    fd_a = (L(a=alpha * (1 + eps)) - L(a=alpha * (1 - eps))) / (2 * eps * alpha)
    fd_c = (L(c=co * (1 + eps)) - L(c=co * (1 - eps))) / (2 * eps * co)
    dv = torch.as_tensor(rng.normal(0, 1, tuple(v0.shape)), device='cuda')
    ev = 1e-6
    fd_v = (L(v=v0 + ev * dv) - L(v=v0 - ev * dv)) / (2 * ev)
    say('alpha', float(ga), fd_a, 'c0', float(gc), fd_c, 'v0', float(torch.sum(gv * dv)), fd_v)
    assert _rel(float(ga), fd_a) < 1e-4
    assert _rel(float(gc), fd_c) < 1e-4
    assert _rel(float(torch.sum(gv * dv)), fd_v) < 1e-4
    # initial positions and densities of the fluid (through the saved initial state)
    for k, scale in (('x', 1e-7), ('rho', 1e-4)):
        d = torch.zeros(run.n, dtype=torch.float64, device='cuda')
        d[run.fluid] = torch.as_tensor(rng.normal(0, 1, run.n_fluid), device='cuda')
        base = run._base[k].clone()
        run._base[k] = base + scale * d
        lp = L()
        run._base[k] = base - scale * d
        lm = L()
        run._base[k] = base
        fd = (lp - lm) / (2 * scale)
        an = float(torch.sum(adj[k] * d))
        say('initial', k, an, fd)
        assert _rel(an, fd) < 1e-4, k
    run.close()


def test_adjoint_strict_clamp():
    """PEC(strict=True), the reference's default integrator setting.
    (1) No density near zero: the clamp never acts, and the gradient (and the
    adjoint of every initial field) is bit-identical to strict=False.
    (2) One step whose corrector clamps two densities to exactly zero (fluid
    streaming away from two low-density centres): the gradient of a linear
    functional of the final densities in c0 and the initial velocities
    against central differences (to 1e-5), with the clamp and, on the same
    state, without it (strict=False, where those densities come out
    negative). A clamped density is 0 whatever the inputs, so it passes no
    adjoint back; the two cases' c0 derivatives differ 30-fold (the centres'
    dt drho/dt term through the Courant dt), so the check sees the clamp."""
    from gpu_util import diverging
    sc = scenarios.dam_break(0.4, 0.5, 0.02, tank_w=1.0, tank_h=0.8, name='adj_small')
    m = sc.method()
    rng = np.random.default_rng(3)
    f = sc.pA['label'] == 0
    v0 = torch.as_tensor(rng.normal(0, 0.3, (int(f.sum()), 2)), device='cuda')
    got = []
    for strict in (False, True):
        run = SPHRun(sc.pA, m, nsteps=7, checkpoint_every=3, strict=strict)
        w = _linear_loss_setup(run, 5)
        run.forward(m.alpha, m.co, v0)
        g_xy = torch.stack((w['x'][run.fluid], w['y'][run.fluid]), 1)
        ga, gc, gv = run.backward(g_xy, {'vx': w['vx'], 'vy': w['vy'], 'rho': w['rho']})
        got.append((float(ga), float(gc), gv.cpu().numpy(), {k: v.cpu().numpy() for k, v in run.last_adjoint.items()}))
        run.close()
    assert got[0][0] == got[1][0] and got[0][1] == got[1][1]
    np.testing.assert_array_equal(got[0][2], got[1][2])
    for k in got[0][3]:
        np.testing.assert_array_equal(got[0][3][k], got[1][3][k], err_msg=k)

    fl = np.nonzero(f)[0]
    centres = [int(fl[len(fl) // 3]), int(fl[2 * len(fl) // 3])]
    pA = diverging(sc, centres)
    v0 = torch.as_tensor(np.stack((pA['vx'][f], pA['vy'][f]), 1), device='cuda')
    alpha, co = m.alpha, m.co
    dv = torch.as_tensor(rng.normal(0, 1, tuple(v0.shape)), device='cuda')
    res = {}
    for strict in (True, False):
        run = SPHRun(pA, m, nsteps=1, checkpoint_every=1, strict=strict)
        # the final densities, the centres' weighted like the rest
        w = _linear_loss_setup(run, 7)
        w = {'rho': w['rho']}
        run.forward(alpha, co, v0)
        fin = run.engine.download()
        clamped = bool(np.all(fin['rho'][centres] == 0.0))
        ga, gc, gv = run.backward(torch.zeros((run.n_fluid, 2), dtype=torch.float64, device='cuda'), w)

        def L(c=co, v=v0):
            return _linear_loss(run, w, alpha, c, v)

        eps = 1e-6
        fd_c = (L(c=co * (1 + eps)) - L(c=co * (1 - eps))) / (2 * eps * co)
        fd_v = (L(v=v0 + eps * dv) - L(v=v0 - eps * dv)) / (2 * eps)
        an_v = float(torch.sum(gv * dv))
        say('strict', strict, 'clamped', clamped, 'c0', float(gc), fd_c, 'v0', an_v, fd_v, 'alpha', float(ga))
        assert clamped == strict  # the corrector's clamp took the centres
        assert float(ga) == 0.0  # alpha does not reach a density in one step
        assert _rel(float(gc), fd_c) < 1e-5
        assert _rel(an_v, fd_v) < 1e-5
        res[strict] = (float(gc), an_v)
        run.close()
    # the clamp matters here: without it the centres' dt drho/dt terms
    # change both derivatives
    assert _rel(res[True][0], res[False][0]) > 1.0 and _rel(res[True][1], res[False][1]) > 1e-3


def _target(sc, nsteps, fixed_dt=None):
    m = sc.method()
    run = SPHRun(sc.pA, m, nsteps, checkpoint_every=nsteps, fixed_dt=fixed_dt)
    v0 = torch.zeros((sc.n_fluid, 2), dtype=torch.float64, device='cuda')
    xy = run.forward(0.03, 1.1 * m.co, v0, keep=False).clone()
    run.close()
    return xy


def _mse(xy, target):
    return torch.mean(torch.sum((xy - target) ** 2, 1))


def test_adjoint_checkpoint_interval_is_exact():
    sc = scenarios.dam_break_c5(small=True)
    m = sc.method()
    target = _target(sc, 40)
    v0 = torch.zeros((sc.n_fluid, 2), dtype=torch.float64, device='cuda')
    grads = []
    for k in (40, 7):
        run = SPHRun(sc.pA, m, 40, checkpoint_every=k)
        a = torch.tensor(m.alpha, dtype=torch.float64, requires_grad=True)
        c = torch.tensor(m.co, dtype=torch.float64, requires_grad=True)
        v = v0.clone().requires_grad_(True)
        _mse(simulate(run, a, c, v), target).backward()
        grads.append((float(a.grad), float(c.grad), v.grad.clone()))
        run.close()
    assert grads[0][0] == grads[1][0] and grads[0][1] == grads[1][1]
    assert torch.equal(grads[0][2], grads[1][2])


def test_tape_made_before_the_upload_takes_the_labels():
    """A tape made before any state is on the device holds no fluid mask; a
    recording from slot 0 takes the labels of the state it records, so the
    gradient is the one of a tape made afterwards, bit for bit (walls and
    fluid told apart), with and without hipGraph segments."""
    from hipsph.adjoint import Tape
    sc = scenarios.dam_break_c5(small=True)
    m = sc.method()
    target = _target(sc, 12)
    v0 = torch.zeros((sc.n_fluid, 2), dtype=torch.float64, device='cuda')
    out = []
    for early in (False, True, True):
        run = SPHRun(sc.pA, m, 12, checkpoint_every=6)
        if early:
            run._tape = Tape(run.engine, run.k)
        xt = run.forward(m.alpha, m.co, v0).clone().requires_grad_(True)
        (g,) = torch.autograd.grad(_mse(xt, target), [xt])
        ga, gc, gv = run.backward(g, graphs=len(out) != 2)
        out.append((float(ga), float(gc), gv.clone()))
        run.close()
    for o in out[1:]:
        assert o[0] == out[0][0] and o[1] == out[0][1]
        assert torch.equal(o[2], out[0][2])


def test_adjoint_graph_segments_match_eager_steps():
    """sph_adjoint_segment (hipGraph replays) gives the same bits as
    sph_step_record / sph_adjoint_step launched one by one."""
    sc = scenarios.dam_break_c5(small=True)
    m = sc.method()
    target = _target(sc, 30)
    run = SPHRun(sc.pA, m, 30, checkpoint_every=8)
    v0 = torch.zeros((sc.n_fluid, 2), dtype=torch.float64, device='cuda')
    out = []
    for graphs in (True, False, True):
        xt = run.forward(m.alpha, m.co, v0).clone().requires_grad_(True)
        (g,) = torch.autograd.grad(_mse(xt, target), [xt])
        ga, gc, gv = run.backward(g, graphs=graphs)
        out.append((float(ga), float(gc), gv.clone()))
    run.close()
    for o in out[1:]:
        assert o[0] == out[0][0] and o[1] == out[0][1]
        assert torch.equal(o[2], out[0][2])


def test_adjoint_reused_run_after_new_constants():
    """A run differentiated at (alpha1, c1) and then run forward and backward
    at (alpha2, c2) gives the gradient of a fresh run at (alpha2, c2) bit for
    bit, with graph segments and with one launch per kernel: the segment
    graphs cached in the tape at the first constants are not replayed."""
    sc = scenarios.dam_break_c5(small=True)
    m = sc.method()
    target = _target(sc, 24)
    v0 = torch.zeros((sc.n_fluid, 2), dtype=torch.float64, device='cuda')
    a1, c1 = m.alpha, m.co
    a2, c2 = 0.037, m.co * 1.07

    def grads(run, alpha, co, graphs):
        xt = run.forward(alpha, co, v0).clone().requires_grad_(True)
        (g,) = torch.autograd.grad(_mse(xt, target), [xt])
        ga, gc, gv = run.backward(g, graphs=graphs)
        return float(ga), float(gc), gv.clone()

    fresh = SPHRun(sc.pA, m, 24, checkpoint_every=8)
    ref = grads(fresh, a2, c2, True)
    ref_eager = grads(fresh, a2, c2, False)
    fresh.close()
    reused = SPHRun(sc.pA, m, 24, checkpoint_every=8)
    first = grads(reused, a1, c1, True)
    second = grads(reused, a2, c2, True)
    second_eager = grads(reused, a2, c2, False)
    reused.close()
    say('reused run: alpha grads', first[0], second[0], ref[0], 'c0 grads', first[1], second[1], ref[1])
    assert first[0] != ref[0]  # the two constant sets give different gradients
    for o in (ref_eager, second, second_eager):
        assert o[0] == ref[0] and o[1] == ref[1]
        assert torch.equal(o[2], ref[2])


@pytest.mark.parametrize('dt_mode', ['fixed', 'dynamic'])
def test_adjoint_dam_break_2k_vs_torch_dense(dt_mode):
    """Fixed dt from a generic initial velocity field (v0 ~ N(0, 0.05 m/s)):
    the loss is smooth there and every gradient must match the torch model's.
    Dynamic dt from rest (the differentiable dam break as specified): most
    steps are force-limited and at some the maximum |a|^2 is attained by
    several particles at once (mirror-identical neighbourhoods); and from rest
    many pairs sit exactly on the viscosity switch (v.r = 0, where the
    viscous term has a kink). The loss is not differentiable in v0 there, and
    torch (even split among ties) and the engine (smallest index) return
    different subgradients; alpha and c0 must still match, v0 is reported."""
    from torch_ref import DenseWCSPH
    sc = scenarios.dam_break_c5(small=True)
    m = sc.method()
    K = 200
    fixed = 0.25 * sc.h / m.co if dt_mode == 'fixed' else None
    target = _target(sc, K, fixed)
    run = SPHRun(sc.pA, m, K, checkpoint_every=20, fixed_dt=fixed)
    a = torch.tensor(m.alpha, dtype=torch.float64, requires_grad=True)
    c = torch.tensor(m.co, dtype=torch.float64, requires_grad=True)
    v_init = torch.zeros((sc.n_fluid, 2), dtype=torch.float64)
    if dt_mode == 'fixed':
        v_init = torch.as_tensor(np.random.default_rng(21).normal(0, 0.05, (sc.n_fluid, 2)))
    v = v_init.cuda().requires_grad_(True)
    xy = simulate(run, a, c, v)
    loss = _mse(xy, target)
    loss.backward()
    say('2k', dt_mode, 'force-criterion steps', run.force_steps, 'of', K)
    run.close()
    # reference: torch autograd through the dense-pair model on the CPU
    model = DenseWCSPH(sc.pA, m, sc.h)
    ra = torch.tensor(m.alpha, dtype=torch.float64, requires_grad=True)
    rc = torch.tensor(m.co, dtype=torch.float64, requires_grad=True)
    rv = v_init.clone().requires_grad_(True)
    x, y, *_ = model.run(ra, rc, rv[:, 0], rv[:, 1], K, fixed_dt=fixed)
    rxy = torch.stack((x, y), 1)
    np.testing.assert_allclose(rxy.detach().numpy(), xy.detach().cpu().numpy(), rtol=0, atol=1e-10)
    rloss = _mse(rxy, target.cpu())
    rloss.backward()
    gv, rgv = v.grad.cpu().numpy(), rv.grad.numpy()
    dv = float(np.linalg.norm(gv - rgv) / np.linalg.norm(rgv))
    say('2k', dt_mode, 'loss', float(loss.detach()), float(rloss.detach()), 'alpha', float(a.grad), float(ra.grad),
        'c0', float(c.grad), float(rc.grad), 'v0 rel diff', dv)
    assert _rel(float(a.grad), float(ra.grad)) < 1e-6
    assert _rel(float(c.grad), float(rc.grad)) < 1e-6
    if dt_mode == 'fixed':
        assert dv < 1e-6


def _grads_20k(sc, target, precision, K, fixed=None):
    m = sc.method()
    run = SPHRun(sc.pA, m, K, checkpoint_every=20, precision=precision, fixed_dt=fixed)
    v0 = torch.zeros((sc.n_fluid, 2), dtype=torch.float64, device='cuda')
    xt = run.forward(m.alpha, m.co, v0).clone().requires_grad_(True)
    loss = _mse(xt, target)
    (g_xy,) = torch.autograd.grad(loss, [xt])
    ga, gc, gv = run.backward(g_xy)
    say('20k', precision, 'force-criterion steps', run.force_steps, 'of', K)
    return run, float(loss), float(ga), float(gc), gv


@pytest.mark.parametrize('dt_mode', ['dynamic', 'fixed'])
def test_adjoint_dam_break_20k_vs_fd_and_fp32_pairs(dt_mode):
    """Central differences, eps = 1e-6 relative: alpha and c0 in both dt modes;
    the initial-velocity direction with fixed dt (smooth loss; see the 2k test
    for the dynamic case). The FP32-pair forward's gradient matches FP64's."""
    sc = scenarios.dam_break_c5()
    m = sc.method()
    K = 200
    fixed = 0.25 * sc.h / m.co if dt_mode == 'fixed' else None
    target = _target(sc, K, fixed)
    run, loss, ga, gc, gv = _grads_20k(sc, target, 'fp64', K, fixed)
    v0 = torch.zeros((sc.n_fluid, 2), dtype=torch.float64, device='cuda')

    def L(a=m.alpha, c=m.co, v=v0):
        return float(_mse(run.forward(a, c, v, keep=False), target))

    eps = 1e-6
    fd_a = (L(a=m.alpha * (1 + eps)) - L(a=m.alpha * (1 - eps))) / (2 * eps * m.alpha)
    fd_c = (L(c=m.co * (1 + eps)) - L(c=m.co * (1 - eps))) / (2 * eps * m.co)
    rng = np.random.default_rng(9)
    dv = torch.as_tensor(rng.normal(0, 1, tuple(v0.shape)), device='cuda')
    ev = 1e-6 * np.sqrt(9.81 * sc.height)
    fd_v = (L(v=v0 + ev * dv) - L(v=v0 - ev * dv)) / (2 * ev)
    an_v = float(torch.sum(gv * dv))
    say('20k fp64', dt_mode, 'loss', loss, 'alpha', ga, fd_a, 'c0', gc, fd_c, 'v0', an_v, fd_v)
    assert _rel(ga, fd_a) < 1e-3
    assert _rel(gc, fd_c) < 1e-3
    if dt_mode == 'fixed':
        assert _rel(an_v, fd_v) < 1e-3
    run.close()
    # FP32-pair forward (FP64 integrator), adjoint on its own trajectory
    run32, loss32, ga32, gc32, gv32 = _grads_20k(sc, target, 'fp32c', K, fixed)
    rv = float(torch.linalg.norm(gv32 - gv) / torch.linalg.norm(gv))
    say('20k fp32c', dt_mode, 'loss', loss32, 'alpha', ga32, 'c0', gc32, 'v0 rel diff', rv)
    assert _rel(loss32, loss) < 1e-3
    assert _rel(ga32, ga) < 1e-3 and _rel(gc32, gc) < 1e-3
    if dt_mode == 'fixed':
        assert rv < 1e-3
    run32.close()
