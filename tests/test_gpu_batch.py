"""Batched ensemble (sph_batch_*): every instance evolves exactly as alone.

Each check runs a batch of small dam breaks of different geometry,
resolution, height (hence c0, B, D) and alpha, and compares every instance
with a single-simulation engine (which is itself bit-exact against the CPU
oracle) on the same initial state:
  * FP64 pair pass: bit-identical state and clock after 40 steps, eager and
    hipGraph replay, with the pair pass one thread per slot and as the
    persistent work-queue kernel;
  * FP32C pair pass: bit-identical as well, batch and single engine taking
    the same form of the float pass;
  * both in the uniform form of the pair pass (every instance has one h and
    one fluid mass) and in the general one (a second fluid mass);
  * an instance whose duration ends early stops while the others go on,
    exactly as a single engine with that duration.
"""
import numpy as np
import pytest

from gpu_util import FIELDS, make_engine, say
from hipsph import scenarios
from hipsph.batch import BatchEngine, autotune, params_for
from hipsph.device import soa_from_particles

pytestmark = pytest.mark.gpu


def _members():
    geo = [(0.4, 0.5, 0.02, 0.01), (0.3, 0.6, 0.02, 0.03), (0.2, 0.2, 0.01, 0.02), (0.5, 0.3, 0.025, 0.005),
           (0.25, 0.4, 0.0125, 0.045)]
    out = []
    for k, (w, hgt, dx, alpha) in enumerate(geo):
        sc = scenarios.dam_break(w, hgt, dx, tank_w=1.0, tank_h=0.8, name=f'member{k}')
        out.append((sc, alpha))
    return out


def _two_masses(pA):
    """A copy with a second fluid mass (the first fluid particle's 0.1 %
    heavier): the general form of the pair pass."""
    pA = pA.copy()
    i = int(np.flatnonzero(pA['label'] == 0)[0])
    pA['m'][i] *= 1.001
    return pA


def _single(sc, alpha, steps, precision, duration=np.inf, pA=None):
    m = sc.method()
    m.alpha = alpha
    e = make_engine(sc.pA if pA is None else pA, m, precision=precision, duration=duration)
    e.step(steps)
    d = e.download()
    st = e.status()
    form = e.pair_form()
    e.close()
    return d, st, form


@pytest.mark.parametrize('precision', ['fp64', 'fp32c'])
@pytest.mark.parametrize('graph', [False, True])
@pytest.mark.parametrize('persistent', [False, True])
@pytest.mark.parametrize('uniform', [True, False])
def test_batch_instances_match_single_engines(precision, graph, persistent, uniform):
    mem = _members()
    K = 40
    states = [sc.pA if uniform else _two_masses(sc.pA) for sc, _ in mem]
    b = BatchEngine([params_for(sc, alpha, precision=precision) for sc, alpha in mem])
    b.set_tuning(256, 0, persistent)
    b.upload([soa_from_particles(pA) for pA in states])
    form = b.pair_form()
    assert form == precision + ('-uniform' if uniform else '')
    b.step(K, graph=graph)
    got = b.download()
    st = b.status()
    b.close()
    say('batch', form, 'graph', graph, 'persistent', persistent, 'sizes', [sc.n for sc, _ in mem])
    for k, (sc, alpha) in enumerate(mem):
        want, wst, wform = _single(sc, alpha, K, precision, pA=states[k])
        assert wform == form
        assert st[k]['steps'] == wst['steps'] == K
        assert st[k]['t'] == wst['t'] and st[k]['dt'] == wst['dt']
        for fld in FIELDS + ('p',):
            np.testing.assert_array_equal(got[k][fld], want[fld], err_msg=f'{k} {fld}')
        # m, h and label never change: the download returns the uploaded ones
        # in reference order (the uniform form keeps them where the upload put them)
        for fld in ('m', 'h', 'label'):
            np.testing.assert_array_equal(got[k][fld], states[k][fld].astype(got[k][fld].dtype), err_msg=f'{k} {fld}')


@pytest.mark.parametrize('precision', ['fp64', 'fp32c'])
@pytest.mark.parametrize('persistent', [False, True])
def test_batch_gaussian_matches_single_engines(precision, persistent):
    """The Gaussian kernel (3h support, the fullest contributor lists) in
    the uniform form of the batch: every instance bit-identical to its
    single engine after 40 eager steps."""
    mem = _members()[:3]
    K = 40
    b = BatchEngine([params_for(sc, alpha, precision=precision, kernel='gaussian') for sc, alpha in mem])
    b.set_tuning(256, 0, persistent)
    b.upload([soa_from_particles(sc.pA) for sc, _ in mem])
    assert b.pair_form() == precision + '-uniform'
    b.step(K)
    got = b.download()
    b.close()
    say('batch gaussian', precision, 'persistent', persistent)
    for k, (sc, alpha) in enumerate(mem):
        m = sc.method()
        m.alpha = alpha
        e = make_engine(sc.pA, m, kernel='gaussian', precision=precision)
        assert e.pair_form() == precision + '-uniform'
        e.step(K)
        want = e.download()
        e.close()
        for fld in FIELDS + ('p',):
            np.testing.assert_array_equal(got[k][fld], want[fld], err_msg=f'{k} {fld}')


@pytest.mark.parametrize('precision', ['fp64', 'fp32c'])
def test_batch_border_bounds_match_full_scan(precision):
    """Each instance's neighbour update after a fused PEC step takes the grid's
    bounds from its border cells (the default) or scans all its particles:
    the same states and clocks after 40 graph-replayed steps, one instance
    stopping after about 10 (its updates end there); every update but each
    instance's first takes the border path."""
    mem = _members()
    sc1, a1 = mem[1]
    _, probe, _ = _single(sc1, a1, 10, precision)
    durations = [np.inf, probe['t'] * 0.999] + [np.inf] * (len(mem) - 2)
    out = []
    for border in (True, False):
        ps = []
        for (sc, alpha), dur in zip(mem, durations):
            p = params_for(sc, alpha, precision=precision)
            p.duration = dur
            ps.append(p)
        b = BatchEngine(ps)
        b.set_border_bounds(border)
        b.upload([soa_from_particles(sc.pA) for sc, _ in mem])
        assert b.pair_form() == precision + '-uniform'
        b.step(40, graph=True)
        got = b.download()
        st = b.status()
        cnt = b.border_count()
        b.close()
        say('batch border bounds', precision, border, cnt)
        assert [x['steps'] for x in st] == [40, 10] + [40] * (len(mem) - 2)
        if border:
            assert cnt == [39, 9] + [39] * (len(mem) - 2), cnt
        else:
            assert cnt == [0] * len(mem)
        out.append((got, st))
    for k in range(len(mem)):
        assert out[0][1][k] == out[1][1][k]
        for fld in FIELDS + ('p',):
            np.testing.assert_array_equal(out[0][0][k][fld], out[1][0][k][fld], err_msg=f'{k} {fld}')


def test_batch_duration_per_instance():
    mem = _members()[:3]
    # instance 1 stops after about 10 steps' time
    sc1, a1 = mem[1]
    _, probe, _ = _single(sc1, a1, 10, 'fp64')
    durations = [np.inf, probe['t'] * 0.999, np.inf]
    ps = []
    for (sc, alpha), dur in zip(mem, durations):
        p = params_for(sc, alpha, precision='fp64')
        p.duration = dur
        ps.append(p)
    b = BatchEngine(ps)
    b.upload([soa_from_particles(sc.pA) for sc, _ in mem])
    b.step(25, graph=True)
    got = b.download()
    st = b.status()
    b.close()
    assert st[0]['steps'] == 25 and st[2]['steps'] == 25
    assert st[1]['steps'] == 10
    for k, ((sc, alpha), dur) in enumerate(zip(mem, durations)):
        want, wst, _ = _single(sc, alpha, 25, 'fp64', duration=dur)
        assert st[k]['steps'] == wst['steps'] and st[k]['t'] == wst['t']
        for fld in FIELDS:
            np.testing.assert_array_equal(got[k][fld], want[fld], err_msg=f'{k} {fld}')


def test_batch_autotune_keeps_results():
    """The launch shape (workgroup size, contributor-list length, one thread
    per slot or the persistent work-queue kernel) changes no result: a list
    flushes in listing order, a lane's target is its own, whichever
    workgroup takes its tile. The persistent runs replay their graph 12
    times, so the queue's self-reset is exercised too."""
    mem = _members()
    states = [_two_masses(sc.pA) for sc, _ in mem]  # the general pass: every launch shape applies
    b = BatchEngine([params_for(sc, alpha, precision='fp32c') for sc, alpha in mem])
    runs = []
    for threads, ll, per in ((256, 24, False), (64, 16, False), (128, 32, False), (256, 24, True),
                             (64, 32, True), (128, 16, True)):
        b.upload([soa_from_particles(pA) for pA in states])
        b.set_tuning(threads, ll, per)
        b.step(12, graph=True)
        runs.append(b.download())
    for other in runs[1:]:
        for k, (got, want) in enumerate(zip(other, runs[0])):
            for fld in ('x', 'y', 'vx', 'vy', 'rho', 'ax', 'ay', 'drho'):
                np.testing.assert_array_equal(got[fld], want[fld], err_msg=f'{k} {fld}')
    b.upload([soa_from_particles(pA) for pA in states])
    best, table = autotune(b, steps=2)
    say('batch autotune', best, table)
    assert best[0] in (64, 128, 256) and best[1] in (16, 24, 32) and best[2] in (False, True)
    assert len(table) == 18
    # the uniform pass has one launch shape: only the schedule is timed, and
    # the two schedules agree
    outs = []
    for per in (False, True):
        b.upload([soa_from_particles(sc.pA) for sc, _ in mem])
        assert b.pair_form() == 'fp32c-uniform'
        b.set_tuning(256, 0, per)
        b.step(12, graph=True)
        outs.append(b.download())
    for k, (got, want) in enumerate(zip(*outs)):
        for fld in ('x', 'y', 'vx', 'vy', 'rho', 'ax', 'ay', 'drho'):
            np.testing.assert_array_equal(got[fld], want[fld], err_msg=f'{k} {fld}')
    best, table = autotune(b, steps=2)
    assert len(table) == 2 and best[:2] == (256, 32)
    b.close()
