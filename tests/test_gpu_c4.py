"""The batched ensemble on its own workload, against the CPU oracle.

Members of the 256-instance ensemble (scenarios.ensemble_c4: column width
U(0.5, 1.5) m, height U(0.5, 2.5) m, dx drawn from {0.005, 0.01, 0.02} m, a
4.0 x 3.0 tank, alpha ~ U(0.005, 0.05), each from its own seed) are advanced
together in one BatchEngine and each is compared with oracle.Solver, the CPU
restatement of the reference's loop body, run alone with the same constants:
  * FP64 pair pass: every field and the clock bit-identical after 40 steps,
    with the pair stage one thread per slot and as the persistent work-queue
    kernel, both replayed from a hipGraph;
  * FP32C pair pass: within the float pair arithmetic's error (the bounds
    are stated in the test).
The members cover all three resolutions, including the largest instance of
the ensemble (136,948 fluid particles at dx = 0.005).
"""
import numpy as np
import pytest

import oracle
from gpu_util import FIELDS, say
from hipsph import scenarios
from hipsph.batch import BatchEngine, params_for
from hipsph.device import soa_from_particles

pytestmark = pytest.mark.gpu

# instance numbers: dx = 0.005 (largest and smallest of the ensemble), 0.01
# (largest, smallest), 0.02 (largest, smallest) and the first two
MEMBERS = (107, 41, 131, 77, 33, 166, 0, 1)
STEPS = 40
_CACHE = {}


def _members():
    if 'mem' not in _CACHE:
        scs, alphas = scenarios.ensemble_c4(members=MEMBERS)
        _CACHE['mem'] = (scs, alphas)
    return _CACHE['mem']


def _oracle_runs():
    if 'oracle' not in _CACHE:
        scs, alphas = _members()
        out = []
        for sc, a in zip(scs, alphas):
            m = sc.method()
            m.alpha = a
            s = oracle.Solver(sc.pA.copy(), m, 'cubic', 'pec', h=sc.h, nthreads=16)  # it steps in place
            hist = s.run(STEPS)
            out.append((s.pA.copy(), s.t, hist))
        _CACHE['oracle'] = out
    return _CACHE['oracle']


def _batch(precision, persistent, graph=True):
    scs, alphas = _members()
    b = BatchEngine([params_for(sc, a, precision=precision) for sc, a in zip(scs, alphas)])
    b.set_tuning(256, 0, persistent)
    b.upload([soa_from_particles(sc.pA) for sc in scs])
    assert b.pair_form() == precision + '-uniform'  # each member: one h, one fluid mass
    b.step(STEPS, graph=graph)
    got = b.download()
    st = b.status()
    b.close()
    return got, st


def test_c4_members_cover_the_three_resolutions():
    scs, alphas = _members()
    assert sorted({sc.dx for sc in scs}) == [0.005, 0.01, 0.02]
    assert all(0.005 <= a <= 0.05 for a in alphas)
    assert max(sc.n_fluid for sc in scs) == 136948  # the largest member of the ensemble
    assert all(sc.meta['tank_w'] == 4.0 and sc.meta['tank_h'] == 3.0 for sc in scs)


@pytest.mark.parametrize('persistent', [False, True])
def test_c4_batch_fp64_bitexact_vs_oracle(persistent):
    scs, _ = _members()
    say('c4 fp64 batch, persistent', persistent, 'members', MEMBERS, 'sizes', [sc.n for sc in scs])
    got, st = _batch('fp64', persistent)
    ref = _oracle_runs()
    for k, sc in enumerate(scs):
        pA, t, hist = ref[k]
        assert st[k]['steps'] == STEPS and (st[k]['flags'] & ~4) == 0, (MEMBERS[k], st[k])  # 4: an aliased key is no halt
        assert st[k]['t'] == t, (MEMBERS[k], st[k]['t'], t)
        assert st[k]['dt'] == hist[-1, 0]
        for fld in FIELDS:
            np.testing.assert_array_equal(got[k][fld], pA[fld], err_msg=f'member {MEMBERS[k]} {fld}')


def test_c4_batch_fp32c_vs_oracle():
    """Float pairs with compensated sums, FP64 integrator and dt: after 40
    steps the density within 1e-6 relative, positions within 1e-5 dx,
    velocities within 1e-6 c0, the clock within 1e-5 relative (dt follows
    max |a|^2 of the float pass), the rates within 1e-4 of their largest
    magnitude; walls (which do not move) are identical."""
    scs, alphas = _members()
    got, st = _batch('fp32c', False)
    ref = _oracle_runs()
    for k, sc in enumerate(scs):
        pA, t, _ = ref[k]
        m = sc.method()
        f = sc.pA['label'] == 0
        g = got[k]
        e_rho = float(np.max(np.abs(g['rho'][f] - pA['rho'][f]) / pA['rho'][f]))
        e_x = float(np.max(np.hypot(g['x'][f] - pA['x'][f], g['y'][f] - pA['y'][f]))) / sc.dx
        e_v = float(np.max(np.hypot(g['vx'][f] - pA['vx'][f], g['vy'][f] - pA['vy'][f]))) / m.co
        e_t = abs(st[k]['t'] - t) / t
        e_r = {fld: float(np.max(np.abs(g[fld][f] - pA[fld][f])) / np.max(np.abs(pA[fld][f])))
               for fld in ('ax', 'ay', 'drho')}
        say('c4 fp32c member', MEMBERS[k], 'n', sc.n, 'rho', e_rho, 'x/dx', e_x, 'v/c0', e_v, 't', e_t, 'rates', e_r)
        assert st[k]['steps'] == STEPS
        assert e_rho <= 1e-6 and e_x <= 1e-5 and e_v <= 1e-6 and e_t <= 1e-5, MEMBERS[k]
        assert all(v <= 1e-4 for v in e_r.values()), (MEMBERS[k], e_r)
        for fld in ('x', 'y', 'vx', 'vy', 'rho'):
            np.testing.assert_array_equal(g[fld][~f], pA[fld][~f])


def test_c4_batch_out_of_grid_key_flags_its_instance():
    """A particle whose flat key falls past its instance's grid (the corner
    (4, 2) of a 4 x 2 grid of 1 m cells: cx = ncx, cy = ncy) raises
    SPH_FLAG_OOB_KEY on that instance only, as it does for a single engine;
    the other instance is unflagged."""
    from hipsph._lib import SPH_FLAG_OOB_KEY
    scs, alphas = _members()
    sc = scs[5]  # the smallest member
    pA = sc.pA.copy()
    x = np.array([0.0, 1.5, 2.5, 4.0, 0.2, 4.0, 3.9])
    y = np.array([0.0, 0.5, 1.5, 0.5, 1.6, 2.0, 1.9])
    odd = np.zeros(len(x), dtype=pA.dtype)
    odd['x'], odd['y'], odd['h'], odd['rho'], odd['m'] = x, y, 0.5, 1000.0, 1.0
    m = scs[5].method()
    b = BatchEngine([params_for(sc, alphas[5]), params_for(scenarios.Scenario('odd', odd, 2.0, 0.5), 0.01)])
    b.upload([soa_from_particles(pA), soa_from_particles(odd)])
    b.step(1, graph=False)
    st = b.status()
    b.close()
    say('c4 oob flags', st[0]['flags'], st[1]['flags'], m.co)
    assert st[0]['flags'] & SPH_FLAG_OOB_KEY == 0
    assert st[1]['flags'] & SPH_FLAG_OOB_KEY
