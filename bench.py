"""WCSPH step throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Workload: the scaled dam break at dx = 0.002 m, 1M-particle variant (2.0 m x
2.5 m column -> 1,250,000 fluid particles, 3 wall layers of a 4.0 m x 3.0 m
tank -> 15,018 wall particles), cubic spline, h = 1.3 dx, PEC, dynamic dt.
A "step" is one pass of the Solver loop body on the device: dt reduction,
predict, neighbour search (bounds, keys, counting sort, reorder), fused pair
pass (Tait EOS, continuity, momentum + artificial viscosity + gravity,
Lennard-Jones wall force), correct. The state is resident in HBM before the
timed region; W untimed steps (which capture the step as a hipGraph) precede
K timed ones replayed from the graphs with no per-launch timing, bracketed by
device synchronisation.

`value` is particle-steps/s (fluid particles x steps / s) with the FP32
compensated pair pass and the FP64 integrator; the FP64 pair pass is
reported beside it. `roofline` is for the dominant kernel (the fused pair
pass): algorithmic bytes per launch (21 N + 12 N_f + 8 C for FP32C,
41 N + 24 N_f + 8 C for FP64; N all particles, N_f fluid, C grid cells)
over its mean launch time, measured with HIP events on its launch stream in
a separate pass of eager steps right after the timed region. `traffic` and
`valu_issue_frac` come from the committed counter summaries in profiles/ and
are reported only when those were made with this library build (the source
hash sph_version() reports). `cpu_baseline` is the CPU oracle (a C port of
the reference step, one thread like the reference's serial numba loop) on the
same initial state, over a bounded number of steps. The default run also
reports the 4M stress lattice's stage times (`stress_4M`) and the 256-
instance ensemble's throughput (`ensemble_256`).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, 'offshore-sph_amd'))

HBM_PEAK_GBS = 8000.0
METRIC = 'WCSPH particle-steps/sec at 1M particles (FP32 vs FP64); fwd+adjoint/fwd time'


def algorithmic_pair_bytes(precision, n, nf, cells, corrector=False):
    """Bytes the pair pass must move once per launch: per particle its
    position, velocity, density and label (float: 8 + 8 + 4 + 1, double:
    16 + 16 + 8 + 1), per fluid particle the three rates it writes (float 12,
    double 24), per cell its offset (8). With the corrector fused in, per
    fluid particle also the start-of-step x, y, vx, vy, rho and previous ax,
    ay it reads (56, double) and the x, y, vx, vy, rho it writes (40)."""
    if precision == 'fp32c':
        b = 21 * n + 12 * nf + 8 * cells
    else:
        b = 41 * n + 24 * nf + 8 * cells
    return b + (96 * nf if corrector else 0)


def run_gpu(sc, precision, steps, warmup, timing_steps=50, host_reps=0):
    """The headline timing: W warm-up steps (the first captures the step as a
    hipGraph per buffer parity), then K steps replayed from the graphs with
    no per-launch timing, bracketed by device synchronisation. Then, apart
    from that timed region, `timing_steps` eager steps with HIP events around
    every pair pass and neighbour update (on the stream they run on) for the
    per-stage times and the roofline."""
    import torch
    from hipsph.device import Engine, EngineParams, soa_from_particles
    m = sc.method()
    e = Engine(EngineParams.from_method(sc.n, m, kernel='cubic', precision=precision, integrator='pec'))
    e.upload(soa_from_particles(sc.pA))
    e.step(max(warmup, 3), graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.step(steps, graph=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st = e.status()
    grid = e.grid()
    res = {'elapsed_s': t1 - t0, 'ms_per_step': (t1 - t0) / steps * 1e3, 'status': st, 'grid': grid,
           'pair_form': e.pair_form(),
           # neighbour updates so far (one per step) and those whose bounds came from the border cells
           'updates': st['steps'], 'border_bounds_updates': e.border_count()}
    # per-stage times, outside the timed region
    e.timing_enable(True)
    torch.cuda.synchronize()
    q0 = time.perf_counter()
    e.step(timing_steps)
    torch.cuda.synchronize()
    q1 = time.perf_counter()
    tm = e.timing_read()
    e.timing_enable(False)
    res.update({'pair_ms': tm['pair_ms'], 'nn_ms': tm['nn_ms'], 'timed_launches': tm['count'],
                'eager_events_ms_per_step': (q1 - q0) / timing_steps * 1e3})
    if host_reps:
        # a caller that hands over host arrays every step (the reference's
        # boundary): upload from numpy, one step, download to numpy; not `value`
        host = soa_from_particles(sc.pA)
        e.upload(host, reset_clock=False)
        e.step(1)
        e.download()
        torch.cuda.synchronize()
        h0 = time.perf_counter()
        for _ in range(host_reps):
            e.upload(host, reset_clock=False)
            e.step(1)
            e.download()
        torch.cuda.synchronize()
        res['host_roundtrip_ms_per_step'] = (time.perf_counter() - h0) / host_reps * 1e3
    if precision == 'fp64':
        d = e.nn_dump(with_qr=False)
        import numpy as np
        f = sc.pA['label'] == 0
        res['pairs_per_step'] = int(np.diff(d['off'])[f].sum())
    e.close()
    return res


def evolved_timing(sc, before=4000, steps=200, timing_steps=50):
    """The FP32C step on a flow that has run `before` steps from the initial
    lattice (the flow disorders it: more candidates per target, uneven
    contributor counts): K steps replayed from the graphs, then eager steps
    with HIP events around the pair pass and the neighbour update."""
    import torch
    from hipsph.device import Engine, EngineParams, soa_from_particles
    e = Engine(EngineParams.from_method(sc.n, sc.method(), kernel='cubic', precision='fp32c', integrator='pec'))
    e.upload(soa_from_particles(sc.pA))
    done = 0
    while done < before:
        k = min(1000, before - done)
        e.step(k, graph=True)
        torch.cuda.synchronize()
        done += k
    t0 = time.perf_counter()
    e.step(steps, graph=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    e.timing_enable(True)
    e.step(timing_steps)
    tm = e.timing_read()
    e.timing_enable(False)
    st = e.status()
    border = e.border_count()
    e.close()
    return {'steps_before': before, 'sim_time_s': st['t'], 'steps': steps, 'ms_per_step': (t1 - t0) / steps * 1e3,
            'pair_ms': tm['pair_ms'], 'nn_ms': tm['nn_ms'], 'flags': st['flags'],
            'border_bounds_updates': border, 'updates': before + steps + timing_steps}


def run_cpu(sc, budget_s=12.0, max_steps=20, nthreads=1):
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle
    s = oracle.Solver(sc.pA, sc.method(), 'cubic', 'pec', h=sc.h, nthreads=nthreads)
    t0 = time.perf_counter()
    k = 0
    while k < max_steps:
        s.run(1)
        k += 1
        if time.perf_counter() - t0 > budget_s and k >= 2:
            break
    el = time.perf_counter() - t0
    return {'steps': k, 'elapsed_s': el, 'value': sc.n_fluid * k / el}


def profile_pair(sc, launches):
    """Counter-collection driver: per precision 3 + `launches` eager steps of
    the 1M workload (the last `launches` are the profiled ones: the step's
    pair pass with the corrector fused in), then `launches` plain pair passes
    (`sph_compute`, no corrector) on the state they leave; then two kernels of
    known bytes that calibrate the read and write counters: a plain 256 MiB
    device copy (torch copy_, streaming requests) and a gather of 32M doubles
    through a permutation that is random inside each block of 16 elements
    (torch index_select: every 128-byte line of the source is read whole, in
    narrow requests, as the reorder and corrector gather)."""
    import torch
    from hipsph.device import Engine, EngineParams, soa_from_particles
    print(f'# build {build_id()}', file=sys.stderr, flush=True)
    for prec in ('fp32c', 'fp64'):
        e = Engine(EngineParams.from_method(sc.n, sc.method(), precision=prec))
        e.upload(soa_from_particles(sc.pA))
        e.step(3 + launches, graph=False)
        for _ in range(launches):
            e.compute()
        torch.cuda.synchronize()
        print(f'# profiled {launches} steps and {launches} pair passes, {prec} ({e.pair_form()})', file=sys.stderr,
              flush=True)
        e.close()
    n = 32 * 1024 * 1024
    src = torch.empty(n, dtype=torch.float64, device='cuda')
    dst = torch.empty_like(src)
    src.fill_(1.0)
    for _ in range(3):
        dst.copy_(src)
    blk = 16
    g = torch.Generator(device='cpu').manual_seed(7)
    local = torch.argsort(torch.rand(n // blk, blk, generator=g), dim=1)
    perm = (local + torch.arange(0, n, blk).unsqueeze(1)).reshape(-1).to('cuda')
    for _ in range(3):
        torch.index_select(src, 0, perm, out=dst)
    torch.cuda.synchronize()
    print(f'# calibration copy: {n * 8} bytes read and written per launch; gather: {n * 8} bytes of doubles and '
          f'{n * 8} bytes of int64 indices read, {n * 8} bytes written per launch', file=sys.stderr, flush=True)


def precision_report(sc, checkpoints=(100, 500, 1000)):
    """FP32C pair pass (FP64 integrator) against the FP64 pair pass, both on
    the device from the same initial state: max and RMS over the fluid of
    |drho|/rho, |dx|/dx and |dv|/c0 at each checkpoint step."""
    import numpy as np
    from hipsph.device import Engine, EngineParams, soa_from_particles
    m = sc.method()
    eng = {}
    for prec in ('fp64', 'fp32c'):
        eng[prec] = Engine(EngineParams.from_method(sc.n, m, precision=prec))
        eng[prec].upload(soa_from_particles(sc.pA))
    f = sc.pA['label'] == 0
    out, done = [], 0
    for k in checkpoints:
        for e in eng.values():
            e.step(k - done, graph=True)
        done = k
        a, b = eng['fp64'].download(), eng['fp32c'].download()
        drho = np.abs(b['rho'][f] - a['rho'][f]) / a['rho'][f]
        dxn = np.hypot(b['x'][f] - a['x'][f], b['y'][f] - a['y'][f]) / sc.dx
        dvn = np.hypot(b['vx'][f] - a['vx'][f], b['vy'][f] - a['vy'][f]) / m.co
        st = eng['fp64'].status()
        out.append({'step': k, 't': st['t'],
                    'rho_rel_max': float(drho.max()), 'rho_rel_rms': float(np.sqrt(np.mean(drho ** 2))),
                    'x_over_dx_max': float(dxn.max()), 'x_over_dx_rms': float(np.sqrt(np.mean(dxn ** 2))),
                    'v_over_c0_max': float(dvn.max()), 'v_over_c0_rms': float(np.sqrt(np.mean(dvn ** 2)))})
        print(f'# precision at step {k}: rho {out[-1]["rho_rel_max"]:.3g} max, x/dx {out[-1]["x_over_dx_max"]:.3g}'
              f' max, v/c0 {out[-1]["v_over_c0_max"]:.3g} max', file=sys.stderr, flush=True)
    for e in eng.values():
        e.close()
    return out


def adjoint_timing(repeats=3, nsteps=200, every=20):
    """The differentiable dam break (0.5 m x 1.0 m column at dx 0.005: 20,000
    fluid particles, 200 PEC steps, checkpoints every 20): wall time of the
    forward alone and of forward (with checkpoints) + adjoint, which
    recomputes each segment while recording and reverses it. Loss: mean
    squared distance of the final positions to a run at alpha 0.03, c0 x 1.1."""
    import torch
    from hipsph import scenarios
    from hipsph.adjoint import SPHRun
    sc = scenarios.dam_break_c5()
    m = sc.method()
    run = SPHRun(sc.pA, m, nsteps, checkpoint_every=every)
    v0 = torch.zeros((sc.n_fluid, 2), dtype=torch.float64, device=run.device)
    target = run.forward(0.03, 1.1 * m.co, v0, keep=False).clone()

    def fwd_bwd():
        xt = run.forward(m.alpha, m.co, v0, keep=True).clone().requires_grad_(True)
        loss = torch.mean(torch.sum((xt - target) ** 2, 1))
        (g,) = torch.autograd.grad(loss, [xt])
        return run.backward(g)

    fwd_bwd()
    t_f, t_fb = [], []
    for _ in range(repeats):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run.forward(m.alpha, m.co, v0, keep=False)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ga, gc, gv = fwd_bwd()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        t_f.append(t1 - t0)
        t_fb.append(t2 - t1)
    run.close()
    f, fb = min(t_f), min(t_fb)
    print(f'# adjoint: forward {f * 1e3:.2f} ms, forward+adjoint {fb * 1e3:.2f} ms ({fb / f:.2f}x)', file=sys.stderr,
          flush=True)
    return {'workload': sc.name, 'n_particles': sc.n, 'n_fluid': sc.n_fluid, 'steps': nsteps,
            'checkpoint_every': every, 'forward_ms': f * 1e3, 'forward_adjoint_ms': fb * 1e3, 'ratio': fb / f,
            'force_limited_steps': run.force_steps // (repeats + 1),
            'dL_dalpha': float(ga), 'dL_dc0': float(gc), 'dL_dv0_norm': float(torch.linalg.norm(gv))}


def ensemble(steps=500, count=256, seq_steps=10, seq_instances=64):
    """The batched ensemble: `count` dam breaks (width U(0.5, 1.5), height
    U(0.5, 2.5), dx from {0.005, 0.01, 0.02}, alpha U(0.005, 0.05), seeded per
    instance), FP32C pairs and FP64 integrator, `steps` steps each in one batch
    (one launch per stage, hipGraph replay, pair workgroup size,
    contributor-list length and schedule autotuned).
    Beside it: the same instances as separate engines one after another
    (graph replay each), timed on the first `seq_instances` for `seq_steps`."""
    import torch
    import numpy as np
    from hipsph import scenarios
    from hipsph.batch import BatchEngine, autotune, params_for
    from hipsph.device import Engine, soa_from_particles
    scs, alphas = scenarios.ensemble_c4(count)
    nf = sum(s.n_fluid for s in scs)
    n = sum(s.n for s in scs)
    print(f'# ensemble: {count} instances, {n} particles ({nf} fluid)', file=sys.stderr, flush=True)
    b = BatchEngine([params_for(s, a, precision='fp32c') for s, a in zip(scs, alphas)])
    b.upload([soa_from_particles(s.pA) for s in scs])
    best, table = autotune(b)
    b.upload([soa_from_particles(s.pA) for s in scs])
    b.step(2, graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.step(steps, graph=True)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    st = b.status()
    b.timing_enable(True)
    b.step(5, graph=False)
    pair_ms = b.timing_read()['pair_ms']
    b.timing_enable(False)
    b.close()
    steps_done = [s['steps'] for s in st]
    halted = sum(1 for s in st if s['flags'] & 3)
    # sequential baseline on a subset
    sub = list(zip(scs, alphas))[:seq_instances]
    seq = 0.0
    for s, a in sub:
        m = s.method()
        m.alpha = a
        from hipsph.device import EngineParams
        e = Engine(EngineParams.from_method(s.n, m, precision='fp32c'))
        e.upload(soa_from_particles(s.pA))
        e.step(3, graph=True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        e.step(seq_steps, graph=True)
        torch.cuda.synchronize()
        seq += (time.perf_counter() - t1) / seq_steps
        e.close()
    sub_nf = sum(s.n_fluid for s, _ in sub)
    seq_full_ms = seq * 1e3 * nf / sub_nf  # scaled by fluid count to the whole ensemble
    res = {'instances': count, 'n_particles': n, 'n_fluid': nf, 'steps_each': steps,
           'min_steps': int(min(steps_done)), 'halted_instances': halted,
           'ms_per_step': el / steps * 1e3,
           'particle_steps_per_s': nf * steps / el, 'pair_ms': pair_ms,
           'pair_threads': best[0], 'pair_list_len': best[1], 'pair_persistent': best[2], 'autotune': table,
           'sequential_engines_ms_per_step_est': seq_full_ms,
           'sequential_sample': f'{seq_instances} instances x {seq_steps} steps, scaled by fluid count',
           'batch_speedup_vs_sequential': seq_full_ms / (el / steps * 1e3)}
    print(f'# ensemble: {res["ms_per_step"]:.3f} ms/step ({res["particle_steps_per_s"]:.3g} particle-steps/s), '
          f'pair {pair_ms:.3f} ms, sequential engines ~{seq_full_ms:.2f} ms/step', file=sys.stderr, flush=True)
    return res


STRESS_BYTES = {  # algorithmic bytes per launch (N all, N_f fluid, C cells)
    ('fused', 'fp32c'): (21, 12, 8), ('fused', 'fp64'): (41, 24, 8),
    ('continuity', 'fp32c'): (17, 4, 8), ('continuity', 'fp64'): (33, 8, 8),
    ('momentum', 'fp32c'): (21, 8, 8), ('momentum', 'fp64'): (41, 16, 8),
}


def stress(reps, emit=True):
    """Pair-kernel stress: 4,000,000 jittered particles (2000 x 2000, dx 1e-3)
    in a closed box; per kernel and precision, the neighbour update and the
    continuity, momentum and fused pair passes timed apart (HIP events, mean
    over `reps` launches, each stage after a neighbour update of its own)."""
    import numpy as np
    from hipsph import scenarios
    from hipsph.device import Engine, EngineParams, soa_from_particles
    sc = scenarios.stress_lattice()
    n, nf = sc.n, sc.n_fluid
    print(f'# stress workload {sc.name}: {n} particles ({nf} fluid)', file=sys.stderr, flush=True)
    rows = []
    for kernel in ('cubic', 'gaussian'):
        for prec in ('fp32c', 'fp64'):
            e = Engine(EngineParams.from_method(n, sc.method(), kernel=kernel, precision=prec))
            e.upload(soa_from_particles(sc.pA))
            e.nn_update()
            pairs = int(np.sum(np.maximum(e.compute_counts(), 0)))
            cells = e.grid()['n_cells']
            row = {'kernel': kernel, 'precision': prec, 'pair_form': e.pair_form(), 'pairs': pairs,
                   'cells': cells}
            e.timing_enable(True)
            for _ in range(reps):
                e.nn_update()
            row['neighbour_update_ms'] = e.timing_read()['nn_ms']
            for stage, fn in (('continuity', e.continuity), ('momentum', e.momentum), ('fused', e.compute)):
                e.nn_update()
                fn()
                e.timing_read()
                for _ in range(reps):
                    fn()
                ms = e.timing_read()['pair_ms']
                a, b, c = STRESS_BYTES[(stage, prec)]
                byts = a * n + b * nf + c * cells
                row[f'{stage}_ms'] = ms
                row[f'{stage}_hbm_frac'] = byts / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9)
            row['fused_pair_interactions_per_s'] = pairs / (row['fused_ms'] * 1e-3)
            e.timing_enable(False)
            e.close()
            rows.append(row)
            print(f'# stress {kernel} {prec} ({row["pair_form"]}): nn {row["neighbour_update_ms"]:.3f} ms, '
                  f'continuity {row["continuity_ms"]:.3f}, momentum {row["momentum_ms"]:.3f}, fused '
                  f'{row["fused_ms"]:.3f} ms', file=sys.stderr, flush=True)
    out = {'workload': sc.name, 'n_particles': n, 'n_fluid': nf, 'h': sc.h, 'reps': reps, 'stages': rows}
    if emit:
        print(json.dumps({'metric': 'pair-kernel stress stage times (4M particles)', 'unit': 'ms', 'n_gpus': 1,
                          'higher_is_better': False, 'data': 'synthetic', **out}), flush=True)
    return out


# ---------------------------------------------------------------------------
# summaries of rocprofv3 runs (CSV output) for profiles/
# ---------------------------------------------------------------------------
# Per-request sizes at the L2 -> memory-fabric interface (TCC_EA0) of MI355X,
# checked against two kernels of known traffic in every counter run (a
# 256 MiB device copy and a 256 MiB gather through 16-element-local random
# indices; both within 0.1 %): every read request that is not a 32-byte one
# moves a 128-byte line, every write request is 64 or 32 bytes. rocprofv3's
# derived FETCH_SIZE books the 128-byte reads at 64 bytes (it reads half of
# the copy's bytes), so it is not used.
RD_COUNTERS = ('TCC_EA0_RDREQ_sum', 'TCC_EA0_RDREQ_32B_sum', 'TCC_BUBBLE_sum', 'TCC_EA0_RDREQ_DRAM_sum')
WR_COUNTERS = ('TCC_EA0_WRREQ_sum', 'TCC_EA0_WRREQ_64B_sum', 'TCC_EA0_WRREQ_DRAM_sum')
SQ_COUNTERS = ('SQ_WAVES', 'SQ_INSTS_VALU', 'SQ_ACTIVE_INST_VALU', 'SQ_THREAD_CYCLES_VALU', 'SQ_INSTS_SALU',
               'SQ_INSTS_VMEM_RD', 'SQ_INSTS_LDS', 'SQ_WAIT_INST_ANY')
SIMDS, CLOCK_HZ = 1024, 2.4e9  # MI355X: 256 CUs x 4 SIMDs, peak engine clock


def read_bytes(c):
    return 128.0 * (c['TCC_EA0_RDREQ_sum'] - c['TCC_EA0_RDREQ_32B_sum']) + 32.0 * c['TCC_EA0_RDREQ_32B_sum']


def write_bytes(c):
    return 64.0 * c['TCC_EA0_WRREQ_64B_sum'] + 32.0 * (c['TCC_EA0_WRREQ_sum'] - c['TCC_EA0_WRREQ_64B_sum'])


def _csv(dirname, suffix):
    import csv
    import glob
    hits = sorted(glob.glob(os.path.join(dirname, '**', '*' + suffix), recursive=True))
    if not hits:
        raise FileNotFoundError(f'no *{suffix} under {dirname}')
    with open(hits[0]) as f:
        return list(csv.DictReader(f))


def _dispatches(rows):
    """{dispatch id: (kernel name, {counter: value})} of a counter CSV."""
    out = {}
    for r in rows:
        d = out.setdefault(int(r['Dispatch_Id']), (r['Kernel_Name'], {}))
        d[1][r['Counter_Name']] = float(r['Counter_Value'])
    return out


def _short(name):
    base = name.replace('(anonymous namespace)::', '').replace('hipsph::', '').replace('void ', '')
    return base.split('(')[0].strip()


def _pair_prec(name):
    if 'k_pair' not in name or 'batch' in name:
        return None
    return 'fp64' if ('uniform64' in name or 'k_pair<double' in name) else 'fp32c'


def summarize(raw, launches=5):
    """Write profiles/pair_traffic.json, pair_valu.json and the three Markdown
    summaries from the CSV directories of one counter-run series over
    `bench.py --profile-pair K` (kt_pp: kernel trace; pmc_rd / pmc_wr / pmc_sq:
    the counters above) and a kernel trace of a short headline run (kt_step).
    The pair figures are means over the last `launches` dispatches of each
    pair kernel (the profiled passes)."""
    import statistics as stt
    build = None
    for log in ('kt_pp.log', 'pmc_rd.log'):
        p = os.path.join(raw, log)
        if os.path.exists(p):
            for line in open(p):
                if line.startswith('# build '):
                    build = line.split()[2]
    trace = _csv(os.path.join(raw, 'kt_pp'), '_kernel_trace.csv')
    dur = {int(r['Dispatch_Id']): (int(r['End_Timestamp']) - int(r['Start_Timestamp'])) * 1e-3 for r in trace}
    rd = _dispatches(_csv(os.path.join(raw, 'pmc_rd'), '_counter_collection.csv'))
    wr = _dispatches(_csv(os.path.join(raw, 'pmc_wr'), '_counter_collection.csv'))
    sq = _dispatches(_csv(os.path.join(raw, 'pmc_sq'), '_counter_collection.csv'))
    tr_names = {int(r['Dispatch_Id']): r['Kernel_Name'] for r in trace}

    def last(d, name):
        ids = sorted(i for i, (k, _) in d.items() if _short(k) == name)
        return ids[-launches:]

    def mean(d, ids, f):
        return stt.mean(f(d[i][1]) for i in ids)

    # calibration: the three largest copies and the gathers
    cal = {}
    for tag, key, known_r, known_w in (('copy', '__amd_rocclr_copyBuffer', 1 << 28, 1 << 28),
                                       ('gather', '_scatter_gather_elementwise_kernel', 1 << 29, 1 << 28)):
        ids = sorted((i for i, (k, _) in rd.items() if key in k), key=lambda i: -rd[i][1]['TCC_EA0_RDREQ_sum'])[:3]
        idw = sorted((i for i, (k, _) in wr.items() if key in k), key=lambda i: -wr[i][1]['TCC_EA0_WRREQ_sum'])[:3]
        r_b, w_b = mean(rd, ids, read_bytes), mean(wr, idw, write_bytes)
        fetch_kib = mean(rd, ids, lambda c: (c['TCC_BUBBLE_sum'] * 128 + (c['TCC_EA0_RDREQ_sum'] - c['TCC_BUBBLE_sum']
                                                                           - c['TCC_EA0_RDREQ_32B_sum']) * 64
                                             + c['TCC_EA0_RDREQ_32B_sum'] * 32) / 1024)
        cal[tag] = {'kernel': _short(rd[ids[0]][0]), 'dispatches': len(ids), 'known_read_bytes': known_r,
                    'known_write_bytes': known_w, 'read_bytes': r_b, 'write_bytes': w_b,
                    'read_over_known': r_b / known_r, 'write_over_known': w_b / known_w,
                    'FETCH_SIZE_kib': fetch_kib, 'FETCH_SIZE_over_known': fetch_kib * 1024 / known_r}
    traffic = {'build': build, 'model': 'read = 128 B x (TCC_EA0_RDREQ_sum - TCC_EA0_RDREQ_32B_sum) + 32 B x '
                                        'TCC_EA0_RDREQ_32B_sum; write = 64 B x TCC_EA0_WRREQ_64B_sum + 32 B x '
                                        '(TCC_EA0_WRREQ_sum - TCC_EA0_WRREQ_64B_sum)',
               'calibration': cal}
    valu = {'build': build, 'peak': {'simds': SIMDS, 'clock_hz': CLOCK_HZ, 'cycles_per_wave64_valu': 4}}
    # every single-simulation pair kernel of the runs; 'fp32c' / 'fp64' name
    # the step's (the pass with the corrector fused in when it ran)
    names = sorted({_short(k) for k, _ in rd.values() if _pair_prec(k)})
    traffic['kernels'], valu['kernels'] = {}, {}
    for name in names:
        ids_r, ids_w, ids_s = last(rd, name), last(wr, name), last(sq, name)
        ids_t = sorted(i for i, k in tr_names.items() if _short(k) == name)[-launches:]
        if not (ids_r and ids_w and ids_s and ids_t):
            continue
        us = stt.mean(dur[i] for i in ids_t)
        r_b, w_b = mean(rd, ids_r, read_bytes), mean(wr, ids_w, write_bytes)
        traffic['kernels'][name] = {'kernel': name, 'dispatches': len(ids_r), 'read_bytes': r_b, 'write_bytes': w_b,
                                    'hbm_bytes_per_launch': r_b + w_b, 'avg_launch_us': us,
                                    'achieved_GBps': (r_b + w_b) / (us * 1e-6) / 1e9}
        c = {k: mean(sq, ids_s, lambda d, k=k: d[k]) for k in SQ_COUNTERS}
        valu['kernels'][name] = {'kernel': name, 'dispatches': len(ids_s),
                                 **{k + '_mean': v for k, v in c.items()}, 'avg_launch_us': us,
                                 'valu_instructions_per_wave': c['SQ_INSTS_VALU'] / c['SQ_WAVES'],
                                 'valu_issue_frac': c['SQ_INSTS_VALU'] * 4 / (SIMDS * CLOCK_HZ * us * 1e-6),
                                 'valu_lane_utilization': c['SQ_THREAD_CYCLES_VALU'] / (c['SQ_ACTIVE_INST_VALU'] * 64)}
    for prec in ('fp32c', 'fp64'):
        mine = [k for k in traffic['kernels'] if _pair_prec(k) == prec]
        step = [k for k in mine if '_pec' in k] or mine
        if step:
            traffic[prec] = traffic['kernels'][step[0]]
            valu[prec] = valu['kernels'][step[0]]
    # other kernels of the counter runs: mean per dispatch (all dispatches)
    others = {}
    for i, (k, c) in rd.items():
        if k.startswith('k_') or 'hipsph' in k:
            others.setdefault(_short(k), {'r': [], 'w': []})['r'].append(read_bytes(c))
    for i, (k, c) in wr.items():
        if _short(k) in others:
            others[_short(k)]['w'].append(write_bytes(c))
    traffic['other_kernels_MB'] = {k: {'dispatches': len(v['r']), 'read_MB_mean': stt.mean(v['r']) / 1e6,
                                       'write_MB_mean': stt.mean(v['w']) / 1e6 if v['w'] else None,
                                       'read_MB_last': v['r'][-1] / 1e6}
                                   for k, v in sorted(others.items())}
    pdir = os.path.join(ROOT, 'profiles')
    with open(os.path.join(pdir, 'pair_traffic.json'), 'w') as f:
        json.dump(traffic, f, indent=1)
    with open(os.path.join(pdir, 'pair_valu.json'), 'w') as f:
        json.dump(valu, f, indent=1)
    # the headline run's kernel statistics, split at the switch from the
    # float to the double pair pass
    tr = sorted(_csv(os.path.join(raw, 'kt_step'), '_kernel_trace.csv'), key=lambda r: int(r['Start_Timestamp']))
    first64 = next((j for j, r in enumerate(tr) if _pair_prec(r['Kernel_Name']) == 'fp64'), len(tr))
    phases = {}
    for tag, part in (('fp32c', tr[:first64]), ('fp64', tr[first64:])):
        st = {}
        for r in part:
            st.setdefault(_short(r['Kernel_Name']), []).append(
                (int(r['End_Timestamp']) - int(r['Start_Timestamp'])) * 1e-3)
        phases[tag] = st
    return traffic, valu, phases, build


NN_KERNELS = ('k_bounds', 'k_grid', 'k_keys', 'k_scan_reduce', 'k_scan_apply', 'k_scatter', 'k_cellsort', 'k_reorder')
STEP_KERNELS = NN_KERNELS + ('k_correct', 'k_dt_final')
PROFILE_COMMANDS = """    cd /tmp && export TMPDIR=/tmp      # OUT: an output directory, R: the repository
    rocprofv3 --kernel-trace --stats -f csv -d OUT/kt_pp -o kt -- python3 -u R/bench.py --profile-pair 5 2> OUT/kt_pp.log
    rocprofv3 --pmc TCC_EA0_RDREQ_sum TCC_EA0_RDREQ_32B_sum TCC_BUBBLE_sum TCC_EA0_RDREQ_DRAM_sum -f csv \\
        -d OUT/pmc_rd -o rd -- python3 -u R/bench.py --profile-pair 5
    rocprofv3 --pmc TCC_EA0_WRREQ_sum TCC_EA0_WRREQ_64B_sum TCC_EA0_WRREQ_DRAM_sum -f csv \\
        -d OUT/pmc_wr -o wr -- python3 -u R/bench.py --profile-pair 5
    rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU SQ_ACTIVE_INST_VALU SQ_THREAD_CYCLES_VALU SQ_INSTS_SALU \\
        SQ_INSTS_VMEM_RD SQ_INSTS_LDS SQ_WAIT_INST_ANY -f csv -d OUT/pmc_sq -o sq -- python3 -u R/bench.py --profile-pair 5
    rocprofv3 --kernel-trace --stats -f csv -d OUT/kt_step -o kt -- python3 -u R/bench.py --steps 200 --warmup 10 \\
        --no-cpu --no-precision --no-adjoint --stress-reps 0 --ensemble-steps 0 --evolved-steps 0
    python3 R/bench.py --summarize OUT"""


def write_summaries(traffic, valu, phases, build, n=None, nf=None, cells=None):
    """The Markdown summaries of summarize()'s figures (profiles/*.md)."""
    pdir = os.path.join(ROOT, 'profiles')
    cal = traffic['calibration']
    head = (f'Library build `{build}` (the source hash `sph_version()` reports). MI355X, ROCm 7.2. Commands:\n\n'
            f'{PROFILE_COMMANDS}\n')
    wl = ('1M dam break (1,265,021 particles, 1,250,000 fluid, cubic spline, one h and one fluid mass: the '
          'uniform pair passes). `--profile-pair 5`: per precision 8 eager steps (the last 5 profiled: the '
          'step\'s pair pass with the PEC corrector fused in, `k_pair_uniform*_pec`), then 5 plain pair passes '
          '(`sph_compute`, `k_pair_uniform*`) on the state they leave; then the calibration kernels.')
    L = ['# HBM traffic of the fused pair pass — 1M dam break', '', wl, '', head,
         'Bytes per dispatch from the request counters at the L2 -> memory interface:',
         f'`{traffic["model"]}`.', '',
         'Calibration (the same runs; three dispatches each):', '',
         '| kernel | known read | model read | ratio | known write | model write | ratio | FETCH_SIZE / known read |',
         '|---|---|---|---|---|---|---|---|']
    for tag, c in cal.items():
        L.append(f'| {tag}: `{c["kernel"]}` | {c["known_read_bytes"] / 1e6:.1f} MB | {c["read_bytes"] / 1e6:.1f} MB | '
                 f'{c["read_over_known"]:.4f} | {c["known_write_bytes"] / 1e6:.1f} MB | {c["write_bytes"] / 1e6:.1f} MB'
                 f' | {c["write_over_known"]:.4f} | {c["FETCH_SIZE_over_known"]:.3f} |')
    L += ['', 'The model reads both kernels to within 0.1 %; rocprofv3\'s derived FETCH_SIZE (which books each '
          'non-32-byte read request at 64 bytes) reads half, so it is not used.', '',
          '| kernel | dispatches | read MB | write MB | HBM MB / launch | avg us (trace) | GB/s | % of 8 TB/s |',
          '|---|---|---|---|---|---|---|---|']
    for name, t in traffic['kernels'].items():
        prec = _pair_prec(name)
        if t:
            L.append(f'| `{t["kernel"]}` ({prec}) | {t["dispatches"]} | {t["read_bytes"] / 1e6:.2f} | '
                     f'{t["write_bytes"] / 1e6:.2f} | **{t["hbm_bytes_per_launch"] / 1e6:.2f}** | {t["avg_launch_us"]:.1f}'
                     f' | {t["achieved_GBps"]:.0f} | {100 * t["achieved_GBps"] / HBM_PEAK_GBS:.1f} |')
    L += ['', 'Reading: the plain uniform passes read 16-byte position records, 16-byte (float) or 32-byte '
          '(double) velocity/density/pressure records and 8-byte cell coordinates per particle, plus the cell '
          'offsets; the writes are ax, ay, drho as doubles per fluid particle (24 B x 1.25e6 = 30 MB). The '
          'measured reads equal one pass over those arrays: neighbour-cell records re-read by other workgroups '
          'hit in L2. The fused passes also gather the start-of-step x, y, vx, vy, rho and the previous ax, ay '
          '(56 B per fluid particle, through the reorder\'s permutation) and write the corrected x, y, vx, vy, '
          'rho (40 B).', '',
          'Other kernels of the same runs (MB per dispatch; the first step after an upload gathers from an '
          'unsorted state, the later ones from a nearly sorted one — `last` is the last dispatch):', '',
          '| kernel | dispatches | read MB (mean) | read MB (last) | write MB (mean) |', '|---|---|---|---|---|']
    for k, v in traffic['other_kernels_MB'].items():
        if _pair_prec(k):
            continue
        w = f'{v["write_MB_mean"]:.2f}' if v['write_MB_mean'] is not None else '-'
        L.append(f'| `{k}` | {v["dispatches"]} | {v["read_MB_mean"]:.2f} | {v["read_MB_last"]:.2f} | {w} |')
    with open(os.path.join(pdir, 'pair_traffic_dam_break_1M.md'), 'w') as f:
        f.write('\n'.join(L) + '\n')

    L = ['# VALU issue of the fused pair pass — 1M dam break', '', wl, '', head,
         '| kernel | SQ_WAVES | SQ_INSTS_VALU | VALU instr. / wave | avg us (trace) | VALU issue fraction | '
         'VALU lane utilization |', '|---|---|---|---|---|---|---|']
    for name, v in valu['kernels'].items():
        prec = _pair_prec(name)
        if v:
            L.append(f'| `{v["kernel"]}` ({prec}) | {v["SQ_WAVES_mean"]:.0f} | {v["SQ_INSTS_VALU_mean"]:.4g} | '
                     f'{v["valu_instructions_per_wave"]:.0f} | {v["avg_launch_us"]:.1f} | **{v["valu_issue_frac"]:.2f}**'
                     f' | {v["valu_lane_utilization"]:.2f} |')
    L += ['', 'VALU issue fraction = SQ_INSTS_VALU x 4 cycles (a wave64 VALU instruction holds a 16-lane SIMD for 4 '
          'cycles) / (1,024 SIMDs x 2.4 GHz x the average duration): a lower bound on the share of SIMD cycles '
          'spent issuing VALU work (the clock is at most 2.4 GHz; transcendental and double divide steps take '
          'more than 4 cycles). Lane utilization = SQ_THREAD_CYCLES_VALU / (SQ_ACTIVE_INST_VALU x 64), rocprofv3\'s '
          'VALUUtilization: the share of lanes active in the VALU instructions issued (divergence of the walk '
          'and the flush).', '', 'Other counters (means per dispatch):', '']
    for name, v in valu['kernels'].items():
        L.append(f'* `{name}`: ' + ', '.join(f'{k} {v[k + "_mean"]:.4g}' for k in SQ_COUNTERS))
    with open(os.path.join(pdir, 'pair_valu_dam_break_1M.md'), 'w') as f:
        f.write('\n'.join(L) + '\n')

    L = ['# Kernel statistics of the 1M dam break step', '',
         'The headline run, shortened: `bench.py --steps 200 --warmup 10` (FP32C run, then FP64 run; each 10 '
         'warm-up steps, 200 steps replayed from hipGraphs, then 50 eager steps with HIP events around the pair '
         'pass and the neighbour update, and the host round trip). Kernels of the engine only; the split between '
         'the runs is the first double pair dispatch.', '', head]
    for tag in ('fp32c', 'fp64'):
        st = phases.get(tag, {})
        rows = sorted(((k, v) for k, v in st.items() if k.startswith('k_') or 'k_pair' in k),
                      key=lambda kv: -sum(kv[1]))
        L += [f'## {tag}', '', '| kernel | calls | avg us | min us | max us | total ms |', '|---|---|---|---|---|---|']
        for k, v in rows:
            L.append(f'| `{k}` | {len(v)} | {sum(v) / len(v):.2f} | {min(v):.2f} | {max(v):.2f} | {sum(v) * 1e-3:.2f} |')
        per_step = {k: sum(v) / len(v) for k, v in rows if k in STEP_KERNELS or _pair_prec(k)}
        nn = sum(per_step.get(k, 0.0) for k in NN_KERNELS)
        L += ['', f'One step = one call of each of {len(per_step)} kernels: {sum(per_step.values()):.1f} us of kernel '
                  f'time (the neighbour update, {" + ".join(NN_KERNELS)}: {nn:.1f} us), before the gaps between '
                  f'launches.', '']
    with open(os.path.join(pdir, 'step_dam_break_1M_kernel_stats.md'), 'w') as f:
        f.write('\n'.join(L) + '\n')


STATS_RUNS = {  # kernel-trace runs summarised into profiles/ (directory, command, file, title)
    'kt_stress': ('--stress 20', 'stress_4M_kernel_stats.md',
                  'Kernel statistics — pair-kernel stress lattice (4,024,030 particles)'),
    'kt_ens': ('--ensemble', 'ensemble_256_dam_breaks_kernel_stats.md',
               'Kernel statistics — batched ensemble of 256 dam breaks (7,677,758 particles)'),
    'kt_adj': ('--adjoint', 'adjoint_dam_break_20k_kernel_stats.md',
               'Kernel statistics — forward + adjoint, differentiable dam break (23,021 particles)'),
}
STATS_NOTES = {
    'kt_stress': 'Workload: 2000 x 2000 jittered fluid particles at dx = 1e-3 m (h = 1.3 dx, velocities '
                 'N(0, (0.05 c0)^2)) in a closed box of 3 wall layers: 595,984 cells, 168,747,172 neighbour '
                 'pairs. Per kernel (cubic, Gaussian) and precision, 20 launches each of the neighbour update, '
                 'the continuity pass, the momentum pass and the fused pass, all in the uniform form '
                 '`k_pair_uniform<kernel, mode>` / `k_pair_uniform64<kernel, mode>` (kernel 0 cubic, 1 Gaussian; '
                 'mode 0 fused, 1 continuity, 2 momentum).',
    'kt_ens': 'Workload: 256 dam breaks (column width U(0.5, 1.5) m, height U(0.5, 2.5) m, dx from '
              '{0.005, 0.01, 0.02} m, alpha U(0.005, 0.05), 4.0 x 3.0 m tank), 961 to 136,948 fluid particles '
              'each, FP32C pairs, FP64 integrator, PEC, per-instance dt; every stage one launch over all '
              'instances, the step replayed as a hipGraph. Every instance has one h and one fluid mass, so the '
              'pair pass is the uniform one with the corrector fused in (`k_pair_batch_uniform_pec`, one '
              'workgroup per 256-slot tile of one instance); the autotuner times it against the persistent '
              'work-queue schedule (`k_pair_batch_uniform_queue` + `kb_correct`).',
    'kt_adj': 'Workload: 0.5 m x 1.0 m column at dx = 0.005 m (20,000 fluid, 3,021 wall particles), cubic, PEC, '
              'dynamic dt, FP64, 200 steps, checkpoints every 20; backward per segment: restore, recompute '
              'while recording (the fused pair + corrector pass, then `k_tape_pair`), reverse '
              '(`k_adj_correct`, `k_adj_pair` with the predictor\'s reverse folded in, `k_adj_final`). The run: '
              'one target forward, one warm-up forward + backward, three timed pairs.',
}


def write_stats(raw, build):
    """profiles/*_kernel_stats.md from the kernel-trace runs under raw (see
    STATS_RUNS), each with the JSON line the run printed."""
    for d, (args, fname, title) in STATS_RUNS.items():
        if not os.path.isdir(os.path.join(raw, d)):
            continue
        rows = _csv(os.path.join(raw, d), '_kernel_stats.csv')
        line = None
        logp = os.path.join(raw, d + '.log')
        if os.path.exists(logp):
            for ln in open(logp):
                if ln.startswith('{'):
                    line = ln.strip()
        L = [f'# {title}', '', STATS_NOTES[d], '',
             f'Library build `{build}`. MI355X, ROCm 7.2. Command (R: the repository):', '',
             '    cd /tmp && export TMPDIR=/tmp',
             f'    rocprofv3 --kernel-trace --stats -f csv -d OUT/{d} -o kt -- python3 -u R/bench.py {args} '
             f'> OUT/{d}.log 2>&1',
             f'    python3 R/bench.py --summarize OUT', '']
        if line:
            L += ['The run\'s own result line (timings under the profiler):', '', '```', line[:3000], '```', '']
        L += ['| kernel | calls | avg us | min us | max us | total ms | % |', '|---|---|---|---|---|---|---|']
        for r in rows[:28]:
            L.append(f'| `{_short(r["Name"])[:70]}` | {r["Calls"]} | {float(r["AverageNs"]) / 1e3:.2f} | '
                     f'{float(r["MinNs"]) / 1e3:.2f} | {float(r["MaxNs"]) / 1e3:.2f} | '
                     f'{float(r["TotalDurationNs"]) / 1e6:.2f} | {float(r["Percentage"]):.2f} |')
        with open(os.path.join(ROOT, 'profiles', fname), 'w') as f:
            f.write('\n'.join(L) + '\n')


def _profile(name):
    p = os.path.join(ROOT, 'profiles', name)
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f)
    except Exception:
        return None


def profile_for_build(name, build):
    """A committed counter summary, if it was made with this library build
    (the source hash sph_version() reports); else None and the reason."""
    d = _profile(name)
    if d is None:
        return None, f'profiles/{name} is missing'
    if d.get('build') != build:
        return None, f'profiles/{name} is from build {d.get("build")}, this library is {build}'
    return d, None


def build_id():
    from hipsph import _lib
    v = _lib.load().sph_version().decode()
    return v.split(' source ')[1].split()[0] if ' source ' in v else v


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=1000)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--timing-steps', type=int, default=50,
                    help='eager steps with HIP events around each stage, after the timed region')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--cpu-budget', type=float, default=12.0)
    ap.add_argument('--small', action='store_true', help='C1 (dx 0.02) instead of the 1M workload')
    ap.add_argument('--cpu-threads', type=int, default=16,
                    help='threads of the second (OpenMP) CPU baseline; 0 skips it')
    ap.add_argument('--profile-pair', type=int, default=0, metavar='K',
                    help='only run K pair passes per precision and the calibration kernels (for counter runs)')
    ap.add_argument('--stress', type=int, default=0, metavar='REPS',
                    help='only run the 4M-particle pair-kernel stress timings (REPS launches per stage)')
    ap.add_argument('--stress-reps', type=int, default=10, help='launches per stage of the stress lattice in '
                    'the default run (0 skips it)')
    ap.add_argument('--ensemble-steps', type=int, default=200, help='steps of the 256-instance ensemble in the '
                    'default run (0 skips it)')
    ap.add_argument('--no-precision', action='store_true', help='skip the FP32C-vs-FP64 error report')
    ap.add_argument('--evolved-steps', type=int, default=4000, help='steps run before timing the FP32C step on '
                    'the evolved flow (0: skip)')
    ap.add_argument('--no-adjoint', action='store_true', help='skip the forward+adjoint timing')
    ap.add_argument('--adjoint', action='store_true', help='only run the forward+adjoint timing')
    ap.add_argument('--ensemble', action='store_true', help='only run the 256-instance batched ensemble')
    ap.add_argument('--summarize', metavar='DIR', help='write profiles/ summaries from the rocprofv3 CSV '
                    'directories under DIR (no GPU)')
    a = ap.parse_args()
    if a.summarize:
        res = summarize(a.summarize)
        write_summaries(*res)
        write_stats(a.summarize, res[3])
        return
    if a.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) != 1:
        print('bench.py: this build runs on one GPU; the multi-GPU (slab-decomposed) path is not built. '
              'Run with --gpus 1.', file=sys.stderr)
        sys.exit(2)

    if a.stress:
        stress(a.stress)
        return
    if a.ensemble:
        print(json.dumps({'metric': 'batched ensemble particle-steps/s (256 dam breaks)', 'n_gpus': 1,
                          'higher_is_better': True, 'data': 'synthetic', 'dtype': 'f32',
                          'ensemble': ensemble()}), flush=True)
        return
    if a.adjoint:
        print(json.dumps({'metric': 'forward+adjoint / forward wall time', 'n_gpus': 1, 'higher_is_better': False,
                          'data': 'synthetic', 'adjoint': adjoint_timing()}), flush=True)
        return
    from hipsph import scenarios
    sc = scenarios.dam_break_c1() if a.small else scenarios.dam_break_c2(million=True)
    print(f'# workload {sc.name}: {sc.n} particles ({sc.n_fluid} fluid)', file=sys.stderr, flush=True)
    if a.profile_pair:
        profile_pair(sc, a.profile_pair)
        return

    r32 = run_gpu(sc, 'fp32c', a.steps, a.warmup, a.timing_steps, host_reps=10)
    print(f'# fp32c {r32["ms_per_step"]:.4f} ms/step ({r32["pair_form"]}), pair {r32["pair_ms"]:.4f} ms, '
          f'neighbour update {r32["nn_ms"]:.4f} ms', file=sys.stderr, flush=True)
    r64 = run_gpu(sc, 'fp64', a.steps, a.warmup, a.timing_steps)
    print(f'# fp64 {r64["ms_per_step"]:.4f} ms/step ({r64["pair_form"]}), pair {r64["pair_ms"]:.4f} ms',
          file=sys.stderr, flush=True)

    n, nf = sc.n, sc.n_fluid
    cells = r32['grid']['n_cells']
    value = nf * a.steps / r32['elapsed_s']
    # one h and one fluid mass, PEC without XSPH: the step runs the uniform
    # pair pass with the corrector fused in
    fused32, fused64 = r32['pair_form'].endswith('uniform'), r64['pair_form'].endswith('uniform')
    b32 = algorithmic_pair_bytes('fp32c', n, nf, cells, corrector=fused32)
    b64 = algorithmic_pair_bytes('fp64', n, nf, cells, corrector=fused64)
    ach32 = b32 / (r32['pair_ms'] * 1e-3) / 1e9
    ach64 = b64 / (r64['pair_ms'] * 1e-3) / 1e9
    pairs = r64['pairs_per_step']

    build = build_id()
    traffic, traffic_note = profile_for_build('pair_traffic.json', build)
    valu, valu_note = profile_for_build('pair_valu.json', build)

    def prof(d, prec, field):
        return d.get(prec, {}).get(field) if d else None

    prec_report = None if a.no_precision else precision_report(sc)
    evolved = evolved_timing(sc, a.evolved_steps) if a.evolved_steps else None
    if evolved:
        print(f'# fp32c after {evolved["steps_before"]} steps (t = {evolved["sim_time_s"]:.4f} s): '
              f'{evolved["ms_per_step"]:.4f} ms/step, pair {evolved["pair_ms"]:.4f} ms, neighbour update '
              f'{evolved["nn_ms"]:.4f} ms', file=sys.stderr, flush=True)
    adj = None if a.no_adjoint else adjoint_timing()
    st4m = stress(a.stress_reps, emit=False) if a.stress_reps else None
    ens = ensemble(steps=a.ensemble_steps, seq_instances=16) if a.ensemble_steps else None

    cpu = cpu_omp = None
    if not a.no_cpu:
        c1 = run_cpu(sc, a.cpu_budget, nthreads=1)
        print(f'# cpu oracle 1 thread: {c1["value"]:.4g} particle-steps/s over {c1["steps"]} steps',
              file=sys.stderr, flush=True)
        cpu = {'value': c1['value'], 'unit': 'particle-steps/s', 'cores': 1, 'kind': 'port',
               'sample': f'{c1["steps"]} steps of the same {n}-particle dam break from the same initial state '
                         f'({c1["elapsed_s"]:.1f} s), serial C port of the reference step (oracle/refcpu.c)'}
        if a.cpu_threads > 1:
            cm = run_cpu(sc, a.cpu_budget, nthreads=a.cpu_threads)
            print(f'# cpu oracle {a.cpu_threads} threads: {cm["value"]:.4g} particle-steps/s over {cm["steps"]} '
                  f'steps', file=sys.stderr, flush=True)
            cpu_omp = {'value': cm['value'], 'unit': 'particle-steps/s', 'cores': a.cpu_threads, 'kind': 'port',
                       'sample': f'{cm["steps"]} steps ({cm["elapsed_s"]:.1f} s), OpenMP pair loop of the same port'}

    kname = {'fp32c-uniform': 'k_pair_uniform_pec<cubic> (pair pass + PEC corrector)', 'fp32c': 'k_pair<float, cubic>',
             'fp64-uniform': 'k_pair_uniform64_pec<cubic> (pair pass + PEC corrector)', 'fp64': 'k_pair<double, cubic>'}
    line = {
        'metric': METRIC,
        'value': value,
        'unit': 'particle-steps/s',
        'n_gpus': 1,
        'steps': a.steps,
        'warmup': a.warmup,
        'ms_per_step': r32['ms_per_step'],
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic',
        'config': {'workload': sc.name, 'n_particles': n, 'n_fluid': nf, 'dx': sc.dx, 'kernel': 'cubic',
                   'integrator': 'pec', 'pair_precision': 'fp32c (float pairs, compensated sums)',
                   'pair_form': r32['pair_form'], 'integrator_dtype': 'f64', 'h': sc.h, 'cells': cells,
                   'timed_region': 'K steps replayed from hipGraphs, no per-launch timing',
                   'parallelism': 'single GPU'},
        'roofline': {'bound': 'hbm', 'achieved': ach32, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': ach32 / HBM_PEAK_GBS, 'traffic': prof(traffic, 'fp32c', 'hbm_bytes_per_launch'),
                     'traffic_note': traffic_note, 'kernel': kname[r32['pair_form']],
                     'algorithmic_bytes_per_launch': b32, 'avg_launch_ms': r32['pair_ms'],
                     'launches': r32['timed_launches'],
                     'timing': f'HIP events around each pair launch over {a.timing_steps} eager steps after the '
                               f'timed region',
                     'valu_issue_frac': prof(valu, 'fp32c', 'valu_issue_frac'), 'valu_note': valu_note,
                     'build': build},
        'cpu_baseline': cpu,
        'speedup_vs_cpu': (value / cpu['value']) if cpu else None,
        'cpu_baseline_omp': cpu_omp,
        'speedup_vs_cpu_omp': (value / cpu_omp['value']) if cpu_omp else None,
        'pairs_per_step': pairs,
        'pair_interactions_per_s': pairs * a.steps / r32['elapsed_s'],
        'stage_ms': {('pair_and_corrector' if fused32 else 'pair'): r32['pair_ms'],
                     'neighbour_update': r32['nn_ms'],
                     'rest': r32['eager_events_ms_per_step'] - r32['pair_ms'] - r32['nn_ms']},
        'neighbour_updates': {'steps': r32['updates'], 'border_bounds': r32['border_bounds_updates']},
        'eager_events_ms_per_step': r32['eager_events_ms_per_step'],
        'pcie_inclusive': ({'ms_per_step': r32['host_roundtrip_ms_per_step'],
                            'particle_steps_per_s': nf / (r32['host_roundtrip_ms_per_step'] * 1e-3),
                            'note': 'state uploaded from host numpy arrays (pageable) and downloaded every step; '
                                    'not the headline value'}
                           if 'host_roundtrip_ms_per_step' in r32 else None),
        'fp64': {'value': nf * a.steps / r64['elapsed_s'], 'ms_per_step': r64['ms_per_step'],
                 'pair_form': r64['pair_form'],
                 'roofline': {'bound': 'hbm', 'achieved': ach64, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                              'frac': ach64 / HBM_PEAK_GBS, 'traffic': prof(traffic, 'fp64', 'hbm_bytes_per_launch'),
                              'kernel': kname[r64['pair_form']], 'algorithmic_bytes_per_launch': b64,
                              'avg_launch_ms': r64['pair_ms'],
                              'valu_issue_frac': prof(valu, 'fp64', 'valu_issue_frac')},
                 'stage_ms': {('pair_and_corrector' if fused64 else 'pair'): r64['pair_ms'],
                              'neighbour_update': r64['nn_ms']}},
        'precision': prec_report,
        'evolved': evolved,
        'fwd_adjoint_over_fwd': adj['ratio'] if adj else None,
        'adjoint': adj,
        'stress_4M': st4m,
        'ensemble_256': ens,
        'sim_time_s': r32['status']['t'],
    }
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
