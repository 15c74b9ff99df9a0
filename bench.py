"""WCSPH step throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Workload: the scaled dam break at dx = 0.002 m, 1M-particle variant (2.0 m x
2.5 m column -> 1,250,000 fluid particles, 3 wall layers of a 4.0 m x 3.0 m
tank -> 15,018 wall particles), cubic spline, h = 1.3 dx, PEC, dynamic dt.
A "step" is one pass of the Solver loop body on the device: dt reduction,
predict, neighbour search (bounds, keys, counting sort, reorder), fused pair
pass (Tait EOS, continuity, momentum + artificial viscosity + gravity,
Lennard-Jones wall force), correct. The state is resident in HBM before the
timed region; W untimed steps precede K timed ones, bracketed by device
synchronisation.

`value` is particle-steps/s (fluid particles x steps / s) with the FP32
compensated pair pass and the FP64 integrator; the FP64 pair pass is
reported beside it. `roofline` is for the dominant kernel (the fused pair
pass): algorithmic bytes per launch (21 N + 12 N_f + 8 C for FP32C,
41 N + 24 N_f + 8 C for FP64; N all particles, N_f fluid, C grid cells)
over its mean launch time, measured with HIP events on its launch stream
inside the timed region. `cpu_baseline` is the CPU oracle (a C port of the
reference step, one thread like the reference's serial numba loop) on the
same initial state, over a bounded number of steps.
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


def algorithmic_pair_bytes(precision, n, nf, cells):
    if precision == 'fp32c':
        return 21 * n + 12 * nf + 8 * cells
    return 41 * n + 24 * nf + 8 * cells


def run_gpu(sc, precision, steps, warmup, graph_steps):
    import torch
    from hipsph.device import Engine, EngineParams, soa_from_particles
    m = sc.method()
    e = Engine(EngineParams.from_method(sc.n, m, kernel='cubic', precision=precision, integrator='pec'))
    e.upload(soa_from_particles(sc.pA))
    e.step(warmup)
    torch.cuda.synchronize()
    e.timing_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.step(steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    tm = e.timing_read()
    e.timing_enable(False)
    st = e.status()
    grid = e.grid()
    res = {'elapsed_s': t1 - t0, 'ms_per_step': (t1 - t0) / steps * 1e3, 'pair_ms': tm['pair_ms'],
           'nn_ms': tm['nn_ms'], 'timed_launches': tm['count'], 'status': st, 'grid': grid}
    if graph_steps:
        e.step(2, graph=True)  # capture both buffer parities
        torch.cuda.synchronize()
        g0 = time.perf_counter()
        e.step(graph_steps, graph=True)
        torch.cuda.synchronize()
        g1 = time.perf_counter()
        res['graph_ms_per_step'] = (g1 - g0) / graph_steps * 1e3
    if precision == 'fp64':
        d = e.nn_dump(with_qr=False)
        import numpy as np
        f = sc.pA['label'] == 0
        res['pairs_per_step'] = int(np.diff(d['off'])[f].sum())
    e.close()
    return res


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
    """Counter-collection driver: `launches` pair passes per precision on the
    1M workload (after 3 steps), then a plain 256 MiB device copy whose
    FETCH_SIZE / WRITE_SIZE calibrate the counters to bytes."""
    import torch
    from hipsph.device import Engine, EngineParams, soa_from_particles
    for prec in ('fp32c', 'fp64'):
        e = Engine(EngineParams.from_method(sc.n, sc.method(), precision=prec))
        e.upload(soa_from_particles(sc.pA))
        e.step(3)
        e.nn_update()
        for _ in range(launches):
            e.compute()
        torch.cuda.synchronize()
        print(f'# profiled {launches} pair passes, {prec}', file=sys.stderr, flush=True)
        e.close()
    src = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device='cuda')
    dst = torch.empty_like(src)
    src.fill_(1.0)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f'# calibration copy: {src.numel() * 8} bytes read and written per launch', file=sys.stderr, flush=True)


def traffic_from_profiles(precision):
    p = os.path.join(ROOT, 'profiles', 'pair_traffic.json')
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            d = json.load(f)
        return d.get(precision, {}).get('hbm_bytes_per_launch')
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=1000)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--graph-steps', type=int, default=200)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--cpu-budget', type=float, default=12.0)
    ap.add_argument('--small', action='store_true', help='C1 (dx 0.02) instead of the 1M workload')
    ap.add_argument('--cpu-threads', type=int, default=16,
                    help='threads of the second (OpenMP) CPU baseline; 0 skips it')
    ap.add_argument('--profile-pair', type=int, default=0, metavar='K',
                    help='only run K pair passes per precision and a calibration copy (for counter runs)')
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) != 1:
        print('bench.py: this build runs on one GPU; the multi-GPU (slab-decomposed) path is not built. '
              'Run with --gpus 1.', file=sys.stderr)
        sys.exit(2)

    from hipsph import scenarios
    sc = scenarios.dam_break_c1() if a.small else scenarios.dam_break_c2(million=True)
    print(f'# workload {sc.name}: {sc.n} particles ({sc.n_fluid} fluid)', file=sys.stderr, flush=True)
    if a.profile_pair:
        profile_pair(sc, a.profile_pair)
        return

    r32 = run_gpu(sc, 'fp32c', a.steps, a.warmup, a.graph_steps)
    print(f'# fp32c {r32["ms_per_step"]:.4f} ms/step, pair {r32["pair_ms"]:.4f} ms', file=sys.stderr, flush=True)
    r64 = run_gpu(sc, 'fp64', a.steps, a.warmup, a.graph_steps)
    print(f'# fp64 {r64["ms_per_step"]:.4f} ms/step, pair {r64["pair_ms"]:.4f} ms', file=sys.stderr, flush=True)

    n, nf = sc.n, sc.n_fluid
    cells = r32['grid']['n_cells']
    value = nf * a.steps / r32['elapsed_s']
    b32 = algorithmic_pair_bytes('fp32c', n, nf, cells)
    b64 = algorithmic_pair_bytes('fp64', n, nf, cells)
    ach32 = b32 / (r32['pair_ms'] * 1e-3) / 1e9
    ach64 = b64 / (r64['pair_ms'] * 1e-3) / 1e9
    pairs = r64['pairs_per_step']

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
                   'integrator': 'pec', 'pair_precision': 'fp32c (float pairs, TwoSum-compensated sums)',
                   'integrator_dtype': 'f64', 'h': sc.h, 'cells': cells, 'parallelism': 'single GPU'},
        'roofline': {'bound': 'hbm', 'achieved': ach32, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': ach32 / HBM_PEAK_GBS, 'traffic': traffic_from_profiles('fp32c'),
                     'kernel': 'k_pair_fp32c<cubic>', 'algorithmic_bytes_per_launch': b32,
                     'avg_launch_ms': r32['pair_ms'], 'launches': r32['timed_launches']},
        'cpu_baseline': cpu,
        'speedup_vs_cpu': (value / cpu['value']) if cpu else None,
        'cpu_baseline_omp': cpu_omp,
        'speedup_vs_cpu_omp': (value / cpu_omp['value']) if cpu_omp else None,
        'pairs_per_step': pairs,
        'pair_interactions_per_s': pairs * a.steps / r32['elapsed_s'],
        'stage_ms': {'pair': r32['pair_ms'], 'neighbour_update': r32['nn_ms'],
                     'rest': r32['ms_per_step'] - r32['pair_ms'] - r32['nn_ms']},
        'graph_ms_per_step': r32.get('graph_ms_per_step'),
        'fp64': {'value': nf * a.steps / r64['elapsed_s'], 'ms_per_step': r64['ms_per_step'],
                 'graph_ms_per_step': r64.get('graph_ms_per_step'),
                 'roofline': {'bound': 'hbm', 'achieved': ach64, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                              'frac': ach64 / HBM_PEAK_GBS, 'traffic': traffic_from_profiles('fp64'),
                              'kernel': 'k_pair_fp64<cubic>', 'algorithmic_bytes_per_launch': b64,
                              'avg_launch_ms': r64['pair_ms']},
                 'stage_ms': {'pair': r64['pair_ms'], 'neighbour_update': r64['nn_ms']}},
        'fwd_adjoint_over_fwd': None,
        'sim_time_s': r32['status']['t'],
    }
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
