// sph_common.h — device code shared by the single-simulation engine and the
// batched ensemble: workgroup reductions, the PEC half-step expressions, the
// Tait pressure, and the cell-sort pipeline (prefix scan of bin counts,
// scatter, per-bin rank sort). Included by more than one translation unit:
// everything is in an anonymous namespace, so each gets its own copy.
#pragma once
#include <type_traits>

#include "sph_internal.h"

namespace hipsph {
namespace {

// Host: calls f with the kernel id as a compile-time constant
// (std::integral_constant<int, id>), for launchers instantiated per kernel.
template <typename F>
inline void with_kernel(int kernel, F &&f) {
    switch (kernel) {
        case SPH_KERNEL_WENDLAND: f(std::integral_constant<int, SPH_KERNEL_WENDLAND>{}); break;
        case SPH_KERNEL_GAUSSIAN: f(std::integral_constant<int, SPH_KERNEL_GAUSSIAN>{}); break;
        default: f(std::integral_constant<int, SPH_KERNEL_CUBIC>{}); break;
    }
}

// ---------------------------------------------------------------------------
// reductions (wavefront shuffles, then LDS across the 4 waves of a block)
// ---------------------------------------------------------------------------
template <typename T, typename Op>
__device__ inline T wave_reduce(T v, Op op) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o, 64));
    return v;
}

template <typename T, typename Op>
__device__ inline T block_reduce(T v, Op op, T *lds) {
    v = wave_reduce(v, op);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) lds[w] = v;
    __syncthreads();
    if (w == 0) {
        v = (lane < (int)(blockDim.x >> 6)) ? lds[lane] : Op::identity();
        v = wave_reduce(v, op);
    }
    __syncthreads();
    return v;  // valid in wave 0
}

struct MinOp {
    __device__ static double identity() { return INFINITY; }
    __device__ double operator()(double a, double b) const { return b < a ? b : a; }
};
struct MaxOp {
    __device__ static double identity() { return -INFINITY; }
    __device__ double operator()(double a, double b) const { return b > a ? b : a; }
};
struct AddOp {
    __device__ static double identity() { return 0.0; }
    __device__ double operator()(double a, double b) const { return a + b; }
};

__device__ inline double fold_op(int kind, double a, double b) {
    return kind == 0 ? (b < a ? b : a) : kind == 1 ? (b > a ? b : a) : a + b;
}

// Fold NQ quantities over a whole workgroup (up to 16 wavefronts): a shuffle
// tree inside each wavefront, one LDS round, then wavefront 0 folds the
// per-wave values of four quantities at a time, one quantity per group of 16
// lanes (the single-workgroup finalisers are latency-bound: no serial loop
// over the waves). kind: 0 min, 1 max, 2 sum; every sum folded here is a
// count, exact in any order. The result is valid in thread 0. lds holds
// NQ * 16 doubles.
template <int NQ>
__device__ inline void block_fold(double (&v)[NQ], const int (&kind)[NQ], double *lds) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[q] = fold_op(kind[q], v[q], __shfl_xor(v[q], o, 64));
        if (lane == 0) lds[q * 16 + w] = v[q];
    }
    __syncthreads();
    if (w != 0) return;
#pragma unroll
    for (int q0 = 0; q0 < NQ; q0 += 4) {
        const int q = q0 + (lane >> 4), ww = lane & 15;
        int kd = 2;
#pragma unroll
        for (int j = 0; j < NQ; ++j) kd = q == j ? kind[j] : kd;
        double x = kd == 0 ? INFINITY : (kd == 1 ? -INFINITY : 0.0);
        if (q < NQ && ww < nw) x = lds[q * 16 + ww];
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) x = fold_op(kd, x, __shfl_xor(x, o, 64));
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (q0 + j < NQ) v[q0 + j] = __shfl(x, j * 16, 64);
    }
}

// Per-workgroup partials of NQ quantities in one LDS round (block_fold):
// thread 0 stores part[q * R + k]. lds holds NQ * 16 doubles.
template <int NQ>
__device__ inline void block_partials(double (&v)[NQ], const int (&kind)[NQ], double *lds, double *part, int R,
                                      int k) {
    block_fold<NQ>(v, kind, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) part[q * R + k] = v[q];
    }
}

// src/Integrators/PEC.py:32-59 (useXSPH = False branch), one particle. The
// fused step evaluates the same expressions inside the neighbour update.
__device__ inline double pec_half_pos(double x, double v, double dt) { return x + 0.5 * dt * v; }
__device__ inline double pec_half_vel(double v, double a, double dt, double damping) {
#pragma clang fp contract(off)  // the integrator's arithmetic in every unit that inlines it
    return (v + 0.5 * dt * a) / (1 + 0.5 * damping);
}

// src/Integrators/PEC.py:61-89, one fluid particle from its start-of-step
// state (x0 ... rho0), the velocity that moved it in the predictor (mvx, mvy:
// the half-step velocity, or the XSPH velocity with useXSPH) and this step's
// rates: writes x, y, vx, vy, rho of slot j.
__device__ inline void pec_correct(const State &s, int j, double x0, double y0, double vx0, double vy0, double rho0,
                                   double mvx, double mvy, double ax, double ay, double drho, double dt,
                                   double damping, int strict) {
#pragma clang fp contract(off)
    double mid_x = x0 + 0.5 * dt * mvx;
    double mid_y = y0 + 0.5 * dt * mvy;
    double mid_vx = (vx0 + 0.5 * dt * ax) / (1 + 0.5 * damping);
    double mid_vy = (vy0 + 0.5 * dt * ay) / (1 + 0.5 * damping);
    s.x[j] = 2 * mid_x - x0;
    s.y[j] = 2 * mid_y - y0;
    s.vx[j] = 2 * mid_vx - vx0;
    s.vy[j] = 2 * mid_vy - vy0;
    double mid_rho = rho0 + 0.5 * dt * drho;
    double r = 2 * mid_rho - rho0;
    if (strict && r < 0.0) r = 0.0;
    s.rho[j] = r;
}

// The corrector of fluid slot j after a lean reorder: the start-of-step state
// from the previous buffer at the particle's previous slot, the half-step
// velocity formed again there (the predictor's expression). Loaded by
// pec_start (before the pair walk, so that the gathers are in flight under
// it), applied by pec_finish.
struct PecStart {
    double x0, y0, vx0, vy0, rho0, vxh, vyh;
};
__device__ inline PecStart pec_start(const PecArgs &C, int j, double dt, double damping) {
    const int q = C.perm[j];
    PecStart p;
    p.x0 = C.o.x[q];
    p.y0 = C.o.y[q];
    p.vx0 = C.o.vx[q];
    p.vy0 = C.o.vy[q];
    p.rho0 = C.o.rho[q];
    p.vxh = pec_half_vel(p.vx0, C.o.ax[q], dt, damping);
    p.vyh = pec_half_vel(p.vy0, C.o.ay[q], dt, damping);
    return p;
}
__device__ inline void pec_finish(const PecArgs &C, int j, const PecStart &p, double ax, double ay, double drho,
                                  double dt, double damping) {
    pec_correct(C.s, j, p.x0, p.y0, p.vx0, p.vy0, p.rho0, p.vxh, p.vyh, ax, ay, drho, dt, damping, C.strict);
}

// time step: src/Equations/TimeStep.py:58-90 computeVars over the fluid
// (min h, max |a|^2, count, non-finite count), one workgroup's partials.
struct DtVars {
    double mn = INFINITY, mx = -INFINITY, cnt = 0.0, bad = 0.0;
    __device__ void add(double h, double ax, double ay) {
#pragma clang fp contract(off)
        double a2 = ax * ax + ay * ay;
        if (h < mn) mn = h;
        if (a2 > mx) mx = a2;
        if (!(a2 == a2)) bad += 1.0;
        cnt += 1.0;
    }
    // this workgroup's partials as entry k of R (one LDS round, plain
    // stores: part[q * R + k]); lds holds kDtQ * 16 doubles
    __device__ void store(double *part, int R, int k, double *lds) {
        const int kind[kDtQ] = {0, 1, 2, 2};
        double v[kDtQ] = {mn, mx, cnt, bad};
        block_partials<kDtQ>(v, kind, lds, part, R, k);
    }
    __device__ void publish(double *part, double *lds) { store(part, gridDim.x, blockIdx.x, lds); }
};
__device__ inline double pec_half_rho(double rho, double drho, double dt, int strict) {
    double r = rho + 0.5 * dt * drho;
    if (strict && r < 0.0) r = 0.0;
    return r;
}

// src/Integrators/NewmarkBeta.py:41-66, one coupled particle: the predictor
// x + v dt + a (1/2 - beta) dt^2, v + a (1 - gamma) dt and the corrector
// x + a beta dt^2, v + a gamma dt, left to right, dt^2 the product dt dt (a
// power with the integer literal 2 in the numba source).
__device__ inline double nm_pred_pos(double x, double v, double a, double beta, double dt) {
    return x + v * dt + a * (0.5 - beta) * (dt * dt);
}
__device__ inline double nm_pred_vel(double v, double a, double gamma, double dt) { return v + a * (1 - gamma) * dt; }
__device__ inline double nm_corr_pos(double x, double a, double beta, double dt) { return x + a * beta * (dt * dt); }
__device__ inline double nm_corr_vel(double v, double a, double gamma, double dt) { return v + a * gamma * dt; }

// Exclusive prefix scan of counts[0 .. ncells] into cellStart, in tiles of
// kScanTile bins: the tile sums, then each tile's scan offset by the sum of
// the tile sums before it (a few hundred integers, read by that tile's
// workgroup: no separate pass over the tile sums).
__device__ inline int block_scan_excl(int v, int *lds, int *total) {
    // inclusive scan inside each wave
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int s = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int u = __shfl_up(s, o, 64);
        if (lane >= o) s += u;
    }
    if (lane == 63) lds[w] = s;
    __syncthreads();
    int nw = blockDim.x >> 6;
    int off = 0, tot = 0;
    for (int k = 0; k < nw; ++k) {
        int ws = lds[k];
        if (k < w) off += ws;
        tot += ws;
    }
    __syncthreads();
    *total = tot;
    return off + s - v;
}

__global__ void k_scan_reduce(const Ctrl *ctrl, const int *counts, int *tileSums) {
    if (!ctrl->active) return;
    long long nbins = ctrl->grid.ncells + 1;
    long long ntiles = (nbins + kScanTile - 1) / kScanTile;
    __shared__ double lds[8];
    for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        long long base = tile * kScanTile;
        double s = 0.0;
        for (int k = 0; k < kScanItems; ++k) {
            long long b = base + (long long)threadIdx.x * kScanItems + k;
            if (b < nbins) s += (double)counts[b];
        }
        s = block_reduce(s, AddOp(), lds);
        if (threadIdx.x == 0) tileSums[tile] = (int)s;
    }
}

__global__ void k_scan_apply(const Ctrl *ctrl, int *counts, const int *tileSums, int *cellStart) {
    if (!ctrl->active) return;
    long long nbins = ctrl->grid.ncells + 1;
    long long ntiles = (nbins + kScanTile - 1) / kScanTile;
    __shared__ int lds[16];
    for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        long long base = tile * kScanTile;
        int v[kScanItems];
        int sum = 0;
        long long b0 = base + (long long)threadIdx.x * kScanItems;
        for (int k = 0; k < kScanItems; ++k) {
            long long b = b0 + k;
            v[k] = b < nbins ? counts[b] : 0;
            sum += v[k];
        }
        int pre = 0;
        for (long long j = threadIdx.x; j < tile; j += blockDim.x) pre += tileSums[j];
        int tot;
        block_scan_excl(pre, lds, &pre);
        int ex = block_scan_excl(sum, lds, &tot) + pre;
        for (int k = 0; k < kScanItems; ++k) {
            long long b = b0 + k;
            if (b < nbins) {
                cellStart[b] = ex;
                counts[b] = 0;  // ready for the next update (no memset node in a graph)
            }
            ex += v[k];
        }
        if (tile == ntiles - 1 && threadIdx.x == blockDim.x - 1) cellStart[nbins] = ex;
    }
}

__global__ void k_scatter(const Ctrl *ctrl, const int *cellStart, const int *bin, const int *rank, const int *id,
                          int *perm, int *permKey, int n) {
    if (!ctrl->active) return;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int pos = cellStart[bin[t]] + rank[t];
    perm[pos] = t;
    permKey[pos] = id[t];
}

// Sort the slots [s, e) of one bin by key (the reference index; keys are
// distinct) with one wavefront: each element's destination is the number of
// smaller keys in the bin. Bins of up to 64 sort in registers (one load and
// one store round); larger ones rank through global memory and stage the
// result in tmpKey/tmpVal before copying it back.
__device__ void wave_sort_bin(int s, int e, int *perm, int *permKey, int *tmpKey, int *tmpVal) {
    const int lane = threadIdx.x & 63;
    const int len = e - s;
    if (len <= 64) {
        const int key = lane < len ? permKey[s + lane] : 0x7fffffff;
        const int val = lane < len ? perm[s + lane] : 0;
        int r = 0;
        for (int j = 0; j < len; ++j) r += __shfl(key, j, 64) < key;
        if (lane < len) {
            permKey[s + r] = key;
            perm[s + r] = val;
        }
        return;
    }
    for (int i = s + lane; i < e; i += 64) {
        const int key = permKey[i];
        int r = 0;
        for (int j = s; j < e; ++j) r += permKey[j] < key;
        tmpKey[s + r] = key;
        tmpVal[s + r] = perm[i];
    }
    __threadfence_block();
    for (int i = s + lane; i < e; i += 64) {
        permKey[i] = tmpKey[i];
        perm[i] = tmpVal[i];
    }
}

// Order each bin by the reference index, which is what the reference's
// head-inserted lists are ordered by (descending there; the pair pass walks
// the bin backwards). Only bins flagged by k_keys (filled by more than one
// run) can be out of order; most of them are cells whose run crossed a
// wavefront boundary in k_keys. Each wavefront reads the flags of 64
// consecutive bins and clears them; every lane sorts its own flagged bin.
constexpr int kLaneSort = 16;  // bins up to this size are sorted by one lane

__global__ void k_cellsort(const Ctrl *ctrl, const int *cellStart, int *perm, int *permKey, int *binflag,
                           int *tmpKey, int *tmpVal) {
    if (!ctrl->active) return;
    const long long nbins = ctrl->grid.ncells + 1;
    const int lane = threadIdx.x & 63;
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long b0 = wave * 64; b0 < nbins; b0 += nwaves * 64) {
        const long long bl = b0 + lane;
        int f = 0;
        if (bl < nbins) {
            f = binflag[bl];
            if (f) binflag[bl] = 0;
        }
        // Each lane sorts its own flagged bin of up to kLaneSort slots in
        // registers (rank = number of smaller keys; cells hold ~3-9
        // particles); longer bins go to the whole wavefront afterwards.
        const bool own = f != 0;
        int s = 0, len = 0;
        if (own) {
            s = cellStart[bl];
            len = cellStart[bl + 1] - s;
        }
        if (own && len <= kLaneSort) {
            int key[kLaneSort], val[kLaneSort];
#pragma unroll
            for (int i = 0; i < kLaneSort; ++i) {
                key[i] = i < len ? permKey[s + i] : 0x7fffffff;
                val[i] = i < len ? perm[s + i] : 0;
            }
#pragma unroll
            for (int i = 0; i < kLaneSort; ++i) {
                int r = 0;
#pragma unroll
                for (int j = 0; j < kLaneSort; ++j) r += key[j] < key[i];
                if (i < len) {
                    permKey[s + r] = key[i];
                    perm[s + r] = val[i];
                }
            }
        }
        unsigned long long m = __ballot(own && len > kLaneSort);
        while (m) {
            const int k = __ffsll((long long)m) - 1;
            m &= m - 1;
            const long long b = b0 + k;
            wave_sort_bin(cellStart[b], cellStart[b + 1], perm, permKey, tmpKey, tmpVal);
        }
    }
}

// _bin's cell of a position (src/Tools/NNLinkedList.py:146-176): floor of
// the true quotient per axis, clamped only for the int conversion (any
// clamped value is far outside the grid). The keys kernels and the reorders
// both call it, so a particle's stored cell is the one it was binned by.
__device__ inline int2 cell_of(double x, double y, const Grid &g) {
    double fx = floor((x - g.xmin) / g.cs);
    double fy = floor((y - g.ymin) / g.cs);
    // clamped so the integer conversion is defined; NaN goes below the grid
    fx = !(fx >= -1.0e9) ? -1.0e9 : (fx > 1.0e9 ? 1.0e9 : fx);
    fy = !(fy >= -1.0e9) ? -1.0e9 : (fy > 1.0e9 ? 1.0e9 : fy);
    return make_int2((int)(long long)fx, (int)(long long)fy);
}

// The border-cell bounds (after a fused PEC step; see border_bounds_ok in
// the engine): the margin test, 2 disp under the cell size by far more than
// the rounding of positions and cell arithmetic (false for +inf and NaN), on
// a grid of at least 3 x 3 cells.
__device__ inline bool border_margin_ok(const Grid &g, double disp) {
    if (g.ncx < 3 || g.ncy < 3) return false;
    const double slack = 1e-12 * (fabs(g.xmin) + fabs(g.xmax) + fabs(g.ymin) + fabs(g.ymax));
    return 2.0 * disp * (1.0 + 1e-6) + slack < g.cs * (1.0 - 1e-6);
}

// The slots of the border cells of grid g (bins from cb, offsets in
// cellStart), shared by nthreads threads (tid in [0, nthreads), nthreads a
// multiple of 64): rows 0, ncy - 2, ncy - 1 and the out-of-grid bin, one
// slot range, thread-strided (it also holds what a flat key carries past the
// last row); and for each row r <= ncy - 3 the run of columns ncx - 2,
// ncx - 1 of row r and column 0 of row r + 1 (consecutive keys: one slot
// range, one wavefront; it also holds what a flat key carries past the last
// column into the next row). take(slot) for each.
template <class F>
__device__ inline void border_walk(const int *cellStart, long long cb, const Grid &g, int tid, int nthreads,
                                   F &&take) {
    const long long ncx = g.ncx, ncy = g.ncy;
    const int a0 = cellStart[cb], a1 = cellStart[cb + ncx];
    const int b0 = cellStart[cb + (ncy - 2) * ncx], b1 = cellStart[cb + g.ncells + 1];
    const int la = a1 - a0, lab = la + (b1 - b0);
    for (int j = tid; j < lab; j += nthreads) take(j < la ? a0 + j : b0 + (j - la));
    const int lane = tid & 63;
    for (long long r = tid >> 6; r <= ncy - 3; r += nthreads >> 6) {
        const int s0 = cellStart[cb + r * ncx + ncx - 2], s1 = cellStart[cb + (r + 1) * ncx + 1];
        for (int i = s0 + lane; i < s1; i += 64) take(i);
    }
}

// Tait pressure + Pb (src/Equations/TaitEOS.py:6-31; WCSPH.py:131-149).
__device__ inline double tait_p(double rho, int8_t label, double gamma, double B, double rho0, double Pb) {
    double p = 0.0;
    if (label == kFluid) {
        double ratio = ipow(rho / rho0, gamma);
        p = (ratio - 1.0) * B;
    }
    return p + Pb;
}

// The packed pair-pass records of slot t in format fmt (kRecFp64,
// kRecFp32, kRecUniform): w = m for fluid and -label otherwise; p / rho^2
// with p the Tait pressure + Pb (src/Tools/SolverTools.py:122-127,
// src/Equations/Momentum.py:28,36). Returns p / rho^2.
__device__ inline double write_pair_records(int t, double x, double y, double vx, double vy, double rho, double m,
                                            double h, int8_t lab, const Phys &ph, int fmt, void *recA, void *recB) {
    const double p = tait_p(rho, lab, ph.gamma, ph.B, ph.rho0, ph.Pb);
    const double pr = p / (rho * rho);
    const double w = lab == kFluid ? m : -(double)lab;
    if (fmt == kRecUniform) {
        ((double2 *)recA)[t] = make_double2(x, y);
        ((float4 *)recB)[t] = make_float4((float)vx, (float)vy, (float)rho,
                                          lab == kFluid ? (float)pr : __int_as_float(0x7fc00000));
    } else if (fmt == kRecUniform64) {
        ((double2 *)recA)[t] = make_double2(x, y);
        ((double4 *)recB)[t] = make_double4(vx, vy, rho, lab == kFluid ? pr : __longlong_as_double(0x7ff8000000000000ll));
    } else if (fmt == kRecFp32) {
        Rec<float>::A ra;
        ra.x = x;
        ra.y = y;
        ra.h = (float)h;
        ra.w = (float)w;
        ra.vx = (float)vx;
        ra.vy = (float)vy;
        ((Rec<float>::A *)recA)[t] = ra;
        ((Rec<float>::B *)recB)[t] = Rec<float>::B{(float)rho, (float)pr};
    } else {
        ((Rec<double>::A *)recA)[t] = Rec<double>::A{x, y, h, w};
        ((Rec<double>::B *)recB)[t] = Rec<double>::B{vx, vy, rho, pr};
    }
    return pr;
}

// ---------------------------------------------------------------------------
// staged walks of the uniform pair passes
// ---------------------------------------------------------------------------
// Staged walk: the A records of the slot ranges a workgroup's targets walk,
// one range per row offset, are copied to LDS once (coalesced) and the walk
// and the flush read them there (the scattered 16-byte loads of the walk
// kept the vector memory path of the float pass 80 % busy); a workgroup
// whose ranges hold more than STAGE_CAP records walks global memory instead.
// 1024 (16 KB; the 1M dam break's workgroups need about 800): 5 workgroups
// per CU, as many as the fused float pass's 81 registers allow. Staging B as
// well (3 workgroups per CU) or a shorter list (a flush mid-walk in most
// threads) cost more than the walk saves. Workgroups of at most 256 threads.
#ifndef HIPSPH_STAGE_CAP
#define HIPSPH_STAGE_CAP 1024
#endif
constexpr int kStageCap = HIPSPH_STAGE_CAP;
static_assert(kStageCap > 0 && kStageCap < 65536, "LDS indices are 16-bit");

// Which diagonal neighbour cells of a target at (X, Y) in cell c lie wholly
// beyond the listing radius (r^2 > lim2 for every point of the cell): bit 0
// below left, 1 below right, 2 above left, 3 above right. With cells 2h wide
// and the radius 2h, about a fifth of them (the target near its cell's
// centre); their particles could only be candidates that fail the test, so
// skipping them leaves every sum as it was. Distances to the cell edges are
// taken 1e-9 cs short of the computed ones, far more than the rounding of
// cell_of's floor((x - xmin) / cs) can move a particle. (The staged float
// walk skips them; in the double walk the extra registers cost more than
// the skipped candidates save.)
__device__ inline unsigned far_corners(double X, double Y, int2 c, const Grid &g, double lim2) {
    const double e = 1e-9 * g.cs;
    const double x0 = g.xmin + g.cs * c.x, y0 = g.ymin + g.cs * c.y;
    double l = X - x0 - e, r = x0 + g.cs - X - e, b = Y - y0 - e, a = y0 + g.cs - Y - e;
    l = l > 0.0 ? l * l : 0.0;
    r = r > 0.0 ? r * r : 0.0;
    b = b > 0.0 ? b * b : 0.0;
    a = a > 0.0 ? a * a : 0.0;
    return (l + b > lim2 ? 1u : 0u) | (r + b > lim2 ? 2u : 0u) | (l + a > lim2 ? 4u : 0u) | (r + a > lim2 ? 8u : 0u);
}

// The runs of slots target c walks in rows c.y - 1, c.y, c.y + 1 (index d =
// 0, 1, 2): [rs[d], re[d]), empty for a row outside the grid; the diagonal
// cells in `far` (far_corners) are left out of the runs of rows 0 and 2.
template <typename Args>
__device__ inline void row_runs(const Args &P, int ncx, int ncy, long long base, int2 c, unsigned far, int rs[3],
                                int re[3]) {
    const int lo = c.x - 1 > 0 ? c.x - 1 : 0;
    const int hi = c.x + 1 < ncx - 1 ? c.x + 1 : ncx - 1;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int cy = c.y + d - 1;
        rs[d] = 0;
        re[d] = 0;
        if (lo <= hi && ncy > cy && cy > -1) {
            const unsigned f = d == 0 ? far : (d == 2 ? far >> 2 : 0u);
            const int l = (f & 1u) && lo == c.x - 1 ? lo + 1 : lo;
            const int h = (f & 2u) && hi == c.x + 1 ? hi - 1 : hi;
            const long long row = base + (long long)ncx * cy;
            rs[d] = P.cellStart[row + l];
            re[d] = P.cellStart[row + h + 1];
        }
    }
}

// The workgroup's staging plan: row range d is [smin[d], smin[d] + len[d])
// of the slots, held from LDS record off[d] on.
struct Stage {
    int smin[3], len[3], off[3], total;
};

// Folds every thread's row runs (empty for a thread without a walk) into
// the workgroup's ranges. Ends with a barrier; every thread returns the same
// plan.
__device__ inline Stage stage_plan(const int rs[3], const int re[3], int *s_red) {
    int v[6];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const bool any = rs[d] < re[d];
        v[d] = any ? rs[d] : 0x7fffffff;
        v[3 + d] = any ? re[d] : -0x7fffffff;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int a = __shfl_xor(v[d], o, 64), b = __shfl_xor(v[3 + d], o, 64);
            v[d] = a < v[d] ? a : v[d];
            v[3 + d] = b > v[3 + d] ? b : v[3 + d];
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < 6; ++q) s_red[q * 4 + w] = v[q];
    __syncthreads();
    Stage S;
    int off = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int mn = 0x7fffffff, mx = -0x7fffffff;
        for (int k = 0; k < nw; ++k) {
            const int a = s_red[d * 4 + k], b = s_red[(3 + d) * 4 + k];
            mn = a < mn ? a : mn;
            mx = b > mx ? b : mx;
        }
        S.smin[d] = mn;
        S.len[d] = mx > mn ? mx - mn : 0;
        S.off[d] = off;
        off += S.len[d];
    }
    S.total = off;
    return S;
}

// Copies the planned ranges of A into LDS. Ends with a barrier.
template <typename Args>
__device__ inline void stage_copy(const Args &P, const Stage &S, double2 *sA) {
#pragma unroll
    for (int d = 0; d < 3; ++d)
        for (int i = threadIdx.x; i < S.len[d]; i += blockDim.x)
            sA[S.off[d] + i] = at32(P.A, (unsigned)(S.smin[d] + i));
    __syncthreads();
}

}  // namespace
}  // namespace hipsph
