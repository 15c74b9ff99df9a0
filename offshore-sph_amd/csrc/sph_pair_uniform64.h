// sph_pair_uniform64.h — the double pair pass for one smoothing length for
// every particle and one mass for every fluid particle, bit-identical to the
// general double pass (sph_pair_impl.h) and so to the reference's loop
// (src/Tools/SolverTools.py:120-175). Included by sph_pair_fp64.hip only
// (-ffp-contract=off).
//
// Every contribution keeps the reference's operation order; what the uniform
// case changes is everything around it:
//   * h_ij = 0.5 (h + h) = h, the re-averaged h of the viscosity likewise,
//     m_j = m: the kernel normalisation 10 / (7 pi) / (h h) (Wendland's,
//     the Gaussian's (1 / pi) / (h h)), 0.01 h h and c_ij = c0 / 2 are formed
//     once on the host with the same operations, so their bits are the ones
//     the reference forms per pair;
//   * the kernel support (2h; 3h for the Gaussian) lies within the cut-off
//     q_max h and the wall range r0 inside it: the walk lists every particle
//     within max(support, r0) (a double r^2 test, no cut-off test, no square
//     root), and a listed pair outside the support or the wall range adds
//     nothing, as in the reference (a Gaussian pair at q <= 3 is a neighbour
//     whenever q_max >= 3: the same double q decides both);
//   * records: A = (x, y), B = (vx, vy, rho, p / rho^2) with a NaN in place of
//     p / rho^2 for particles that are not fluid;
//   * the walk visits the reference's nine cells in its order (x shift outer,
//     y shift inner, each cell from its highest slot down), UNROLL records
//     loaded before they are tested.
#pragma once
#include "sph_common.h"

namespace hipsph {
namespace {

#ifndef HIPSPH_PEC64_EARLY
#define HIPSPH_PEC64_EARLY 0  // 1: the corrector's gathers before the walk (137 VGPRs, 3 waves: 10 % slower)
#endif

#ifndef HIPSPH_UNIFORM64_UNROLL
#define HIPSPH_UNIFORM64_UNROLL 4
#endif
#ifndef HIPSPH_UNIFORM64_LIST
#define HIPSPH_UNIFORM64_LIST 32
#endif
// The Gaussian's lists (about 42 contributors) fill at least once: the
// longest list that keeps the 4 workgroups per CU its registers allow (40 KB
// of LDS each; 38 beside the corrector's 512 B) flushes least often (on a 4M
// jittered lattice 40 entries take 1.59 ms, 32 1.84 ms, 24 1.96 ms).
#ifndef HIPSPH_UNIFORM64_GAUSS_LIST
#define HIPSPH_UNIFORM64_GAUSS_LIST 40
#endif
#ifndef HIPSPH_UNIFORM64_GAUSS_PEC_LIST
#define HIPSPH_UNIFORM64_GAUSS_PEC_LIST 38
#endif
constexpr int kUniform64List = HIPSPH_UNIFORM64_LIST;  // contributor-list entries per thread (4 B each)
template <int KERNEL>
__host__ __device__ constexpr int uniform64_list() {
    return KERNEL == SPH_KERNEL_GAUSSIAN ? HIPSPH_UNIFORM64_GAUSS_LIST : kUniform64List;
}
template <int KERNEL>
__host__ __device__ constexpr int uniform64_pec_list() {
    return KERNEL == SPH_KERNEL_GAUSSIAN ? HIPSPH_UNIFORM64_GAUSS_PEC_LIST : kUniform64List;
}

struct Uniform64Args {
    const double2 *A;  // (x, y)
    const double4 *B;  // (vx, vy, rho, p / rho^2); p / rho^2 is NaN for particles that are not fluid
    const int2 *cxy;
    const int *cellStart;
    double *ax, *ay, *drho;
    int n;
};

struct Uniform64Const {
    double lim2;           // max((2h)^2, r0^2) (1 + 1e-12): the listed pairs
    double h, m, alpha_k;  // alpha_k: the kernel normalisation 10 / (7 pi) / (h h) (Wendland: 9 / (4 pi) / (h h))
    double c01;            // 0.01 h h, in the reference's order (0.01 h) h
    double cij, alpha, beta;
    double r0, D, p1, p2, g, qmax;
};

__device__ inline bool not_fluid64(double prs) { return prs != prs; }

// The rates of fluid target t (bt: its B record) over a grid of ncx x ncy
// cells starting at bin `base` (0 for a single simulation, the instance's
// first bin in a batch); false when t has no neighbour and keeps its
// previous rates. MODE kContinuity sums drho only, kMomentum ax and ay only
// (the other outputs are then zero); each sum keeps the order of the fused
// pass, so the split passes give its bits.
template <int KERNEL, int MODE = kFused, int LIST = uniform64_list<KERNEL>()>
__device__ inline bool uniform64_rates(const Uniform64Args &P, const Uniform64Const &K, int ncx, int ncy,
                                       long long base, int t, const double4 &bt, unsigned *s_list, double &drho,
                                       double &ax, double &ay) {
    const double2 at = P.A[t];
    const double X = at.x, Y = at.y;
    const int2 c = P.cxy[t];
    const unsigned stride = blockDim.x;
    unsigned *const list0 = s_list + threadIdx.x;
    unsigned *lp = list0;
    double arho = 0.0, a0 = 0.0, a1 = 0.0, f0 = 0.0, f1 = 0.0;
    // The reference's expressions (Continuity.py:13-16, Momentum.py:36-56,
    // BoundaryForce.py:31-41, CubicSpline.py:53-70 / Gaussian.py:41-60 /
    // Wendland.py:32-64) for each listed pair, in listing order.
    auto flush = [&]() {
        for (unsigned *q = list0; q != lp; q += stride) {
            const unsigned u = *q;
            const double2 a = at32(P.A, u);
            const double4 b = at32(P.B, u);
            const double dx = X - a.x, dy = Y - a.y;
            const double xij2 = dx * dx + dy * dy;
            const double r = sqrt(xij2);
            if (!not_fluid64(b.w)) {
                const double qq = r / K.h;
                double dwx, dwy;
                if (KERNEL == SPH_KERNEL_GAUSSIAN) {
                    const double tmp = r * K.h;
                    if (!(tmp > 1e-12 && qq <= 3)) continue;
                    const double dwdq = -2 * qq * K.alpha_k * exp(-qq * qq);
                    const double f = dwdq / tmp;
                    dwx = f * dx;
                    dwy = f * dy;
                } else {
                    double grad;
                    if (KERNEL == SPH_KERNEL_WENDLAND) {
                        if ((qq >= 2.0) || (r < 1e-10)) continue;
                        const double inner = 1.0 - 0.5 * qq;
                        grad = ipow(inner, 5.0) * (-14.0 / 3.0) * qq * (1 + 2.5 * qq);
                    } else {
                        if ((qq > 2.0) || (r < 1e-10)) continue;
                        grad = (qq > 1.0) ? -0.75 * ((2.0 - qq) * (2.0 - qq)) : -3.0 * qq * (1.0 - 0.75 * qq);
                    }
                    const double norm = K.alpha_k * grad / (K.h * r);
                    dwx = norm * dx;
                    dwy = norm * dy;
                }
                const double vx = bt.x - b.x, vy = bt.y - b.y;
                const double dot = vx * dwx + vy * dwy;
                if constexpr (MODE != kMomentum) arho += K.m * dot;
                if constexpr (MODE == kContinuity) continue;
                const double dv = vx * dx + vy * dy;
                double PI_ij = 0.0;
                if (dv < 0) {
                    const double rhoij = 0.5 * (bt.z + b.z);
                    const double muij = K.h * dv / (r * r + K.c01);
                    const double ppij = muij * (K.beta * muij - K.alpha * K.cij);
                    PI_ij = ppij / rhoij;
                }
                const double factor = bt.w + b.w + PI_ij;
                a0 += -K.m * factor * dwx;
                a1 += -K.m * factor * dwy;
            } else if (MODE != kContinuity && !(r > K.r0) && r > 1e-12) {
                const double frac = K.r0 / r;
                const double tmp = ipow(frac, K.p1) - ipow(frac, K.p2);
                const double fac = K.D * tmp;
                f0 += fac * dx / (r * r);
                f1 += fac * dy / (r * r);
            }
        }
        lp = list0;
    };
    constexpr int U = HIPSPH_UNIFORM64_UNROLL;
    static_assert(U >= 1 && U < LIST, "unroll within the list");
    unsigned *const full = list0 + (LIST - U) * stride;  // past this, U appends could overflow
    // the Gaussian's lists all fill: every active lane flushes when one must,
    // one flush loop for the wave instead of one per lane
    auto overflow = [&]() {
        if constexpr (KERNEL == SPH_KERNEL_GAUSSIAN) return __ballot(lp > full) != 0;
        return lp > full;
    };
    auto test = [&](double2 a, int k) {
        const double dx = X - a.x, dy = Y - a.y;
        const double r2 = dx * dx + dy * dy;
        // written whether or not it is kept (within the list's bounds either
        // way): no branch per candidate
        *lp = (unsigned)k;
        lp += (r2 <= K.lim2 && r2 > 0.0) ? stride : 0u;
    };
    // src/Tools/NNLinkedList.py:50-56: x shift outer, y shift inner, cells at
    // the grid's edge dropped; each cell from its highest slot down (:129-141)
    for (int da = -1; da <= 1; ++da) {
        const int cx = c.x + da;
        if (!(ncx > cx && cx > -1)) continue;
        for (int db = -1; db <= 1; ++db) {
            const int cy = c.y + db;
            if (!(ncy > cy && cy > -1)) continue;
            const long long cell = base + cx + (long long)ncx * cy;
            const int s = P.cellStart[cell], e = P.cellStart[cell + 1];
            int k = e - 1;
            for (; k - U + 1 >= s; k -= U) {
                if (overflow()) flush();
                double2 a[U];
#pragma unroll
                for (int j = 0; j < U; ++j) a[j] = at32j(P.A, (unsigned)k, -j);
#pragma unroll
                for (int j = 0; j < U; ++j) test(a[j], k - j);
            }
            if (k >= s) {
                if (overflow()) flush();
                for (; k >= s; --k) test(at32(P.A, (unsigned)k), k);  // fewer than U
            }
        }
    }
    flush();
    // a target whose own cell is in the grid finds itself (q = 0): it has a
    // neighbour; one outside counts them (SolverTools.py:166-169)
    if (!(c.x >= 0 && c.x < ncx && c.y >= 0 && c.y < ncy)) {
        const double lim = (K.qmax * K.h) * (K.qmax * K.h);
        int cnt = 0;
        for (int da = -1; da <= 1; ++da) {
            const int cx = c.x + da;
            if (!(ncx > cx && cx > -1)) continue;
            for (int db = -1; db <= 1; ++db) {
                const int cy = c.y + db;
                if (!(ncy > cy && cy > -1)) continue;
                const long long cell = base + cx + (long long)ncx * cy;
                for (int k = P.cellStart[cell]; k < P.cellStart[cell + 1]; ++k) {
                    const double2 a = at32(P.A, (unsigned)k);
                    const double ddx = X - a.x, ddy = Y - a.y;
                    const double r2 = ddx * ddx + ddy * ddy;
                    bool in = r2 < lim * (1.0 - 1e-12);
                    if (!in && !(r2 > lim * (1.0 + 1e-12))) in = exact_q_test(X, Y, K.h, a.x, a.y, K.h, K.qmax);
                    cnt += in;
                }
            }
        }
        if (cnt == 0) return false;
    }
    drho = arho;
    ax = a0 + f0;
    ay = MODE == kContinuity ? 0.0 : (a1 - K.g) + f1;  // WCSPH.py:169, then SolverTools.py:170-171
    return true;
}

// One target t of the pass: its rates, or nothing for a wall or a target
// that keeps its previous rates; MODE selects the rates written.
template <int KERNEL, int MODE = kFused>
__device__ inline void uniform64_target(const Uniform64Args &P, const Uniform64Const &K, int ncx, int ncy,
                                        long long base, int t, unsigned *s_list) {
    const double4 bt = P.B[t];
    if (not_fluid64(bt.w)) return;  // walls are not targets
    double drho, ax, ay;
    if (!uniform64_rates<KERNEL, MODE>(P, K, ncx, ncy, base, t, bt, s_list, drho, ax, ay)) return;
    if constexpr (MODE != kMomentum) P.drho[t] = drho;
    if constexpr (MODE != kContinuity) {
        P.ax[t] = ax;
        P.ay[t] = ay;
    }
}

template <int KERNEL, int MODE>
__global__ __launch_bounds__(256) void k_pair_uniform64(const Ctrl *ctrl, Uniform64Args P, Uniform64Const K) {
    __shared__ unsigned s_list[uniform64_list<KERNEL>() * 256];
    if (!ctrl->active) return;
    const int t = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= P.n) return;
    const Grid &g = ctrl->grid;
    uniform64_target<KERNEL, MODE>(P, K, (int)g.ncx, (int)g.ncy, 0, t, s_list);
}

// A batch target t: its instance's grid and constants. The lanes of a wave
// may hold targets of two (rarely more) instances; each pass of the loop
// takes the first remaining lane's instance, whose grid and constants are
// then the same for the whole pass (scalar registers, not one copy per
// lane), and runs the lanes of that instance. Every pass retires at least
// the lane it was chosen from.
template <int KERNEL>
__device__ inline void uniform64_target_batch(const BInst *inst, const int *instOf, const Uniform64Args &P, const Uniform64Const *Ks, int t,
                                   unsigned *s_list) {
    const int k = instOf[t];
    for (;;) {
        const int k0 = __builtin_amdgcn_readfirstlane(k);
        if (k == k0) {
            const BInst &I = inst[k0];
            if (I.active) uniform64_target<KERNEL>(P, Ks[k0], (int)I.grid.ncx, (int)I.grid.ncy, I.cellBase, t, s_list);
            break;
        }
    }
}

// The pass with PEC's corrector and the next dt's partials fused in, as the
// float pass's k_pair_uniform_pec (the corrector's arithmetic is the
// engine's, bit for bit).
template <int KERNEL>
__global__ __launch_bounds__(256) void k_pair_uniform64_pec(Ctrl *ctrl, Uniform64Args P, Uniform64Const K,
                                                            PecArgs C) {
    __shared__ unsigned s_list[uniform64_pec_list<KERNEL>() * 256];
    __shared__ double lds[kDtQ * 16];
    if (!ctrl->active) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctrl->dt_parts = gridDim.x;
        ctrl->pec_parts = 1;
    }
    const int t = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    DtVars v;
    if (t < P.n) {
        const double4 bt = P.B[t];
        if (!not_fluid64(bt.w)) {
            const Grid &g = ctrl->grid;
            const double dt = ctrl->dt, damping = ctrl->damping;
#if HIPSPH_PEC64_EARLY
            const PecStart p0 = pec_start(C, t, dt, damping);
#endif
            double drho, ax, ay;
            if (uniform64_rates<KERNEL, kFused, uniform64_pec_list<KERNEL>()>(P, K, (int)g.ncx, (int)g.ncy, 0, t, bt,
                                                                              s_list, drho, ax, ay)) {
                P.drho[t] = drho;
                P.ax[t] = ax;
                P.ay[t] = ay;
            } else {
                drho = P.drho[t];
                ax = P.ax[t];
                ay = P.ay[t];
            }
            v.add(C.hc.on ? C.hc.h : C.s.h[t], ax, ay);
#if !HIPSPH_PEC64_EARLY
            const PecStart p0 = pec_start(C, t, dt, damping);
#endif
            pec_finish(C, t, p0, ax, ay, drho, dt, damping);
        }
    }
    v.publish(C.part, lds);
}

// The batched ensemble (one thread per slot, and the persistent work-queue
// form), as the float pass's k_pair_batch_uniform[_queue].
template <int KERNEL>
__global__ __launch_bounds__(256) void k_pair_batch_uniform64(const BInst *inst, const int *instOf, Uniform64Args P,
                                                              const Uniform64Const *Ks) {
    __shared__ unsigned s_list[uniform64_list<KERNEL>() * 256];
    const int t = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= P.n) return;
    uniform64_target_batch<KERNEL>(inst, instOf, P, Ks, t, s_list);
}

template <int KERNEL>
__global__ __launch_bounds__(256) void k_pair_batch_uniform64_queue(const BInst *inst, const int *instOf,
                                                                    Uniform64Args P, const Uniform64Const *Ks,
                                                                    int *queue) {
    __shared__ unsigned s_list[uniform64_list<KERNEL>() * 256];
    __shared__ int s_tile;
    const int ntiles = (int)((P.n + (long long)blockDim.x - 1) / blockDim.x);
    const int xcd = (int)(blockIdx.x % kXcds);
    const int per = (ntiles + kXcds - 1) / kXcds;
    const int lo = xcd * per;
    const int hi = lo + per < ntiles ? lo + per : ntiles;
    int *ctr = queue + xcd * kQueueStride;
    for (;;) {
        if (threadIdx.x == 0) s_tile = lo + atomicAdd(ctr, 1);
        __syncthreads();
        const int tile = s_tile;
        __syncthreads();  // every lane has read s_tile before lane 0 writes the next
        if (tile >= hi) break;
        const int t = tile * (int)blockDim.x + (int)threadIdx.x;
        if (t < P.n) uniform64_target_batch<KERNEL>(inst, instOf, P, Ks, t, s_list);
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(queue + kXcds * kQueueStride, 1) == (int)gridDim.x - 1)
            for (int k = 0; k <= kXcds; ++k) atomicExch(queue + k * kQueueStride, 0);
    }
}


// The batched ensemble with PEC's corrector and the dt partials fused in:
// one workgroup per chunk (slots cs[c] .. cs[c + 1], at most 256 of one
// instance), so the instance's grid, constants and integrator fields are the
// same for the whole workgroup. The chunk's dt partials go to
// C.part[q * NC + c], where the batch's corrector stores them. An instance
// that is not active keeps its state; its fluid still enters the partials.
template <int KERNEL>
__global__ __launch_bounds__(256) void k_pair_batch_uniform64_pec(const BInst *inst, const int *cs, const int *chunkInst, int NC,
                                                       Uniform64Args P, const Uniform64Const *Ks, PecArgs C) {
    __shared__ unsigned s_list[uniform64_pec_list<KERNEL>() * 256];
    __shared__ double lds[kDtQ * 16];
    const int c = xcd_remap(blockIdx.x, gridDim.x);
    if (c >= NC) return;  // the whole workgroup
    const int k = chunkInst[c];
    const BInst &I = inst[k];
    DtVars v;
    const int t = cs[c] + (int)threadIdx.x;
    if (t < cs[c + 1]) {
        const double4 bt = P.B[t];
        if (!not_fluid64(bt.w)) {
            double drho, ax, ay;
            if (I.active) {
                const double dt = I.dt, damping = I.damping;
                PecArgs Ci = C;
                Ci.strict = I.strict;
#if HIPSPH_PEC64_EARLY
                const PecStart p0 = pec_start(Ci, t, dt, damping);
#endif
                if (uniform64_rates<KERNEL, kFused, uniform64_pec_list<KERNEL>()>(
                        P, Ks[k], (int)I.grid.ncx, (int)I.grid.ncy, I.cellBase, t, bt, s_list, drho, ax, ay)) {
                    P.drho[t] = drho;
                    P.ax[t] = ax;
                    P.ay[t] = ay;
                } else {
                    drho = P.drho[t];
                    ax = P.ax[t];
                    ay = P.ay[t];
                }
#if !HIPSPH_PEC64_EARLY
                const PecStart p0 = pec_start(Ci, t, dt, damping);
#endif
                pec_finish(Ci, t, p0, ax, ay, drho, dt, damping);
            } else {
                ax = P.ax[t];
                ay = P.ay[t];
            }
            v.add(I.uh_on ? I.uh : C.s.h[t], ax, ay);
        }
    }
    v.store(C.part, NC, c, lds);
    if (c == 0 && threadIdx.x == 0 && C.mark) *C.mark = 1;
}

}  // namespace
}  // namespace hipsph
