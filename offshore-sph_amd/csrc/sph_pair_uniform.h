// sph_pair_uniform.h — the float pair pass for the common case of one
// smoothing length for every particle and one mass for every fluid particle
// (every dam break and stress lattice here): the fused pass of
// sph_pair_impl.h (Tait p/rho^2, continuity, momentum + artificial
// viscosity + gravity, Lennard-Jones wall force) with the cubic spline or
// Wendland C4, in fewer instructions per candidate and per contributor.
// Included by sph_pair_fp32.hip only.
//
// What the uniform case removes:
//   * h_ij = h and m_j = m are constants: the candidate test compares r^2
//     with one constant, max(2h, r0)^2; the kernel's normalisation, 1/h and
//     the viscosity's 0.01 h^2 are hoisted.
//   * Both limits lie inside the cut-off 3h, so no candidate takes the cut-off
//     test (a target whose own cell is outside the grid counts its neighbours
//     in a separate walk, only to decide "no neighbours, keep the previous
//     rates", src/Tools/SolverTools.py:166-169).
//   * Records are 16 bytes each: A = (x, y) in double, B = (vx, vy, rho,
//     p/rho^2) in float with a NaN in place of p/rho^2 for every particle
//     that is not fluid. The walk reads A only and lists every particle
//     within max(2h, r0) of the target; a contributor reads both and takes
//     the wall force for the non-fluid ones.
//   * The candidate test runs in double (no conversions): on CDNA4 a double
//     add or FMA issues like a float one.
//   * The density and force sums take each float term into a double
//     accumulator: the term is exact in double, so the sum's error is the
//     double rounding of each add (2^-53 relative), tighter than TwoSum's
//     float compensation, in two instructions per term instead of seven.
//     The wall force joins the momentum sum (the general pass keeps it apart
//     until the end; in float the order is not the reference's anyway).
//   * The walk loads UNROLL records before testing them (no rotating
//     registers), and the list of contributors advances by pointer.
#pragma once
#include "sph_common.h"

namespace hipsph {
namespace {

#ifndef HIPSPH_PEC_EARLY
#define HIPSPH_PEC_EARLY 1  // the corrector's gathers issued before the walk (1 % faster than after it)
#endif

#ifndef HIPSPH_UNIFORM_UNROLL
#define HIPSPH_UNIFORM_UNROLL 4
#endif
// Contributor-list entries per thread; a list is flushed before a group of
// UNROLL candidates would overflow it, so a thread with more than LIST -
// UNROLL candidates flushes mid-walk while the rest of its wave waits: the
// list must hold the contributors of most targets (20 on the initial
// lattice, 18-25 once the flow has disordered it). 28 for the staged passes
// (16-bit entries beside the staged records, 5 workgroups per CU either
// way): on the 1M dam break after 4,000 steps 24 took 133 us against 102,
// the same at the start. 26 for the batch's unstaged rate passes (26 KB of
// LDS, 6 workgroups per CU).
#ifndef HIPSPH_UNIFORM_LIST
#define HIPSPH_UNIFORM_LIST 28
#endif
#ifndef HIPSPH_UNIFORM_BATCH_LIST
#define HIPSPH_UNIFORM_BATCH_LIST 26
#endif
#ifndef HIPSPH_UNIFORM_GAUSS_LIST
#define HIPSPH_UNIFORM_GAUSS_LIST 22
#endif
#ifndef HIPSPH_UNIFORM_FLUSH_ALL
#define HIPSPH_UNIFORM_FLUSH_ALL 0
#endif
constexpr int kUniformList = HIPSPH_UNIFORM_LIST;
// The Gaussian's 3h support holds about 42 contributors: every list fills
// at least once, whatever its length, so occupancy decides: 22 entries (7
// workgroups per CU, as many as its registers allow) take 0.65 ms on a 4M
// jittered lattice, 18 0.74, 26 0.68, 36 0.85, 48 1.05.
constexpr int kUniformGaussList = HIPSPH_UNIFORM_GAUSS_LIST;
template <int KERNEL>
__host__ __device__ constexpr int uniform_list() {
    return KERNEL == SPH_KERNEL_GAUSSIAN ? kUniformGaussList : kUniformList;
}
template <int KERNEL>
__host__ __device__ constexpr int uniform_batch_list() {
    return KERNEL == SPH_KERNEL_GAUSSIAN ? kUniformGaussList : HIPSPH_UNIFORM_BATCH_LIST;
}

struct UniformArgs {
    const double2 *A;  // (x, y)
    const float4 *B;   // (vx, vy, rho, p / rho^2); p / rho^2 is NaN for particles that are not fluid
    const int2 *cxy;
    const int *cellStart;
    double *ax, *ay, *drho;
    int n;
};

struct UniformConst {
    // The listed pairs. Cubic spline and Wendland: r^2 <= lim2 = max((2h)^2,
    // r0^2) (1 + 1e-5), a superset of the contributors (the terms test the
    // support). Gaussian: the reference's q = r / h <= 3, the exact double
    // test for r^2 within [lo, lim2] = (3h)^2 (1 -+ 1e-12).
    double lim2, lo;
    double hd, qmax, g;  // h, the cut-off in units of h, gravity
    float h, hinv, m, ck;  // ck: the kernel's normalisation / h^3 (dW/dr / r = ck g(q) / r)
    float c01, ac, beta;   // 0.01 h^2, alpha c0 / 2, beta
    float r0, D;
    float p1, p2;
    int stage;  // the single-simulation passes stage each workgroup's records in LDS
};

__device__ inline bool not_fluid(float prs) { return prs != prs; }

// The gradient factor g(q) of the kernel's dW/dq (zero outside the support;
// the Gaussian's listed pairs all lie inside its support).
template <int KERNEL>
__device__ inline float grad_q(float q) {
    if (KERNEL == SPH_KERNEL_GAUSSIAN) {
        return -2.0f * q * __expf(-q * q);  // src/Kernels/Gaussian.py:54
    } else if (KERNEL == SPH_KERNEL_WENDLAND) {
        const float inner = 1.0f - 0.5f * q;
        const float i2 = inner * inner;
        const float gq = i2 * i2 * inner * (-14.0f / 3.0f) * q * (1.0f + 2.5f * q);
        return q < 2.0f ? gq : 0.0f;
    } else {
        const float t = 2.0f - q;
        const float g2 = -0.75f * t * t;
        const float g1 = -3.0f * q * (1.0f - 0.75f * q);
        return q > 2.0f ? 0.0f : (q > 1.0f ? g2 : g1);
    }
}

// The terms of one listed pair (dx, dy: x_i - x_j): (m v_ij . grad W,
// momentum + viscosity along x and y) for a fluid neighbour, (0, the
// Lennard-Jones force) for a wall within r0, zero otherwise.
template <int KERNEL>
__device__ inline float3 uniform_terms(const UniformConst &K, float dx, float dy, float vxi, float vyi, float rhoi,
                                       float slf, const float4 &b) {
    const float r2 = dx * dx + dy * dy;
    const float rinv = __builtin_amdgcn_rsqf(r2);  // listed pairs have r > 0
    const float r = r2 * rinv;
    if (!not_fluid(b.w)) {
        const float q = r * K.hinv;
        float norm = K.ck * grad_q<KERNEL>(q) * rinv;
        if (KERNEL == SPH_KERNEL_GAUSSIAN && !(r * K.h > 1e-12f)) norm = 0.0f;  // Gaussian.py:48-49
        const float dwx = norm * dx, dwy = norm * dy;
        const float vx = vxi - b.x, vy = vyi - b.y;
        const float dv = vx * dx + vy * dy;
        // mu (beta mu - alpha c_ij) / rho_ij, mu = h dv / (r^2 + 0.01 h^2), one reciprocal
        auto av = [&]() {
            const float rhoij = 0.5f * (rhoi + b.z);
            const float dn = r2 + K.c01;
            const float hv = K.h * dv;
            return hv * (K.beta * hv - K.ac * dn) * __builtin_amdgcn_rcpf(dn * dn * rhoij);
        };
        float pi = 0.0f;
        // The Gaussian (no branch in its gradient) takes the term for every
        // pair and keeps it for the approaching ones (3-5 % faster there);
        // the cubic spline and Wendland branch (selects here and in the
        // support test made the cubic pass 20 % slower).
        if constexpr (KERNEL == SPH_KERNEL_GAUSSIAN) {
            const float v = av();
            pi = dv < 0.0f ? v : 0.0f;
        } else if (dv < 0.0f) {
            pi = av();
        }
        const float fm = -K.m * (slf + b.w + pi);
        return make_float3(K.m * (vx * dwx + vy * dwy), fm * dwx, fm * dwy);
    }
    if (r <= K.r0 && r > 1e-12f) {
        const float frac = K.r0 * rinv;
        const float tmp = ipow(frac, K.p1) - ipow(frac, K.p2);
        const float fac = K.D * tmp * (rinv * rinv);
        return make_float3(0.0f, fac * dx, fac * dy);
    }
    return make_float3(0.0f, 0.0f, 0.0f);
}

// The reference's neighbour count over the 3 x 3 cells of an out-of-grid
// target (q = r / h <= qmax, the exact double test in the rounding band).
__device__ inline int uniform_count(const UniformArgs &P, const UniformConst &K, double X, double Y, int2 c, int ncx,
                                    int ncy, long long base) {
    const int lo = c.x - 1 > 0 ? c.x - 1 : 0;
    const int hi = c.x + 1 < ncx - 1 ? c.x + 1 : ncx - 1;
    const double lim = (K.qmax * K.hd) * (K.qmax * K.hd);
    int cnt = 0;
    for (int db = -1; db <= 1 && lo <= hi; ++db) {
        const int cy = c.y + db;
        if (!(ncy > cy && cy > -1)) continue;
        const long long row = base + (long long)ncx * cy;
        const int s = P.cellStart[row + lo], e = P.cellStart[row + hi + 1];
        for (int k = s; k < e; ++k) {
            const double2 a = at32(P.A, (unsigned)k);
            const double dx = X - a.x, dy = Y - a.y;
            const double r2 = dx * dx + dy * dy;
            bool in = r2 < lim * (1.0 - 1e-12);
            if (!in && !(r2 > lim * (1.0 + 1e-12))) in = exact_q_test(X, Y, K.hd, a.x, a.y, K.hd, K.qmax);
            cnt += in;
        }
    }
    return cnt;
}

// The rates of fluid target t (bt: its B record) over a grid of ncx x ncy
// cells starting at bin `base` (0 for a single simulation, the instance's
// first bin in a batch); false when t has no neighbour and keeps its
// previous rates. MODE kContinuity sums drho only, kMomentum ax and ay only
// (the other outputs are then zero).
template <int KERNEL, int MODE = kFused, int LIST = uniform_list<KERNEL>()>
__device__ inline bool uniform_rates(const UniformArgs &P, const UniformConst &K, int ncx, int ncy, long long base,
                                     int t, const float4 &bt, unsigned *s_list, double &drho, double &ax, double &ay) {
    const double2 at = P.A[t];
    const double X = at.x, Y = at.y;
    const int2 c = P.cxy[t];
    const unsigned stride = blockDim.x;
    unsigned *const list0 = s_list + threadIdx.x;
    unsigned *lp = list0;
    double srho = 0.0, sax = 0.0, say = 0.0;
    const int lo = c.x - 1 > 0 ? c.x - 1 : 0;
    const int hi = c.x + 1 < ncx - 1 ? c.x + 1 : ncx - 1;
    auto flush = [&]() {
        for (unsigned *q = list0; q != lp; q += stride) {
            const unsigned u = *q;
            const double2 a = at32(P.A, u);
            const float4 b = at32(P.B, u);
            const float3 f = uniform_terms<KERNEL>(K, (float)(X - a.x), (float)(Y - a.y), bt.x, bt.y, bt.z, bt.w, b);
            if constexpr (MODE != kMomentum) srho += (double)f.x;
            if constexpr (MODE != kContinuity) {
                sax += (double)f.y;
                say += (double)f.z;
            }
        }
        lp = list0;
    };
    constexpr int U = HIPSPH_UNIFORM_UNROLL;
    static_assert(U >= 1 && U < LIST, "unroll within the list");
    unsigned *const full = list0 + (LIST - U) * stride;  // past this, U appends could overflow
    auto test = [&](double2 a, int k) {
        const double dx = X - a.x, dy = Y - a.y;
        const double r2 = dx * dx + dy * dy;
        bool in = r2 <= K.lim2 && r2 > 0.0;
        if (KERNEL == SPH_KERNEL_GAUSSIAN && in && !(r2 < K.lo)) in = exact_q_test(X, Y, K.hd, a.x, a.y, K.hd, 3.0);
        // written whether or not it is kept (within the list's bounds either
        // way): no branch per candidate
        *lp = (unsigned)k;
        lp += in ? stride : 0u;
    };
    // A list about to overflow is flushed; with FLUSH_ALL (always for the
    // Gaussian) every active lane of the wave flushes with it, so the wave
    // runs one flush loop instead of one per lane that fills at its own
    // candidate. A lane's sums keep their order either way.
    auto overflow = [&]() {
        if constexpr (HIPSPH_UNIFORM_FLUSH_ALL || KERNEL == SPH_KERNEL_GAUSSIAN) return __ballot(lp > full) != 0;
        return lp > full;
    };
    for (int db = -1; db <= 1 && lo <= hi; ++db) {
        const int cy = c.y + db;
        if (!(ncy > cy && cy > -1)) continue;
        const long long row = base + (long long)ncx * cy;
        const int s = P.cellStart[row + lo], e = P.cellStart[row + hi + 1];
        int k = s;
        for (; k + U <= e; k += U) {
            if (overflow()) flush();
            double2 a[U];
#pragma unroll
            for (int j = 0; j < U; ++j) a[j] = at32j(P.A, (unsigned)k, j);
#pragma unroll
            for (int j = 0; j < U; ++j) test(a[j], k + j);
        }
        if (k < e) {
            if (overflow()) flush();
            for (; k < e; ++k) test(at32(P.A, (unsigned)k), k);  // fewer than U
        }
    }
    flush();
    const bool inside = c.x >= 0 && c.x < ncx && c.y >= 0 && c.y < ncy;  // it finds itself: a neighbour
    if (!inside && uniform_count(P, K, X, Y, c, ncx, ncy, base) == 0) return false;  // SolverTools.py:166-169
    drho = srho;
    ax = sax;
    ay = MODE == kContinuity ? 0.0 : say - K.g;  // WCSPH.py:169
    return true;
}

// ---------------------------------------------------------------------------
// staged walk
// ---------------------------------------------------------------------------
// uniform_rates over the staged records: the same walk (rows c.y - 1 ..
// c.y + 1, each run of slots ascending, UNROLL candidates loaded before they
// are tested) and the same flushes in the same order, so the same sums. The
// list holds LDS record indices (16-bit).
template <int KERNEL, int MODE, int LIST>
__device__ inline bool uniform_rates_staged(const UniformArgs &P, const UniformConst &K, int ncx, int ncy,
                                            long long base, int t, const float4 &bt, const int rs[3], const int re[3],
                                            const Stage &S, const double2 *sA, unsigned short *s_list,
                                            double &drho, double &ax, double &ay) {
    const double2 at = P.A[t];
    const double X = at.x, Y = at.y;
    const int2 c = P.cxy[t];
    const unsigned stride = blockDim.x;
    unsigned short *const list0 = s_list + threadIdx.x;
    unsigned short *lp = list0;
    double srho = 0.0, sax = 0.0, say = 0.0;
    // LDS record i holds slot i - sh[d] of the range d it lies in
    const int sh0 = S.off[0] - S.smin[0], sh1 = S.off[1] - S.smin[1], sh2 = S.off[2] - S.smin[2];
    auto flush = [&]() {
        for (unsigned short *q = list0; q != lp; q += stride) {
            const int i = *q;
            const double2 a = sA[i];
            const int shi = i >= S.off[2] ? sh2 : (i >= S.off[1] ? sh1 : sh0);
            const float4 b = at32(P.B, (unsigned)(i - shi));
            const float3 f = uniform_terms<KERNEL>(K, (float)(X - a.x), (float)(Y - a.y), bt.x, bt.y, bt.z, bt.w, b);
            if constexpr (MODE != kMomentum) srho += (double)f.x;
            if constexpr (MODE != kContinuity) {
                sax += (double)f.y;
                say += (double)f.z;
            }
        }
        lp = list0;
    };
    constexpr int U = HIPSPH_UNIFORM_UNROLL;
    static_assert(U >= 1 && U < LIST, "unroll within the list");
    unsigned short *const full = list0 + (LIST - U) * stride;  // past this, U appends could overflow
    auto test = [&](double2 a, int i) {
        const double dx = X - a.x, dy = Y - a.y;
        const double r2 = dx * dx + dy * dy;
        bool in = r2 <= K.lim2 && r2 > 0.0;
        if (KERNEL == SPH_KERNEL_GAUSSIAN && in && !(r2 < K.lo)) in = exact_q_test(X, Y, K.hd, a.x, a.y, K.hd, 3.0);
        *lp = (unsigned short)i;  // written whether or not it is kept (within the list's bounds)
        lp += in ? stride : 0u;
    };
    auto overflow = [&]() {
        if constexpr (HIPSPH_UNIFORM_FLUSH_ALL || KERNEL == SPH_KERNEL_GAUSSIAN) return __ballot(lp > full) != 0;
        return lp > full;
    };
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (!(rs[d] < re[d])) continue;
        const int shd = d == 0 ? sh0 : (d == 1 ? sh1 : sh2);
        const int e = re[d] + shd;  // the run in LDS record indices
        int i = rs[d] + shd;
        for (; i + U <= e; i += U) {
            if (overflow()) flush();
            double2 a[U];
#pragma unroll
            for (int j = 0; j < U; ++j) a[j] = sA[i + j];
#pragma unroll
            for (int j = 0; j < U; ++j) test(a[j], i + j);
        }
        if (i < e) {
            if (overflow()) flush();
            for (; i < e; ++i) test(sA[i], i);  // fewer than U
        }
    }
    flush();
    const bool inside = c.x >= 0 && c.x < ncx && c.y >= 0 && c.y < ncy;  // it finds itself: a neighbour
    if (!inside && uniform_count(P, K, X, Y, c, ncx, ncy, base) == 0) return false;  // SolverTools.py:166-169
    drho = srho;
    ax = sax;
    ay = MODE == kContinuity ? 0.0 : say - K.g;  // WCSPH.py:169
    return true;
}

// LDS of a staged pass with a LIST-entry list: the staged records and a
// 16-bit list, or (a workgroup that walks global memory) a 32-bit list.
template <int LIST>
constexpr int stage_bytes() {
    constexpr int staged = kStageCap * 16 + LIST * 256 * 2;
    constexpr int global = LIST * 256 * 4;
    return staged > global ? staged : global;
}

// One workgroup's walk, staged or not: every thread of the workgroup
// constructs it (the plan and the copy have barriers); rates() then gives a
// walking thread's rates as uniform_rates does.
template <int KERNEL, int MODE, int LIST>
struct StagedWalk {
    unsigned char *mem;
    int rs[3] = {0, 0, 0}, re[3] = {0, 0, 0};
    Stage S{};
    bool staged = false;

    long long base;  // the grid's first bin (a batch instance's cellBase)

    __device__ StagedWalk(const UniformArgs &P, const UniformConst &K, const Grid &g, int ncx, int ncy,
                          long long base_, int t, bool walks, unsigned char *s_mem, int *s_red)
        : mem(s_mem), base(base_) {
        if (!K.stage) return;  // uniform over the workgroup
        if (walks) {
            const double2 at = P.A[t];
            const int2 c = P.cxy[t];
            row_runs(P, ncx, ncy, base, c, far_corners(at.x, at.y, c, g, K.lim2), rs, re);
        }
        S = stage_plan(rs, re, s_red);
        staged = S.total <= kStageCap;
        if (staged) stage_copy(P, S, reinterpret_cast<double2 *>(mem));
    }

    __device__ bool rates(const UniformArgs &P, const UniformConst &K, int ncx, int ncy, int t, const float4 &bt,
                          double &drho, double &ax, double &ay) {
        if (staged)
            return uniform_rates_staged<KERNEL, MODE, LIST>(P, K, ncx, ncy, base, t, bt, rs, re, S,
                                                            reinterpret_cast<const double2 *>(mem),
                                                            reinterpret_cast<unsigned short *>(mem + kStageCap * 16),
                                                            drho, ax, ay);
        return uniform_rates<KERNEL, MODE, LIST>(P, K, ncx, ncy, base, t, bt, reinterpret_cast<unsigned *>(mem),
                                                 drho, ax, ay);
    }
};

// One target t of the batch's unstaged rate pass: its rates, or nothing for
// a wall or a target that keeps its previous rates; MODE selects the rates
// written.
template <int KERNEL, int MODE = kFused>
__device__ inline void uniform_target(const UniformArgs &P, const UniformConst &K, int ncx, int ncy, long long base,
                                      int t, unsigned *s_list) {
    const float4 bt = P.B[t];
    if (not_fluid(bt.w)) return;  // walls are not targets
    double drho, ax, ay;
    if (!uniform_rates<KERNEL, MODE, uniform_batch_list<KERNEL>()>(P, K, ncx, ncy, base, t, bt, s_list, drho, ax, ay))
        return;
    if constexpr (MODE != kMomentum) P.drho[t] = drho;
    if constexpr (MODE != kContinuity) {
        P.ax[t] = ax;
        P.ay[t] = ay;
    }
}

template <int KERNEL, int MODE>
__global__ __launch_bounds__(256) void k_pair_uniform(const Ctrl *ctrl, UniformArgs P, UniformConst K) {
    constexpr int LIST = uniform_list<KERNEL>();
    __shared__ __align__(16) unsigned char s_mem[stage_bytes<LIST>()];
    __shared__ int s_red[6 * 4];
    if (!ctrl->active) return;
    const int t = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const Grid &g = ctrl->grid;
    const int ncx = (int)g.ncx, ncy = (int)g.ncy;
    float4 bt = make_float4(0.0f, 0.0f, 0.0f, __builtin_nanf(""));
    if (t < P.n) bt = P.B[t];
    const bool walks = t < P.n && !not_fluid(bt.w);  // walls are not targets
    StagedWalk<KERNEL, MODE, LIST> W(P, K, g, ncx, ncy, 0, t, walks, s_mem, s_red);
    double drho, ax, ay;
    if (!walks || !W.rates(P, K, ncx, ncy, t, bt, drho, ax, ay)) return;
    if constexpr (MODE != kMomentum) P.drho[t] = drho;
    if constexpr (MODE != kContinuity) {
        P.ax[t] = ax;
        P.ay[t] = ay;
    }
}

// A batch target t: its instance's grid and constants. The lanes of a wave
// may hold targets of two (rarely more) instances; each pass of the loop
// takes the first remaining lane's instance, whose grid and constants are
// then the same for the whole pass (scalar registers, not one copy per
// lane), and runs the lanes of that instance. Every pass retires at least
// the lane it was chosen from.
template <int KERNEL>
__device__ inline void uniform_target_batch(const BInst *inst, const int *instOf, const UniformArgs &P, const UniformConst *Ks, int t,
                                   unsigned *s_list) {
    const int k = instOf[t];
    for (;;) {
        const int k0 = __builtin_amdgcn_readfirstlane(k);
        if (k == k0) {
            const BInst &I = inst[k0];
            if (I.active) uniform_target<KERNEL>(P, Ks[k0], (int)I.grid.ncx, (int)I.grid.ncy, I.cellBase, t, s_list);
            break;
        }
    }
}

// The pass with PEC's corrector and the next dt's partials fused in (a PEC
// step whose reorder ran lean, no Newmark-beta particles): each fluid target
// takes its rates (or keeps the previous ones, which the reorder left in its
// slot), then the corrector's update of its slot (src/Integrators/PEC.py:61-
// 89, double arithmetic without contraction) and the time-step reduction's
// terms of the new rates (TimeStep.py:58-90). The workgroup's dt partials go
// to C.part[q * gridDim.x + blockIdx.x]; ctrl->dt_parts records the count
// and ctrl->pec_parts that the corrector moved the fluid. With K.stage the
// workgroup stages its walk's records in LDS first (see stage_plan); a
// workgroup whose ranges exceed the stage, or K.stage = 0, walks global
// memory. The sums are the same either way.
template <int KERNEL>
__global__ __launch_bounds__(256) void k_pair_uniform_pec(Ctrl *ctrl, UniformArgs P, UniformConst K, PecArgs C) {
    constexpr int LIST = uniform_list<KERNEL>();
    __shared__ __align__(16) unsigned char s_mem[stage_bytes<LIST>()];
    __shared__ double lds[kDtQ * 16];
    __shared__ int s_red[6 * 4];
    if (!ctrl->active) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctrl->dt_parts = gridDim.x;
        ctrl->pec_parts = 1;
    }
    const int t = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const Grid &g = ctrl->grid;
    const int ncx = (int)g.ncx, ncy = (int)g.ncy;
    float4 bt = make_float4(0.0f, 0.0f, 0.0f, __builtin_nanf(""));
    if (t < P.n) bt = P.B[t];
    const bool walks = t < P.n && !not_fluid(bt.w);
    StagedWalk<KERNEL, kFused, LIST> W(P, K, g, ncx, ncy, 0, t, walks, s_mem, s_red);
    DtVars v;
    if (t < P.n) {
        if (walks) {
            const double dt = ctrl->dt, damping = ctrl->damping;
#if HIPSPH_PEC_EARLY
            const PecStart p0 = pec_start(C, t, dt, damping);
#endif
            double drho, ax, ay;
            if (W.rates(P, K, ncx, ncy, t, bt, drho, ax, ay)) {
                P.drho[t] = drho;
                P.ax[t] = ax;
                P.ay[t] = ay;
            } else {
                drho = P.drho[t];
                ax = P.ax[t];
                ay = P.ay[t];
            }
            v.add(C.hc.on ? C.hc.h : C.s.h[t], ax, ay);
#if !HIPSPH_PEC_EARLY
            const PecStart p0 = pec_start(C, t, dt, damping);
#endif
            pec_finish(C, t, p0, ax, ay, drho, dt, damping);
        }
    }
    v.publish(C.part, lds);
}

// The batched ensemble: each target with its own instance's grid and
// constants (Ks[k] for instance k); targets of finished or halted instances
// do nothing.
template <int KERNEL>
__global__ __launch_bounds__(256) void k_pair_batch_uniform(const BInst *inst, const int *instOf, UniformArgs P,
                                                            const UniformConst *Ks) {
    __shared__ unsigned s_list[uniform_batch_list<KERNEL>() * 256];
    const int t = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= P.n) return;
    uniform_target_batch<KERNEL>(inst, instOf, P, Ks, t, s_list);
}

// The same as a persistent kernel draining a per-XCD work queue of tiles of
// blockDim.x slots (as k_pair_batch_queue; 9 counters kQueueStride ints
// apart, zero before the first launch, cleared by the last workgroup out).
template <int KERNEL>
__global__ __launch_bounds__(256) void k_pair_batch_uniform_queue(const BInst *inst, const int *instOf, UniformArgs P,
                                                                  const UniformConst *Ks, int *queue) {
    __shared__ unsigned s_list[uniform_batch_list<KERNEL>() * 256];
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
        if (t < P.n) uniform_target_batch<KERNEL>(inst, instOf, P, Ks, t, s_list);
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
// The tile's walk is staged in LDS as the single-simulation pass's (one
// instance per tile: its grid and bins).
template <int KERNEL>
__global__ __launch_bounds__(256) void k_pair_batch_uniform_pec(const BInst *inst, const int *cs, const int *chunkInst, int NC,
                                                       UniformArgs P, const UniformConst *Ks, PecArgs C) {
    constexpr int LIST = uniform_list<KERNEL>();
    __shared__ __align__(16) unsigned char s_mem[stage_bytes<LIST>()];
    __shared__ double lds[kDtQ * 16];
    __shared__ int s_red[6 * 4];
    const int c = xcd_remap(blockIdx.x, gridDim.x);
    if (c >= NC) return;  // the whole workgroup
    const int k = chunkInst[c];
    const BInst &I = inst[k];
    const Grid &g = I.grid;
    const int ncx = (int)g.ncx, ncy = (int)g.ncy;
    const int t = cs[c] + (int)threadIdx.x;
    const bool in = t < cs[c + 1];
    float4 bt = make_float4(0.0f, 0.0f, 0.0f, __builtin_nanf(""));
    if (in) bt = P.B[t];
    StagedWalk<KERNEL, kFused, LIST> W(P, Ks[k], g, ncx, ncy, I.cellBase, t, in && I.active && !not_fluid(bt.w),
                                       s_mem, s_red);
    DtVars v;
    if (in) {
        if (!not_fluid(bt.w)) {
            double drho, ax, ay;
            if (I.active) {
                const double dt = I.dt, damping = I.damping;
                PecArgs Ci = C;
                Ci.strict = I.strict;
#if HIPSPH_PEC_EARLY
                const PecStart p0 = pec_start(Ci, t, dt, damping);
#endif
                if (W.rates(P, Ks[k], ncx, ncy, t, bt, drho, ax, ay)) {
                    P.drho[t] = drho;
                    P.ax[t] = ax;
                    P.ay[t] = ay;
                } else {
                    drho = P.drho[t];
                    ax = P.ax[t];
                    ay = P.ay[t];
                }
#if !HIPSPH_PEC_EARLY
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
