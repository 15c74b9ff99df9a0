// sph_internal.h — layouts shared by the translation units of libhipsph.
//
// Device memory layout (one context, n active particles):
//   * State is structure-of-arrays in CELL ORDER: after every neighbour update
//     the whole state is permuted so that particles of one grid cell are
//     contiguous and, inside a cell, sorted by their index in the reference's
//     active array (`id`). Two such buffer sets exist (ping-pong); `cur` names
//     the live one. Nothing on the step path reads the reference order; only
//     upload/download convert.
//   * Cell offsets: cellStart[0 .. ncells+1], an exclusive prefix sum of the
//     per-cell counts. Bin `ncells` collects particles whose flat key lies
//     outside the grid (the reference writes those out of bounds); no query
//     ever visits it.
//   * Control block (Ctrl): time, dt, step count, flags and the grid of the
//     last update live on the device, so a whole step runs without a host
//     round trip and can be captured in a hipGraph.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hipsph.h"

namespace hipsph {

constexpr int kFluid = 0;
constexpr int kCoupled = 3;
constexpr int kBlock = 256;        // threads per block of streaming kernels
// bins per thread in the cell scan (tiles of 1,024 bins: 438 tiles on the
// 1M dam break's 447k cells; against 2,048-bin tiles the apply kernel takes
// 5.13 us instead of 5.85 and the step 1.5 us less; 512 no better)
constexpr int kScanItems = 4;
constexpr int kScanTile = kBlock * kScanItems;
constexpr int kRedBlocks = 2048;   // max workgroups of a grid-wide reduction

struct Grid {
    double xmin, xmax, ymin, ymax, cs;
    long long ncx, ncy, ncells;
    int uniform_h;  // min h == max h over all particles
};

struct Ctrl {
    double t, dt, dt_c, dt_f;
    long long step;
    int active;          // 1 while the current step may run
    unsigned flags;      // SPH_FLAG_* bits
    Grid grid;
    double duration, fixed_dt, damping;
    double alpha, beta, co, B, Pb;
    int advance_t;       // 0 while settling (t does not advance)
    int nn_valid;        // cell order and offsets match the positions
    int nm;              // Newmark-beta on the coupled particles
    double nm_beta, nm_gamma;
    int dt_parts;        // workgroups whose dt partials the partials buffer holds (set by their kernel)
    int pec_parts;       // 1 when those partials come from the fused PEC pass (cleared by k_dt_final)
    double disp;         // bound on how far any particle's predicted position lies from where it was
                         // last binned, per axis (k_dt_final; +inf when unknown)
    double vbound;       // bound on every particle's |vx|, |vy| at the start of the step (+inf: unknown)
    double abound;       // bound on every fluid particle's |a| in the rates the step's predictor uses
    long long border_updates;  // neighbour updates whose bounds came from the border cells
};

// Grid-wide reductions (bounds, dt) run over at most kRedBlocks workgroups in
// grid-stride loops; each workgroup stores its partial result with plain
// stores into part[q * R + blockIdx.x] (q = quantity, R = workgroups of that
// launch) and a single-workgroup kernel folds the R partials. Same-address
// device-scope atomics from every workgroup serialise across the XCDs and
// cost ~10 ns each, far more than the second launch.
constexpr int kBoundsQ = 7;  // min/max x, min/max y, min/max h, non-finite count
constexpr int kDtQ = 4;      // min h, max |a|^2, fluid count, non-finite count
constexpr int kBoundsVQ = 8;  // the single engine's bounds: kBoundsQ, then the largest |vx|, |vy|

// One cell-ordered buffer set.
struct State {
    double *x, *y, *vx, *vy, *rho, *ax, *ay, *drho, *x0, *y0, *vx0, *vy0, *rho0, *m, *h;
    double *xsphx, *xsphy;  // XSPH velocity (particle_dtype xsphx, xsphy)
    int8_t *label;
    int *id;
};

// Packed pair-pass records, one per particle in cell order (written by the
// reorder kernel). w = m for fluid and -label otherwise; prs = p / rho^2 with
// p the Tait pressure + Pb (src/Equations/Momentum.py:28,36).
template <typename R>
struct Rec;
template <>
struct Rec<double> {
    struct __align__(16) A { double x, y, h, w; };
    struct __align__(16) B { double vx, vy, rho, prs; };
};
template <>
struct Rec<float> {
    struct __align__(16) A {
        double x, y;
        float h, w, vx, vy;
    };
    struct __align__(8) B { float rho, prs; };
};

// One h and one mass for every particle of the state (and no dynamic h):
// both buffers then hold them in every slot, so the reorder need not permute
// them and the bounds and dt reductions need not read them.
struct HmConst {
    int on;
    double h, m;
};

// PEC's corrector fused into the uniform pair passes (a step whose reorder
// ran lean): slot j of s (this step's cell order) came from slot perm[j] of
// o (the start-of-step buffer), which holds the start-of-step state and the
// previous rates; part receives the next dt's partials (kDtQ per workgroup).
struct PecArgs {
    State s, o;
    const int *perm;
    double *part;
    HmConst hc;
    int strict;
    int *mark = nullptr;  // when set, the pass stores 1 there (its partials come from a corrector that moved the fluid)
};

// The equations of a pair pass: all of them (continuity, momentum, wall
// force, gravity), the continuity equation alone (drho) or the rest (ax, ay).
constexpr int kFused = 0, kContinuity = 1, kMomentum = 2;

// Record formats written by the reorder: the double pass's, the general
// float pass's, and the float pass's for one h and one fluid mass
// (A = double2 (x, y), B = float4 (vx, vy, rho, p / rho^2), p / rho^2 NaN for
// particles that are not fluid).
constexpr int kRecFp64 = 0, kRecFp32 = 1, kRecUniform = 2;
// the double pass's for one h and one fluid mass: A = double2 (x, y),
// B = double4 (vx, vy, rho, p / rho^2), p / rho^2 NaN when not fluid
constexpr int kRecUniform64 = 3;
static_assert(kRecFp64 == SPH_PAIR_FP64 && kRecFp32 == SPH_PAIR_FP32C && kRecUniform == SPH_PAIR_FP32C_UNIFORM &&
                  kRecUniform64 == SPH_PAIR_FP64_UNIFORM,
              "record formats are the pair forms of the C ABI");

// Constants of the uniform float pass (one h, one fluid mass), in double.
struct UniformParams {
    // max(support^2, r0^2) (1 + 1e-5; 1 + 1e-12 for the double pass); the
    // float Gaussian pass: (3h)^2 (1 + 1e-12), with lo = (3h)^2 (1 - 1e-12)
    // the start of the band of its exact q <= 3 test (elsewhere lo = lim2)
    double lim2, lo;
    double h, m, qmax, g;
    double ck;     // the kernel's normalisation / h^3
    double ac;     // alpha c0 / 2
    double beta, r0, D, p1, p2;
    // the double pass's constants, formed with the reference's operations
    double alpha_k;  // 10 / (7 pi) / (h h), Wendland 9 / (4 pi) / (h h), Gaussian (1 / pi) / (h h)
    double c01;      // (0.01 h) h
    double cij;      // 0.5 (c0 + 0)
    double alpha;
    int stage = 1;   // the single-simulation float passes stage each workgroup's records in LDS
};

// Inputs of the pair pass, cell ordered.
struct PairIn {
    const void *A, *B;     // Rec<double> or Rec<float> records
    const double *H;       // double h
    const int2 *cxy;       // raw cell coordinates (cx, cy) of each particle
    const int *cellStart;
    // XSPH only: every particle's mass (walls included) and the half-step
    // velocity in double
    const double *M = nullptr, *VX = nullptr, *VY = nullptr;
};

struct PairOut {
    double *ax, *ay, *drho;
    int *count;  // optional: neighbour count per particle (cell order), -1 for walls
    int mode;    // 0 fused, 1 continuity only (drho), 2 momentum only (ax, ay)
    double *xsx = nullptr, *xsy = nullptr;  // non-null: XSPH velocity out (fused and momentum modes)
};

// Summation density (state arrays of the current buffer and the B records).
struct SdOut {
    double *rho, *drho;
    void *B;
};

struct Phys {
    double rho0, gamma, B, Pb, co, alpha, beta, r0, D, p1, p2, g, qmax;
    int kernel;
    double eps;  // XSPH factor
};

struct BInst;

// Launchers implemented in the pair translation units (threads: workgroup
// size of the pair pass, 64, 128 or 256; chosen by the autotuner).
hipError_t launch_pair_fp64(const Ctrl *ctrl, PairIn in, PairOut out, Phys ph, int n, int threads, hipStream_t s);
hipError_t launch_pair_fp32c(const Ctrl *ctrl, PairIn in, PairOut out, Phys ph, int n, int threads, hipStream_t s);
// The fused float pass for one h and one fluid mass (kRecUniform records;
// cubic spline or Wendland).
hipError_t launch_pair_fp32u(const Ctrl *ctrl, PairIn in, PairOut out, const UniformParams &u, int kernel, int n,
                             hipStream_t s);
// The fused double pass for one h and one fluid mass (kRecUniform64 records;
// cubic spline or Wendland), bit-identical to launch_pair_fp64's.
hipError_t launch_pair_fp64u(const Ctrl *ctrl, PairIn in, PairOut out, const UniformParams &u, int kernel, int n,
                             hipStream_t s);
// The same passes with PEC's corrector and the next dt's partials fused in
// (no Newmark-beta particles): they write ax, ay, drho and the corrected x,
// y, vx, vy, rho of every fluid slot.
hipError_t launch_pair_fp32u_pec(Ctrl *ctrl, PairIn in, PairOut out, const UniformParams &u, int kernel, int n,
                                 const PecArgs &pec, hipStream_t s);
hipError_t launch_pair_fp64u_pec(Ctrl *ctrl, PairIn in, PairOut out, const UniformParams &u, int kernel, int n,
                                 const PecArgs &pec, hipStream_t s);
// Whether parameters P with one h for every particle allow the uniform
// passes (the cubic spline or Wendland, the plain loop, 2h and r0 inside the
// cut-off), and their constants for that h and fluid mass m.
bool uniform_eligible(const sph_params_t &P, double h);
constexpr long long kUniformMaxN = 1LL << 27;  // 32-B records of this many slots stay below 4 GiB
UniformParams make_uniform_params(const sph_params_t &P, double h, double m);
// The uniform passes of a batch: each instance's constants are a block of
// uniform_const_size_*() bytes in the device array Ks (instance order),
// written on the host by uniform_const_*. queue / grid as for
// launch_pair_batch_*.
size_t uniform_const_size_fp32();
size_t uniform_const_size_fp64();
void uniform_const_fp32(const UniformParams &u, void *out);
void uniform_const_fp64(const UniformParams &u, void *out);
hipError_t launch_pair_batch_fp32u(const BInst *inst, const int *instOf, PairIn in, PairOut out, const void *Ks,
                                   int kernel, int n, int *queue, int grid, hipStream_t s);
hipError_t launch_pair_batch_fp64u(const BInst *inst, const int *instOf, PairIn in, PairOut out, const void *Ks,
                                   int kernel, int n, int *queue, int grid, hipStream_t s);
// The batch's uniform passes with PEC's corrector and the dt partials
// fused in: one 256-thread workgroup per chunk of one instance (slots
// cs[c] .. cs[c + 1], c < NC, at most 256 each; chunkInst: the chunk's
// instance); the partials go to pec.part[q * NC + c] (pec.strict is unused:
// each instance's own).
hipError_t launch_pair_batch_fp32u_pec(const BInst *inst, const int *cs, const int *chunkInst, int NC, PairIn in,
                                       PairOut out, const void *Ks, int kernel, int n, const PecArgs &pec,
                                       hipStream_t s);
hipError_t launch_pair_batch_fp64u_pec(const BInst *inst, const int *cs, const int *chunkInst, int NC, PairIn in,
                                       PairOut out, const void *Ks, int kernel, int n, const PecArgs &pec,
                                       hipStream_t s);
int pair_batch_uniform_occupancy_fp32(int kernel);
int pair_batch_uniform_occupancy_fp64(int kernel);
// Summation density before the pair pass (src/Tools/SolverTools.py:129-140).
hipError_t launch_sumden_fp64(const Ctrl *ctrl, PairIn in, SdOut out, Phys ph, int n, int threads, hipStream_t s);
hipError_t launch_sumden_fp32c(const Ctrl *ctrl, PairIn in, SdOut out, Phys ph, int n, int threads, hipStream_t s);

// One simulation of a batched ensemble (device array, one per instance).
// Instance k owns the particle slots [off, off + n) of the concatenated
// cell-ordered state and the bins [cellBase, cellBase + grid.ncells] of the
// segmented cell grid (the last one collects keys outside its grid; an
// inactive instance has ncells = 0 and parks every particle there).
struct BInst {
    double t, dt, dt_c, dt_f;
    long long step;
    int active;
    unsigned flags;
    Grid grid;
    long long cellBase;
    int off, n;
    int advance_t;
    int nfluid;
    long long cap;  // most cells this instance's grid may have
    double duration, fixed_dt, damping;
    double rho0, gamma, co, B, Pb, alpha, beta, r0, D, p1, p2, g, qmax, scale, gamma_c, gamma_f;
    int strict, kernel;
    // one h for every particle of the instance (the uniform pass): uh; m and
    // h then stay where the upload put them (slot = id) and are not permuted
    int uh_on;
    double uh;
    // the bounds of the border-cell path, as Ctrl's disp, vbound, abound
    double disp, vbound, abound;
    long long border_updates;
};

__host__ __device__ inline Phys inst_phys(const BInst &I) {
    Phys ph;
    ph.rho0 = I.rho0;
    ph.gamma = I.gamma;
    ph.B = I.B;
    ph.Pb = I.Pb;
    ph.co = I.co;
    ph.alpha = I.alpha;
    ph.beta = I.beta;
    ph.r0 = I.r0;
    ph.D = I.D;
    ph.p1 = I.p1;
    ph.p2 = I.p2;
    ph.g = I.g;
    ph.qmax = I.qmax;
    ph.kernel = I.kernel;
    ph.eps = 0.0;  // no XSPH in a batch
    return ph;
}

// The pair pass over a batch: instOf[t] is the instance of slot t.
// list: the per-thread contributor list length (16, 24, 32; 0 = default).
// queue (9 counters, 32 ints apart, zero before the first launch): the
// persistent work-queue form over about `grid` workgroups; null: one thread
// per slot.
hipError_t launch_pair_batch_fp64(const BInst *inst, const int *instOf, PairIn in, PairOut out, int kernel, int n,
                                  int threads, int list, int *queue, int grid, hipStream_t s);
hipError_t launch_pair_batch_fp32c(const BInst *inst, const int *instOf, PairIn in, PairOut out, int kernel, int n,
                                   int threads, int list, int *queue, int grid, hipStream_t s);
// Resident workgroups per CU of the persistent form (not to be called under
// stream capture).
int pair_batch_queue_occupancy_fp64(int kernel, int list, int threads);
int pair_batch_queue_occupancy_fp32c(int kernel, int list, int threads);

// Number of blocks for a grid over n items, padded to a multiple of the XCD
// count so that the block remap below is a bijection.
constexpr int kXcds = 8;
// The persistent batch pair pass's work queue: one counter per XCD and one
// for the workgroups that finished, 128 B apart.
constexpr int kQueueStride = 32;
constexpr int kQueueInts = (kXcds + 1) * kQueueStride;
inline int xcd_blocks(long long n, int bs) {
    long long b = (n + bs - 1) / bs;
    if (b < 1) b = 1;
    return (int)((b + kXcds - 1) / kXcds * kXcds);
}

// The dispatcher hands consecutive workgroups to the 8 XCDs in turn, and
// each XCD has its own L2. Remap so that each XCD walks one contiguous range
// of logical blocks: neighbouring cells (which share neighbour data) then
// meet in the same L2.
__device__ inline int xcd_remap(int b, int nb) {
    int per = nb / kXcds;
    return (b % kXcds) * per + (b / kXcds);
}

// Element i of a device array as the array's base plus a 32-bit byte offset
// (for arrays below 4 GiB): the base stays in scalar registers and each load
// takes one 32-bit vector offset instead of 64-bit address arithmetic per lane.
#ifndef HIPSPH_AT32
#define HIPSPH_AT32 1
#endif
template <typename T>
__device__ inline const T &at32(const T *base, unsigned i) {
#if HIPSPH_AT32
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + i * (unsigned)sizeof(T));
#else
    return base[i];
#endif
}

// base[i + j] for a constant j: the 32-bit byte offset of i, then j records
// further in 64-bit address space, so that j folds into the load's
// immediate offset (one offset computation for a run of loads).
template <typename T>
__device__ inline const T &at32j(const T *base, unsigned i, int j) {
    return reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + i * (unsigned)sizeof(T))[j];
}

// Integer power as a left-to-right product for small integer exponents
// (what the reference's pow(x, 4.0) etc. evaluate to up to rounding).
template <typename R>
__device__ inline R ipow(R x, R p) {
    if (p >= R(1) && p <= R(16) && p == floor(p)) {
        int n = (int)p;
        R r = x;
        for (int k = 1; k < n; ++k) r = r * x;
        return r;
    }
    if (p == R(0)) return R(1);
    return pow(x, p);
}

}  // namespace hipsph
