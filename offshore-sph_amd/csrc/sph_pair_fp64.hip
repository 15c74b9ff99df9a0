// sph_pair_fp64.hip — the fused pair pass in double precision.
//
// Instantiates sph_pair_impl.h with R = double in the reference's order
// (x-shift outer, y-shift inner, descending reference index inside a cell).
// Built with -ffp-contract=off and no fast-math: every operation is the
// reference's, in its order, so for the cubic spline and the Wendland kernel
// the results are bit-identical to the CPU oracle; the Gaussian differs only
// through exp(). Terms the pass skips are exactly +-0 in the reference
// (fluid pairs outside the kernel support) or absent (walls beyond r0), and
// adding +-0 to a sum that starts at +0 never changes its bits.
//
// Neighbour predicate: the reference tests q = sqrt(r^2) / h_ij <= 3 with
// correctly rounded sqrt and division (src/Tools/NNLinkedList.py:63-71). Pairs
// with r^2 > (3 h_ij)^2 (1 + 1e-12) cannot pass that test and pairs with
// r^2 < (3 h_ij)^2 (1 - 1e-12) cannot fail it (the three roundings are below
// 1e-15 relative); only pairs in between evaluate the exact expression.
// The form for one h and one fluid mass is in sph_pair_uniform64.h.
#include "sph_pair_impl.h"
#include "sph_pair_uniform64.h"

namespace hipsph {

hipError_t launch_pair_fp64(const Ctrl *ctrl, PairIn in, PairOut out, Phys ph, int n, int threads, hipStream_t s) {
    return launch_pair<double, true>(ctrl, pair_args<double>(in, &out, n), ph, threads, out.mode, s);
}

size_t uniform_const_size_fp64() { return sizeof(Uniform64Const); }

void uniform_const_fp64(const UniformParams &u, void *out) {
    Uniform64Const K;
    K.lim2 = u.lim2;
    K.h = u.h;
    K.m = u.m;
    K.alpha_k = u.alpha_k;
    K.c01 = u.c01;
    K.cij = u.cij;
    K.alpha = u.alpha;
    K.beta = u.beta;
    K.r0 = u.r0;
    K.D = u.D;
    K.p1 = u.p1;
    K.p2 = u.p2;
    K.g = u.g;
    K.qmax = u.qmax;
    *(Uniform64Const *)out = K;
}

static Uniform64Args uniform64_args(const PairIn &in, const PairOut &out, int n) {
    Uniform64Args P;
    P.A = (const double2 *)in.A;
    P.B = (const double4 *)in.B;
    P.cxy = in.cxy;
    P.cellStart = in.cellStart;
    P.ax = out.ax;
    P.ay = out.ay;
    P.drho = out.drho;
    P.n = n;
    return P;
}

hipError_t launch_pair_fp64u(const Ctrl *ctrl, PairIn in, PairOut out, const UniformParams &u, int kernel, int n,
                             hipStream_t s) {
    const Uniform64Args P = uniform64_args(in, out, n);
    Uniform64Const K;
    uniform_const_fp64(u, &K);
    const int nb = xcd_blocks(n, 256);
    with_kernel(kernel, [&](auto kk) {
        constexpr int KN = decltype(kk)::value;
        if (out.mode == kContinuity)
            k_pair_uniform64<KN, kContinuity><<<nb, 256, 0, s>>>(ctrl, P, K);
        else if (out.mode == kMomentum)
            k_pair_uniform64<KN, kMomentum><<<nb, 256, 0, s>>>(ctrl, P, K);
        else
            k_pair_uniform64<KN, kFused><<<nb, 256, 0, s>>>(ctrl, P, K);
    });
    return hipGetLastError();
}

hipError_t launch_pair_fp64u_pec(Ctrl *ctrl, PairIn in, PairOut out, const UniformParams &u, int kernel, int n,
                                 const PecArgs &pec, hipStream_t s) {
    const Uniform64Args P = uniform64_args(in, out, n);
    Uniform64Const K;
    uniform_const_fp64(u, &K);
    const int nb = xcd_blocks(n, 256);  // its dt partials: one per workgroup
    with_kernel(kernel, [&](auto kk) {
        k_pair_uniform64_pec<decltype(kk)::value><<<nb, 256, 0, s>>>(ctrl, P, K, pec);
    });
    return hipGetLastError();
}

hipError_t launch_pair_batch_fp64u(const BInst *inst, const int *instOf, PairIn in, PairOut out, const void *Ks,
                                   int kernel, int n, int *queue, int grid, hipStream_t s) {
    const Uniform64Args P = uniform64_args(in, out, n);
    const Uniform64Const *K = (const Uniform64Const *)Ks;
    const int nb = xcd_blocks(n, 256);
    if (queue) {  // a multiple of the XCD count, no more workgroups than tiles
        if (grid > nb) grid = nb;
        grid = grid / kXcds * kXcds;
        if (grid < kXcds) grid = kXcds;
        with_kernel(kernel, [&](auto kk) {
            k_pair_batch_uniform64_queue<decltype(kk)::value><<<grid, 256, 0, s>>>(inst, instOf, P, K, queue);
        });
    } else {
        with_kernel(kernel, [&](auto kk) {
            k_pair_batch_uniform64<decltype(kk)::value><<<nb, 256, 0, s>>>(inst, instOf, P, K);
        });
    }
    return hipGetLastError();
}

int pair_batch_uniform_occupancy_fp64(int kernel) {
    int per = 0;
    hipError_t e = hipSuccess;
    with_kernel(kernel, [&](auto kk) {
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_pair_batch_uniform64_queue<decltype(kk)::value>, 256,
                                                         0);
    });
    return e == hipSuccess && per > 0 ? per : 1;
}

hipError_t launch_sumden_fp64(const Ctrl *ctrl, PairIn in, SdOut out, Phys ph, int n, int threads, hipStream_t s) {
    return launch_sumden<double, true>(ctrl, pair_args<double>(in, nullptr, n), out, ph, threads, s);
}

hipError_t launch_pair_batch_fp64(const BInst *inst, const int *instOf, PairIn in, PairOut out, int kernel, int n,
                                  int threads, int list, int *queue, int grid, hipStream_t s) {
    return launch_pair_batch<double, true>(inst, instOf, pair_args<double>(in, &out, n), kernel, threads, list, queue,
                                           grid, s);
}

int pair_batch_queue_occupancy_fp64(int kernel, int list, int threads) {
    return queue_occupancy<double, true>(kernel, list, threads);
}

hipError_t launch_pair_batch_fp64u_pec(const BInst *inst, const int *cs, const int *chunkInst, int NC, PairIn in, PairOut out,
                                       const void *Ks, int kernel, int n, const PecArgs &pec, hipStream_t s) {
    const Uniform64Args P = uniform64_args(in, out, n);
    const Uniform64Const *K = (const Uniform64Const *)Ks;
    const int nb = (NC + kXcds - 1) / kXcds * kXcds;  // xcd_remap maps [0, nb) onto itself
    with_kernel(kernel, [&](auto kk) {
        k_pair_batch_uniform64_pec<decltype(kk)::value><<<nb, 256, 0, s>>>(inst, cs, chunkInst, NC, P, K, pec);
    });
    return hipGetLastError();
}

}  // namespace hipsph
