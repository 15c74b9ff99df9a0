// sph_pair_fp32.hip — the fused pair pass with float pair arithmetic.
//
// Instantiates sph_pair_impl.h with R = float. Same physics and neighbour set
// as the double pass. Differences:
//   * the separation x_i - x_j is formed from the double positions and rounded
//     once to float (relative error 6e-8 at any tank size; an absolute float
//     position would carry ~1e-4 dx of error on a 4 m tank);
//   * the rest of the pair arithmetic is float with native reciprocal and
//     reciprocal square root (~1 ulp);
//   * the three sums (drho, a_x, a_y) and the wall force are accumulated with
//     TwoSum (Knuth) compensation and returned as double (sum + compensation);
//   * the neighbour predicate stays exact: pairs whose float r^2 lies within
//     1e-5 (relative) of (3 h_ij)^2 are re-tested with the reference's double
//     expression under contraction off.
// Built with -ffp-contract=fast-honor-pragmas (the pragmas below switch
// contraction off where exactness matters). The uniform-h, uniform-mass form
// of the same pass is in sph_pair_uniform.h.
#include "sph_pair_impl.h"
#include "sph_pair_uniform.h"

namespace hipsph {

hipError_t launch_pair_fp32c(const Ctrl *ctrl, PairIn in, PairOut out, Phys ph, int n, int threads, hipStream_t s) {
    return launch_pair<float, false>(ctrl, pair_args<float>(in, &out, n), ph, threads, out.mode, s);
}

hipError_t launch_sumden_fp32c(const Ctrl *ctrl, PairIn in, SdOut out, Phys ph, int n, int threads, hipStream_t s) {
    return launch_sumden<float, false>(ctrl, pair_args<float>(in, nullptr, n), out, ph, threads, s);
}

hipError_t launch_pair_batch_fp32c(const BInst *inst, const int *instOf, PairIn in, PairOut out, int kernel, int n,
                                   int threads, int list, int *queue, int grid, hipStream_t s) {
    return launch_pair_batch<float, false>(inst, instOf, pair_args<float>(in, &out, n), kernel, threads, list, queue,
                                           grid, s);
}

size_t uniform_const_size_fp32() { return sizeof(UniformConst); }

void uniform_const_fp32(const UniformParams &u, void *out) {
    UniformConst K;
    K.lim2 = u.lim2;
    K.lo = u.lo;
    K.hd = u.h;
    K.qmax = u.qmax;
    K.g = u.g;
    K.h = (float)u.h;
    K.hinv = (float)(1.0 / u.h);
    K.m = (float)u.m;
    K.ck = (float)u.ck;
    K.c01 = (float)(0.01 * u.h * u.h);
    K.ac = (float)u.ac;
    K.beta = (float)u.beta;
    K.r0 = (float)u.r0;
    K.D = (float)u.D;
    K.p1 = (float)u.p1;
    K.p2 = (float)u.p2;
    K.stage = u.stage;
    *(UniformConst *)out = K;
}

static UniformArgs uniform_args(const PairIn &in, const PairOut &out, int n) {
    UniformArgs P;
    P.A = (const double2 *)in.A;
    P.B = (const float4 *)in.B;
    P.cxy = in.cxy;
    P.cellStart = in.cellStart;
    P.ax = out.ax;
    P.ay = out.ay;
    P.drho = out.drho;
    P.n = n;
    return P;
}

hipError_t launch_pair_fp32u(const Ctrl *ctrl, PairIn in, PairOut out, const UniformParams &u, int kernel, int n,
                             hipStream_t s) {
    const UniformArgs P = uniform_args(in, out, n);
    UniformConst K;
    uniform_const_fp32(u, &K);
    const int nb = xcd_blocks(n, 256);
    with_kernel(kernel, [&](auto kk) {
        constexpr int KN = decltype(kk)::value;
        if (out.mode == kContinuity)
            k_pair_uniform<KN, kContinuity><<<nb, 256, 0, s>>>(ctrl, P, K);
        else if (out.mode == kMomentum)
            k_pair_uniform<KN, kMomentum><<<nb, 256, 0, s>>>(ctrl, P, K);
        else
            k_pair_uniform<KN, kFused><<<nb, 256, 0, s>>>(ctrl, P, K);
    });
    return hipGetLastError();
}

hipError_t launch_pair_fp32u_pec(Ctrl *ctrl, PairIn in, PairOut out, const UniformParams &u, int kernel, int n,
                                 const PecArgs &pec, hipStream_t s) {
    const UniformArgs P = uniform_args(in, out, n);
    UniformConst K;
    uniform_const_fp32(u, &K);
    const int nb = xcd_blocks(n, 256);  // its dt partials: one per workgroup
    with_kernel(kernel, [&](auto kk) {
        k_pair_uniform_pec<decltype(kk)::value><<<nb, 256, 0, s>>>(ctrl, P, K, pec);
    });
    return hipGetLastError();
}

hipError_t launch_pair_batch_fp32u(const BInst *inst, const int *instOf, PairIn in, PairOut out, const void *Ks,
                                   int kernel, int n, int *queue, int grid, hipStream_t s) {
    const UniformArgs P = uniform_args(in, out, n);
    const UniformConst *K = (const UniformConst *)Ks;
    const int nb = xcd_blocks(n, 256);
    if (queue) {  // a multiple of the XCD count, no more workgroups than tiles
        if (grid > nb) grid = nb;
        grid = grid / kXcds * kXcds;
        if (grid < kXcds) grid = kXcds;
        with_kernel(kernel, [&](auto kk) {
            k_pair_batch_uniform_queue<decltype(kk)::value><<<grid, 256, 0, s>>>(inst, instOf, P, K, queue);
        });
    } else {
        with_kernel(kernel, [&](auto kk) {
            k_pair_batch_uniform<decltype(kk)::value><<<nb, 256, 0, s>>>(inst, instOf, P, K);
        });
    }
    return hipGetLastError();
}

int pair_batch_uniform_occupancy_fp32(int kernel) {
    int per = 0;
    hipError_t e = hipSuccess;
    with_kernel(kernel, [&](auto kk) {
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_pair_batch_uniform_queue<decltype(kk)::value>, 256, 0);
    });
    return e == hipSuccess && per > 0 ? per : 1;
}

int pair_batch_queue_occupancy_fp32c(int kernel, int list, int threads) {
    return queue_occupancy<float, false>(kernel, list, threads);
}

hipError_t launch_pair_batch_fp32u_pec(const BInst *inst, const int *cs, const int *chunkInst, int NC, PairIn in, PairOut out,
                                       const void *Ks, int kernel, int n, const PecArgs &pec, hipStream_t s) {
    const UniformArgs P = uniform_args(in, out, n);
    const UniformConst *K = (const UniformConst *)Ks;
    const int nb = (NC + kXcds - 1) / kXcds * kXcds;  // xcd_remap maps [0, nb) onto itself
    with_kernel(kernel, [&](auto kk) {
        k_pair_batch_uniform_pec<decltype(kk)::value><<<nb, 256, 0, s>>>(inst, cs, chunkInst, NC, P, K, pec);
    });
    return hipGetLastError();
}

}  // namespace hipsph
