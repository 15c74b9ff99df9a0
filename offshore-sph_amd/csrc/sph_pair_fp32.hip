// sph_pair_fp32.hip — the fused pair pass with float pair arithmetic.
//
// Same physics and neighbour set as sph_pair_fp64.hip. Differences:
//   * positions come in as float offsets from each particle's raw cell origin
//     (written by the reorder kernel from the double state), so x_i - x_j is
//     (xl_i - xl_j) - (cell shift) * cs with errors ~1e-7 of a cell, never the
//     ~1e-4 dx an absolute float position would give on a 4 m tank;
//   * the three sums (drho, a_x, a_y) are accumulated with TwoSum (Knuth)
//     compensation and returned as double (sum + compensation);
//   * the neighbour predicate stays exact: pairs whose float r^2 lies within
//     1e-5 (relative) of (3 h_ij)^2 are re-tested with the reference's double
//     expression, under contraction off. Particles whose raw cell is not their
//     bin (the reference's key aliasing at the grid edge) carry NaN offsets,
//     which sends every pair they take part in through that double path.
// Built with -ffp-contract=fast; the exact predicate disables it locally.
#include "sph_internal.h"

namespace hipsph {
namespace {

constexpr float kPIf = 3.14159265358979f;

// v_rcp_f32 / v_rsq_f32: ~1 ulp, one instruction each.
__device__ inline float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

__device__ inline void two_sum(float &s, float &c, float x) {
    float t = s + x;
    float z = t - s;
    c += (s - (t - z)) + (x - z);
    s = t;
}

// The reference's neighbour test in double (NNLinkedList.py:63-71); also
// returns the double separation for the caller's float arithmetic.
__device__ __noinline__ bool exact_pair(const PairIn &in, int t, int u, double qmax, float &dxf, float &dyf) {
#pragma clang fp contract(off)
    const double dx = in.x[t] - in.x[u], dy = in.y[t] - in.y[u];
    const double xij2 = dx * dx + dy * dy;
    const double r = sqrt(xij2);
    const double h_ij = 0.5 * (in.h[t] + in.h[u]);
    const double q = r / h_ij;
    dxf = (float)dx;
    dyf = (float)dy;
    return q <= qmax;
}

template <int KERNEL>
__device__ inline float grad_norm(float r, float rinv, float q, float hinv) {
    if (KERNEL == SPH_KERNEL_GAUSSIAN) {
        // src/Kernels/Gaussian.py:39-60: dW/dq / (r h)
        if (!(q <= 3.0f) || !(r > 1e-12f)) return 0.0f;
        const float alpha_c = (1.0f / kPIf) * hinv * hinv;
        return -2.0f * q * alpha_c * __expf(-q * q) * rinv * hinv;
    } else if (KERNEL == SPH_KERNEL_WENDLAND) {
        // src/Kernels/Wendland.py:32-64
        if (q >= 2.0f || r < 1e-10f) return 0.0f;
        const float inner = 1.0f - 0.5f * q;
        const float i2 = inner * inner;
        const float grad = i2 * i2 * inner * (-14.0f / 3.0f) * q * (1.0f + 2.5f * q);
        return (9.0f / (4.0f * kPIf)) * hinv * hinv * grad * hinv * rinv;
    } else {
        // src/Kernels/CubicSpline.py:38-70
        if (q > 2.0f || r < 1e-10f) return 0.0f;
        const float grad = (q > 1.0f) ? -0.75f * (2.0f - q) * (2.0f - q) : -3.0f * q * (1.0f - 0.75f * q);
        return (10.0f / (7.0f * kPIf)) * hinv * hinv * grad * hinv * rinv;
    }
}

template <int KERNEL>
__global__ __launch_bounds__(256) void k_pair_fp32c(const Ctrl *ctrl, PairIn in, PairOut out, Phys ph, int n) {
    if (!ctrl->active) return;
    const int t = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= n || in.label[t] != kFluid) return;
    const Grid g = ctrl->grid;
    const float xi = in.xl[t], yi = in.yl[t], vxi = in.vxf[t], vyi = in.vyf[t], rhoi = in.rhof[t], hi = in.hf[t];
    const float slf = in.prsf[t];
    const int2 c = in.cxy[t];
    const float csf = (float)g.cs;
    const float qmax = (float)ph.qmax;
    const float r0 = (float)ph.r0, D = (float)ph.D, p1 = (float)ph.p1, p2 = (float)ph.p2;
    const float alpha = (float)ph.alpha, beta = (float)ph.beta, cij = 0.5f * (float)ph.co;
    float arho = 0.f, carho = 0.f, a0 = 0.f, ca0 = 0.f, a1 = 0.f, ca1 = 0.f, f0 = 0.f, cf0 = 0.f, f1 = 0.f, cf1 = 0.f;
    int cnt = 0;
    for (int da = -1; da <= 1; ++da) {
        const long long cx = (long long)c.x + da;
        if (!(g.ncx > cx && cx > -1)) continue;
        const float offx = (float)da * csf;
        for (int db = -1; db <= 1; ++db) {
            const long long cy = (long long)c.y + db;
            if (!(g.ncy > cy && cy > -1)) continue;
            const float offy = (float)db * csf;
            const int cell = (int)(cx + g.ncx * cy);
            const int s = in.cellStart[cell], e = in.cellStart[cell + 1];
            for (int u = s; u < e; ++u) {
                float dx = (xi - in.xl[u]) - offx;
                float dy = (yi - in.yl[u]) - offy;
                const float r2 = dx * dx + dy * dy;
                const float hij = 0.5f * (hi + in.hf[u]);
                const float lim = (qmax * hij) * (qmax * hij);
                if (r2 > lim * 1.00001f) continue;
                if (!(r2 < lim * 0.99999f)) {
                    if (!exact_pair(in, t, u, ph.qmax, dx, dy)) continue;
                }
                ++cnt;
                const float rr2 = dx * dx + dy * dy;
                const float rinv = rr2 > 0.f ? frsq(rr2) : 0.f;
                const float r = rr2 * rinv;
                if (in.label[u] == kFluid) {
                    const float hinv = frcp(hij);
                    const float q = r * hinv;
                    const float norm = grad_norm<KERNEL>(r, rinv, q, hinv);
                    if (norm == 0.0f) continue;
                    const float dwx = norm * dx, dwy = norm * dy;
                    const float vx = vxi - in.vxf[u], vy = vyi - in.vyf[u];
                    const float mj = in.mf[u];
                    two_sum(arho, carho, mj * (vx * dwx + vy * dwy));
                    const float dv = vx * dx + vy * dy;
                    float PI_ij = 0.0f;
                    if (dv < 0.f) {
                        const float hm = 0.5f * (hi + hij);
                        const float rhoij = 0.5f * (rhoi + in.rhof[u]);
                        const float muij = hm * dv * frcp(r * r + 0.01f * hm * hm);
                        PI_ij = muij * (beta * muij - alpha * cij) * frcp(rhoij);
                    }
                    const float fm = -mj * (slf + in.prsf[u] + PI_ij);
                    two_sum(a0, ca0, fm * dwx);
                    two_sum(a1, ca1, fm * dwy);
                } else if (!(r > r0) && r > 1e-12f) {
                    const float frac = r0 * rinv;
                    const float tmp = ipow(frac, p1) - ipow(frac, p2);
                    const float fac = D * tmp * (rinv * rinv);
                    two_sum(f0, cf0, fac * dx);
                    two_sum(f1, cf1, fac * dy);
                }
            }
        }
    }
    if (cnt == 0) return;
    out.drho[t] = (double)arho + (double)carho;
    out.ax[t] = ((double)a0 + (double)ca0) + ((double)f0 + (double)cf0);
    out.ay[t] = (((double)a1 + (double)ca1) - ph.g) + ((double)f1 + (double)cf1);
}

}  // namespace

hipError_t launch_pair_fp32c(const Ctrl *ctrl, PairIn in, PairOut out, Phys ph, int n, hipStream_t s) {
    const int nb = xcd_blocks(n, 256);
    switch (ph.kernel) {
        case SPH_KERNEL_GAUSSIAN:
            k_pair_fp32c<SPH_KERNEL_GAUSSIAN><<<nb, 256, 0, s>>>(ctrl, in, out, ph, n);
            break;
        case SPH_KERNEL_WENDLAND:
            k_pair_fp32c<SPH_KERNEL_WENDLAND><<<nb, 256, 0, s>>>(ctrl, in, out, ph, n);
            break;
        default:
            k_pair_fp32c<SPH_KERNEL_CUBIC><<<nb, 256, 0, s>>>(ctrl, in, out, ph, n);
    }
    return hipGetLastError();
}

}  // namespace hipsph
