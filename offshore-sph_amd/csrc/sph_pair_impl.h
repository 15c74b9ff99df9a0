// sph_pair_impl.h — the fused pair pass (Tait-derived p/rho^2, continuity,
// momentum + artificial viscosity + gravity, Lennard-Jones wall force).
// Included by sph_pair_fp64.hip (R = double, reference order,
// -ffp-contract=off) and sph_pair_fp32.hip (R = float, -ffp-contract=fast-honor-pragmas).
//
// The pass is bound by the VALU, not by HBM (1M dam break, FP32C cubic: VALU
// instructions x 4 cycles fill at least 78 % of the SIMD cycles once the
// cached loads are software-pipelined, ~57 % before; HBM traffic ~0.65 TB/s),
// so the design minimises instructions per candidate and keeps loads in
// flight:
//   * One thread per particle of the cell-ordered state; fluid particles are
//     targets. Consecutive threads hold neighbouring particles, so the
//     candidates of a wavefront overlap and are served by L1/L2.
//   * Each particle's pair inputs are two packed records written by the
//     reorder kernel (one address computation, wide loads):
//       double: A = (x, y, h, w), B = (vx, vy, rho, p/rho^2)
//       float:  A = (x, y as double; h, w, vx, vy as float), B = (rho, p/rho^2)
//     with w = m for fluid and w = -label for walls (the pair equations use no
//     wall mass). The float path forms x_i - x_j in double and rounds it once:
//     on this GPU a double add costs the same issue slot as a float add, and
//     the separation keeps full accuracy at any tank size.
//   * Pass 1 walks the 3x3 cells and appends, per target, the candidates that
//     are neighbours and contribute (inside the kernel support, or within r0
//     of a wall) to a list in LDS (interleaved by lane: conflict-free). With a
//     support inside the cut-off (cubic, Wendland) a fluid candidate inside
//     the support needs no cut-off test (-6 % on the FP64 pass). The float
//     pass walks the three cells of each row as one run of slots, and the
//     walk loads the next candidate's record before testing the current one
//     (-11 % cubic, -25 % Gaussian on the float pass); the float cubic and
//     Wendland list loops do the same (-4.5 %). Pass 2
//     runs the costly pair arithmetic over the list only, so lanes of a wave
//     stay busy instead of idling through candidates that add nothing. The list
//     is flushed when full, which keeps the order of the sums.
//
// MODE selects the equations: kFused (continuity + momentum + wall force +
// gravity), kContinuity (drho only) or kMomentum (ax, ay only). The split
// passes evaluate the same expressions in the same order as the fused one.
#pragma once
#include "sph_internal.h"

namespace hipsph {
namespace {

constexpr int kPairThreads = 256;
constexpr int kListLen = 32;  // list entries per thread before a flush (largest case)
// The float pass with a 2h support (~21 contributors) keeps 24 entries:
// 24 KB of LDS per workgroup lets 6 workgroups share a CU instead of 5
// (-3 % on the 1M dam break). With 16, lanes flush mid-walk at different
// candidates and the pass takes twice as long; the double pass is
// register-bound at 4 workgroups per CU either way.
template <typename R, int KERNEL>
__host__ __device__ constexpr int list_len() {
    return (sizeof(R) == 4 && KERNEL != SPH_KERNEL_GAUSSIAN) ? 24 : kListLen;
}
constexpr double kPI = 3.141592653589793;

__device__ inline double rec_vx(const Rec<double>::A &, const Rec<double>::B &b) { return b.vx; }
__device__ inline double rec_vy(const Rec<double>::A &, const Rec<double>::B &b) { return b.vy; }
__device__ inline float rec_vx(const Rec<float>::A &a, const Rec<float>::B &) { return a.vx; }
__device__ inline float rec_vy(const Rec<float>::A &a, const Rec<float>::B &) { return a.vy; }

template <typename R>
struct PairArgs {
    const typename Rec<R>::A *A;
    const typename Rec<R>::B *B;
    const double *H;  // double h (exact neighbour re-test of the float path)
    const int2 *cxy;  // raw cell of each particle
    const int *cellStart;
    double *ax, *ay, *drho;
    int *count;  // optional diagnostic output
    int n;
    // XSPH (XS instantiations only): masses of all particles, the double
    // half-step velocity, and the XSPH velocity out
    const double *M, *VX, *VY;
    double *xsx, *xsy;
};

// PairArgs from the engine-side descriptors (out may be null for passes that
// write no rates).
template <typename R>
inline PairArgs<R> pair_args(const PairIn &in, const PairOut *out, int n) {
    PairArgs<R> P;
    P.A = (const typename Rec<R>::A *)in.A;
    P.B = (const typename Rec<R>::B *)in.B;
    P.H = in.H;
    P.cxy = in.cxy;
    P.cellStart = in.cellStart;
    P.ax = out ? out->ax : nullptr;
    P.ay = out ? out->ay : nullptr;
    P.drho = out ? out->drho : nullptr;
    P.count = out ? out->count : nullptr;
    P.n = n;
    P.M = in.M;
    P.VX = in.VX;
    P.VY = in.VY;
    P.xsx = out ? out->xsx : nullptr;
    P.xsy = out ? out->xsy : nullptr;
    return P;
}

template <typename R>
struct Acc;

template <>
struct Acc<double> {  // plain sums in the reference's order
    double s = 0.0;
    __device__ void add(double v) { s += v; }
    __device__ double get() const { return s; }
};

template <>
struct Acc<float> {  // TwoSum (Knuth) compensated float sum
    float s = 0.f, c = 0.f;
    __device__ void add(float v) {
#pragma clang fp contract(off)
        float t = s + v;
        float z = t - s;
        c += (s - (t - z)) + (v - z);
        s = t;
    }
    __device__ double get() const { return (double)s + (double)c; }
};

// Kernel support radius in units of h (the gradient is exactly zero beyond).
template <int KERNEL>
__device__ constexpr double support_q() {
    return KERNEL == SPH_KERNEL_GAUSSIAN ? 3.0 : 2.0;
}

// Whether pass 1 may skip the cut-off test for fluid candidates of this
// target (scan_cell's `implied`): no count is wanted, the kernel support lies
// well inside the cut-off, and the target finds itself (its own cell is in
// the grid and h > 0, so q = 0), so "no neighbours" (SolverTools.py:166-169)
// cannot happen and the count is not needed for it either.
template <int KERNEL>
__device__ inline bool cutoff_implied(const int *count, double qmax, int2 c, int ncx, int ncy, double h) {
    return !count && support_q<KERNEL>() < 0.99 * qmax && c.x >= 0 && c.x < ncx && c.y >= 0 && c.y < ncy &&
           h > 0.0;
}

// Double gradient in the reference's expression order
// (src/Kernels/CubicSpline.py:38-70, Gaussian.py:39-60, Wendland.py:32-64),
// with h = h_ij, r = r_ij and q = r / h (the reference recomputes q = r / h in
// the kernel: the same value). Returns false where the reference returns 0.
template <int KERNEL>
__device__ inline bool grad_exact(double dx, double dy, double r, double q, double h, double &dwx, double &dwy) {
    if (KERNEL == SPH_KERNEL_GAUSSIAN) {
        double tmp = r * h;
        if (!(tmp > 1e-12 && q <= 3)) return false;
        double alpha_c = (1.0 / kPI) / (h * h);
        double dwdq = -2 * q * alpha_c * exp(-q * q);
        double f = dwdq / tmp;
        dwx = f * dx;
        dwy = f * dy;
        return true;
    } else if (KERNEL == SPH_KERNEL_WENDLAND) {
        double alpha = 9.0 / (4.0 * kPI) / (h * h);
        if ((q >= 2.0) || (r < 1e-10)) return false;
        double inner = 1.0 - 0.5 * q;
        double grad = ipow(inner, 5.0) * (-14.0 / 3.0) * q * (1 + 2.5 * q);
        double norm = alpha * grad / (h * r);
        dwx = norm * dx;
        dwy = norm * dy;
        return true;
    } else {
        double alpha = 10.0 / (7.0 * kPI) / (h * h);
        if ((q > 2.0) || (r < 1e-10)) return false;
        double grad = (q > 1.0) ? -0.75 * ((2.0 - q) * (2.0 - q)) : -3.0 * q * (1.0 - 0.75 * q);
        double norm = alpha * grad / (h * r);
        dwx = norm * dx;
        dwy = norm * dy;
        return true;
    }
}

// Float gradient factor (dW/dr) / r with native rcp/rsq (~1 ulp).
template <int KERNEL>
__device__ inline float grad_norm_f(float r, float rinv, float q, float hinv) {
    if (KERNEL == SPH_KERNEL_GAUSSIAN) {
        if (!(q <= 3.0f) || !(r > 1e-12f)) return 0.0f;
        return -2.0f * q * ((float)(1.0 / kPI) * hinv * hinv) * __expf(-q * q) * rinv * hinv;
    } else if (KERNEL == SPH_KERNEL_WENDLAND) {
        if (q >= 2.0f || r < 1e-10f) return 0.0f;
        const float inner = 1.0f - 0.5f * q;
        const float i2 = inner * inner;
        const float grad = i2 * i2 * inner * (-14.0f / 3.0f) * q * (1.0f + 2.5f * q);
        return (float)(9.0 / (4.0 * kPI)) * hinv * hinv * grad * hinv * rinv;
    } else {
        if (q > 2.0f || r < 1e-10f) return 0.0f;
        const float grad = (q > 1.0f) ? -0.75f * (2.0f - q) * (2.0f - q) : -3.0f * q * (1.0f - 0.75f * q);
        return (float)(10.0 / (7.0 * kPI)) * hinv * hinv * grad * hinv * rinv;
    }
}

// The reference's neighbour test in double (src/Tools/NNLinkedList.py:63-71),
// contraction off whatever the translation unit's default.
__device__ inline bool exact_q_test(double xi, double yi, double hi, double xj, double yj, double hj, double qmax) {
#pragma clang fp contract(off)
    const double dx = xi - xj, dy = yi - yj;
    const double xij2 = dx * dx + dy * dy;
    const double r = sqrt(xij2);
    const double h_ij = 0.5 * (hi + hj);
    const double q = r / h_ij;
    return q <= qmax;
}

template <typename R>
struct Target {
    double X, Y, H;            // double position and h
    R h, vx, vy, rho, slf;     // slf = p_i / rho_i^2 (Momentum.py:28)
    int slot;                  // the target's own slot (XSPH: other particles at r = 0 count)
};

template <typename R>
struct Sums {
    Acc<R> arho, a0, a1, f0, f1;
    Acc<R> xs0, xs1;  // XSPH
};

// Kernel values W (src/Kernels/CubicSpline.py:10-36, Gaussian.py:16-25,
// Wendland.py:10-30) in the reference's expression order, for XSPH and the
// summation density; zero outside the support.
template <int KERNEL>
__device__ inline double w_exact(double r, double h) {
    if (KERNEL == SPH_KERNEL_GAUSSIAN) {
        const double q = r / h;
        if (q <= 3.0) {
            const double alpha_c = (1.0 / kPI) / (h * h);
            return alpha_c * exp(-q * q);
        }
        return 0.0;
    } else if (KERNEL == SPH_KERNEL_WENDLAND) {
        const double alpha = 9.0 / (4.0 * kPI) / (h * h);
        const double q = r / h;
        if (q >= 2.0) return 0.0;
        const double inner = 1.0 - 0.5 * q;
        return alpha * (ipow(inner, 6.0) * (35.0 / 12.0 * q * q + 3.0 * q + 1.0));
    } else {
        const double alpha = 10.0 / (7.0 * kPI) / (h * h);
        const double q = r / h;
        if (q > 2.0) return 0.0;
        if (q > 1.0) return alpha * 0.25 * ipow(2.0 - q, 3.0);
        return alpha * (1.0 - 1.5 * ipow(q, 2.0) * (1.0 - 0.5 * q));
    }
}

// Float W with native exp / reciprocal (q = r / h, hinv = 1 / h).
template <int KERNEL>
__device__ inline float w_f(float q, float hinv) {
    if (KERNEL == SPH_KERNEL_GAUSSIAN) {
        if (!(q <= 3.0f)) return 0.0f;
        return (float)(1.0 / kPI) * hinv * hinv * __expf(-q * q);
    } else if (KERNEL == SPH_KERNEL_WENDLAND) {
        if (q >= 2.0f) return 0.0f;
        const float inner = 1.0f - 0.5f * q;
        const float i2 = inner * inner, i6 = i2 * i2 * i2;
        return (float)(9.0 / (4.0 * kPI)) * hinv * hinv * (i6 * ((35.0f / 12.0f) * q * q + 3.0f * q + 1.0f));
    } else {
        if (q > 2.0f) return 0.0f;
        const float a = (float)(10.0 / (7.0 * kPI)) * hinv * hinv;
        if (q > 1.0f) {
            const float t = 2.0f - q;
            return a * 0.25f * t * t * t;
        }
        return a * (1.0f - 1.5f * q * q * (1.0f - 0.5f * q));
    }
}

// Pass 2: the pair arithmetic for candidate u. XS adds the XSPH term of
// every listed neighbour, walls included (src/Equations/XSPH.py:26-31:
// fac = -epsilon m_j W_ij / rho_ij times v_i - v_j).
template <typename R, int KERNEL, int MODE, bool XS = false>
__device__ inline void contribute(const PairArgs<R> &P, const Target<R> &T, const Phys &ph, unsigned u,
                                  const typename Rec<R>::A &a, Sums<R> &S) {
    if constexpr (sizeof(R) == 8) {
        // Reference expression order: Continuity.py:13-16, Momentum.py:36-56,
        // BoundaryForce.py:31-41.
        const double dx = T.X - a.x, dy = T.Y - a.y;
        const double xij2 = dx * dx + dy * dy;
        const double r = sqrt(xij2);
        const double hij = 0.5 * ((double)T.h + (double)a.h);
        if constexpr (XS) {
            const typename Rec<R>::B bx = P.B[u];
            const double w = w_exact<KERNEL>(r, hij);
            const double mj = a.w >= 0 ? (double)a.w : P.M[u];
            const double rho_ij = 0.5 * ((double)T.rho + (double)bx.rho);
            const double fac = -ph.eps * mj * w / rho_ij;
            S.xs0.add(fac * ((double)T.vx - (double)rec_vx(a, bx)));
            S.xs1.add(fac * ((double)T.vy - (double)rec_vy(a, bx)));
        }
        if (a.w >= 0) {
            const double q = r / hij;
            double dwx, dwy;
            if (!grad_exact<KERNEL>(dx, dy, r, q, hij, dwx, dwy)) return;  // the reference adds exactly +-0
            const typename Rec<R>::B b = P.B[u];
            const double vx = (double)T.vx - (double)rec_vx(a, b), vy = (double)T.vy - (double)rec_vy(a, b);
            const double mj = (double)a.w;
            if constexpr (MODE != kMomentum) {
                const double dot = vx * dwx + vy * dwy;
                S.arho.add((R)(mj * dot));
            }
            if constexpr (MODE == kContinuity) return;
            const double dv = vx * dx + vy * dy;
            double PI_ij = 0.0;
            if (dv < 0) {
                const double hm = 0.5 * ((double)T.h + hij);
                const double cij = 0.5 * (ph.co + 0.0);  // c_j is never filled: c_ij = c_i / 2
                const double rhoij = 0.5 * ((double)T.rho + (double)b.rho);
                const double muij = hm * dv / (r * r + 0.01 * hm * hm);
                const double ppij = muij * (ph.beta * muij - ph.alpha * cij);
                PI_ij = ppij / rhoij;
            }
            const double factor = (double)T.slf + (double)b.prs + PI_ij;
            S.a0.add((R)(-mj * factor * dwx));
            S.a1.add((R)(-mj * factor * dwy));
        } else if (MODE != kContinuity && !(r > ph.r0) && r > 1e-12) {
            const double frac = ph.r0 / r;
            const double tmp = ipow(frac, ph.p1) - ipow(frac, ph.p2);
            const double fac = ph.D * tmp;
            S.f0.add((R)(fac * dx / (r * r)));
            S.f1.add((R)(fac * dy / (r * r)));
        }
    } else {
        const float dx = (float)(T.X - a.x), dy = (float)(T.Y - a.y);
        const float r2 = dx * dx + dy * dy;
        const float rinv = r2 > 0.f ? __builtin_amdgcn_rsqf(r2) : 0.f;
        const float r = r2 * rinv;
        const float hij = 0.5f * ((float)T.h + (float)a.h);
        if constexpr (XS) {
            const typename Rec<R>::B bx = P.B[u];
            const float w = w_f<KERNEL>(r * __builtin_amdgcn_rcpf(hij), __builtin_amdgcn_rcpf(hij));
            const float mj = a.w >= 0.f ? (float)a.w : (float)P.M[u];
            const float rho_ij = 0.5f * ((float)T.rho + (float)bx.rho);
            const float fac = -(float)ph.eps * mj * w * __builtin_amdgcn_rcpf(rho_ij);
            S.xs0.add(fac * ((float)T.vx - (float)rec_vx(a, bx)));
            S.xs1.add(fac * ((float)T.vy - (float)rec_vy(a, bx)));
        }
        if (a.w >= 0.f) {
            const float hinv = __builtin_amdgcn_rcpf(hij);
            const float q = r * hinv;
            const float norm = grad_norm_f<KERNEL>(r, rinv, q, hinv);
            if (norm == 0.0f) return;
            const typename Rec<R>::B b = P.B[u];
            const float dwx = norm * dx, dwy = norm * dy;
            const float vx = (float)T.vx - (float)rec_vx(a, b), vy = (float)T.vy - (float)rec_vy(a, b);
            const float mj = (float)a.w;
            if constexpr (MODE != kMomentum) S.arho.add(mj * (vx * dwx + vy * dwy));
            if constexpr (MODE == kContinuity) return;
            const float dv = vx * dx + vy * dy;
            float PI_ij = 0.0f;
            if (dv < 0.f) {
                // mu (beta mu - alpha c_ij) / rho_ij with mu = hm dv / Dn, with one reciprocal:
                // hm dv (beta hm dv - alpha c_ij Dn) / (Dn^2 rho_ij)
                const float hm = 0.5f * ((float)T.h + hij);
                const float rhoij = 0.5f * ((float)T.rho + (float)b.rho);
                const float Dn = r * r + 0.01f * hm * hm;
                const float hv = hm * dv;
                PI_ij = hv * ((float)ph.beta * hv - (float)ph.alpha * (0.5f * (float)ph.co) * Dn) *
                        __builtin_amdgcn_rcpf(Dn * Dn * rhoij);
            }
            const float fm = -mj * ((float)T.slf + (float)b.prs + PI_ij);
            S.a0.add(fm * dwx);
            S.a1.add(fm * dwy);
        } else if (MODE != kContinuity && !(r > (float)ph.r0) && r > 1e-12f) {
            const float frac = (float)ph.r0 * rinv;
            const float tmp = ipow(frac, (float)ph.p1) - ipow(frac, (float)ph.p2);
            const float fac = (float)ph.D * tmp * (rinv * rinv);
            S.f0.add(fac * dx);
            S.f1.add(fac * dy);
        }
    }
}

template <typename R, int KERNEL, int MODE, bool XS = false>
__device__ inline void flush(const PairArgs<R> &P, const Target<R> &T, const Phys &ph, const unsigned *list, int nlist,
                             Sums<R> &S) {
    // The double pass is register-bound, and the Gaussian float pass (42
    // contributors, lists flushed together) lost 2.5 % to the lower occupancy:
    // no second record in flight there.
    if constexpr (sizeof(R) == 8 || KERNEL == SPH_KERNEL_GAUSSIAN) {
        for (int j = 0; j < nlist; ++j) {
            const unsigned u = list[j * blockDim.x];
            contribute<R, KERNEL, MODE, XS>(P, T, ph, u, P.A[u], S);
        }
        return;
    }
    // software-pipelined as the walk: the next listed record is in flight
    unsigned un = 0;
    typename Rec<R>::A an{};
    if (nlist > 0) {
        un = list[0];
        an = P.A[un];
    }
    for (int j = 0; j < nlist; ++j) {
        const unsigned u = un;
        const typename Rec<R>::A a = an;
        if (j + 1 < nlist) {
            un = list[(j + 1) * blockDim.x];
            an = P.A[un];
        }
        contribute<R, KERNEL, MODE, XS>(P, T, ph, u, a, S);
    }
}

// Pass 1 over the candidates [s, e) of one cell (descending for the reference
// order): counts neighbours and appends contributors to this thread's list,
// flushing it through pass 2 when full. `implied`: no count is wanted and the
// kernel support lies well inside the neighbour cut-off (2h against 3h for
// the cubic spline and Wendland), so a fluid candidate inside the support is
// a neighbour and the cut-off test is skipped for it.
template <typename R, int KERNEL, bool DESC, int MODE, bool XS = false, int LIST = list_len<R, KERNEL>()>
__device__ inline int scan_cell(const PairArgs<R> &P, const Target<R> &T, const Phys &ph, int s, int e, unsigned *list,
                                int &nlist, Sums<R> &S, bool implied) {
    constexpr double kEps = sizeof(R) == 8 ? 1e-12 : 1e-5;
    const R qmax = (R)ph.qmax;
    const R ks = (R)support_q<KERNEL>();
    const R r0sq = (R)ph.r0 * (R)ph.r0 * R(1.0 + kEps);
    int cnt = 0;
    const int len = e - s;
    // software-pipelined: the next candidate's record (the float pass: the
    // next two) is in flight while this one is tested
    constexpr int kAhead = sizeof(R) == 4 ? 2 : 1;
    typename Rec<R>::A nxt{}, nxt2{};
    if (len > 0) nxt = P.A[(unsigned)(DESC ? e - 1 : s)];
    if (kAhead == 2 && len > 1) nxt2 = P.A[(unsigned)(DESC ? e - 2 : s + 1)];
    for (int k = 0; k < len; ++k) {
        const unsigned u = (unsigned)(DESC ? e - 1 - k : s + k);
        const typename Rec<R>::A a = nxt;
        if constexpr (kAhead == 2) {
            nxt = nxt2;
            if (k + 2 < len) nxt2 = P.A[(unsigned)(DESC ? e - 3 - k : s + k + 2)];
        } else {
            if (k + 1 < len) nxt = P.A[(unsigned)(DESC ? e - 2 - k : s + k + 1)];
        }
        const R dx = (R)(T.X - a.x), dy = (R)(T.Y - a.y);
        const R r2 = dx * dx + dy * dy;
        const R hij = R(0.5) * (T.h + (R)a.h);
        const R sh = ks * hij;
        // XSPH also needs the walls inside the kernel support
        const R sup2 = a.w >= R(0) ? sh * sh * R(1.0 + kEps)
                                   : (XS ? fmax(sh * sh * R(1.0 + kEps), r0sq) : r0sq);
        bool in = true;
        if (!implied || !(a.w >= R(0))) {
            const R qh = qmax * hij;
            const R lim = qh * qh;
            in = r2 < lim * R(1.0 - kEps);
            if (!in && !(r2 > lim * R(1.0 + kEps))) {
                // within the rounding band of the cut-off: the reference's exact test
                const double hj = sizeof(R) == 8 ? (double)a.h : P.H[u];
                in = exact_q_test(T.X, T.Y, T.H, a.x, a.y, hj, ph.qmax);
            }
            cnt += in;
        }
        // r = 0 pairs contribute exactly nothing (to XSPH they do, W(0) != 0,
        // unless it is the target itself); walls add nothing to continuity
        const bool apart = XS ? (r2 > R(0) || (int)u != T.slot) : r2 > R(0);
        if (in && apart && !(r2 > sup2) && (MODE != kContinuity || a.w >= R(0))) {
            list[nlist * blockDim.x] = (unsigned)u;
            ++nlist;
            if constexpr (KERNEL != SPH_KERNEL_GAUSSIAN) {
                // 2h support: ~21 contributors, the list rarely fills
                if (nlist == LIST) {
                    flush<R, KERNEL, MODE, XS>(P, T, ph, list, nlist, S);
                    nlist = 0;
                }
            }
        }
        if constexpr (KERNEL == SPH_KERNEL_GAUSSIAN) {
            // 3h support: ~42 contributors, every list fills at least once.
            // When one lane's list is full, every active lane flushes
            // together; per-lane flushes at different candidates would run
            // the flush loop once per lane. A lane's sums keep their order.
            if (__ballot(nlist == LIST) != 0) {
                flush<R, KERNEL, MODE, XS>(P, T, ph, list, nlist, S);
                nlist = 0;
            }
        }
    }
    return cnt;
}

// Pass 1 over the 3x3 cells around the raw cell c of a grid of ncx x ncy
// cells starting at bin `base`. EXACT (the double pass): the reference's
// order, x shift outer and y shift inner (src/Tools/NNLinkedList.py:50-56),
// each cell from its highest index down (head-inserted lists, :129-141).
// Otherwise (the float pass needs only a fixed order) the three cells of each
// row, which are consecutive bins, as one run of slots.
template <typename R, int KERNEL, bool EXACT, int MODE, bool XS = false, int LIST = list_len<R, KERNEL>()>
__device__ inline int walk_cells(const PairArgs<R> &P, const Target<R> &T, const Phys &ph, int2 c, int ncx, int ncy,
                                 long long base, unsigned *list, int &nlist, Sums<R> &S, bool implied) {
    int cnt = 0;
    if constexpr (EXACT) {
        for (int da = -1; da <= 1; ++da) {
            const int cx = c.x + da;
            if (!(ncx > cx && cx > -1)) continue;
            for (int db = -1; db <= 1; ++db) {
                const int cy = c.y + db;
                if (!(ncy > cy && cy > -1)) continue;
                const long long k = base + cx + (long long)ncx * cy;
                cnt += scan_cell<R, KERNEL, true, MODE, XS, LIST>(P, T, ph, P.cellStart[k], P.cellStart[k + 1], list,
                                                                   nlist, S, implied);
            }
        }
    } else {
        const int lo = c.x - 1 > 0 ? c.x - 1 : 0;
        const int hi = c.x + 1 < ncx - 1 ? c.x + 1 : ncx - 1;
        if (lo > hi) return 0;
        for (int db = -1; db <= 1; ++db) {
            const int cy = c.y + db;
            if (!(ncy > cy && cy > -1)) continue;
            const long long row = base + (long long)ncx * cy;
            cnt += scan_cell<R, KERNEL, false, MODE, XS, LIST>(P, T, ph, P.cellStart[row + lo],
                                                                P.cellStart[row + hi + 1], list, nlist, S, implied);
        }
    }
    return cnt;
}

// The pass for target slot t: the 3x3 cells of its grid (ncx x ncy cells
// starting at bin `base`), sums, outputs.
template <typename R, int KERNEL, bool EXACT, int MODE, int LIST = list_len<R, KERNEL>()>
__device__ inline void pair_target(const PairArgs<R> &P, const Phys &ph, int ncx, int ncy, long long base, int t,
                                   unsigned *s_list) {
    const typename Rec<R>::A at = P.A[t];
    if (!(at.w >= R(0))) {  // walls are not targets
        if (P.count) P.count[t] = -1;
        return;
    }
    const typename Rec<R>::B bt = P.B[t];
    Target<R> T;
    T.X = at.x;
    T.Y = at.y;
    T.H = sizeof(R) == 8 ? (double)at.h : P.H[t];
    T.h = (R)at.h;
    T.vx = (R)rec_vx(at, bt);
    T.vy = (R)rec_vy(at, bt);
    T.rho = (R)bt.rho;
    T.slf = (R)bt.prs;
    const int2 c = P.cxy[t];
    unsigned *list = s_list + threadIdx.x;
    int nlist = 0, cnt = 0;
    Sums<R> S;
    const bool implied = cutoff_implied<KERNEL>(P.count, ph.qmax, c, ncx, ncy, T.H);
    cnt = walk_cells<R, KERNEL, EXACT, MODE, false, LIST>(P, T, ph, c, ncx, ncy, base, list, nlist, S, implied);
    flush<R, KERNEL, MODE>(P, T, ph, list, nlist, S);
    if (P.count) P.count[t] = cnt;
    if (!implied && cnt == 0) return;  // SolverTools.py:166-169: no neighbours, keep the previous values
    if constexpr (MODE != kMomentum) P.drho[t] = S.arho.get();
    if constexpr (MODE != kContinuity) {
        const double a0 = S.a0.get(), a1 = S.a1.get(), f0 = S.f0.get(), f1 = S.f1.get();
        P.ax[t] = a0 + f0;
        P.ay[t] = (a1 - ph.g) + f1;  // WCSPH.py:169, then SolverTools.py:170-171
    }
}

// The single-simulation pass keeps its own copy of the body: sharing
// pair_target with the batched kernel changed its register allocation and
// cost ~11% (measured on the 1M dam break).
template <typename R, int KERNEL, bool EXACT, int MODE, bool XS>
__global__ __launch_bounds__(kPairThreads) void k_pair(const Ctrl *ctrl, PairArgs<R> P, Phys ph) {
    __shared__ unsigned s_list[list_len<R, KERNEL>() * kPairThreads];
    if (!ctrl->active) return;
    const int t = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= P.n) return;
    const typename Rec<R>::A at = P.A[t];
    if (!(at.w >= R(0))) {  // walls are not targets
        if (P.count) P.count[t] = -1;
        return;
    }
    const typename Rec<R>::B bt = P.B[t];
    const Grid g = ctrl->grid;
    const int ncx = (int)g.ncx, ncy = (int)g.ncy;
    Target<R> T;
    T.X = at.x;
    T.Y = at.y;
    T.H = sizeof(R) == 8 ? (double)at.h : P.H[t];
    T.h = (R)at.h;
    T.vx = (R)rec_vx(at, bt);
    T.vy = (R)rec_vy(at, bt);
    T.rho = (R)bt.rho;
    T.slf = (R)bt.prs;
    T.slot = t;
    const int2 c = P.cxy[t];
    unsigned *list = s_list + threadIdx.x;
    int nlist = 0, cnt = 0;
    Sums<R> S;
    const bool implied = cutoff_implied<KERNEL>(P.count, ph.qmax, c, ncx, ncy, T.H);
    cnt = walk_cells<R, KERNEL, EXACT, MODE, XS>(P, T, ph, c, ncx, ncy, 0, list, nlist, S, implied);
    flush<R, KERNEL, MODE, XS>(P, T, ph, list, nlist, S);
    if (P.count) P.count[t] = cnt;
    if (!implied && cnt == 0) return;  // SolverTools.py:166-169: no neighbours, keep the previous values
    if constexpr (MODE != kMomentum) P.drho[t] = S.arho.get();
    if constexpr (MODE != kContinuity) {
        const double a0 = S.a0.get(), a1 = S.a1.get(), f0 = S.f0.get(), f1 = S.f1.get();
        P.ax[t] = a0 + f0;
        P.ay[t] = (a1 - ph.g) + f1;  // WCSPH.py:169, then SolverTools.py:170-171
    }
    if constexpr (XS) {  // WCSPH.compute_velocity: xsph = v + correction (WCSPH.py:180-186)
        const double vxi = sizeof(R) == 8 ? (double)T.vx : P.VX[t];
        const double vyi = sizeof(R) == 8 ? (double)T.vy : P.VY[t];
        P.xsx[t] = vxi + S.xs0.get();
        P.xsy[t] = vyi + S.xs1.get();
    }
}

// Summation density (src/Tools/SolverTools.py:129-140, src/Equations/
// SummationDensity.py): for each fluid target, rho = the sum over its fluid
// neighbours (same walk and neighbour rule as the pair pass, the target itself
// included) of m_j W(r_ij, h_ij), in the reference's order for the double
// pass. The loop's pressure is the one of the density before the sum
// (:122-127 run first), so the target's record gets rho_new and
// p(rho_old) / rho_new^2 for the pair pass that follows; only the A records
// are read here, so rewriting B in place is safe. A target with neighbours
// gets drho = 0 (WCSPH.compute_density_change, WCSPH.py:200-201); one with
// none keeps its drho and, as in the reference, takes the empty sum 0 as its
// density.
// One run of candidate slots [s0, e0) of the summation density's walk (DESC:
// from the highest index down), next record in flight. `implied` as in
// scan_cell: a fluid candidate inside the support is a neighbour, one outside
// adds W = 0 exactly, walls add nothing, and the count is not formed.
template <typename R, int KERNEL, bool DESC>
__device__ inline int sumden_run(const PairArgs<R> &P, double X, double Y, double H, R hi, const Phys &ph, int s0,
                                 int e0, Acc<R> &rho, bool implied) {
    constexpr double kEps = sizeof(R) == 8 ? 1e-12 : 1e-5;
    const R qmax = (R)ph.qmax;
    const R ks = (R)support_q<KERNEL>();
    int cnt = 0;
    const int len = e0 - s0;
    typename Rec<R>::A nxt{};
    if (len > 0) nxt = P.A[(unsigned)(DESC ? e0 - 1 : s0)];
    for (int i = 0; i < len; ++i) {
        const unsigned u = (unsigned)(DESC ? e0 - 1 - i : s0 + i);
        const typename Rec<R>::A a = nxt;
        if (i + 1 < len) nxt = P.A[(unsigned)(DESC ? e0 - 2 - i : s0 + i + 1)];
        const R dx = (R)(X - a.x), dy = (R)(Y - a.y);
        const R r2 = dx * dx + dy * dy;
        const R hij = R(0.5) * (hi + (R)a.h);
        if (implied) {
            if (!(a.w >= R(0))) continue;
            const R sh = ks * hij;
            if (r2 > sh * sh * R(1.0 + kEps)) continue;
        } else {
            const R qh = qmax * hij;
            const R lim = qh * qh;
            bool in = r2 < lim * R(1.0 - kEps);
            if (!in && !(r2 > lim * R(1.0 + kEps))) {
                const double hj = sizeof(R) == 8 ? (double)a.h : P.H[u];
                in = exact_q_test(X, Y, H, a.x, a.y, hj, ph.qmax);
            }
            cnt += in;
            if (!in || !(a.w >= R(0))) continue;
        }
        if constexpr (sizeof(R) == 8) {
            const double ex = X - a.x, ey = Y - a.y;
            const double r = sqrt(ex * ex + ey * ey);
            rho.add((double)a.w * w_exact<KERNEL>(r, 0.5 * (H + (double)a.h)));
        } else {
            const float rinv = r2 > 0.f ? __builtin_amdgcn_rsqf(r2) : 0.f;
            const float hinv = __builtin_amdgcn_rcpf(hij);
            rho.add((float)a.w * w_f<KERNEL>(r2 * rinv * hinv, hinv));
        }
    }
    return cnt;
}

template <typename R, int KERNEL, bool EXACT>
__global__ __launch_bounds__(kPairThreads) void k_sumden(const Ctrl *ctrl, PairArgs<R> P, Phys ph, SdOut O) {
    if (!ctrl->active) return;
    const int t = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= P.n) return;
    const typename Rec<R>::A at = P.A[t];
    if (!(at.w >= R(0))) return;  // fluid only
    const Grid g = ctrl->grid;
    const int ncx = (int)g.ncx, ncy = (int)g.ncy;
    const double X = at.x, Y = at.y, H = sizeof(R) == 8 ? (double)at.h : P.H[t];
    const R hi = (R)at.h;
    Acc<R> rho;
    int cnt = 0;
    const int2 c = P.cxy[t];
    const bool implied = cutoff_implied<KERNEL>(nullptr, ph.qmax, c, ncx, ncy, H);
    if constexpr (EXACT) {  // the reference's cell order (as walk_cells)
        for (int da = -1; da <= 1; ++da) {
            const int cx = c.x + da;
            if (!(ncx > cx && cx > -1)) continue;
            for (int db = -1; db <= 1; ++db) {
                const int cy = c.y + db;
                if (!(ncy > cy && cy > -1)) continue;
                const int k = cx + ncx * cy;
                cnt += sumden_run<R, KERNEL, true>(P, X, Y, H, hi, ph, P.cellStart[k], P.cellStart[k + 1], rho,
                                                   implied);
            }
        }
    } else {  // row runs
        const int lo = c.x - 1 > 0 ? c.x - 1 : 0;
        const int hi_x = c.x + 1 < ncx - 1 ? c.x + 1 : ncx - 1;
        for (int db = -1; db <= 1 && lo <= hi_x; ++db) {
            const int cy = c.y + db;
            if (!(ncy > cy && cy > -1)) continue;
            const int row = ncx * cy;
            cnt += sumden_run<R, KERNEL, false>(P, X, Y, H, hi, ph, P.cellStart[row + lo],
                                                P.cellStart[row + hi_x + 1], rho, implied);
        }
    }
    const double rho_new = rho.get();
    const double ratio = ipow(O.rho[t] / ph.rho0, ph.gamma);
    const double p_old = (ratio - 1.0) * ph.B + ph.Pb;  // Tait + Pb of the density before the sum
    O.rho[t] = rho_new;
    typename Rec<R>::B *B = (typename Rec<R>::B *)O.B;
    B[t].rho = (R)rho_new;
    B[t].prs = (R)(p_old / (rho_new * rho_new));
    if (implied || cnt > 0) O.drho[t] = 0.0;
}

template <typename R, bool EXACT>
hipError_t launch_sumden(const Ctrl *ctrl, const PairArgs<R> &P, const SdOut &O, const Phys &ph, int threads,
                         hipStream_t s) {
    if (threads != 64 && threads != 128 && threads != 256) threads = kPairThreads;
    const int nb = xcd_blocks(P.n, threads);
    switch (ph.kernel) {
        case SPH_KERNEL_GAUSSIAN:
            k_sumden<R, SPH_KERNEL_GAUSSIAN, EXACT><<<nb, threads, 0, s>>>(ctrl, P, ph, O);
            break;
        case SPH_KERNEL_WENDLAND:
            k_sumden<R, SPH_KERNEL_WENDLAND, EXACT><<<nb, threads, 0, s>>>(ctrl, P, ph, O);
            break;
        default:
            k_sumden<R, SPH_KERNEL_CUBIC, EXACT><<<nb, threads, 0, s>>>(ctrl, P, ph, O);
    }
    return hipGetLastError();
}

// Batched ensemble: each target walks its own instance's cells with its own
// constants; targets of finished or halted instances do nothing. LIST (the
// per-thread contributor list, LIST x 4 B x the workgroup in LDS) is a tile
// size the batch's autotuner chooses.
template <typename R, int KERNEL, bool EXACT, int LIST>
__global__ __launch_bounds__(kPairThreads) void k_pair_batch(const BInst *inst, const int *instOf, PairArgs<R> P) {
    __shared__ unsigned s_list[LIST * kPairThreads];
    const int t = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= P.n) return;
    const BInst &I = inst[instOf[t]];
    if (!I.active) return;
    pair_target<R, KERNEL, EXACT, kFused, LIST>(P, inst_phys(I), (int)I.grid.ncx, (int)I.grid.ncy, I.cellBase, t,
                                                s_list);
}

// The same pass as a persistent kernel: a grid of about one resident
// occupancy (CUs x workgroups per CU) takes tiles of blockDim.x consecutive
// slots from a work queue until it is drained. The queue is one counter per
// XCD, each over a contiguous eighth of the tiles, and the workgroups the
// dispatcher hands to an XCD (blockIdx % 8) take from that XCD's counter, so
// neighbouring tiles still share an L2. Every workgroup ends once its
// counter has passed the end of its range; none waits for another. The last
// workgroup out (a separate counter) clears all counters, so the next launch,
// graph replays included, starts from zero without a memset. Each target's
// sums do not depend on which workgroup runs it: results are those of
// k_pair_batch bit for bit.
template <typename R, int KERNEL, bool EXACT, int LIST>
__global__ __launch_bounds__(kPairThreads) void k_pair_batch_queue(const BInst *inst, const int *instOf,
                                                                   PairArgs<R> P, int *queue) {
    __shared__ unsigned s_list[LIST * kPairThreads];
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
        if (t < P.n) {
            const BInst &I = inst[instOf[t]];
            if (I.active)
                pair_target<R, KERNEL, EXACT, kFused, LIST>(P, inst_phys(I), (int)I.grid.ncx, (int)I.grid.ncy,
                                                            I.cellBase, t, s_list);
        }
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(queue + kXcds * kQueueStride, 1) == (int)gridDim.x - 1)
            for (int k = 0; k <= kXcds; ++k) atomicExch(queue + k * kQueueStride, 0);
    }
}

// The batch's list lengths; 0 picks the single-simulation pass's. queue:
// the persistent form over `grid` workgroups (a multiple of 8), else one
// thread per slot.
template <typename R, int KERNEL, bool EXACT>
void launch_batch_list(const BInst *inst, const int *instOf, const PairArgs<R> &P, int list, int nb, int threads,
                       int *queue, int grid, hipStream_t s) {
    if (list == 0) list = list_len<R, KERNEL>();
    if (queue) {
        if (list == 16)
            k_pair_batch_queue<R, KERNEL, EXACT, 16><<<grid, threads, 0, s>>>(inst, instOf, P, queue);
        else if (list == 24)
            k_pair_batch_queue<R, KERNEL, EXACT, 24><<<grid, threads, 0, s>>>(inst, instOf, P, queue);
        else
            k_pair_batch_queue<R, KERNEL, EXACT, 32><<<grid, threads, 0, s>>>(inst, instOf, P, queue);
        return;
    }
    if (list == 16)
        k_pair_batch<R, KERNEL, EXACT, 16><<<nb, threads, 0, s>>>(inst, instOf, P);
    else if (list == 24)
        k_pair_batch<R, KERNEL, EXACT, 24><<<nb, threads, 0, s>>>(inst, instOf, P);
    else
        k_pair_batch<R, KERNEL, EXACT, 32><<<nb, threads, 0, s>>>(inst, instOf, P);
}

// Resident workgroups per CU of the persistent batch pass (hipOccupancy), for
// sizing its grid outside any stream capture.
template <typename R, int KERNEL, bool EXACT>
int queue_occupancy_list(int list, int threads) {
    if (list == 0) list = list_len<R, KERNEL>();
    int per = 0;
    hipError_t e;
    if (list == 16)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_pair_batch_queue<R, KERNEL, EXACT, 16>, threads, 0);
    else if (list == 24)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_pair_batch_queue<R, KERNEL, EXACT, 24>, threads, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_pair_batch_queue<R, KERNEL, EXACT, 32>, threads, 0);
    return e == hipSuccess && per > 0 ? per : 1;
}

template <typename R, bool EXACT>
int queue_occupancy(int kernel, int list, int threads) {
    switch (kernel) {
        case SPH_KERNEL_GAUSSIAN: return queue_occupancy_list<R, SPH_KERNEL_GAUSSIAN, EXACT>(list, threads);
        case SPH_KERNEL_WENDLAND: return queue_occupancy_list<R, SPH_KERNEL_WENDLAND, EXACT>(list, threads);
        default: return queue_occupancy_list<R, SPH_KERNEL_CUBIC, EXACT>(list, threads);
    }
}

template <typename R, bool EXACT, int MODE, bool XS = false>
void launch_pair_mode(const Ctrl *ctrl, const PairArgs<R> &P, const Phys &ph, int nb, int threads, hipStream_t s) {
    switch (ph.kernel) {
        case SPH_KERNEL_GAUSSIAN:
            k_pair<R, SPH_KERNEL_GAUSSIAN, EXACT, MODE, XS><<<nb, threads, 0, s>>>(ctrl, P, ph);
            break;
        case SPH_KERNEL_WENDLAND:
            k_pair<R, SPH_KERNEL_WENDLAND, EXACT, MODE, XS><<<nb, threads, 0, s>>>(ctrl, P, ph);
            break;
        default:
            k_pair<R, SPH_KERNEL_CUBIC, EXACT, MODE, XS><<<nb, threads, 0, s>>>(ctrl, P, ph);
    }
}

template <typename R, bool EXACT>
hipError_t launch_pair_batch(const BInst *inst, const int *instOf, const PairArgs<R> &P, int kernel, int threads,
                             int list, int *queue, int grid, hipStream_t s) {
    if (threads != 64 && threads != 128 && threads != 256) threads = kPairThreads;
    const int nb = xcd_blocks(P.n, threads);
    if (queue) {  // no more workgroups than tiles; a multiple of the XCD count, so every queue has takers
        if (grid > nb) grid = nb;
        grid = grid / kXcds * kXcds;
        if (grid < kXcds) grid = kXcds;
    }
    switch (kernel) {
        case SPH_KERNEL_GAUSSIAN:
            launch_batch_list<R, SPH_KERNEL_GAUSSIAN, EXACT>(inst, instOf, P, list, nb, threads, queue, grid, s);
            break;
        case SPH_KERNEL_WENDLAND:
            launch_batch_list<R, SPH_KERNEL_WENDLAND, EXACT>(inst, instOf, P, list, nb, threads, queue, grid, s);
            break;
        default:
            launch_batch_list<R, SPH_KERNEL_CUBIC, EXACT>(inst, instOf, P, list, nb, threads, queue, grid, s);
    }
    return hipGetLastError();
}

template <typename R, bool EXACT>
hipError_t launch_pair(const Ctrl *ctrl, const PairArgs<R> &P, const Phys &ph, int threads, int mode, hipStream_t s) {
    if (threads != 64 && threads != 128 && threads != 256) threads = kPairThreads;
    const int nb = xcd_blocks(P.n, threads);
    const bool xs = P.xsx != nullptr;  // XSPH velocity requested (fused and momentum passes only)
    if (mode == kContinuity)
        launch_pair_mode<R, EXACT, kContinuity>(ctrl, P, ph, nb, threads, s);
    else if (mode == kMomentum)
        xs ? launch_pair_mode<R, EXACT, kMomentum, true>(ctrl, P, ph, nb, threads, s)
           : launch_pair_mode<R, EXACT, kMomentum>(ctrl, P, ph, nb, threads, s);
    else
        xs ? launch_pair_mode<R, EXACT, kFused, true>(ctrl, P, ph, nb, threads, s)
           : launch_pair_mode<R, EXACT, kFused>(ctrl, P, ph, nb, threads, s);
    return hipGetLastError();
}

}  // namespace
}  // namespace hipsph
