// sph_pair_fp64.hip — the fused pair pass in double precision.
//
// One thread per cell-ordered particle; fluid particles gather over the 3x3
// block of cells around their raw cell, in the reference's order (x-shift
// outer, y-shift inner, descending reference index inside a cell), and sum
// continuity, momentum + artificial viscosity and the Lennard-Jones wall force
// exactly as src/Tools/SolverTools.py:143-173 composes them. Built with
// -ffp-contract=off and no fast-math: every operation is the reference's,
// in its order, so for the cubic spline the results are bit-identical to the
// CPU oracle; the Gaussian differs only through exp().
//
// Neighbour predicate: the reference tests q = sqrt(r^2) / h_ij <= 3 with
// correctly rounded sqrt and division (src/Tools/NNLinkedList.py:63-71). Pairs
// with r^2 > (3 h_ij)^2 (1 + 1e-12) cannot pass that test (the rounding of the
// three operations is below 1e-15 relative), so they are rejected without the
// sqrt/division; every other pair evaluates the reference's exact expression.
#include "sph_internal.h"

namespace hipsph {
namespace {

constexpr double kPI = 3.141592653589793;

// Kernel gradient with h = h_ij, r = r_ij and q = r / h (the reference
// recomputes q = r / h inside the kernel; it is the same value).
// CubicSpline: src/Kernels/CubicSpline.py:38-70.
// Gaussian:    src/Kernels/Gaussian.py:39-60.
// Wendland:    src/Kernels/Wendland.py:32-64.
template <int KERNEL>
__device__ inline void grad_w(double dx, double dy, double r, double q, double h, double &dwx, double &dwy) {
    dwx = 0.0;
    dwy = 0.0;
    if (KERNEL == SPH_KERNEL_GAUSSIAN) {
        double tmp = r * h;
        if (tmp > 1e-12 && q <= 3) {
            double alpha_c = (1.0 / kPI) / (h * h);
            double dwdq = -2 * q * alpha_c * exp(-q * q);
            double f = dwdq / tmp;
            dwx = f * dx;
            dwy = f * dy;
        }
    } else if (KERNEL == SPH_KERNEL_WENDLAND) {
        double alpha = 9.0 / (4.0 * kPI) / (h * h);
        if (!((q >= 2.0) || (r < 1e-10))) {
            double inner = 1.0 - 0.5 * q;
            double grad = ipow(inner, 5.0) * (-14.0 / 3.0) * q * (1 + 2.5 * q);
            double norm = alpha * grad / (h * r);
            dwx = norm * dx;
            dwy = norm * dy;
        }
    } else {
        double alpha = 10.0 / (7.0 * kPI) / (h * h);
        if (!((q > 2.0) || (r < 1e-10))) {
            double grad = (q > 1.0) ? -0.75 * ((2.0 - q) * (2.0 - q)) : -3.0 * q * (1.0 - 0.75 * q);
            double norm = alpha * grad / (h * r);
            dwx = norm * dx;
            dwy = norm * dy;
        }
    }
}

template <int KERNEL>
__global__ __launch_bounds__(256) void k_pair_fp64(const Ctrl *ctrl, PairIn in, PairOut out, Phys ph, int n) {
    if (!ctrl->active) return;
    const int t = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= n || in.label[t] != kFluid) return;
    const Grid g = ctrl->grid;
    const double xi = in.x[t], yi = in.y[t], vxi = in.vx[t], vyi = in.vy[t], rhoi = in.rho[t], hi = in.h[t];
    const double slf = in.prs[t];  // p_i / rho_i^2 (Momentum.py:28)
    const int2 c = in.cxy[t];
    const double qmax = ph.qmax;
    double arho = 0.0, a0 = 0.0, a1 = 0.0, f0 = 0.0, f1 = 0.0;
    int cnt = 0;
    for (int da = -1; da <= 1; ++da) {
        const long long cx = (long long)c.x + da;
        if (!(g.ncx > cx && cx > -1)) continue;
        for (int db = -1; db <= 1; ++db) {
            const long long cy = (long long)c.y + db;
            if (!(g.ncy > cy && cy > -1)) continue;
            const int cell = (int)(cx + g.ncx * cy);
            const int s = in.cellStart[cell], e = in.cellStart[cell + 1];
            for (int u = e - 1; u >= s; --u) {
                const double dx = xi - in.x[u], dy = yi - in.y[u];
                const double xij2 = dx * dx + dy * dy;
                const double h_ij = 0.5 * (hi + in.h[u]);
                const double lim = (qmax * h_ij) * (qmax * h_ij);
                if (xij2 > lim * (1.0 + 1e-12)) continue;
                const double r = sqrt(xij2);
                const double q = r / h_ij;
                if (!(q <= qmax)) continue;
                ++cnt;
                if (in.label[u] == kFluid) {
                    double dwx, dwy;
                    grad_w<KERNEL>(dx, dy, r, q, h_ij, dwx, dwy);
                    const double vx = vxi - in.vx[u], vy = vyi - in.vy[u];
                    const double mj = in.m[u];
                    // Continuity.py:13-16
                    const double dot = vx * dwx + vy * dwy;
                    arho += mj * dot;
                    // Momentum.py:36-56 (c_j is never filled: c_ij = 0.5 c_i)
                    const double othr = in.prs[u];
                    const double dv = vx * dx + vy * dy;
                    double PI_ij = 0.0;
                    if (dv < 0) {
                        const double hij = 0.5 * (hi + h_ij);
                        const double cij = 0.5 * (ph.co + 0.0);
                        const double rhoij = 0.5 * (rhoi + in.rho[u]);
                        const double muij = hij * dv / (r * r + 0.01 * hij * hij);
                        const double ppij = muij * (ph.beta * muij - ph.alpha * cij);
                        PI_ij = ppij / rhoij;
                    }
                    const double factor = slf + othr + PI_ij;
                    a0 += -mj * factor * dwx;
                    a1 += -mj * factor * dwy;
                } else if (!(r > ph.r0) && r > 1e-12) {
                    // BoundaryForce.py:31-41
                    const double frac = ph.r0 / r;
                    const double tmp = ipow(frac, ph.p1) - ipow(frac, ph.p2);
                    const double fac = ph.D * tmp;
                    f0 += fac * dx / (r * r);
                    f1 += fac * dy / (r * r);
                }
            }
        }
    }
    if (cnt == 0) return;  // SolverTools.py:166-169: no neighbours, keep previous values
    out.drho[t] = arho;
    out.ax[t] = a0 + f0;
    out.ay[t] = (a1 - ph.g) + f1;  // WCSPH.py:169 then SolverTools.py:170-171
}

}  // namespace

hipError_t launch_pair_fp64(const Ctrl *ctrl, PairIn in, PairOut out, Phys ph, int n, hipStream_t s) {
    const int nb = xcd_blocks(n, 256);
    switch (ph.kernel) {
        case SPH_KERNEL_GAUSSIAN:
            k_pair_fp64<SPH_KERNEL_GAUSSIAN><<<nb, 256, 0, s>>>(ctrl, in, out, ph, n);
            break;
        case SPH_KERNEL_WENDLAND:
            k_pair_fp64<SPH_KERNEL_WENDLAND><<<nb, 256, 0, s>>>(ctrl, in, out, ph, n);
            break;
        default:
            k_pair_fp64<SPH_KERNEL_CUBIC><<<nb, 256, 0, s>>>(ctrl, in, out, ph, n);
    }
    return hipGetLastError();
}

}  // namespace hipsph
