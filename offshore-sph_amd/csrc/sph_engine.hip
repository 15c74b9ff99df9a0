// sph_engine.hip — neighbour search, integrators, time step and the C ABI.
//
// Built with -ffp-contract=off: the neighbour predicate, cell keys, bounds,
// PEC/Euler updates and the dt reduction are IEEE-exact jitclass code in the
// reference (no fastmath), so every + - * / sqrt here is one correctly
// rounded double operation in the reference's order, and the results are
// bit-identical to it given identical inputs.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <algorithm>
#include <mutex>
#include <vector>

#include "sph_ctx.h"
#include "sph_common.h"

using namespace hipsph;

// ---------------------------------------------------------------------------
// error handling
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

namespace hipsph {
int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
}  // namespace hipsph

Phys hipsph::make_phys(const sph_params_t &P) {
    Phys ph;
    ph.rho0 = P.rho0;
    ph.gamma = P.gamma;
    ph.B = P.B;
    ph.Pb = P.Pb;
    ph.co = P.co;
    ph.alpha = P.alpha;
    ph.beta = P.beta;
    ph.r0 = P.r0;
    ph.D = P.D;
    ph.p1 = P.p1;
    ph.p2 = P.p2;
    ph.g = P.gravity;
    ph.qmax = P.qmax;
    ph.kernel = P.kernel;
    ph.eps = P.epsilon;
    return ph;
}

// ---------------------------------------------------------------------------
// time step: src/Equations/TimeStep.py:58-90 computeVars over the fluid
// (per-workgroup partials: DtVars), then Solver._minTimeStep
// ---------------------------------------------------------------------------
// Dt reduction over the fluid of the current state (used when no corrector
// has left the reduction of this state behind). Grid-stride over R workgroups.
__global__ void k_dt_partial(Ctrl *ctrl, const int8_t *label, const double *h, const double *ax, const double *ay,
                             int n, double *part) {
    __shared__ double lds[kDtQ * 16];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctrl->dt_parts = gridDim.x;
        ctrl->pec_parts = 0;
    }
    DtVars v;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (label[i] == kFluid) v.add(h[i], ax[i], ay[i]);
    v.publish(part, lds);
}

// Solver._minTimeStep (src/Solver.py:210-230): courant (TimeStep.py:47-49),
// force (:51-56), min(c, f) (:45), from the partials of the fluid reduction
// (ctrl->dt_parts of them, as the kernel that wrote them recorded). Also opens the step: the reference's loop condition
// `while self.t < self.duration` (src/Solver.py:366) and `self.t += self.dt`
// (:409) move here, so a captured step needs no host. It also bounds how far
// this step's predictor puts any particle from where the last update binned
// it (ctrl->disp, for k_bounds): after the fused PEC pass of the last step
// (ctrl->pec_parts; dt0 its dt, V = ctrl->vbound its velocity bound at the
// start, A0 = ctrl->abound its predictor's rates, A1 this fold's), a fluid
// particle binned at x + dt0 v / 2 was corrected to x + dt0 vh, with |vh| <=
// V + dt0 A0 / 2 (damping only shrinks it), and its new velocity v' has
// |v'| <= V + dt0 A1, so per axis it now lies within dt0 (V + dt0 A0 / 2) +
// dt0 V / 2 of where it was binned and this step's predictor moves it dt |v'|
// / 2 further; V + dt0 A1 bounds the next step's start. Anything else (other
// partials, non-finite rates, a halted step, a second call on the same
// partials) leaves +inf: the next update scans every particle. One workgroup.
__global__ void k_dt_final(Ctrl *ctrl, double gamma_c, double gamma_f, double *hist, const double *part) {
    __shared__ double lds[kDtQ * 16];
    const int R = ctrl->dt_parts;
    // control fields are loaded with the partials (used after the fold)
    unsigned flags = ctrl->flags;
    const double t0 = ctrl->t, duration = ctrl->duration, fixed_dt = ctrl->fixed_dt, co = ctrl->co;
    const long long step = ctrl->step;
    const int advance_t = ctrl->advance_t;
    const int kind[kDtQ] = {0, 1, 2, 2};
    double v[kDtQ] = {INFINITY, -INFINITY, 0.0, 0.0};
    // eight partials per thread per round, their loads issued together (the
    // fused pair pass leaves one per workgroup: ~5,000 at 1M particles, one
    // round; four per round: 6.64 against 6.38 us; 512 threads in two
    // rounds 6.85, 256 threads with sixteen per round 12.0)
    constexpr int U = 8;
    for (int k0 = threadIdx.x; k0 < R; k0 += U * blockDim.x) {
        double u[U][kDtQ];
#pragma unroll
        for (int j = 0; j < U; ++j) {
            const int k = k0 + j * blockDim.x;
#pragma unroll
            for (int q = 0; q < kDtQ; ++q) u[j][q] = k < R ? part[q * R + k] : (q == 0 ? INFINITY : (q == 1 ? -INFINITY : 0.0));
        }
#pragma unroll
        for (int j = 0; j < U; ++j) {
            v[0] = u[j][0] < v[0] ? u[j][0] : v[0];
            v[1] = u[j][1] > v[1] ? u[j][1] : v[1];
            v[2] += u[j][2];
            v[3] += u[j][3];
        }
    }
    block_fold<kDtQ>(v, kind, lds);
    if (threadIdx.x != 0) return;
    const int pec = ctrl->pec_parts;
    const double dt0 = ctrl->dt, V = ctrl->vbound, A0 = ctrl->abound;
    ctrl->pec_parts = 0;  // these partials are used once
    ctrl->disp = INFINITY;
    ctrl->vbound = INFINITY;
    ctrl->abound = INFINITY;
    const double mn = v[0], mx = v[1], cnt = v[2], bad = v[3];
    const unsigned halt = SPH_FLAG_CAPACITY | SPH_FLAG_NONFINITE | SPH_FLAG_NO_FLUID | SPH_FLAG_ZERO_DENSITY;
    if ((flags & halt) || !(t0 < duration)) {
        ctrl->active = 0;
        return;
    }
    if (cnt == 0) {
        ctrl->flags = flags | SPH_FLAG_NO_FLUID;
        ctrl->active = 0;
        return;
    }
    if (bad != 0) flags |= SPH_FLAG_NONFINITE;
    double dt, c, f;
    if (fixed_dt > 0.0) {
        dt = fixed_dt;  // src/Solver.py:370-372
        c = 0.0;
        f = 0.0;
    } else {
        const double max_c = co;  // c = co for every fluid particle (WCSPH.py:110-129)
        c = gamma_c * mn / max_c;
        f = (mx < 1e-12) ? 1e10 : gamma_f * sqrt(mn / mx);
        dt = (f < c) ? f : c;
        if (dt < 1e-6) flags |= SPH_FLAG_DT_SMALL;  // src/Solver.py:220-224
    }
    const long long k = step % SPH_DT_HISTORY;
    hist[3 * k + 0] = dt;
    hist[3 * k + 1] = c;
    hist[3 * k + 2] = f;
    ctrl->flags = flags;
    ctrl->dt = dt;
    ctrl->dt_c = c;
    ctrl->dt_f = f;
    if (advance_t) ctrl->t = t0 + dt;
    ctrl->step = step + 1;
    ctrl->active = 1;
    if (bad == 0) {
        const double A1 = sqrt(mx);
        ctrl->abound = A1;
        if (pec) {
            const double V1 = V + dt0 * A1;
            const double d = dt0 * (V + 0.5 * dt0 * A0) + 0.5 * dt0 * V + 0.5 * dt * V1;
            if (d == d) ctrl->disp = d;  // NaN: +inf
            ctrl->vbound = V1;
        }
    }
}

// KineticEnergy (src/Equations/KineticEnergy.py:5-12) over the fluid:
// per-workgroup partials, then one workgroup folds them into part[R].
__global__ void k_ke_partial(const State s, int n, double *part) {
    __shared__ double lds[8];
    double k = 0.0;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
        if (s.label[j] == kFluid) {
            const double vx = s.vx[j], vy = s.vy[j];
            k += 0.5 * s.m[j] * (vx * vx + vy * vy);
        }
    k = block_reduce(k, AddOp(), lds);
    if (threadIdx.x == 0) part[blockIdx.x] = k;
}

__global__ void k_ke_final(double *part, int R) {
    __shared__ double lds[16];
    double k = 0.0;
    for (int b = threadIdx.x; b < R; b += blockDim.x) k += part[b];
    k = block_reduce(k, AddOp(), lds);
    if (threadIdx.x == 0) part[R] = k;
}

// ---------------------------------------------------------------------------
// integrators (fluid only: src/Solver.py:380,396 pass pA[f_indexes])
// ---------------------------------------------------------------------------
__global__ void k_pec_predict(const Ctrl *ctrl, State s, int n, int strict, int xs) {
    if (!ctrl->active) return;
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n || s.label[j] != kFluid) return;
    const double dt = ctrl->dt, damping = ctrl->damping;
    double x = s.x[j], y = s.y[j], vx = s.vx[j], vy = s.vy[j], rho = s.rho[j];
    s.x0[j] = x;
    s.y0[j] = y;
    s.vx0[j] = vx;
    s.vy0[j] = vy;
    s.x[j] = pec_half_pos(x, xs ? s.xsphx[j] : vx, dt);  // PEC.py:43-48
    s.y[j] = pec_half_pos(y, xs ? s.xsphy[j] : vy, dt);
    s.vx[j] = pec_half_vel(vx, s.ax[j], dt, damping);
    s.vy[j] = pec_half_vel(vy, s.ay[j], dt, damping);
    s.rho0[j] = rho;
    s.rho[j] = pec_half_rho(rho, s.drho[j], dt, strict);
}

// src/Integrators/PEC.py:61-89 and src/Integrators/Euler.py:17-26. Also
// leaves the next step's dt partials (h and a of the fluid do not change
// until the next pair pass) in `part`. Grid-stride over R workgroups.
// PEC's start-of-step values come from x0 ... rho0 of the same buffer, or,
// with `perm` (a step whose reorder did not write them), from the buffer
// before the reorder at the particle's previous slot perm[j]: the same values.
// The half-step velocity likewise comes from vx, vy or is formed again from
// the start-of-step velocity and the previous rates there.
// xs: PEC moves x by the XSPH velocity of this step's loop (PEC.py:66-71).
// With Newmark-beta on, the coupled particles' corrector
// (NewmarkBeta.py:54-66, src/Solver.py:397-398) with the accelerations the
// state holds now (the coupling hook's).
__global__ void k_correct(Ctrl *ctrl, State s, State o, const int *perm, int n, int integrator, int strict, int xs,
                          double *part, HmConst hc) {
    if (!ctrl->active) return;
    __shared__ double lds[kDtQ * 16];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctrl->dt_parts = gridDim.x;
        ctrl->pec_parts = 0;
    }
    const double dt = ctrl->dt, damping = ctrl->damping;
    const int nm = ctrl->nm;
    DtVars v;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const int8_t lab = s.label[j];
        if (lab != kFluid) {
            if (nm && lab == kCoupled) {
                const double nb = ctrl->nm_beta, ng = ctrl->nm_gamma, ax = s.ax[j], ay = s.ay[j];
                s.x[j] = nm_corr_pos(s.x[j], ax, nb, dt);
                s.y[j] = nm_corr_pos(s.y[j], ay, nb, dt);
                s.vx[j] = nm_corr_vel(s.vx[j], ax, ng, dt);
                s.vy[j] = nm_corr_vel(s.vy[j], ay, ng, dt);
            }
            continue;
        }
        const double ax = s.ax[j], ay = s.ay[j];
        v.add(hc.on ? hc.h : s.h[j], ax, ay);
        if (integrator == SPH_INTEGRATOR_PEC) {
            double x0, y0, vx0, vy0, rho0, vxh, vyh;
            if (perm) {
                // the half-step velocity again from the same inputs (the
                // predictor's expression, as in k_reorder)
                const int q = perm[j];
                x0 = o.x[q];
                y0 = o.y[q];
                vx0 = o.vx[q];
                vy0 = o.vy[q];
                rho0 = o.rho[q];
                vxh = pec_half_vel(vx0, o.ax[q], dt, damping);
                vyh = pec_half_vel(vy0, o.ay[q], dt, damping);
            } else {
                x0 = s.x0[j];
                y0 = s.y0[j];
                vx0 = s.vx0[j];
                vy0 = s.vy0[j];
                rho0 = s.rho0[j];
                vxh = s.vx[j];
                vyh = s.vy[j];
            }
            pec_correct(s, j, x0, y0, vx0, vy0, rho0, xs ? s.xsphx[j] : vxh, xs ? s.xsphy[j] : vyh, ax, ay, s.drho[j],
                        dt, damping, strict);
        } else {
            double vx = s.vx[j], vy = s.vy[j];
            s.x[j] = s.x[j] + dt * vx + 0.5 * dt * dt * ax;
            s.y[j] = s.y[j] + dt * vy + 0.5 * dt * dt * ay;
            s.vx[j] = vx + dt * ax;
            s.vy[j] = vy + dt * ay;
            s.rho[j] = s.rho[j] + dt * s.drho[j];
        }
    }
    v.publish(part, lds);
}

// ---------------------------------------------------------------------------
// neighbour search: src/Tools/NNLinkedList.py
// ---------------------------------------------------------------------------
// Position the neighbour update sees: the predicted one when the PEC
// predictor is fused into the update (fluid only), else the stored one.
// pred: 0 none, 1 PEC, 2 PEC moving x by the XSPH velocity (PEC.py:43-48).
// With Newmark-beta on (ctrl->nm) the coupled particles' predicted position
// (NewmarkBeta.py:46-47).
__device__ inline void upd_pos(const Ctrl *ctrl, const State &a, int i, int pred, double dt, double &x,
                               double &y) {
    x = a.x[i];
    y = a.y[i];
    if (pred) {
        // the velocity is loaded with the label, not after it, and selected
        const double vx = pred == 2 ? a.xsphx[i] : a.vx[i], vy = pred == 2 ? a.xsphy[i] : a.vy[i];
        const bool f = a.label[i] == kFluid;
        x = f ? pec_half_pos(x, vx, dt) : x;
        y = f ? pec_half_pos(y, vy, dt) : y;
    }
    if (ctrl->nm && a.label[i] == kCoupled) {
        const double b = ctrl->nm_beta;
        x = nm_pred_pos(x, a.vx[i], a.ax[i], b, dt);
        y = nm_pred_pos(y, a.vy[i], a.ay[i], b, dt);
    }
}

// Whether this update's bounds can come from the border cells alone (the
// fused PEC step's update, with one h): the state is in the cell order of
// the grid it was last binned on (ctrl->grid, the offsets in cellStart) and
// no particle's predicted position lies more than ctrl->disp from where it
// was binned, per axis, with 2 disp under the cell size by a margin far
// above the rounding of the positions and of the cell arithmetic. Then the
// particle binned at the old minimum x (column 0) is now at most disp to its
// right, and one of a column >= 1 was binned at least a cell further right
// and is at most disp to its left: the minimum lies in column 0. The one
// binned at the old maximum x lies beyond ncx - 1 cells and one of a column
// <= ncx - 3 within ncx - 2, so the maximum lies in columns ncx - 2 or
// ncx - 1; y likewise with rows 0, ncy - 2 and ncy - 1.
__device__ inline bool border_bounds_ok(const Ctrl *ctrl, const int *cellStart, int pred, const HmConst &hc) {
    if (!cellStart || pred != 1 || !hc.on || ctrl->nm || !ctrl->nn_valid) return false;
    return border_margin_ok(ctrl->grid, ctrl->disp);
}

// _init :86-101 bounds and _get_cell_size :103-110 min h, over all active
// particles: per-workgroup partials, folded by k_grid. Grid-stride over at
// most kBoundsBlocks workgroups, four particles per thread per round with
// their loads issued together (one workgroup per 256 particles measured 40 %
// slower, the per-workgroup fold paid in more rounds; 2048 workgroups with
// one particle per round 3 % slower). With cellStart (the fused PEC step)
// and border_bounds_ok, only the particles of the border cells
// (border_walk), over the first kBorderBlocks workgroups. The extremes are
// the same values the full scan finds. With PEC's
// predictor the full scan also finds the largest |vx|, |vy| (k_grid keeps it
// as ctrl->vbound; +inf if one is not finite).
constexpr int kBoundsBlocks = 1024;
// workgroups that walk the border cells (one per CU: 8.4 us on the 1M dam
// break, against 9.6 with all 1,024 and 10.1 with 64)
constexpr int kBorderBlocks = 256;
__global__ void k_bounds(Ctrl *ctrl, State a, int n, int pred, double *part, HmConst hc, const int *cellStart) {
    if (!ctrl->active) return;
    __shared__ double lds[kBoundsVQ * 16];
    double xmn = INFINITY, xmx = -INFINITY, ymn = INFINITY, ymx = -INFINITY, hmn = INFINITY, hmx = -INFINITY;
    double bad = 0.0, vmx = -INFINITY;
    const double dt = ctrl->dt;
    auto take = [&](int i) {
        double xi, yi;
        upd_pos(ctrl, a, i, pred, dt, xi, yi);
        const double hi = hc.on ? hc.h : a.h[i];
        if (!(xi - xi == 0.0) || !(yi - yi == 0.0)) bad += 1.0;  // NaN or Inf
        xmn = xi < xmn ? xi : xmn;
        xmx = xi > xmx ? xi : xmx;
        ymn = yi < ymn ? yi : ymn;
        ymx = yi > ymx ? yi : ymx;
        hmn = hi < hmn ? hi : hmn;
        hmx = hi > hmx ? hi : hmx;
    };
    const int stride = gridDim.x * blockDim.x;
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    if (border_bounds_ok(ctrl, cellStart, pred, hc)) {
        if (blockIdx.x >= kBorderBlocks) {  // the rest store empty partials
            if (threadIdx.x == 0) {
                const double e[kBoundsVQ] = {INFINITY, -INFINITY, INFINITY, -INFINITY, hc.h, hc.h, 0.0, -INFINITY};
#pragma unroll
                for (int q = 0; q < kBoundsVQ; ++q) part[q * gridDim.x + blockIdx.x] = e[q];
            }
            return;
        }
        const int fstride = (kBorderBlocks < (int)gridDim.x ? kBorderBlocks : (int)gridDim.x) * blockDim.x;
        if (gt == 0) ctrl->border_updates += 1;
        hmn = hmx = hc.h;  // every particle's h
        border_walk(cellStart, 0, ctrl->grid, gt, fstride, take);
    } else {
        for (int i0 = gt; i0 < n; i0 += 4 * stride) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int i = i0 + u * stride;
                if (i < n) {
                    take(i);
                    if (pred == 1) {
                        const double ux = fabs(a.vx[i]), uy = fabs(a.vy[i]);
                        const double w = (isfinite(ux) && isfinite(uy)) ? (ux > uy ? ux : uy) : INFINITY;
                        vmx = w > vmx ? w : vmx;
                    }
                }
            }
        }
    }
    const int kind[kBoundsVQ] = {0, 1, 0, 1, 0, 1, 2, 1};
    double v[kBoundsVQ] = {xmn, xmx, ymn, ymx, hmn, hmx, bad, vmx};
    block_partials<kBoundsVQ>(v, kind, lds, part, gridDim.x, blockIdx.x);
}

// _get_cell_size :103-110 and _get_number_of_cells :112-127 from the folded
// bounds; writes the grid (or the halt flags) to the control block. Returns
// false when the step must halt.
__device__ bool derive_grid(Ctrl *ctrl, double scale, long long cap, const double *v) {
    const double xmn = v[0], xmx = v[1], ymn = v[2], ymx = v[3], hmn = v[4], hmx = v[5], bad = v[6];
    Grid g;
    g.xmin = xmn;
    g.xmax = xmx;
    g.ymin = ymn;
    g.ymax = ymx;
    double cell_size = hmn * scale;
    if (cell_size < 1e-6) cell_size = 1.0;
    g.cs = cell_size;
    g.uniform_h = hmn == hmx;
    double cell_size1 = 1. / cell_size;
    double fx = ceil(cell_size1 * (xmx - xmn));
    double fy = ceil(cell_size1 * (ymx - ymn));
    if (bad != 0.0 || !(fx < 2.0e9) || !(fy < 2.0e9)) {
        ctrl->flags |= SPH_FLAG_NONFINITE;
        ctrl->active = 0;
        ctrl->nn_valid = 0;
        return false;
    }
    long long ncx = (long long)fx, ncy = (long long)fy;
    if (ncx < 1) ncx = 1;
    if (ncy < 1) ncy = 1;
    g.ncx = ncx;
    g.ncy = ncy;
    g.ncells = ncx * ncy;
    ctrl->grid = g;
    if (g.ncells > cap) {
        ctrl->flags |= SPH_FLAG_CAPACITY;
        ctrl->active = 0;
        ctrl->nn_valid = 0;
        return false;
    }
    return true;
}

// Fold the R bounds partials and derive the grid; a velocity bound that the
// scan found (>= 0) becomes ctrl->vbound. One workgroup of kGridThreads
// (256: 6.6 us on the 1M dam break against 7.9 with 1,024, four partials per
// thread); every load, the active flag's included, is issued before the
// fold.
constexpr int kGridThreads = 256;
__global__ void k_grid(Ctrl *ctrl, const double *part, int R, double scale, long long cap) {
    const int active = ctrl->active;
    __shared__ double lds[kBoundsVQ * 16];
    const int kind[kBoundsVQ] = {0, 1, 0, 1, 0, 1, 2, 1};
    double v[kBoundsVQ] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY, 0.0, -INFINITY};
    for (int k = threadIdx.x; k < R; k += blockDim.x) {
        double u[kBoundsVQ];
#pragma unroll
        for (int q = 0; q < kBoundsVQ; ++q) u[q] = part[q * R + k];
#pragma unroll
        for (int q = 0; q < kBoundsVQ; ++q)
            v[q] = kind[q] == 0 ? (u[q] < v[q] ? u[q] : v[q]) : kind[q] == 1 ? (u[q] > v[q] ? u[q] : v[q]) : v[q] + u[q];
    }
    block_fold<kBoundsVQ>(v, kind, lds);
    if (threadIdx.x == 0 && active) {
        if (v[7] >= 0.0) ctrl->vbound = v[7];
        derive_grid(ctrl, scale, cap, v);
    }
}

// _bin :129-141: cid = floor((x - xmin) / cs) per axis (:146-176), flat key
// cx + ncx * cy (:200-208). Counts per bin; the rank inside a bin comes from
// one atomic per run of equal bins within a wavefront (the state is already
// in the previous step's cell order, so ~7 consecutive lanes share a bin).
// k_cellsort fixes the order inside each bin afterwards.
__global__ void k_keys(Ctrl *ctrl, State a, int n, int pred, int *counts, int *bin, int *rank, int *binflag) {
    if (!ctrl->active) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const Grid g = ctrl->grid;
    int b = -1;
    if (t < n) {
        double x, y;
        upd_pos(ctrl, a, t, pred, ctrl->dt, x, y);
        const int2 c = cell_of(x, y, g);
        const long long key = (long long)c.x + g.ncx * (long long)c.y;
        if (key >= 0 && key < g.ncells) {
            b = (int)key;
        } else {
            b = (int)g.ncells;
            atomicOr(&ctrl->flags, (unsigned)SPH_FLAG_OOB_KEY);
        }
    }
    // Runs of equal bins in lane order, one atomic per run. A run also breaks
    // where the reference index drops (consecutive slots from two old cells),
    // so every run is in reference-index order; a bin that receives two runs
    // is listed for k_cellsort below.
    const int id = t < n ? a.id[t] : 0x7fffffff;
    const int prev = __shfl_up(b, 1, 64);
    const int prev_id = __shfl_up(id, 1, 64);
    const unsigned long long heads = __ballot(lane == 0 || b != prev || id < prev_id);
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const int leader = 63 - __clzll(heads & upto);
    const unsigned long long after = heads & ~upto;
    const int next = after ? __ffsll((long long)after) - 1 : 64;
    int base = 0;
    if (lane == leader && b >= 0) {
        base = atomicAdd(&counts[b], next - leader);
        // A bin filled by more than one run may be out of reference-index
        // order: flag it for k_cellsort, which clears the flag again. Every
        // run but the first to arrive sees base != 0, so such a bin is always
        // flagged; the store is idempotent, so no atomic is needed.
        if (base != 0) binflag[b] = 1;
    }
    base = __shfl(base, leader, 64);
    if (t < n) {
        bin[t] = b;
        rank[t] = base + (lane - leader);
    }
}

// Permute the whole state into the new cell order and write the packed
// pair-pass records: p / rho^2 per particle (the momentum equation's slf and
// othr terms, src/Equations/Momentum.py:28,36) and w = m for fluid, -label
// otherwise.
// lean (a PEC step whose fused pair pass and corrector follow at once): for
// the fluid, x0 ... rho0 are not written (the corrector reads them from this
// kernel's source buffer), nor the half-step x, y, vx, vy, rho (the records
// carry them to the pair pass; the corrector forms the half-step velocity
// again and overwrites the rest), nor the rates of a
// particle inside the grid (its own cell is walked, so it has at least one
// neighbour, q = 0, and the pair pass writes all three). A particle outside
// the grid may find none and keep its previous rates (SolverTools.py:166-169).
// The raw cell of each slot (for the pair pass's 3x3 walk) is formed again
// from the position the keys were computed from.
// pred as for upd_pos; xs: carry the XSPH velocity with the state; dyn:
// dynamic h for the fluid, formed here, after the keys were binned with the
// previous h (Solver._compute: nn.update :238, then computeH :242-246) and
// from the density the loop sees (the predicted one with PEC).
__global__ void k_reorder(Ctrl *ctrl, State a, State b, const int *perm, int2 *cxy_s, Phys ph, int n, int fmt,
                          void *recA, void *recB, int pred, int strict, int lean, int xs, int dyn, HmConst hc) {
    // An inactive step (halted, or past the duration) still flips the
    // buffers on the host, so copy the state unpermuted.
    const int act = ctrl->active;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if (t == 0 && act) ctrl->nn_valid = 1;
    int s = act ? perm[t] : t;
    double x = a.x[s], y = a.y[s], vx = a.vx[s], vy = a.vy[s], rho = a.rho[s];
    double m = hc.on ? hc.m : a.m[s], h = hc.on ? hc.h : a.h[s];
    const double ax = a.ax[s], ay = a.ay[s], drho = a.drho[s];
    int8_t lab = a.label[s];
    const bool skip = lean && act && pred && lab == kFluid;
    if (act && pred && lab == kFluid) {
        // PEC.predict (src/Integrators/PEC.py:32-59) fused into the permutation
        const double dt = ctrl->dt, damping = ctrl->damping;
        if (!skip) {
            b.x0[t] = x;
            b.y0[t] = y;
            b.vx0[t] = vx;
            b.vy0[t] = vy;
            b.rho0[t] = rho;
        }
        x = pec_half_pos(x, pred == 2 ? a.xsphx[s] : vx, dt);
        y = pec_half_pos(y, pred == 2 ? a.xsphy[s] : vy, dt);
        vx = pec_half_vel(vx, ax, dt, damping);
        vy = pec_half_vel(vy, ay, dt, damping);
        rho = pec_half_rho(rho, drho, dt, strict);
    } else {
        if (act && lab == kCoupled && ctrl->nm) {
            // NewmarkBeta.predict (src/Integrators/NewmarkBeta.py:41-52), src/Solver.py:381-382
            const double dt = ctrl->dt, nb = ctrl->nm_beta, ng = ctrl->nm_gamma;
            x = nm_pred_pos(x, vx, ax, nb, dt);
            y = nm_pred_pos(y, vy, ay, nb, dt);
            vx = nm_pred_vel(vx, ax, ng, dt);
            vy = nm_pred_vel(vy, ay, ng, dt);
        }
        b.x0[t] = a.x0[s];
        b.y0[t] = a.y0[s];
        b.vx0[t] = a.vx0[s];
        b.vy0[t] = a.vy0[s];
        b.rho0[t] = a.rho0[s];
    }
    // SolverTools.computeH :106-118, sigma 1.3, d = 2 (the reference's fastmath
    // (m / rho) ** 0.5 is the square root)
    if (dyn && act && lab == kFluid) h = rho > 1e-12 ? 1.3 * sqrt(m / rho) : 0.0;
    const Grid &g = ctrl->grid;
    const int2 cc = act ? cell_of(x, y, g) : make_int2(0, 0);
    if (!skip) {
        b.x[t] = x;
        b.y[t] = y;
        b.vx[t] = vx;
        b.vy[t] = vy;
        b.rho[t] = rho;
    }
    if (!hc.on) {  // else both buffers hold the one m and h in every slot
        b.m[t] = m;
        b.h[t] = h;
    }
    b.label[t] = lab;
    b.id[t] = a.id[s];
    if (xs || lab != kFluid) {  // without XSPH the fluid's is 0 after every loop (WCSPH.py:187-189)
        b.xsphx[t] = a.xsphx[s];
        b.xsphy[t] = a.xsphy[s];
    }
    if (!skip || !(cc.x >= 0 && cc.x < g.ncx && cc.y >= 0 && cc.y < g.ncy && h > 0.0 && ph.qmax >= 0.0)) {
        b.ax[t] = ax;
        b.ay[t] = ay;
        b.drho[t] = drho;
    }
    if (!act) return;
    cxy_s[t] = cc;
    const double pr = write_pair_records(t, x, y, vx, vy, rho, m, h, lab, ph, fmt, recA, recB);
    if (lab == kFluid && !isfinite(pr)) atomicOr(&ctrl->flags, (unsigned)SPH_FLAG_ZERO_DENSITY);
}

// The records of the current cell-ordered state in format fmt (after a
// neighbour update, when a pass needs another format than the reorder wrote).
__global__ void k_records(const Ctrl *ctrl, State st, Phys ph, int n, int fmt, void *recA, void *recB) {
    if (!ctrl->active) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    write_pair_records(t, st.x[t], st.y[t], st.vx[t], st.vy[t], st.rho[t], st.m[t], st.h[t], st.label[t], ph, fmt,
                       recA, recB);
}

// min and max h over all particles, min and max m over the fluid and over
// all particles (upload, restore and scatter only): per-workgroup partials
// part[q * gridDim.x + b] (grid-stride), folded by k_hm_final into out[0..5].
__global__ void k_hm_range(State st, int n, double *part) {
    __shared__ double lds[6 * 16];
    double v[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double h = st.h[i], m = st.m[i];
        v[0] = h < v[0] ? h : v[0];
        v[1] = h > v[1] ? h : v[1];
        if (st.label[i] == kFluid) {
            v[2] = m < v[2] ? m : v[2];
            v[3] = m > v[3] ? m : v[3];
        }
        v[4] = m < v[4] ? m : v[4];
        v[5] = m > v[5] ? m : v[5];
    }
    const int kind[6] = {0, 1, 0, 1, 0, 1};
    block_partials<6>(v, kind, lds, part, gridDim.x, blockIdx.x);
}

// The R partials of k_hm_range into out[0..5]. One workgroup.
__global__ void k_hm_final(const double *part, int R, double *out) {
    __shared__ double lds[6 * 16];
    const int kind[6] = {0, 1, 0, 1, 0, 1};
    double v[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
    for (int k = threadIdx.x; k < R; k += blockDim.x)
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const double u = part[q * R + k];
            v[q] = kind[q] == 0 ? (u < v[q] ? u : v[q]) : (u > v[q] ? u : v[q]);
        }
    block_fold<6>(v, kind, lds);
    if (threadIdx.x == 0)
        for (int q = 0; q < 6; ++q) out[q] = v[q];
}

// The one h and m into every slot of a buffer.
__global__ void k_fill_hm(State st, int n, double h, double m) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    st.h[i] = h;
    st.m[i] = m;
}

// ---------------------------------------------------------------------------
// upload / download (reference index order <-> cell order)
// ---------------------------------------------------------------------------
__global__ void k_upload(State d, sph_state_t in, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    d.x[i] = d.x0[i] = in.x[i];
    d.y[i] = d.y0[i] = in.y[i];
    d.vx[i] = d.vx0[i] = in.vx[i];
    d.vy[i] = d.vy0[i] = in.vy[i];
    d.rho[i] = d.rho0[i] = in.rho[i];
    d.ax[i] = in.ax ? in.ax[i] : 0.0;
    d.ay[i] = in.ay ? in.ay[i] : 0.0;
    d.drho[i] = in.drho ? in.drho[i] : 0.0;
    d.m[i] = in.m[i];
    d.h[i] = in.h[i];
    d.label[i] = in.label[i];
    d.xsphx[i] = in.xsphx ? in.xsphx[i] : 0.0;
    d.xsphy[i] = in.xsphy ? in.xsphy[i] : 0.0;
    d.id[i] = i;
}

// Clock and flags of a restored state (device side, so a restore needs no
// host round trip): time t, advancing, running, grid to be rebuilt.
__global__ void k_restore_ctrl(Ctrl *ctrl, double t) {
    ctrl->t = t;
    ctrl->flags = 0;
    ctrl->active = 1;
    ctrl->nn_valid = 0;
    ctrl->advance_t = 1;
}

// fz: XSPH is off and a reorder has run since the upload, so the fluid's
// XSPH velocity is not carried and reads as the 0 that WCSPH.compute_velocity
// gives every fluid particle with a neighbour (src/Methods/WCSPH.py:187-189;
// an in-grid particle always has itself). Walls keep theirs.
__device__ inline double xs_read(const State &s, const double *v, int t, int fz) {
    return fz && s.label[t] == kFluid ? 0.0 : v[t];
}

__global__ void k_download(State s, sph_state_t out, Phys ph, int n, int fz) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i = s.id[t];
    if (out.x) out.x[i] = s.x[t];
    if (out.y) out.y[i] = s.y[t];
    if (out.vx) out.vx[i] = s.vx[t];
    if (out.vy) out.vy[i] = s.vy[t];
    if (out.rho) out.rho[i] = s.rho[t];
    if (out.ax) out.ax[i] = s.ax[t];
    if (out.ay) out.ay[i] = s.ay[t];
    if (out.drho) out.drho[i] = s.drho[t];
    if (out.m) out.m[i] = s.m[t];
    if (out.h) out.h[i] = s.h[t];
    if (out.label) out.label[i] = s.label[t];
    if (out.xsphx) out.xsphx[i] = xs_read(s, s.xsphx, t, fz);
    if (out.xsphy) out.xsphy[i] = xs_read(s, s.xsphy, t, fz);
    if (out.p) out.p[i] = tait_p(s.rho[t], s.label[t], ph.gamma, ph.B, ph.rho0, ph.Pb);
}

// ---------------------------------------------------------------------------
// parity dump: NNLinkedList.near for every particle, reference order
// ---------------------------------------------------------------------------
__global__ void k_inverse(const int *id, int *slot_of, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) slot_of[id[t]] = t;
}

// Raw cell of a point (NNLinkedList._find_cell_id_raw :146-176), clamped
// to +-1e9 so the cast is defined (such a cell has no valid neighbour cell).
__device__ inline void raw_cell(const Grid &g, double x, double y, long long &cx0, long long &cy0) {
    double fx = floor((x - g.xmin) / g.cs);
    double fy = floor((y - g.ymin) / g.cs);
    // clamped so the integer conversion is defined; NaN goes below the grid
    fx = !(fx >= -1.0e9) ? -1.0e9 : (fx > 1.0e9 ? 1.0e9 : fx);
    fy = !(fy >= -1.0e9) ? -1.0e9 : (fy > 1.0e9 ? 1.0e9 : fy);
    cx0 = (long long)fx;
    cy0 = (long long)fy;
}

// nearPos (src/Tools/NNLinkedList.py:41-80) of the point (x, y) with length
// h against the last update: the 3 x 3 cells around its raw cell, x shift
// outer and y shift inner, each cell from its highest reference index down
// (the slots of a cell ascend in index), every particle with
// q = r / (0.5 (h + h_j)) <= qmax. f(u, q, r) per neighbour slot u.
template <typename F>
__device__ inline void near_walk(const Grid &g, const State &s, const int *cellStart, double x, double y, double h,
                                 double qmax, F &&f) {
    long long cx0, cy0;
    raw_cell(g, x, y, cx0, cy0);
    for (int a = -1; a <= 1; ++a) {
        for (int b = -1; b <= 1; ++b) {
            long long cx = cx0 + a, cy = cy0 + b;
            if (!(g.ncx > cx && cx > -1) || !(g.ncy > cy && cy > -1)) continue;
            long long cell = cx + g.ncx * cy;
            for (int u = cellStart[cell + 1] - 1; u >= cellStart[cell]; --u) {
                double dx = x - s.x[u], dy = y - s.y[u];
                double xij2 = dx * dx + dy * dy;
                double r = sqrt(xij2);
                double h_ij = 0.5 * (h + s.h[u]);
                double q = r / h_ij;
                if (q <= qmax) f(u, q, r);
            }
        }
    }
}

// near(i) = nearPos of particle i (:82-84). mode 0 counts, mode 1 writes
// indices (and q, r) from off[i].
__global__ void k_nn_dump(const Ctrl *ctrl, State s, const int *cellStart, const int *slot_of, int n, int mode,
                          long long *keys, long long *off, int *idx, double *qo, double *ro, double qmax) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Grid g = ctrl->grid;
    int t = slot_of[i];
    double x = s.x[t], y = s.y[t], h = s.h[t];
    if (mode == 0) {
        long long cx0, cy0;
        raw_cell(g, x, y, cx0, cy0);
        keys[i] = cx0 + g.ncx * cy0;
    }
    long long k = 0, base = mode ? off[i] : 0;
    near_walk(g, s, cellStart, x, y, h, qmax, [&](int u, double q, double r) {
        if (mode) {
            idx[base + k] = s.id[u];
            if (qo) qo[base + k] = q;
            if (ro) ro[base + k] = r;
        }
        ++k;
    });
    if (mode == 0) off[i] = k;
}

// nearPos for npts arbitrary points (px, py, ph), as k_nn_dump.
__global__ void k_near_pos(const Ctrl *ctrl, State s, const int *cellStart, int npts, const double *px,
                           const double *py, const double *ph, int mode, long long *off, int *idx, double *qo,
                           double *ro, double qmax) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    const Grid g = ctrl->grid;
    long long k = 0, base = mode ? off[i] : 0;
    near_walk(g, s, cellStart, px[i], py[i], ph[i], qmax, [&](int u, double q, double r) {
        if (mode) {
            idx[base + k] = s.id[u];
            if (qo) qo[base + k] = q;
            if (ro) ro[base + k] = r;
        }
        ++k;
    });
    if (mode == 0) off[i] = k;
}

// ---------------------------------------------------------------------------
// element-wise plugin kernels (Kernel.evaluate/gradient, TaitEOS)
// ---------------------------------------------------------------------------
__device__ inline double dev_kernel_w(int kernel, double r, double h) {
    const double PI = 3.141592653589793;
    double q = r / h;
    if (kernel == SPH_KERNEL_GAUSSIAN) {  // src/Kernels/Gaussian.py:16-25
        if (q <= 3) return (1.0 / PI) / (h * h) * exp(-q * q);
        return 0.0;
    }
    if (kernel == SPH_KERNEL_WENDLAND) {  // src/Kernels/Wendland.py:10-30
        double alpha = 9.0 / (4.0 * PI) / (h * h);
        if (q >= 2.0) return 0.0;
        double inner = 1.0 - 0.5 * q;
        return alpha * (ipow(inner, 6.0) * (35.0 / 12.0 * q * q + 3.0 * q + 1.0));
    }
    double alpha = 10.0 / (7.0 * PI) / (h * h);  // src/Kernels/CubicSpline.py:10-36
    if (q > 2.0) return 0.0;
    if (q > 1.0) return alpha * 0.25 * ipow(2.0 - q, 3.0);
    return alpha * (1.0 - 1.5 * ipow(q, 2.0) * (1.0 - 0.5 * q));
}

__device__ inline double dev_kernel_g(int kernel, double x, double r, double h) {
    const double PI = 3.141592653589793;
    if (kernel == SPH_KERNEL_GAUSSIAN) {  // src/Kernels/Gaussian.py:39-60
        double tmp = r * h;
        if (tmp > 1e-12) {
            double q = r / h;
            if (q <= 3) {
                double alpha_c = (1.0 / PI) / (h * h);
                double dwdq = -2 * q * alpha_c * exp(-q * q);
                return dwdq / tmp * x;
            }
        }
        return 0.0;
    }
    double q = r / h;
    if (kernel == SPH_KERNEL_WENDLAND) {  // src/Kernels/Wendland.py:32-64
        double alpha = 9.0 / (4.0 * PI) / (h * h);
        if ((q >= 2.0) || (r < 1e-10)) return 0.0;
        double inner = 1.0 - 0.5 * q;
        double grad = ipow(inner, 5.0) * (-14.0 / 3.0) * q * (1 + 2.5 * q);
        return alpha * grad / (h * r) * x;
    }
    double alpha = 10.0 / (7.0 * PI) / (h * h);  // src/Kernels/CubicSpline.py:38-70
    if ((q > 2.0) || (r < 1e-10)) return 0.0;
    double grad = (q > 1.0) ? -0.75 * ipow(2.0 - q, 2.0) : -3.0 * q * (1.0 - 0.75 * q);
    double norm = alpha * grad / (h * r);
    return norm * x;
}

__global__ void k_kernel_eval(int kernel, int n, const double *r, const double *h, double *out) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) out[j] = dev_kernel_w(kernel, r[j], h[j]);
}

__global__ void k_kernel_grad(int kernel, int n, const double *x, const double *r, const double *h, double *out) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) out[j] = dev_kernel_g(kernel, x[j], r[j], h[j]);
}

// src/Kernels/Gaussian.py:27-37
__global__ void k_kernel_deriv(int n, const double *r, const double *h, double *out) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double PI = 3.141592653589793;
    double q = r[j] / h[j];
    double v = 0.0;
    if (q <= 3) v = -2 * q * ((1.0 / PI) / (h[j] * h[j])) * exp(-q * q);
    out[j] = v;
}

__global__ void k_tait(int n, double gamma, double B, double rho0, double Pb, const double *rho, const int8_t *label,
                       double *p) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) p[j] = tait_p(rho[j], label[j], gamma, B, rho0, Pb);
}

// pressure_SPH of the coupled example (examples/IceBreak.py:253-285), one lane
// per point: the Shepard normaliser sum of W(r, h) m / rho over the fluid
// neighbours with rho >= 1e-3 (src/Equations/Shepard.py:5-12), then the
// summation density sum of m (W / normaliser) over the fluid neighbours
// (src/Equations/SummationDensity.py:5-13), both in nearPos order with the
// probe's own h in W, and p = ((rho / rho0)^gamma - 1) B (:282-284).
__global__ void k_probe(const Ctrl *ctrl, State s, const int *cellStart, int npts, const double *px, const double *py,
                        const double *ph, int kernel, double qmax, double rho0, double gamma, double B, double *rho_out,
                        double *p_out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    const Grid g = ctrl->grid;
    const double x = px[i], y = py[i], hp = ph[i];
    double norm = 0.0;
    near_walk(g, s, cellStart, x, y, hp, qmax, [&](int u, double, double r) {
        const double rj = s.rho[u];
        if (rj < 1e-3 || s.label[u] != kFluid) return;
        norm += dev_kernel_w(kernel, r, hp) * s.m[u] / rj;
    });
    double rho = 0.0;
    near_walk(g, s, cellStart, x, y, hp, qmax, [&](int u, double, double r) {
        if (s.label[u] != kFluid) return;
        rho += s.m[u] * (dev_kernel_w(kernel, r, hp) / norm);
    });
    rho_out[i] = rho;
    p_out[i] = (ipow(rho / rho0, gamma) - 1.0) * B;
}

// Records by reference index: ids[j] -> slot slot_of[ids[j]] of the current
// buffer (the coupling hook's particleArray[cInd] reads and writes).
__global__ void k_gather_ids(State s, const int *slot_of, const int *ids, int k, sph_state_t out, Phys ph, int fz) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= k) return;
    const int t = slot_of[ids[j]];
    if (out.x) out.x[j] = s.x[t];
    if (out.y) out.y[j] = s.y[t];
    if (out.vx) out.vx[j] = s.vx[t];
    if (out.vy) out.vy[j] = s.vy[t];
    if (out.rho) out.rho[j] = s.rho[t];
    if (out.ax) out.ax[j] = s.ax[t];
    if (out.ay) out.ay[j] = s.ay[t];
    if (out.drho) out.drho[j] = s.drho[t];
    if (out.m) out.m[j] = s.m[t];
    if (out.h) out.h[j] = s.h[t];
    if (out.label) out.label[j] = s.label[t];
    if (out.xsphx) out.xsphx[j] = xs_read(s, s.xsphx, t, fz);
    if (out.xsphy) out.xsphy[j] = xs_read(s, s.xsphy, t, fz);
    if (out.p) out.p[j] = tait_p(s.rho[t], s.label[t], ph.gamma, ph.B, ph.rho0, ph.Pb);
}

__global__ void k_scatter_ids(State s, const int *slot_of, const int *ids, int k, sph_state_t in) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= k) return;
    const int t = slot_of[ids[j]];
    if (in.x) s.x[t] = in.x[j];
    if (in.y) s.y[t] = in.y[j];
    if (in.vx) s.vx[t] = in.vx[j];
    if (in.vy) s.vy[t] = in.vy[j];
    if (in.rho) s.rho[t] = in.rho[j];
    if (in.ax) s.ax[t] = in.ax[j];
    if (in.ay) s.ay[t] = in.ay[j];
    if (in.drho) s.drho[t] = in.drho[j];
    if (in.m) s.m[t] = in.m[j];
    if (in.h) s.h[t] = in.h[j];
    if (in.xsphx) s.xsphx[t] = in.xsphx[j];
    if (in.xsphy) s.xsphy[t] = in.xsphy[j];
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
static inline int nblk(long long n, int bs = kBlock) {
    long long b = (n + bs - 1) / bs;
    return (int)(b < 1 ? 1 : b);
}

static inline int redblk(long long n) {
    int b = nblk(n);
    return b > kRedBlocks ? kRedBlocks : b;
}

static hipStream_t S(void *s) { return (hipStream_t)s; }

static void free_ctx(sph_ctx *c) {
    if (!c) return;
    for (int k = 0; k < 2; ++k) {
        if (c->arena[k]) hipFree(c->arena[k]);
        if (c->gexec[k]) hipGraphExecDestroy(c->gexec[k]);
    }
    void *ptrs[] = {c->ctrl, c->dt_hist, c->bin, c->rank, c->perm, c->permKey, c->cxy_s,
                    c->counts, c->cellStart, c->tileSums, c->binflag, c->part_b, c->part_d, c->part_k, c->recA, c->recB,
                    c->slot_of};
    for (void *p : ptrs)
        if (p) hipFree(p);
    for (auto e : c->ev_pair) hipEventDestroy(e);
    for (auto e : c->ev_nn) hipEventDestroy(e);
    if (c->own) hipStreamDestroy(c->own);
    delete c;
}

static int alloc_cells(sph_ctx *c, long long cap) {
    if (c->counts) hipFree(c->counts);
    if (c->cellStart) hipFree(c->cellStart);
    if (c->tileSums) hipFree(c->tileSums);
    if (c->binflag) hipFree(c->binflag);
    c->counts = nullptr;
    c->cellStart = nullptr;
    c->tileSums = nullptr;
    c->binflag = nullptr;
    long long tiles = (cap + 1 + kScanTile - 1) / kScanTile;
    HIPCHK(hipMalloc(&c->counts, sizeof(int) * (cap + 1)));
    HIPCHK(hipMalloc(&c->cellStart, sizeof(int) * (cap + 2)));
    HIPCHK(hipMalloc(&c->tileSums, sizeof(int) * (tiles + 1)));
    HIPCHK(hipMemset(c->counts, 0, sizeof(int) * (cap + 1)));
    HIPCHK(hipMemset(c->cellStart, 0, sizeof(int) * (cap + 2)));
    HIPCHK(hipMalloc(&c->binflag, sizeof(int) * (cap + 1)));
    HIPCHK(hipMemset(c->binflag, 0, sizeof(int) * (cap + 1)));
    c->cap_cells = cap;
    return 0;
}

static int check_params(const sph_params_t *P) {
    if (!P) return fail(SPH_EINVAL, "params is NULL");
    if (P->n <= 0 || P->n > (1LL << 30)) return fail(SPH_EINVAL, "n=%lld out of range (1 .. 2^30)", (long long)P->n);
    if (P->kernel < 0 || P->kernel > 2) return fail(SPH_EINVAL, "unknown kernel id %d", P->kernel);
    if (P->precision != SPH_PREC_FP64 && P->precision != SPH_PREC_FP32C)
        return fail(SPH_EINVAL, "unknown precision %d", P->precision);
    if (P->integrator != SPH_INTEGRATOR_PEC && P->integrator != SPH_INTEGRATOR_EULER)
        return fail(SPH_EINVAL, "unknown integrator %d", P->integrator);
    if ((P->use_xsph && !(P->epsilon == P->epsilon)) || P->use_xsph < 0 || P->use_xsph > 1 ||
        P->use_summation_density < 0 || P->use_summation_density > 1 || P->dynamic_h < 0 || P->dynamic_h > 1)
        return fail(SPH_EINVAL, "use_xsph / use_summation_density / dynamic_h must be 0 or 1 (and epsilon finite)");
    if (!(P->scale > 0) || !(P->qmax > 0) || !(P->rho0 > 0))
        return fail(SPH_EINVAL, "scale, qmax and rho0 must be positive");
    if (P->coupled_integrator != SPH_COUPLED_NONE && P->coupled_integrator != SPH_COUPLED_NEWMARK)
        return fail(SPH_EINVAL, "unknown coupled_integrator %d", P->coupled_integrator);
    // NewmarkBeta.__init__ asserts beta * 2 >= gamma (src/Integrators/NewmarkBeta.py:33)
    if (P->coupled_integrator == SPH_COUPLED_NEWMARK &&
        !(P->newmark_beta * 2 >= P->newmark_gamma && P->newmark_beta - P->newmark_beta == 0.0 &&
          P->newmark_gamma - P->newmark_gamma == 0.0))
        return fail(SPH_EINVAL, "Newmark-beta needs finite beta, gamma with 2 beta >= gamma");
    return 0;
}

static int refresh_uniform(sph_ctx *c, hipStream_t s);

static void ctrl_from_params(Ctrl &h, const sph_params_t &P) {
    h.nm = P.coupled_integrator == SPH_COUPLED_NEWMARK;
    h.nm_beta = P.newmark_beta;
    h.nm_gamma = P.newmark_gamma;
    h.duration = P.duration;
    h.fixed_dt = P.fixed_dt;
    h.damping = P.damping;
    h.alpha = P.alpha;
    h.beta = P.beta;
    h.co = P.co;
    h.B = P.B;
    h.Pb = P.Pb;
}

// Calls that index by particle id (steps, downloads, gathers) need a state:
// before the first upload the buffers hold no valid ids.
static int unloaded(sph_ctx *c) {
    return c->loaded ? 0 : fail(SPH_ESTATE, "no state: call sph_state_upload first");
}
// Step calls other than sph_step_end are refused while a split step is open.
static int closed(sph_ctx *c) {
    if (!c->loaded) return unloaded(c);
    return c->mid ? fail(SPH_ESTATE, "a split step is open: call sph_step_end first") : 0;
}

// Whether the fused pass takes its uniform form (sph_pair_uniform.h for
// FP32C, sph_pair_uniform64.h for FP64): the cubic spline or Wendland, the
// plain loop (no XSPH, summation density or dynamic h), one h > 0 for every
// particle and one fluid mass in the state, both the support 2h and the
// wall force's r0 inside the neighbour cut-off qmax h, and at most
// kUniformMaxN particles.
bool hipsph::uniform_eligible(const sph_params_t &P, double h) {
    if (P.n > kUniformMaxN) return false;  // records addressed with 32-bit byte offsets
    if (P.use_xsph || P.use_summation_density || P.dynamic_h) return false;
    if (!(h > 0.0)) return false;
    const double cut = 0.99 * P.qmax * h;
    if (P.kernel == SPH_KERNEL_GAUSSIAN)  // support 3h, cut by the cut-off unless q_max >= 3
        return P.qmax >= 3.0 && P.r0 < 0.99 * 3.0 * h;
    return 2.0 * h < cut && P.r0 < cut;
}

UniformParams hipsph::make_uniform_params(const sph_params_t &P, double h, double m) {
    const double pi = 3.141592653589793;
    UniformParams u;
    const bool gauss = P.kernel == SPH_KERNEL_GAUSSIAN;
    const double sup = (gauss ? 3.0 : 2.0) * h;
    const double lim = sup > P.r0 ? sup : P.r0;
    u.lim2 = lim * lim * (P.precision == SPH_PREC_FP32C ? 1.0 + 1e-5 : 1.0 + 1e-12);
    u.lo = u.lim2;
    if (gauss && P.precision == SPH_PREC_FP32C) {  // the float pass lists exactly q <= 3 (r0 < 3h)
        u.lim2 = sup * sup * (1.0 + 1e-12);
        u.lo = sup * sup * (1.0 - 1e-12);
    }
    u.h = h;
    u.m = m;
    u.qmax = P.qmax;
    u.g = P.gravity;
    const double norm = gauss ? 1.0 / pi : (P.kernel == SPH_KERNEL_WENDLAND ? 9.0 / (4.0 * pi) : 10.0 / (7.0 * pi));
    u.ck = norm / (h * h * h);
    u.ac = P.alpha * (0.5 * P.co);
    // the double pass's constants with the reference's operations
    // (CubicSpline.py:55 / Gaussian.py:47,53 / Wendland.py:34, Momentum.py:43-46)
    u.alpha_k = norm / (h * h);
    u.c01 = 0.01 * h * h;
    u.cij = 0.5 * (P.co + 0.0);
    u.alpha = P.alpha;
    u.beta = P.beta;
    u.r0 = P.r0;
    u.D = P.D;
    u.p1 = P.p1;
    u.p2 = P.p2;
    return u;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *sph_last_error(void) { return g_err.c_str(); }

#ifndef HIPSPH_BUILD_ID
#define HIPSPH_BUILD_ID "unknown"
#endif
const char *sph_version(void) {
    return "hipsph 1.1 gfx950 source " HIPSPH_BUILD_ID " (engine: -ffp-contract=off; pair fp64: -ffp-contract=off; "
           "pair fp32c: -ffp-contract=fast-honor-pragmas)";
}

int sph_ctx_create(const sph_params_t *params, sph_ctx_t **out) {
    if (!out) return fail(SPH_EINVAL, "out is NULL");
    *out = nullptr;
    int rc = check_params(params);
    if (rc) return rc;
    sph_ctx *c = new (std::nothrow) sph_ctx;
    if (!c) return fail(SPH_ENOMEM, "host allocation failed");
    c->P = *params;
    c->n = (int)params->n;
    const size_t n = (size_t)c->n;
    if (hipGetDevice(&c->dev) != hipSuccess) {
        delete c;
        return fail(SPH_EHIP, "no current HIP device");
    }
    // Two cell-ordered buffer sets: 17 double arrays + int8 label + int id.
    const size_t nd = 17;
    const size_t bytes = n * (nd * sizeof(double) + sizeof(int) + 8) + 1024;
    for (int k = 0; k < 2; ++k) {
        if (hipMalloc(&c->arena[k], bytes) != hipSuccess) {
            free_ctx(c);
            return fail(SPH_ENOMEM, "device allocation of %zu bytes failed", bytes);
        }
        char *p = (char *)c->arena[k];
        double **fields[] = {&c->buf[k].x,  &c->buf[k].y,   &c->buf[k].vx,  &c->buf[k].vy,  &c->buf[k].rho,
                             &c->buf[k].ax, &c->buf[k].ay,  &c->buf[k].drho, &c->buf[k].x0, &c->buf[k].y0,
                             &c->buf[k].vx0, &c->buf[k].vy0, &c->buf[k].rho0, &c->buf[k].m, &c->buf[k].h,
                             &c->buf[k].xsphx, &c->buf[k].xsphy};
        for (double **f : fields) {
            *f = (double *)p;
            p += n * sizeof(double);
        }
        c->buf[k].id = (int *)p;
        p += n * sizeof(int);
        c->buf[k].label = (int8_t *)p;
    }
    bool ok = hipMalloc(&c->ctrl, sizeof(Ctrl)) == hipSuccess &&
              hipMalloc(&c->dt_hist, sizeof(double) * 3 * SPH_DT_HISTORY) == hipSuccess &&
              hipMalloc(&c->bin, sizeof(int) * n) == hipSuccess && hipMalloc(&c->rank, sizeof(int) * n) == hipSuccess &&
              hipMalloc(&c->perm, sizeof(int) * n) == hipSuccess &&
              hipMalloc(&c->permKey, sizeof(int) * n) == hipSuccess &&
              hipMalloc(&c->cxy_s, sizeof(int2) * n) == hipSuccess &&

              hipMalloc(&c->part_b, sizeof(double) * kBoundsVQ * kRedBlocks) == hipSuccess &&
              // dt partials: a reduction's workgroups or the fused pair pass's
              hipMalloc(&c->part_d, sizeof(double) * kDtQ * (kRedBlocks > xcd_blocks(n, 256) ? kRedBlocks
                                                                                             : xcd_blocks(n, 256))) ==
                  hipSuccess &&
              hipMalloc(&c->part_k, sizeof(double) * (kRedBlocks + 1)) == hipSuccess &&
              hipMalloc(&c->slot_of, sizeof(int) * n) == hipSuccess;
    if (ok) {
        const bool f32 = params->precision == SPH_PREC_FP32C;
        const size_t ra = f32 ? sizeof(Rec<float>::A) : sizeof(Rec<double>::A);
        // the float pass's B records are 8 bytes, the uniform pass's 16
        const size_t rb = f32 ? sizeof(float4) : sizeof(Rec<double>::B);
        ok = hipMalloc(&c->recA, ra * n) == hipSuccess && hipMalloc(&c->recB, rb * n) == hipSuccess;
    }
    if (!ok) {
        free_ctx(c);
        return fail(SPH_ENOMEM, "device workspace allocation failed");
    }
    // Cell capacity: caller's value, else room for a grid 8x the particle
    // count (the dam-break tanks use ~0.4 cells per particle).
    long long cap = params->cell_capacity > 0 ? params->cell_capacity : (long long)n * 8 + 65536;
    if (cap > (1LL << 31) - 16) cap = (1LL << 31) - 16;
    if ((rc = alloc_cells(c, cap))) {
        free_ctx(c);
        return rc;
    }
    Ctrl h;
    memset(&h, 0, sizeof h);
    ctrl_from_params(h, c->P);
    h.advance_t = 1;
    h.active = 1;
    h.disp = h.vbound = h.abound = INFINITY;
    if (hipMemcpy(c->ctrl, &h, sizeof h, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(c->dt_hist, 0, sizeof(double) * 3 * SPH_DT_HISTORY) != hipSuccess) {
        free_ctx(c);
        return fail(SPH_EHIP, "control block init failed");
    }
    if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
        free_ctx(c);
        return fail(SPH_EHIP, "stream creation failed");
    }
    *out = c;
    return 0;
}

int sph_ctx_destroy(sph_ctx_t *ctx) {
    if (!ctx) return 0;
    hipDeviceSynchronize();
    free_ctx(ctx);
    return 0;
}

int sph_ctx_set_params(sph_ctx_t *ctx, const sph_params_t *params) {
    if (!ctx || !params) return fail(SPH_EINVAL, "NULL argument");
    if (params->n != ctx->P.n || params->kernel != ctx->P.kernel || params->precision != ctx->P.precision ||
        params->use_xsph != ctx->P.use_xsph)
        return fail(SPH_EINVAL, "n, kernel, precision and use_xsph are fixed at creation");
    int rc = check_params(params);
    if (rc) return rc;
    // The control block is rewritten whole: steps still in flight on any
    // stream (the caller's, or the context's own graph stream, which the
    // null stream does not wait for) must finish first.
    HIPCHK(hipDeviceSynchronize());
    ctx->P = *params;
    ctx->dt_fresh = false;
    Ctrl h;
    HIPCHK(hipMemcpy(&h, ctx->ctrl, sizeof h, hipMemcpyDeviceToHost));
    ctrl_from_params(h, ctx->P);
    HIPCHK(hipMemcpy(ctx->ctrl, &h, sizeof h, hipMemcpyHostToDevice));
    for (int k = 0; k < 2; ++k)
        if (ctx->gexec[k]) {
            hipGraphExecDestroy(ctx->gexec[k]);
            ctx->gexec[k] = nullptr;
        }
    ++ctx->gen;  // a tape's cached segment graphs hold the old constants
    return 0;
}

int sph_state_upload(sph_ctx_t *ctx, const sph_state_t *st, int32_t reset_clock, void *stream) {
    if (!ctx || !st || !st->x || !st->y || !st->vx || !st->vy || !st->rho || !st->m || !st->h || !st->label)
        return fail(SPH_EINVAL, "upload needs x, y, vx, vy, rho, m, h and label");
    hipStream_t s = S(stream);
    ctx->cur = 0;
    ctx->dt_fresh = false;
    ctx->mid = false;
    ctx->xs_fluid_zero = false;
    k_upload<<<nblk(ctx->n), kBlock, 0, s>>>(ctx->buf[0], *st, ctx->n);
    HIPCHK(hipGetLastError());
    ctx->loaded = true;
    Ctrl h;
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipMemcpy(&h, ctx->ctrl, sizeof h, hipMemcpyDeviceToHost));
    if (reset_clock) {
        h.t = 0.0;
        h.step = 0;
        h.dt = h.dt_c = h.dt_f = 0.0;
    }
    h.flags = 0;
    h.active = 1;
    h.nn_valid = 0;
    h.pec_parts = 0;
    h.disp = h.vbound = h.abound = INFINITY;
    ctrl_from_params(h, ctx->P);
    HIPCHK(hipMemcpy(ctx->ctrl, &h, sizeof h, hipMemcpyHostToDevice));
    for (int k = 0; k < 2; ++k)
        if (ctx->gexec[k]) {
            hipGraphExecDestroy(ctx->gexec[k]);
            ctx->gexec[k] = nullptr;
        }
    return refresh_uniform(ctx, s);
}

int sph_state_restore(sph_ctx_t *ctx, const sph_state_t *st, double t, void *stream) {
    if (!ctx || !st || !st->x || !st->y || !st->vx || !st->vy || !st->rho || !st->m || !st->h || !st->label)
        return fail(SPH_EINVAL, "restore needs x, y, vx, vy, rho, m, h and label");
    hipStream_t s = S(stream);
    ctx->cur = 0;
    ctx->dt_fresh = false;
    ctx->mid = false;
    ctx->xs_fluid_zero = false;
    k_upload<<<nblk(ctx->n), kBlock, 0, s>>>(ctx->buf[0], *st, ctx->n);
    k_restore_ctrl<<<1, 1, 0, s>>>(ctx->ctrl, t);
    HIPCHK(hipGetLastError());
    ctx->loaded = true;
    return refresh_uniform(ctx, s);
}

int sph_state_download(sph_ctx_t *ctx, const sph_state_t *st, void *stream) {
    if (!ctx || !st) return fail(SPH_EINVAL, "NULL argument");
    if (unloaded(ctx)) return SPH_ESTATE;
    k_download<<<nblk(ctx->n), kBlock, 0, S(stream)>>>(ctx->buf[ctx->cur], *st, make_phys(ctx->P), ctx->n,
                                                        ctx->xs_fluid_zero);
    HIPCHK(hipGetLastError());
    return 0;
}

static int ev_record(sph_ctx *c, std::vector<hipEvent_t> &v, size_t &used, hipStream_t s) {
    if (used >= v.size()) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        v.push_back(e);
    }
    HIPCHK(hipEventRecord(v[used++], s));
    return 0;
}


static bool uniform_pass(const sph_ctx *c) { return c->uni_hm && uniform_eligible(c->P, c->uni_h); }

static int rec_format(const sph_ctx *c) {
    const bool uni = uniform_pass(c);
    if (c->P.precision == SPH_PREC_FP32C) return uni ? kRecUniform : kRecFp32;
    return uni ? kRecUniform64 : kRecFp64;
}

static UniformParams uniform_params(const sph_ctx *c) {
    UniformParams u = make_uniform_params(c->P, c->uni_h, c->uni_m);
    u.stage = c->pair_staging;
    return u;
}

static void drop_graphs(sph_ctx *c) {
    for (int k = 0; k < 2; ++k)
        if (c->gexec[k]) {
            hipGraphExecDestroy(c->gexec[k]);
            c->gexec[k] = nullptr;
        }
    ++c->gen;
}

// The h range of the state and the mass range of its fluid: one h and one
// mass make the fused float pass uniform. Synchronises s; a change of form
// drops the captured steps.
static int refresh_uniform(sph_ctx *c, hipStream_t s) {
    const int before = rec_format(c);
    const bool static_before = c->static_hm;
    const double h_before = c->uni_h, m_before = c->uni_m;
    // partials in the bounds buffer (free between updates), result in part_k
    const int R = redblk(c->n) < kBoundsBlocks ? redblk(c->n) : kBoundsBlocks;
    k_hm_range<<<R, kBlock, 0, s>>>(c->buf[c->cur], c->n, c->part_b);
    k_hm_final<<<1, 1024, 0, s>>>(c->part_b, R, c->part_k);
    HIPCHK(hipGetLastError());
    double v[6];
    HIPCHK(hipMemcpyAsync(v, c->part_k, sizeof v, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    c->uni_hm = v[0] == v[1] && v[2] == v[3];
    c->uni_h = v[0];
    c->uni_m = v[2];
    c->static_hm = !c->P.dynamic_h && v[0] == v[1] && v[4] == v[5];
    if (c->static_hm) {  // both buffers, every slot: the reorder will not write them
        for (int k = 0; k < 2; ++k) k_fill_hm<<<nblk(c->n), kBlock, 0, s>>>(c->buf[k], c->n, v[0], v[4]);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
    }
    if (rec_format(c) != before || c->static_hm != static_before ||
        (c->static_hm && (c->uni_h != h_before || v[4] != m_before)))
        drop_graphs(c);
    return 0;
}

static HmConst hm_const(const sph_ctx *c) {
    HmConst hc{};
    hc.on = c->static_hm && !c->P.dynamic_h;
    hc.h = c->uni_h;
    hc.m = c->uni_m;
    return hc;
}

// pred != 0: the PEC predictor runs inside the update (bounds and keys see
// the predicted positions, the reorder writes the predicted state).
static int nn_update(sph_ctx *c, hipStream_t s, int pred = 0, int lean = 0) {
    if (pred && c->P.use_xsph) pred = 2;  // PEC with XSPH moves x by the XSPH velocity
    const int n = c->n;
    State a = c->buf[c->cur];
    State b = c->buf[1 - c->cur];
    if (c->timing) {
        int rc = ev_record(c, c->ev_nn, c->n_nn, s);
        if (rc) return rc;
    }
    const int R = redblk(n) < kBoundsBlocks ? redblk(n) : kBoundsBlocks;
    // the border-cell bounds only in the step's update with PEC's predictor
    const int *border = c->border_bounds && pred == 1 ? c->cellStart : nullptr;
    k_bounds<<<R, kBlock, 0, s>>>(c->ctrl, a, n, pred, c->part_b, hm_const(c), border);
    k_grid<<<1, kGridThreads, 0, s>>>(c->ctrl, c->part_b, R, c->P.scale, c->cap_cells);
    k_keys<<<nblk(n), kBlock, 0, s>>>(c->ctrl, a, n, pred, c->counts, c->bin, c->rank, c->binflag);
    long long tiles_cap = (c->cap_cells + 1 + kScanTile - 1) / kScanTile;
    int tiles = (int)(tiles_cap < 1024 ? tiles_cap : 1024);
    k_scan_reduce<<<tiles, kBlock, 0, s>>>(c->ctrl, c->counts, c->tileSums);
    k_scan_apply<<<tiles, kBlock, 0, s>>>(c->ctrl, c->counts, c->tileSums, c->cellStart);
    k_scatter<<<nblk(n), kBlock, 0, s>>>(c->ctrl, c->cellStart, c->bin, c->rank, a.id, c->perm, c->permKey, n);
    // bin/rank are free once scattered: they stage the rare bins of more than 64
    k_cellsort<<<redblk(c->cap_cells + 1), kBlock, 0, s>>>(c->ctrl, c->cellStart, c->perm, c->permKey,
                                                           c->binflag, c->bin, c->rank);
    c->rec_fmt = rec_format(c);
    k_reorder<<<nblk(n), kBlock, 0, s>>>(c->ctrl, a, b, c->perm, c->cxy_s, make_phys(c->P), n, c->rec_fmt,
                                         c->recA, c->recB, pred, c->P.strict, lean, c->P.use_xsph, c->P.dynamic_h,
                                         hm_const(c));
    HIPCHK(hipGetLastError());
    if (!c->P.use_xsph) c->xs_fluid_zero = true;
    if (c->timing) {
        int rc = ev_record(c, c->ev_nn, c->n_nn, s);
        if (rc) return rc;
    }
    c->cur = 1 - c->cur;
    return 0;
}

// _loop (src/Tools/SolverTools.py:120-175): the summation density first when
// it is on (then the pass writes no continuity), the pair pass with the XSPH
// velocity when XSPH is on. The split passes (mode 1, 2) are the plain
// equations only; in the uniform form they take the uniform pass of their
// equations.
static int compute(sph_ctx *c, hipStream_t s, int *count = nullptr, int mode = 0) {
    const bool sd = c->P.use_summation_density, xs = c->P.use_xsph;
    if (mode != 0 && (sd || xs))
        return fail(SPH_EUNSUPPORTED, "the split passes have no XSPH or summation density form; use sph_compute");
    State st = c->buf[c->cur];
    PairIn in{c->recA, c->recB, st.h, c->cxy_s, c->cellStart};
    if (xs) {
        in.M = st.m;
        in.VX = st.vx;
        in.VY = st.vy;
    }
    PairOut out{st.ax, st.ay, st.drho, count, sd ? 2 : mode};
    if (xs) {
        out.xsx = st.xsphx;
        out.xsy = st.xsphy;
    }
    const bool f32 = c->P.precision == SPH_PREC_FP32C;
    const Phys ph = make_phys(c->P);
    const bool uni_fmt = c->rec_fmt == kRecUniform || c->rec_fmt == kRecUniform64;
    const bool uni = !count && uni_fmt && rec_format(c) == c->rec_fmt;
    if (!uni && uni_fmt) {  // a pass that reads the general records
        const int fmt = f32 ? kRecFp32 : kRecFp64;
        k_records<<<nblk(c->n), kBlock, 0, s>>>(c->ctrl, st, ph, c->n, fmt, c->recA, c->recB);
        HIPCHK(hipGetLastError());
        c->rec_fmt = fmt;
    }
    if (c->timing) {
        int rc = ev_record(c, c->ev_pair, c->n_pair, s);
        if (rc) return rc;
    }
    if (uni) {
        hipError_t e = f32 ? launch_pair_fp32u(c->ctrl, in, out, uniform_params(c), c->P.kernel, c->n, s)
                           : launch_pair_fp64u(c->ctrl, in, out, uniform_params(c), c->P.kernel, c->n, s);
        if (e != hipSuccess) return fail(SPH_EHIP, "pair launch: %s", hipGetErrorString(e));
        if (c->timing) {
            int rc = ev_record(c, c->ev_pair, c->n_pair, s);
            if (rc) return rc;
        }
        return 0;
    }
    if (sd) {
        const SdOut so{st.rho, st.drho, c->recB};
        hipError_t e = f32 ? launch_sumden_fp32c(c->ctrl, in, so, ph, c->n, c->pair_threads, s)
                           : launch_sumden_fp64(c->ctrl, in, so, ph, c->n, c->pair_threads, s);
        if (e != hipSuccess) return fail(SPH_EHIP, "summation density launch: %s", hipGetErrorString(e));
    }
    hipError_t e = f32 ? launch_pair_fp32c(c->ctrl, in, out, ph, c->n, c->pair_threads, s)
                       : launch_pair_fp64(c->ctrl, in, out, ph, c->n, c->pair_threads, s);
    if (e != hipSuccess) return fail(SPH_EHIP, "pair launch: %s", hipGetErrorString(e));
    if (c->timing) {
        int rc = ev_record(c, c->ev_pair, c->n_pair, s);
        if (rc) return rc;
    }
    return 0;
}

static int dt_step(sph_ctx *c, hipStream_t s) {
    State st = c->buf[c->cur];
    if (!c->dt_fresh) {
        k_dt_partial<<<redblk(c->n), kBlock, 0, s>>>(c->ctrl, st.label, st.h, st.ax, st.ay, c->n, c->part_d);
    }
    k_dt_final<<<1, 1024, 0, s>>>(c->ctrl, c->P.gamma_c, c->P.gamma_f, c->dt_hist, c->part_d);
    HIPCHK(hipGetLastError());
    return 0;
}

static int predict(sph_ctx *c, hipStream_t s) {
    if (c->P.integrator != SPH_INTEGRATOR_PEC) return 0;  // Euler.predict is a no-op
    k_pec_predict<<<nblk(c->n), kBlock, 0, s>>>(c->ctrl, c->buf[c->cur], c->n, c->P.strict, c->P.use_xsph);
    HIPCHK(hipGetLastError());
    return 0;
}

// lean: the step's reorder ran lean (start-of-step values from the previous
// buffer through perm).
static int correct(sph_ctx *c, hipStream_t s, bool lean = false) {
    k_correct<<<redblk(c->n), kBlock, 0, s>>>(c->ctrl, c->buf[c->cur], c->buf[1 - c->cur],
                                              lean ? c->perm : nullptr, c->n, c->P.integrator, c->P.strict,
                                              c->P.use_xsph, c->part_d, hm_const(c));
    HIPCHK(hipGetLastError());
    return 0;
}

// The uniform pair pass with PEC's corrector and the next dt's partials
// fused in, after a lean reorder (no Newmark-beta particles, whose
// corrector k_correct applies). Timed as the pair pass.
static int compute_pec(sph_ctx *c, hipStream_t s) {
    State st = c->buf[c->cur];
    PairIn in{c->recA, c->recB, st.h, c->cxy_s, c->cellStart};
    PairOut out{st.ax, st.ay, st.drho, nullptr, 0};
    const PecArgs pec{st, c->buf[1 - c->cur], c->perm, c->part_d, hm_const(c), c->P.strict};
    if (c->timing) {
        int rc = ev_record(c, c->ev_pair, c->n_pair, s);
        if (rc) return rc;
    }
    hipError_t e = c->P.precision == SPH_PREC_FP32C
                       ? launch_pair_fp32u_pec(c->ctrl, in, out, uniform_params(c), c->P.kernel, c->n, pec, s)
                       : launch_pair_fp64u_pec(c->ctrl, in, out, uniform_params(c), c->P.kernel, c->n, pec, s);
    if (e != hipSuccess) return fail(SPH_EHIP, "pair launch: %s", hipGetErrorString(e));
    if (c->timing) {
        int rc = ev_record(c, c->ev_pair, c->n_pair, s);
        if (rc) return rc;
    }
    return 0;
}

// One Solver.run loop body (src/Solver.py:366-410) with the predictor fused
// into the neighbour update and the next dt reduction into the corrector
// (for one h and one fluid mass, the corrector into the pair pass).
static int one_step(sph_ctx *c, hipStream_t s) {
    int rc;
    const bool pec = c->P.integrator == SPH_INTEGRATOR_PEC;
    // the lean reorder leaves the half-step velocity and density out of the
    // state; XSPH and the summation density read them there
    const bool lean = pec && !c->P.use_xsph && !c->P.use_summation_density;
    if ((rc = dt_step(c, s))) return rc;
    if ((rc = nn_update(c, s, pec, lean))) return rc;
    if (lean && c->P.coupled_integrator != SPH_COUPLED_NEWMARK && uniform_pass(c)) {
        if ((rc = compute_pec(c, s))) return rc;
    } else {
        if ((rc = compute(c, s))) return rc;
        if ((rc = correct(c, s, lean))) return rc;
    }
    c->dt_fresh = true;
    return 0;
}

static int record_step(sph_ctx *ctx, sph_tape *tape, int slot, hipStream_t s);

int sph_step_record(sph_ctx_t *ctx, sph_tape_t *tape, int32_t slot, void *stream) {
    if (!ctx || !tape) return fail(SPH_EINVAL, "NULL argument");
    if (tape->ctx != ctx) return fail(SPH_EINVAL, "tape belongs to another context");
    if (slot < 0 || slot >= tape->nslots) return fail(SPH_EINVAL, "slot %d out of range", slot);
    if (closed(ctx)) return SPH_ESTATE;
    hipStream_t s = S(stream);
    int rc;
    if ((rc = tape_supported(ctx))) return rc;
    if (slot == 0 && (rc = tape_refresh_labels(ctx, tape, s))) return rc;
    return record_step(ctx, tape, slot, s);
}

// One recorded step (as sph_step_record, without the argument checks): the
// ordinary step (the same kernels as one_step, so the same trajectory) with
// the tape written after the pair pass, from the half-step state the pass
// saw and its rates.
static int record_step(sph_ctx *ctx, sph_tape *tape, int slot, hipStream_t s) {
    int rc;
    const bool pec = ctx->P.integrator == SPH_INTEGRATOR_PEC;
    const bool fused = pec && uniform_pass(ctx);  // the tape covers no XSPH, summation density or coupling
    if ((rc = dt_step(ctx, s))) return rc;
    if ((rc = nn_update(ctx, s, pec, fused))) return rc;
    if (fused) {
        if ((rc = compute_pec(ctx, s))) return rc;
        if ((rc = tape_record_pair(ctx, tape, slot, s, true))) return rc;
    } else {
        if ((rc = compute(ctx, s))) return rc;
        if ((rc = tape_record_pair(ctx, tape, slot, s, false))) return rc;
        if ((rc = correct(ctx, s))) return rc;
    }
    ctx->dt_fresh = true;
    return 0;
}

// Capture `body` on stream g into *exec. The capture always ends, and the
// graph is released, whatever fails.
extern "C++" template <typename F>
static int capture_step(hipGraphExec_t *exec, hipStream_t g, F body) {
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(g, hipStreamCaptureModeThreadLocal));
    int rc = body();
    hipError_t ce = hipStreamEndCapture(g, &graph);
    if (!rc && ce != hipSuccess) rc = fail(SPH_EHIP, "end capture: %s", hipGetErrorString(ce));
    if (!rc) {
        ce = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
        if (ce != hipSuccess) {
            *exec = nullptr;
            rc = fail(SPH_EHIP, "graph instantiate: %s", hipGetErrorString(ce));
        }
    }
    if (graph) hipGraphDestroy(graph);
    return rc;
}

// Capture `body` on stream g into *exec (once) and launch it.
extern "C++" template <typename F>
static int graph_launch(hipGraphExec_t *exec, hipStream_t g, F body) {
    if (!*exec) {
        int rc = capture_step(exec, g, body);
        if (rc) return rc;
    }
    HIPCHK(hipGraphLaunch(*exec, g));
    return 0;
}

int sph_adjoint_segment(sph_ctx_t *ctx, sph_tape_t *tape, int32_t nsteps, const sph_adjoint_t *adj, void *stream) {
    if (ctx && closed(ctx)) return SPH_ESTATE;
    if (!ctx || !tape || !adj) return fail(SPH_EINVAL, "NULL argument");
    if (tape->ctx != ctx) return fail(SPH_EINVAL, "tape belongs to another context");
    if (nsteps < 1 || nsteps > tape->nslots) return fail(SPH_EINVAL, "nsteps %d out of 1..%d", nsteps, tape->nslots);
    if (!adj->x || !adj->y || !adj->vx || !adj->vy || !adj->rho || !adj->ax || !adj->ay || !adj->drho ||
        !adj->g_alpha || !adj->g_co)
        return fail(SPH_EINVAL, "every sph_adjoint_t pointer is required");
    int rc = tape_supported(ctx);
    if (rc) return rc;
    if (tape->gen != ctx->gen) {  // constants or launch shape changed since the capture: every graph is stale
        tape_drop_graphs(tape);
        tape->gen = ctx->gen;
    }
    if (std::memcmp(&tape->adj_key, adj, sizeof *adj) != 0) {  // other buffers: the reverse graphs are stale
        for (auto &g : tape->adj_exec)
            if (g) hipGraphExecDestroy(g), g = nullptr;
        tape->adj_key = *adj;
    }
    hipStream_t s = S(stream), g = ctx->own;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreateWithFlags(&e0, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&e1, hipEventDisableTiming));
    HIPCHK(hipEventRecord(e0, s));
    HIPCHK(hipStreamWaitEvent(g, e0, 0));
    rc = tape_refresh_labels(ctx, tape, g);
    for (int k = 0; k < nsteps && !rc; ++k) {
        if (!ctx->dt_fresh) {  // a captured step expects the corrector's dt partials
            rc = record_step(ctx, tape, k, g);
            continue;
        }
        const int start = ctx->cur;
        rc = graph_launch(&tape->rec_exec[2 * (size_t)k + start], g, [&] { return record_step(ctx, tape, k, g); });
        // a capture flips it without running; a launched step ends on the other buffer
        ctx->cur = rc ? start : 1 - start;
    }
    for (int k = nsteps - 1; k >= 0 && !rc; --k)
        rc = graph_launch(&tape->adj_exec[k], g, [&] { return adjoint_launch(ctx, tape, k, adj, g); });
    // the caller's stream waits for whatever was queued, also after an error
    hipEventRecord(e1, g);
    hipStreamWaitEvent(s, e1, 0);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc;
}

int sph_nn_update(sph_ctx_t *ctx, void *stream) {
    if (!ctx) return fail(SPH_EINVAL, "ctx is NULL");
    if (closed(ctx)) return SPH_ESTATE;
    return nn_update(ctx, S(stream));
}

int sph_compute(sph_ctx_t *ctx, void *stream) {
    if (!ctx) return fail(SPH_EINVAL, "ctx is NULL");
    if (closed(ctx)) return SPH_ESTATE;
    ctx->dt_fresh = false;
    return compute(ctx, S(stream));
}

int sph_continuity(sph_ctx_t *ctx, void *stream) {
    if (!ctx) return fail(SPH_EINVAL, "ctx is NULL");
    if (closed(ctx)) return SPH_ESTATE;
    ctx->dt_fresh = false;
    return compute(ctx, S(stream), nullptr, 1);
}

int sph_momentum(sph_ctx_t *ctx, void *stream) {
    if (!ctx) return fail(SPH_EINVAL, "ctx is NULL");
    if (closed(ctx)) return SPH_ESTATE;
    ctx->dt_fresh = false;
    return compute(ctx, S(stream), nullptr, 2);
}

int sph_dt(sph_ctx_t *ctx, void *stream) {
    if (!ctx) return fail(SPH_EINVAL, "ctx is NULL");
    if (closed(ctx)) return SPH_ESTATE;
    return dt_step(ctx, S(stream));
}

int sph_predict(sph_ctx_t *ctx, void *stream) {
    if (!ctx) return fail(SPH_EINVAL, "ctx is NULL");
    if (closed(ctx)) return SPH_ESTATE;
    ctx->dt_fresh = false;  // the positions move: the corrector's displacement terms no longer hold
    return predict(ctx, S(stream));
}

int sph_correct(sph_ctx_t *ctx, void *stream) {
    if (!ctx) return fail(SPH_EINVAL, "ctx is NULL");
    if (closed(ctx)) return SPH_ESTATE;
    int rc = correct(ctx, S(stream));
    if (!rc) ctx->dt_fresh = true;
    return rc;
}

int sph_step(sph_ctx_t *ctx, int64_t nsteps, int32_t use_graph, void *stream) {
    if (!ctx) return fail(SPH_EINVAL, "ctx is NULL");
    if (closed(ctx)) return SPH_ESTATE;
    if (nsteps < 0) return fail(SPH_EINVAL, "nsteps < 0");
    hipStream_t s = S(stream);
    if (!use_graph || ctx->timing) {
        for (int64_t k = 0; k < nsteps; ++k) {
            int rc = one_step(ctx, s);
            if (rc) return rc;
        }
        return 0;
    }
    // Graph path: one captured step per starting buffer (the step flips
    // `cur`), replayed alternately on the context's own stream, ordered after
    // and before the caller's stream with events.
    hipStream_t g = ctx->own;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreateWithFlags(&e0, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&e1, hipEventDisableTiming));
    HIPCHK(hipEventRecord(e0, s));
    HIPCHK(hipStreamWaitEvent(g, e0, 0));
    int rc = 0;
    if (!ctx->dt_fresh && nsteps > 0) {  // the captured step expects the corrector's dt partials
        rc = one_step(ctx, g);
        --nsteps;
    }
    for (int64_t k = 0; k < nsteps && !rc; ++k) {
        const int start = ctx->cur;
        if (!ctx->gexec[start]) {
            rc = capture_step(&ctx->gexec[start], g, [&] { return one_step(ctx, g); });
            ctx->cur = start;  // capture flipped it without running anything
            if (rc) break;
        }
        hipError_t le = hipGraphLaunch(ctx->gexec[start], g);
        if (le != hipSuccess) rc = fail(SPH_EHIP, "graph launch: %s", hipGetErrorString(le));
        else ctx->cur = 1 - start;
    }
    // the caller's stream waits for whatever was queued, also after an error
    hipEventRecord(e1, g);
    hipStreamWaitEvent(s, e1, 0);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc;
}

// The loop body split at the coupling hook (src/Solver.py:389-392). The
// first half writes the whole predicted state (no lean reorder), so what a
// caller changes in between reaches the correctors as in the reference.
int sph_step_begin(sph_ctx_t *ctx, void *stream) {
    if (!ctx) return fail(SPH_EINVAL, "ctx is NULL");
    if (closed(ctx)) return SPH_ESTATE;
    hipStream_t s = S(stream);
    int rc;
    if ((rc = dt_step(ctx, s))) return rc;
    if ((rc = nn_update(ctx, s, ctx->P.integrator == SPH_INTEGRATOR_PEC))) return rc;
    if ((rc = compute(ctx, s))) return rc;
    ctx->mid = true;
    return 0;
}

int sph_step_end(sph_ctx_t *ctx, void *stream) {
    if (!ctx) return fail(SPH_EINVAL, "ctx is NULL");
    if (!ctx->mid) return fail(SPH_ESTATE, "no split step is open: call sph_step_begin first");
    ctx->mid = false;
    int rc = correct(ctx, S(stream));
    if (!rc) ctx->dt_fresh = true;
    return rc;
}

// Stage `ids` (host, validated) in ctx->bin and map every reference index to
// its slot of the current buffer in ctx->slot_of. bin is scratch outside the
// neighbour update.
static int stage_ids(sph_ctx *c, const int64_t *ids, int64_t k, bool distinct, hipStream_t s) {
    if (k < 0 || k > c->n) return fail(SPH_EINVAL, "k=%lld out of range (0 .. n=%d)", (long long)k, c->n);
    if (k > 0 && !ids) return fail(SPH_EINVAL, "ids is NULL");
    std::vector<int> v((size_t)k);
    for (int64_t j = 0; j < k; ++j) {
        if (ids[j] < 0 || ids[j] >= c->n)
            return fail(SPH_EINVAL, "ids[%lld]=%lld outside 0 .. n-1=%d", (long long)j, (long long)ids[j], c->n - 1);
        v[j] = (int)ids[j];
    }
    if (distinct) {
        std::vector<int> w(v);
        std::sort(w.begin(), w.end());
        if (std::adjacent_find(w.begin(), w.end()) != w.end()) return fail(SPH_EINVAL, "ids repeat an index");
    }
    if (k == 0) return 0;
    HIPCHK(hipMemcpyAsync(c->bin, v.data(), sizeof(int) * k, hipMemcpyHostToDevice, s));
    k_inverse<<<nblk(c->n), kBlock, 0, s>>>(c->buf[c->cur].id, c->slot_of, c->n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));  // v is freed on return
    return 0;
}

int sph_state_gather(sph_ctx_t *ctx, const int64_t *ids, int64_t k, const sph_state_t *st, void *stream) {
    if (!ctx || !st) return fail(SPH_EINVAL, "NULL argument");
    if (unloaded(ctx)) return SPH_ESTATE;
    hipStream_t s = S(stream);
    int rc = stage_ids(ctx, ids, k, false, s);
    if (rc || k == 0) return rc;
    k_gather_ids<<<nblk(k), kBlock, 0, s>>>(ctx->buf[ctx->cur], ctx->slot_of, ctx->bin, (int)k, *st,
                                            make_phys(ctx->P), ctx->xs_fluid_zero);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

int sph_state_scatter(sph_ctx_t *ctx, const int64_t *ids, int64_t k, const sph_state_t *st, void *stream) {
    if (!ctx || !st) return fail(SPH_EINVAL, "NULL argument");
    if (unloaded(ctx)) return SPH_ESTATE;
    if (st->label || st->p) return fail(SPH_EINVAL, "label and p cannot be written (p follows from rho)");
    if (!ctx->P.use_xsph && (st->xsphx || st->xsphy))
        return fail(SPH_EUNSUPPORTED, "without XSPH the XSPH velocity is not kept on the device");
    hipStream_t s = S(stream);
    int rc = stage_ids(ctx, ids, k, true, s);
    if (rc || k == 0) return rc;
    k_scatter_ids<<<nblk(k), kBlock, 0, s>>>(ctx->buf[ctx->cur], ctx->slot_of, ctx->bin, (int)k, *st);
    HIPCHK(hipGetLastError());
    ctx->dt_fresh = false;  // h or a may have changed under the corrector's dt partials
    HIPCHK(hipStreamSynchronize(s));
    return (st->h || st->m) ? refresh_uniform(ctx, s) : 0;
}

int sph_near_pos(sph_ctx_t *ctx, int64_t npts, const double *x, const double *y, const double *h, int64_t *off,
                 int32_t *idx, double *q, double *r, int64_t cap, int64_t *total, void *stream) {
    if (!ctx || !off || !total) return fail(SPH_EINVAL, "NULL argument");
    if (npts < 0 || npts > (1LL << 30)) return fail(SPH_EINVAL, "npts=%lld out of range", (long long)npts);
    if (npts > 0 && (!x || !y || !h)) return fail(SPH_EINVAL, "x, y and h are needed");
    hipStream_t s = S(stream);
    int rc = sph_nn_grid(ctx, nullptr, nullptr, stream);
    if (rc) return rc;
    const int np = (int)npts;
    State st = ctx->buf[ctx->cur];
    if (np > 0)
        k_near_pos<<<nblk(np), kBlock, 0, s>>>(ctx->ctrl, st, ctx->cellStart, np, x, y, h, 0, (long long *)off,
                                                nullptr, nullptr, nullptr, ctx->P.qmax);
    HIPCHK(hipGetLastError());
    std::vector<long long> cnt((size_t)np + 1);
    if (np > 0) HIPCHK(hipMemcpyAsync(cnt.data(), off, sizeof(long long) * np, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    long long acc = 0;
    for (int i = 0; i < np; ++i) {
        long long k = cnt[i];
        cnt[i] = acc;
        acc += k;
    }
    cnt[np] = acc;
    *total = acc;
    HIPCHK(hipMemcpyAsync(off, cnt.data(), sizeof(long long) * (np + 1), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    if (acc > cap) return fail(SPH_ECAPACITY, "nearPos needs %lld entries, cap is %lld", acc, (long long)cap);
    if (acc == 0) return 0;
    if (!idx) return fail(SPH_EINVAL, "idx is NULL");
    k_near_pos<<<nblk(np), kBlock, 0, s>>>(ctx->ctrl, st, ctx->cellStart, np, x, y, h, 1, (long long *)off, idx, q,
                                            r, ctx->P.qmax);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

int sph_probe_pressure(sph_ctx_t *ctx, int64_t npts, const double *x, const double *y, const double *h,
                       double *rho, double *p, void *stream) {
    if (!ctx) return fail(SPH_EINVAL, "ctx is NULL");
    if (npts < 0 || npts > (1LL << 30)) return fail(SPH_EINVAL, "npts=%lld out of range", (long long)npts);
    if (npts > 0 && (!x || !y || !h || !rho || !p)) return fail(SPH_EINVAL, "x, y, h, rho and p are needed");
    int rc = sph_nn_grid(ctx, nullptr, nullptr, stream);
    if (rc || npts == 0) return rc;
    const sph_params_t &P = ctx->P;
    k_probe<<<nblk(npts), kBlock, 0, S(stream)>>>(ctx->ctrl, ctx->buf[ctx->cur], ctx->cellStart, (int)npts, x, y, h,
                                                  P.kernel, P.qmax, P.rho0, P.gamma, P.B, rho, p);
    HIPCHK(hipGetLastError());
    return 0;
}

int sph_status(sph_ctx_t *ctx, double out_d[4], int64_t out_i[2], void *stream) {
    if (!ctx) return fail(SPH_EINVAL, "ctx is NULL");
    HIPCHK(hipStreamSynchronize(S(stream)));
    Ctrl h;
    HIPCHK(hipMemcpy(&h, ctx->ctrl, sizeof h, hipMemcpyDeviceToHost));
    if (out_d) {
        out_d[0] = h.t;
        out_d[1] = h.dt;
        out_d[2] = h.dt_c;
        out_d[3] = h.dt_f;
    }
    if (out_i) {
        out_i[0] = h.step;
        out_i[1] = h.flags;
    }
    return 0;
}

int sph_set_clock(sph_ctx_t *ctx, double t, int32_t advance_time, void *stream) {
    if (!ctx) return fail(SPH_EINVAL, "ctx is NULL");
    HIPCHK(hipStreamSynchronize(S(stream)));
    Ctrl h;
    HIPCHK(hipMemcpy(&h, ctx->ctrl, sizeof h, hipMemcpyDeviceToHost));
    h.t = t;
    h.advance_t = advance_time ? 1 : 0;
    HIPCHK(hipMemcpy(ctx->ctrl, &h, sizeof h, hipMemcpyHostToDevice));
    return 0;
}

int sph_kinetic_energy(sph_ctx_t *ctx, double *out, void *stream) {
    if (!ctx || !out) return fail(SPH_EINVAL, "NULL argument");
    if (unloaded(ctx)) return SPH_ESTATE;
    hipStream_t s = S(stream);
    const int R = redblk(ctx->n);
    k_ke_partial<<<R, kBlock, 0, s>>>(ctx->buf[ctx->cur], ctx->n, ctx->part_k);
    k_ke_final<<<1, 1024, 0, s>>>(ctx->part_k, R);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, ctx->part_k + R, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

int sph_dt_history(sph_ctx_t *ctx, double *out3, int64_t count, void *stream) {
    if (!ctx || !out3) return fail(SPH_EINVAL, "NULL argument");
    if (count < 0 || count > SPH_DT_HISTORY) return fail(SPH_EINVAL, "count out of range");
    HIPCHK(hipStreamSynchronize(S(stream)));
    Ctrl h;
    HIPCHK(hipMemcpy(&h, ctx->ctrl, sizeof h, hipMemcpyDeviceToHost));
    if (count > h.step) return fail(SPH_EINVAL, "only %lld steps taken", (long long)h.step);
    std::vector<double> ring(3 * (size_t)SPH_DT_HISTORY);
    HIPCHK(hipMemcpy(ring.data(), ctx->dt_hist, sizeof(double) * ring.size(), hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < count; ++k) {
        long long step = h.step - count + k;
        long long slot = step % SPH_DT_HISTORY;
        for (int c = 0; c < 3; ++c) out3[3 * k + c] = ring[3 * slot + c];
    }
    return 0;
}

int sph_nn_grid(sph_ctx_t *ctx, double out_d[5], int64_t out_i[3], void *stream) {
    if (!ctx) return fail(SPH_EINVAL, "ctx is NULL");
    HIPCHK(hipStreamSynchronize(S(stream)));
    Ctrl h;
    HIPCHK(hipMemcpy(&h, ctx->ctrl, sizeof h, hipMemcpyDeviceToHost));
    if (!h.nn_valid) return fail(SPH_ESTATE, "no valid neighbour update (flags=0x%x)", h.flags);
    if (out_d) {
        out_d[0] = h.grid.xmin;
        out_d[1] = h.grid.xmax;
        out_d[2] = h.grid.ymin;
        out_d[3] = h.grid.ymax;
        out_d[4] = h.grid.cs;
    }
    if (out_i) {
        out_i[0] = h.grid.ncx;
        out_i[1] = h.grid.ncy;
        out_i[2] = h.grid.ncells;
    }
    return 0;
}

int sph_nn_dump(sph_ctx_t *ctx, int64_t *keys, int64_t *off, int32_t *idx, double *q, double *r, int64_t cap,
                int64_t *total, void *stream) {
    if (!ctx || !keys || !off || !total) return fail(SPH_EINVAL, "NULL argument");
    if (unloaded(ctx)) return SPH_ESTATE;
    hipStream_t s = S(stream);
    int rc = sph_nn_grid(ctx, nullptr, nullptr, stream);
    if (rc) return rc;
    const int n = ctx->n;
    State st = ctx->buf[ctx->cur];
    k_inverse<<<nblk(n), kBlock, 0, s>>>(st.id, ctx->slot_of, n);
    k_nn_dump<<<nblk(n), kBlock, 0, s>>>(ctx->ctrl, st, ctx->cellStart, ctx->slot_of, n, 0, (long long *)keys,
                                          (long long *)off, nullptr, nullptr, nullptr, ctx->P.qmax);
    HIPCHK(hipGetLastError());
    std::vector<long long> cnt((size_t)n + 1);
    HIPCHK(hipMemcpyAsync(cnt.data(), off, sizeof(long long) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    long long acc = 0;
    for (int i = 0; i < n; ++i) {
        long long k = cnt[i];
        cnt[i] = acc;
        acc += k;
    }
    cnt[n] = acc;
    *total = acc;
    HIPCHK(hipMemcpyAsync(off, cnt.data(), sizeof(long long) * (n + 1), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));  // cnt is freed on return
    if (acc > cap) return fail(SPH_ECAPACITY, "dump needs %lld entries, cap is %lld", acc, (long long)cap);
    if (!idx) return fail(SPH_EINVAL, "idx is NULL");
    k_nn_dump<<<nblk(n), kBlock, 0, s>>>(ctx->ctrl, st, ctx->cellStart, ctx->slot_of, n, 1, nullptr,
                                          (long long *)off, idx, q, r, ctx->P.qmax);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

__global__ void k_unpermute_int(const int *src, const int *id, int *dst, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[id[t]] = src[t];
}

int sph_compute_counts(sph_ctx_t *ctx, int32_t *counts, void *stream) {
    if (!ctx || !counts) return fail(SPH_EINVAL, "NULL argument");
    if (closed(ctx)) return SPH_ESTATE;
    hipStream_t s = S(stream);
    int rc = sph_nn_grid(ctx, nullptr, nullptr, stream);
    if (rc) return rc;
    // ctx->bin is free between updates: use it as the cell-order count buffer
    ctx->dt_fresh = false;
    if ((rc = compute(ctx, s, ctx->bin))) return rc;
    k_unpermute_int<<<nblk(ctx->n), kBlock, 0, s>>>(ctx->bin, ctx->buf[ctx->cur].id, counts, ctx->n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

int sph_kernel_evaluate(int32_t kernel, int64_t n, const double *r, const double *h, double *out, void *stream) {
    if (n < 0 || n > (1LL << 31) - 1 || kernel < 0 || kernel > 2) return fail(SPH_EINVAL, "bad n or kernel");
    if (n == 0) return 0;
    k_kernel_eval<<<nblk(n), kBlock, 0, S(stream)>>>(kernel, (int)n, r, h, out);
    HIPCHK(hipGetLastError());
    return 0;
}

int sph_kernel_gradient(int32_t kernel, int64_t n, const double *x, const double *r, const double *h, double *out,
                        void *stream) {
    if (n < 0 || n > (1LL << 31) - 1 || kernel < 0 || kernel > 2) return fail(SPH_EINVAL, "bad n or kernel");
    if (n == 0) return 0;
    k_kernel_grad<<<nblk(n), kBlock, 0, S(stream)>>>(kernel, (int)n, x, r, h, out);
    HIPCHK(hipGetLastError());
    return 0;
}

int sph_kernel_derivative(int32_t kernel, int64_t n, const double *r, const double *h, double *out, void *stream) {
    if (kernel != SPH_KERNEL_GAUSSIAN) return fail(SPH_EUNSUPPORTED, "derivative exists for the Gaussian only");
    if (n < 0 || n > (1LL << 31) - 1) return fail(SPH_EINVAL, "bad n");
    if (n == 0) return 0;
    k_kernel_deriv<<<nblk(n), kBlock, 0, S(stream)>>>((int)n, r, h, out);
    HIPCHK(hipGetLastError());
    return 0;
}

int sph_tait_eos(int64_t n, double gamma, double B, double rho0, double Pb, const double *rho, const int8_t *label,
                 double *p_out, void *stream) {
    if (n < 0 || n > (1LL << 31) - 1) return fail(SPH_EINVAL, "bad n");
    if (n == 0) return 0;
    k_tait<<<nblk(n), kBlock, 0, S(stream)>>>((int)n, gamma, B, rho0, Pb, rho, label, p_out);
    HIPCHK(hipGetLastError());
    return 0;
}

int sph_ctx_set_tuning(sph_ctx_t *ctx, int32_t pair_threads) {
    if (!ctx) return fail(SPH_EINVAL, "ctx is NULL");
    if (pair_threads != 64 && pair_threads != 128 && pair_threads != 256)
        return fail(SPH_EINVAL, "pair_threads must be 64, 128 or 256");
    ctx->pair_threads = pair_threads;
    for (int k = 0; k < 2; ++k)
        if (ctx->gexec[k]) {
            hipGraphExecDestroy(ctx->gexec[k]);
            ctx->gexec[k] = nullptr;
        }
    ++ctx->gen;
    return 0;
}

int sph_ctx_set_pair_staging(sph_ctx_t *ctx, int32_t on) {
    if (!ctx) return fail(SPH_EINVAL, "ctx is NULL");
    ctx->pair_staging = on ? 1 : 0;
    drop_graphs(ctx);
    return 0;
}

int sph_ctx_set_border_bounds(sph_ctx_t *ctx, int32_t on) {
    if (!ctx) return fail(SPH_EINVAL, "ctx is NULL");
    ctx->border_bounds = on ? 1 : 0;
    drop_graphs(ctx);
    return 0;
}

int sph_nn_border_count(sph_ctx_t *ctx, int64_t *count, void *stream) {
    if (!ctx || !count) return fail(SPH_EINVAL, "NULL argument");
    HIPCHK(hipStreamSynchronize(S(stream)));
    Ctrl h;
    HIPCHK(hipMemcpy(&h, ctx->ctrl, sizeof h, hipMemcpyDeviceToHost));
    *count = h.border_updates;
    return 0;
}

int sph_pair_form(sph_ctx_t *ctx, int32_t *form) {
    if (!ctx || !form) return fail(SPH_EINVAL, "NULL argument");
    *form = rec_format(ctx);  // kRecFp64 .. kRecUniform64 are SPH_PAIR_* in the same order
    return 0;
}

int sph_ctx_get_tuning(sph_ctx_t *ctx, int32_t *pair_threads) {
    if (!ctx || !pair_threads) return fail(SPH_EINVAL, "NULL argument");
    *pair_threads = ctx->pair_threads;
    return 0;
}

int sph_timing_enable(sph_ctx_t *ctx, int32_t on) {
    if (!ctx) return fail(SPH_EINVAL, "ctx is NULL");
    ctx->timing = on != 0;
    ctx->n_pair = ctx->n_nn = 0;
    return 0;
}

int sph_timing_read(sph_ctx_t *ctx, double *avg_ms_pair, double *avg_ms_nn, int64_t *count, int32_t reset) {
    if (!ctx) return fail(SPH_EINVAL, "ctx is NULL");
    HIPCHK(hipDeviceSynchronize());
    double tp = 0.0, tn = 0.0;
    for (size_t k = 0; k + 1 < ctx->n_pair; k += 2) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ctx->ev_pair[k], ctx->ev_pair[k + 1]));
        tp += ms;
    }
    for (size_t k = 0; k + 1 < ctx->n_nn; k += 2) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ctx->ev_nn[k], ctx->ev_nn[k + 1]));
        tn += ms;
    }
    size_t np = ctx->n_pair / 2, nn = ctx->n_nn / 2;
    if (avg_ms_pair) *avg_ms_pair = np ? tp / np : 0.0;
    if (avg_ms_nn) *avg_ms_nn = nn ? tn / nn : 0.0;
    if (count) *count = (int64_t)np;
    if (reset) ctx->n_pair = ctx->n_nn = 0;
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// nn.update + _loop with no allocation (sph_loop_noalloc)
//
// The workspace of sph_nn_update + sph_compute (cell table, sort arrays,
// cell-ordered pair records, partials, control block) lives in one static
// device array, `scratch`, sized for SPH_NOALLOC_MAX_PARTICLES particles and
// SPH_NOALLOC_MAX_CELLS cells. Host code never passes the array's name to a
// kernel: its device address comes from hipGetSymbolAddress, once per device
// (the first call on that device), and kernels receive that address. Every
// call on a device shares that device's array, so calls that overlap on one
// device race; each call clears what must start at zero.
// ---------------------------------------------------------------------------
namespace {
constexpr long long kNaParticles = SPH_NOALLOC_MAX_PARTICLES;
constexpr long long kNaCells = SPH_NOALLOC_MAX_CELLS;
constexpr size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }
static_assert(sizeof(Rec<double>::A) == 32 && sizeof(Rec<double>::B) == 32 && sizeof(Rec<float>::A) == 32 &&
                  sizeof(Rec<float>::B) <= 32,
              "pair records are sized at 32 bytes in the workspace");
constexpr size_t kNaOffCtrl = 0;
constexpr size_t kNaOffPart = kNaOffCtrl + al256(sizeof(Ctrl));
constexpr size_t kNaOffId = kNaOffPart + al256(sizeof(double) * kBoundsVQ * kBoundsBlocks);
constexpr size_t kNaOffBin = kNaOffId + al256(4 * kNaParticles);
constexpr size_t kNaOffRank = kNaOffBin + al256(4 * kNaParticles);
constexpr size_t kNaOffPerm = kNaOffRank + al256(4 * kNaParticles);
constexpr size_t kNaOffPermKey = kNaOffPerm + al256(4 * kNaParticles);
constexpr size_t kNaOffCxy = kNaOffPermKey + al256(4 * kNaParticles);
constexpr size_t kNaOffRecA = kNaOffCxy + al256(8 * kNaParticles);
constexpr size_t kNaOffRecB = kNaOffRecA + al256(32 * kNaParticles);
constexpr size_t kNaOffH = kNaOffRecB + al256(32 * kNaParticles);
constexpr size_t kNaOffAx = kNaOffH + al256(8 * kNaParticles);
constexpr size_t kNaOffAy = kNaOffAx + al256(8 * kNaParticles);
constexpr size_t kNaOffDrho = kNaOffAy + al256(8 * kNaParticles);
constexpr size_t kNaOffCounts = kNaOffDrho + al256(8 * kNaParticles);
constexpr size_t kNaOffCellStart = kNaOffCounts + al256(4 * (kNaCells + 1));
constexpr size_t kNaOffBinflag = kNaOffCellStart + al256(4 * (kNaCells + 2));
constexpr size_t kNaOffTiles = kNaOffBinflag + al256(4 * (kNaCells + 1));
constexpr size_t kNaBytes = kNaOffTiles + al256(4 * ((kNaCells + 1 + kScanTile - 1) / kScanTile + 1));
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNaBytes];

constexpr int kNaDevices = 64;
static std::mutex g_na_mu;
static void *g_na_ptr[kNaDevices];  // device address of `scratch` per device ordinal

// The current device's address of `scratch` (hipGetSymbolAddress on the
// first call per device, with that device current).
static int na_workspace(unsigned char **out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNaDevices) return fail(SPH_EINVAL, "device ordinal %d beyond %d", dev, kNaDevices);
    std::lock_guard<std::mutex> lock(g_na_mu);
    if (!g_na_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SPH_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_na_ptr[dev] = p;
    }
    *out = (unsigned char *)g_na_ptr[dev];
    return 0;
}

// Start of a call: an active control block with no flags, no predictor and
// no coupled integrator; the cell counts and bin flags cleared over the whole
// capacity (the workspace keeps the last call's contents); the identity as
// each particle's reference index.
__global__ void k_na_init(Ctrl *ctrl, int *counts, int *binflag, int *id, int n, long long ncap) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        Ctrl c{};
        c.active = 1;
        *ctrl = c;
    }
    if (i <= ncap) {
        counts[i] = 0;
        binflag[i] = 0;
    }
    if (i < n) id[i] = (int)i;
}

// The cell-ordered inputs of the pair pass, gathered from the caller's
// arrays through the permutation: records (as k_reorder writes them without
// a predictor), the double h, each slot's raw cell, and the rates a target
// keeps when it finds no neighbour (src/Tools/SolverTools.py:166-169).
__global__ void k_na_reorder(const Ctrl *ctrl, sph_state_t in, const int *perm, int2 *cxy, double *H, double *ax,
                             double *ay, double *drho, Phys ph, int n, int fp32, void *recA, void *recB) {
    if (!ctrl->active) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int s = perm[t];
    const double x = in.x[s], y = in.y[s], vx = in.vx[s], vy = in.vy[s], rho = in.rho[s], m = in.m[s], h = in.h[s];
    const int8_t lab = in.label[s];
    ax[t] = in.ax ? in.ax[s] : 0.0;
    ay[t] = in.ay ? in.ay[s] : 0.0;
    drho[t] = in.drho ? in.drho[s] : 0.0;
    cxy[t] = cell_of(x, y, ctrl->grid);
    H[t] = h;
    write_pair_records(t, x, y, vx, vy, rho, m, h, lab, ph, fp32, recA, recB);
}

// The rates back to the caller's order, and the status word.
__global__ void k_na_finish(const Ctrl *ctrl, const int *perm, const double *ax_s, const double *ay_s,
                            const double *drho_s, double *ax, double *ay, double *drho, int n, int32_t *status) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) *status = (int32_t)ctrl->flags;
    if (!ctrl->active || t >= n) return;
    const int s = perm[t];
    ax[s] = ax_s[t];
    ay[s] = ay_s[t];
    drho[s] = drho_s[t];
}

extern "C" int sph_loop_noalloc(const sph_params_t *params, const sph_state_t *st, double *drho, double *ax,
                                double *ay, int32_t *status, void *stream) {
    if (!params || !st || !drho || !ax || !ay || !status) return fail(SPH_EINVAL, "NULL argument");
    const long long n = params->n;
    if (n < 1 || n > kNaParticles)
        return fail(SPH_EINVAL, "sph_loop_noalloc: n = %lld outside 1 .. %lld (its static workspace)", n,
                    kNaParticles);
    if (!st->x || !st->y || !st->vx || !st->vy || !st->rho || !st->m || !st->h || !st->label)
        return fail(SPH_EINVAL, "sph_loop_noalloc needs x, y, vx, vy, rho, m, h and label");
    int rc = check_params(params);
    if (rc) return rc;
    if (params->use_xsph || params->use_summation_density || params->dynamic_h)
        return fail(SPH_EUNSUPPORTED, "sph_loop_noalloc runs the plain WCSPH loop (no XSPH, summation density or "
                                      "dynamic h)");
    unsigned char *w = nullptr;
    if ((rc = na_workspace(&w))) return rc;
    Ctrl *ctrl = (Ctrl *)(w + kNaOffCtrl);
    double *part = (double *)(w + kNaOffPart);
    int *id = (int *)(w + kNaOffId), *bin = (int *)(w + kNaOffBin), *rank = (int *)(w + kNaOffRank);
    int *perm = (int *)(w + kNaOffPerm), *permKey = (int *)(w + kNaOffPermKey);
    int2 *cxy = (int2 *)(w + kNaOffCxy);
    void *recA = w + kNaOffRecA, *recB = w + kNaOffRecB;
    double *H = (double *)(w + kNaOffH), *ax_s = (double *)(w + kNaOffAx), *ay_s = (double *)(w + kNaOffAy);
    double *drho_s = (double *)(w + kNaOffDrho);
    int *counts = (int *)(w + kNaOffCounts), *cellStart = (int *)(w + kNaOffCellStart);
    int *binflag = (int *)(w + kNaOffBinflag), *tileSums = (int *)(w + kNaOffTiles);
    hipStream_t s = S(stream);
    const int nn = (int)n;
    State a{};  // the caller's arrays, read by the bounds and keys kernels
    a.x = st->x;
    a.y = st->y;
    a.h = st->h;
    a.label = st->label;
    a.id = id;
    const long long ninit = (kNaCells + 1 > n ? kNaCells + 1 : n);
    k_na_init<<<nblk(ninit), kBlock, 0, s>>>(ctrl, counts, binflag, id, nn, kNaCells);
    const int R = redblk(n) < kBoundsBlocks ? redblk(n) : kBoundsBlocks;
    k_bounds<<<R, kBlock, 0, s>>>(ctrl, a, nn, 0, part, HmConst{}, nullptr);
    k_grid<<<1, kGridThreads, 0, s>>>(ctrl, part, R, params->scale, kNaCells);
    k_keys<<<nblk(n), kBlock, 0, s>>>(ctrl, a, nn, 0, counts, bin, rank, binflag);
    const int tiles = (int)((kNaCells + 1 + kScanTile - 1) / kScanTile);
    k_scan_reduce<<<tiles, kBlock, 0, s>>>(ctrl, counts, tileSums);
    k_scan_apply<<<tiles, kBlock, 0, s>>>(ctrl, counts, tileSums, cellStart);
    k_scatter<<<nblk(n), kBlock, 0, s>>>(ctrl, cellStart, bin, rank, id, perm, permKey, nn);
    k_cellsort<<<redblk(kNaCells + 1), kBlock, 0, s>>>(ctrl, cellStart, perm, permKey, binflag, bin, rank);
    const bool f32 = params->precision == SPH_PREC_FP32C;
    const Phys ph = make_phys(*params);
    k_na_reorder<<<nblk(n), kBlock, 0, s>>>(ctrl, *st, perm, cxy, H, ax_s, ay_s, drho_s, ph, nn, f32, recA, recB);
    HIPCHK(hipGetLastError());
    PairIn in{recA, recB, H, cxy, cellStart};
    PairOut out{ax_s, ay_s, drho_s, nullptr, 0};
    hipError_t e = f32 ? launch_pair_fp32c(ctrl, in, out, ph, nn, 256, s) : launch_pair_fp64(ctrl, in, out, ph, nn, 256, s);
    if (e != hipSuccess) return fail(SPH_EHIP, "pair launch: %s", hipGetErrorString(e));
    k_na_finish<<<nblk(n), kBlock, 0, s>>>(ctrl, perm, ax_s, ay_s, drho_s, ax, ay, drho, nn, status);
    HIPCHK(hipGetLastError());
    return 0;
}
