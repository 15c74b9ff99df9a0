// sph_adjoint.hip — reverse mode of the WCSPH step (tape + hand-written adjoint).
//
// The forward step (PEC or Euler, src/Integrators/PEC.py:32-89,
// src/Integrators/Euler.py:13-26; pair pass src/Tools/SolverTools.py:120-175)
// is a map (x, v, rho, a_prev, drho_prev) -> (x', v', rho', a, drho) with the
// pair pass F: (x_half, v_half, rho_half; alpha, c0) -> (a, drho) in the
// middle. sph_step_record runs one ordinary device step and stores what the
// reverse pass needs; sph_adjoint_step maps the adjoint of the step's outputs
// to the adjoint of its inputs and adds the step's share of dL/dalpha and
// dL/dc0 (c0 enters through B = rho0 c0^2 / gamma, the artificial-viscosity
// sound speed c_ij = c0 / 2 and the Courant dt = gamma_c h_min / c0).
//
// Reverse of one PEC step (d = damping, e = 1 + d/2):
//   correct: x' = 2(x + dt/2 v_h) - x, v' = 2(v + dt/2 a)/e - v,
//            rho' = 2(rho + dt/2 drho) - rho
//     -> lam = abar_c + dt/e vbar', kap = dbar_c + dt rhobar', vbar_h = dt xbar',
//        vbar = (2/e - 1) vbar', dtbar += xbar'.v_h + vbar'.a/e + rhobar' drho
//   pair:    (xbar_h, vbar_h, rhobar_h) += F^T (lam, kap)
//   predict: x_h = x + dt/2 v, v_h = (v + dt/2 a_prev)/e, rho_h = rho + dt/2 drho_prev
//     -> xbar += xbar_h, vbar += dt/2 xbar_h + vbar_h/e, rhobar += rhobar_h,
//        abar_prev = dt/2 vbar_h/e, dbar_prev = dt/2 rhobar_h,
//        dtbar += (xbar_h.v + vbar_h.a_prev/e + rhobar_h drho_prev)/2
//   dt:      Courant value gamma_c h_min / c0: c0bar += dtbar (-dt/c0); force
//            value gamma_f sqrt(h_min / a2max), a2max = max over the fluid of
//            |a_prev|^2 (src/Equations/TimeStep.py:40-56): the maximising
//            particle k (recorded; smallest id on ties) gets
//            abar_prev_k += dtbar (-dt / a2max) a_prev_k.
//
// F^T is a gather over the same 3x3 cells and the same neighbour rule as the
// forward pass (no atomics): each fluid particle k visits its neighbours j
// once and evaluates the adjoint of the pair term twice — as the target of
// the pair (k's own sums, weights lam_k, kap_k) and as the neighbour of j's
// pair (j's sums, weights lam_j, kap_j) — so every pair term's derivative with
// respect to k's inputs is summed by k alone. FP64 throughout.
#include <cmath>
#include <cstring>

#include "sph_ctx.h"
#include "sph_common.h"

using namespace hipsph;

namespace {

constexpr double kPI_ = 3.141592653589793;
constexpr int kAdjBlock = 256;
constexpr int kAdjQ = 4;  // partials per workgroup: dtbar, gAlpha, gC0 (+1 spare)

// Kernel gradient factor G(d) with grad W = G r, and dG/dd, for h_ij = hh.
// Zero outside the support (where the forward pass adds exactly nothing).
template <int KERNEL>
__device__ inline bool grad_factor(double d, double hh, double &G, double &Gp) {
    const double q = d / hh;
    if (KERNEL == SPH_KERNEL_GAUSSIAN) {
        if (!(d * hh > 1e-12 && q <= 3.0)) return false;
        const double aG = (1.0 / kPI_) / (hh * hh);
        const double e = exp(-q * q);
        G = -2.0 * aG * e / (hh * hh);
        Gp = 4.0 * aG * q * e / (hh * hh * hh);
        return true;
    } else if (KERNEL == SPH_KERNEL_WENDLAND) {
        if (q >= 2.0 || d < 1e-10) return false;
        const double aW = 9.0 / (4.0 * kPI_) / (hh * hh);
        const double in = 1.0 - 0.5 * q;
        const double in2 = in * in, in4 = in2 * in2;
        G = -(14.0 / 3.0) * aW * in4 * in * (1.0 + 2.5 * q) / (hh * hh);
        Gp = 35.0 * aW * q * in4 / (hh * hh * hh);
        return true;
    } else {
        if (q > 2.0 || d < 1e-10) return false;
        const double aC = 10.0 / (7.0 * kPI_) / (hh * hh);
        double g, gp;
        if (q > 1.0) {
            g = -0.75 * (2.0 - q) * (2.0 - q);
            gp = 1.5 * (2.0 - q);
        } else {
            g = -3.0 * q * (1.0 - 0.75 * q);
            gp = -3.0 + 4.5 * q;
        }
        G = aC * g / (hh * d);
        Gp = aC / (hh * d) * (gp / hh - g / d);
        return true;
    }
}

struct Side {  // one particle of a pair
    double h, rho, m;
    double slf;   // p / rho^2
    double dslf;  // d(p / rho^2) / d rho
    double cslf;  // d(p / rho^2) / d c0 (through B)
};

__device__ inline Side make_side(double h, double rho, double m, const Phys &ph) {
    Side s;
    s.h = h;
    s.rho = rho;
    s.m = m;
    const double ratio = ipow(rho / ph.rho0, ph.gamma);
    const double p = (ratio - 1.0) * ph.B + ph.Pb;
    s.slf = p / (rho * rho);
    const double dp = ph.B * ph.gamma * ratio / rho;
    s.dslf = dp / (rho * rho) - 2.0 * p / (rho * rho * rho);
    s.cslf = 2.0 * (ratio - 1.0) * ph.B / ph.co / (rho * rho);
    return s;
}

// Both adjoint terms of the fluid pair (k, j) for particle k at once, from
// shared geometry (r = x_k - x_j, v = v_k - v_j, G and dG/dd of grad W = G r):
//   l1 = lam_k . (-m_j S_kj grad W) + kap_k m_j (v . grad W)   (k's sums)
//   l2 = lam_j . (-m_k S_jk grad W') + kap_j m_k (v' . grad W') (j's sums,
//        r' = -r, v' = -v, grad W' = -grad W)
// acc += (dl1/dr - dl2/dr', dl1/dv - dl2/dv', dl1/drho_k + dl2/drho_k);
// ga, gc += dl1/dalpha, dl1/dc0 (l2's are summed by particle j).
__device__ inline void pair_vjp2(double rx, double ry, double d, double G, double Gp, double vx, double vy,
                                 const Side &K, const Side &J, double lkx, double lky, double kk, double ljx,
                                 double ljy, double kj, const Phys &ph, double (&acc)[5], double &ga, double &gc) {
    const double d2 = rx * rx + ry * ry;
    const double hh = 0.5 * (K.h + J.h);
    const double wx = G * rx, wy = G * ry;
    const double dv = vx * rx + vy * ry;  // the same in both orientations
    const double cb = 0.5 * ph.co;
    const double rhob = 0.5 * (K.rho + J.rho);
    const bool av = dv < 0.0;
    double PI1 = 0.0, mu1 = 0.0, a11 = 0.0, a12 = 0.0;
    double PI2 = 0.0, mu2 = 0.0, a21 = 0.0, a22 = 0.0;
    if (av) {
        const double hm1 = 0.5 * (K.h + hh), den1 = d2 + 0.01 * hm1 * hm1;
        mu1 = hm1 * dv / den1;
        PI1 = mu1 * (ph.beta * mu1 - ph.alpha * cb) / rhob;
        a11 = hm1 / den1;
        a12 = 2.0 * hm1 * dv / (den1 * den1);
        const double hm2 = 0.5 * (J.h + hh), den2 = d2 + 0.01 * hm2 * hm2;
        mu2 = hm2 * dv / den2;
        PI2 = mu2 * (ph.beta * mu2 - ph.alpha * cb) / rhob;
        a21 = hm2 / den2;
        a22 = 2.0 * hm2 * dv / (den2 * den2);
    }
    const double lw1 = lkx * wx + lky * wy, lw2 = ljx * wx + ljy * wy;
    // orientation 1
    const double S1 = K.slf + J.slf + PI1;
    const double sS1 = -J.m * lw1;
    const double u1x = -J.m * S1 * lkx + kk * J.m * vx, u1y = -J.m * S1 * lky + kk * J.m * vy;
    const double ur1 = (u1x * rx + u1y * ry) * Gp / d;
    double rbx = G * u1x + ur1 * rx, rby = G * u1y + ur1 * ry;
    double vbx = kk * J.m * wx, vby = kk * J.m * wy;
    double rk = sS1 * K.dslf;
    gc += sS1 * (K.cslf + J.cslf);
    // orientation 2 (in terms of r, v: grad W' = -w, v' = -v)
    const double S2 = J.slf + K.slf + PI2;
    const double sS2 = K.m * lw2;
    const double u2x = -K.m * S2 * ljx - kj * K.m * vx, u2y = -K.m * S2 * ljy - kj * K.m * vy;
    const double ur2 = (u2x * rx + u2y * ry) * Gp / d;
    rbx -= G * u2x + ur2 * rx;
    rby -= G * u2y + ur2 * ry;
    vbx += kj * K.m * wx;
    vby += kj * K.m * wy;
    rk += sS2 * K.dslf;
    if (av) {
        const double mb1 = sS1 * (2.0 * ph.beta * mu1 - ph.alpha * cb) / rhob;
        rbx += mb1 * (a11 * vx - a12 * rx);
        rby += mb1 * (a11 * vy - a12 * ry);
        vbx += mb1 * a11 * rx;
        vby += mb1 * a11 * ry;
        rk += sS1 * (-0.5 * PI1 / rhob);
        ga += sS1 * (-mu1 * cb / rhob);
        gc += sS1 * (-mu1 * ph.alpha * 0.5 / rhob);
        const double mb2 = sS2 * (2.0 * ph.beta * mu2 - ph.alpha * cb) / rhob;
        // dmu'/dr' = -(a21 v - a22 r), dmu'/dv' = -a21 r; x_k, v_k enter r', v' with -1
        rbx += mb2 * (a21 * vx - a22 * rx);
        rby += mb2 * (a21 * vy - a22 * ry);
        vbx += mb2 * a21 * rx;
        vby += mb2 * a21 * ry;
        rk += sS2 * (-0.5 * PI2 / rhob);
    }
    acc[0] += rbx;
    acc[1] += rby;
    acc[2] += vbx;
    acc[3] += vby;
    acc[4] += rk;
}

// Adjoint of the Lennard-Jones wall term of fluid i, wall j
// (src/Equations/BoundaryForce.py:31-41): f = D ((r0/d)^p1 - (r0/d)^p2) r / d^2.
__device__ inline void wall_vjp(double rx, double ry, double lamx, double lamy, const Phys &ph, double &gx,
                                double &gy) {
    const double d2 = rx * rx + ry * ry;
    const double d = sqrt(d2);
    if (d > ph.r0 || !(d > 1e-12)) return;
    const double fr = ph.r0 / d;
    const double t1 = ipow(fr, ph.p1), t2 = ipow(fr, ph.p2);
    const double phi = ph.D * (t1 - t2) / d2;
    const double dphi = ph.D / (d2 * d) * (-(ph.p1 + 2.0) * t1 + (ph.p2 + 2.0) * t2);
    const double lr = (lamx * rx + lamy * ry) * dphi / d;
    gx += phi * lamx + lr * rx;
    gy += phi * lamy + lr * ry;
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum NQ values over the workgroup into part[q * R + blockIdx.x].
template <int NQ>
__device__ inline void block_partials(double (&v)[NQ], double *part, int R) {
    __shared__ double lds[NQ * 16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        v[q] = wave_sum(v[q]);
        if (lane == 0) lds[q * 16 + w] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double a = 0.0;
            for (int k = 0; k < nw; ++k) a += lds[q * 16 + k];
            part[q * R + blockIdx.x] = a;
        }
    }
}

// ---------------------------------------------------------------------------
// recording
// ---------------------------------------------------------------------------
__global__ void k_tape_static(State s, int n, int8_t *label) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    label[s.id[t]] = s.label[t];
}

// Argmax of |a_prev|^2 over the fluid at the start of a step, with the
// engine's expression (contraction off): per-workgroup (value, id), then one
// workgroup folds. Ties go to the smaller id.
__device__ inline void amax_merge(double &v, int &i, double v2, int i2) {
    if (v2 > v || (v2 == v && i2 < i)) {
        v = v2;
        i = i2;
    }
}

__global__ void k_tape_amax_final(const double *pv, const int *pi, int R, TapeHdr *hdr) {
    double v = -1.0;
    int id = 0x7fffffff;
    for (int b = threadIdx.x; b < R; b += blockDim.x) amax_merge(v, id, pv[b], pi[b]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) amax_merge(v, id, __shfl_xor(v, o, 64), __shfl_xor(id, o, 64));
    __shared__ double lv[16];
    __shared__ int li[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane == 0) {
        lv[w] = v;
        li[w] = id;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < nw; ++k) amax_merge(v, id, lv[k], li[k]);
        hdr->amax2 = v;
        hdr->amax_id = id;
    }
}

// After the pair pass, everything the reverse step needs: in id order the
// step's start values (velocity and the carried rates, read from the buffer
// before the reorder at the slot's previous position perm[t]) and the pair
// pass's outputs; in the step's cell-slot order the pair pass's inputs as
// records, with each fluid particle's p / rho^2 terms (formed once here, not
// per pair); the walk structures (cell offsets in a grid-stride loop over the
// grid's ncells + 2 entries); dt; and per-workgroup partials of the argmax of
// |a_prev|^2 over the fluid with the engine's expression (contraction off).
// An inactive step is recorded as such only.
// lean: the step's reorder ran lean and the corrector was fused into the
// pair pass, so the fluid's half-step state is not in s: it is formed again
// from the start-of-step values at the particle's previous slot with the
// predictor's expressions (as the reorder formed it for the pair records).
__global__ void k_tape_pair(const Ctrl *ctrl, State s, State o, const int *perm, const int2 *cxy,
                            const int *cellStart, int n, TapeSlot T, TapeHdr *hdr, long long cap, Phys ph,
                            double *pv, int *pi, int lean, int strict) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int act = ctrl->active;
    if (t == 0) {
        hdr->grid = ctrl->grid;
        hdr->active = act;
        hdr->dt = ctrl->dt;
        hdr->dt_c = ctrl->dt_c;
        hdr->dt_f = ctrl->dt_f;
    }
    double v = -1.0;
    int id = 0x7fffffff;
    if (act) {
        const long long m = ctrl->grid.ncells + 2;
        const long long lim = m < cap ? m : cap;
        for (long long k = t; k < lim; k += (long long)gridDim.x * blockDim.x) T.cellStart[k] = cellStart[k];
        if (t < n) {
            const int i = s.id[t];
            const int q = perm[t];
            const double ax = o.ax[q], ay = o.ay[q];
            T.vx0[i] = o.vx[q];
            T.vy0[i] = o.vy[q];
            T.axp[i] = ax;
            T.ayp[i] = ay;
            T.drhop[i] = o.drho[q];
            const int8_t lab = s.label[t];
            const double m = s.m[t], h = s.h[t];
            double x, y, vx, vy, rho;
            if (lean && lab == 0) {
                const double dt = ctrl->dt, damping = ctrl->damping;
                x = pec_half_pos(o.x[q], o.vx[q], dt);
                y = pec_half_pos(o.y[q], o.vy[q], dt);
                vx = pec_half_vel(o.vx[q], ax, dt, damping);
                vy = pec_half_vel(o.vy[q], ay, dt, damping);
                rho = pec_half_rho(o.rho[q], o.drho[q], dt, strict);
            } else {
                x = s.x[t];
                y = s.y[t];
                vx = s.vx[t];
                vy = s.vy[t];
                rho = s.rho[t];
            }
            // PEC's strict clamp (src/Integrators/PEC.py:57-58, 87-88): whether it
            // took the predicted or the corrected density (the adjoint of a
            // clamped density is zero)
            int8_t cl = 0;
            if (strict && lab == 0) {
                const double dt = ctrl->dt, r0 = o.rho[q];
                const double rp = r0 + 0.5 * dt * o.drho[q];
                const double rc = 2 * (r0 + 0.5 * dt * s.drho[t]) - r0;
                cl = (int8_t)((rp < 0.0 ? 1 : 0) | (rc < 0.0 ? 2 : 0));
            }
            T.clamp[i] = cl;
            T.vxh[i] = vx;
            T.vyh[i] = vy;
            T.ax[i] = s.ax[t];
            T.ay[i] = s.ay[t];
            T.drho[i] = s.drho[t];
            T.slot[i] = t;
            T.id[t] = i;
            T.cxy[t] = cxy[t];
            T.pa[t] = make_double4(x, y, h, lab == 0 ? m : -(double)lab);
            T.pb[t] = make_double4(vx, vy, rho, m);
            if (lab == 0) {
                const Side sd = make_side(h, rho, m, ph);
                T.ps[t] = make_double4(sd.slf, sd.dslf, sd.cslf, 0.0);
                amax_merge(v, id, ax * ax + ay * ay, i);
            } else {
                T.ps[t] = make_double4(0.0, 0.0, 0.0, 0.0);
            }
        }
    }
#pragma unroll
    for (int o2 = 32; o2 > 0; o2 >>= 1) amax_merge(v, id, __shfl_xor(v, o2, 64), __shfl_xor(id, o2, 64));
    __shared__ double lv[16];
    __shared__ int li[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane == 0) {
        lv[w] = v;
        li[w] = id;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < nw; ++k) amax_merge(v, id, lv[k], li[k]);
        pv[blockIdx.x] = v;
        pi[blockIdx.x] = id;
    }
}

// ---------------------------------------------------------------------------
// reverse step
// ---------------------------------------------------------------------------
struct Adj {
    double *x, *y, *vx, *vy, *rho, *ax, *ay, *drho;
};

// Corrector (PEC) / Euler update, reversed. Id order, fluid only; the pair
// pass's adjoint inputs (lam, kap) go to the step's cell-slot order.
__global__ void k_adj_correct(const TapeHdr *hdr, TapeSlot T, const int8_t *label, int n, Adj A, double damping,
                              int euler, double4 *lk, double *vbx, double *vby, double *part) {
    if (!hdr->active) return;
    const double dt = hdr->dt, e = 1.0 + 0.5 * damping;
    double v[kAdjQ] = {0.0, 0.0, 0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (label[i] != 0) {
            lk[T.slot[i]] = make_double4(0.0, 0.0, 0.0, 0.0);
            vbx[i] = vby[i] = 0.0;
            A.x[i] = A.y[i] = A.vx[i] = A.vy[i] = A.rho[i] = A.ax[i] = A.ay[i] = A.drho[i] = 0.0;
            continue;
        }
        const double xb = A.x[i], yb = A.y[i], vxb = A.vx[i], vyb = A.vy[i], rb = A.rho[i];
        if (euler) {
            // x' = x + dt v + dt^2/2 a, v' = v + dt a, rho' = rho + dt drho (pair at the start state)
            lk[T.slot[i]] = make_double4(A.ax[i] + dt * vxb + 0.5 * dt * dt * xb, A.ay[i] + dt * vyb + 0.5 * dt * dt * yb,
                                         A.drho[i] + dt * rb, 0.0);
            vbx[i] = 0.0;
            vby[i] = 0.0;
            A.vx[i] = vxb + dt * xb;
            A.vy[i] = vyb + dt * yb;
            const double vx = T.vx0[i], vy = T.vy0[i], ax = T.ax[i], ay = T.ay[i];
            v[0] += xb * (vx + dt * ax) + yb * (vy + dt * ay) + vxb * ax + vyb * ay + rb * T.drho[i];
        } else {
            // a density the strict clamp set to 0 passes no adjoint back
            const bool cc = (T.clamp[i] & 2) != 0;
            const double rbe = cc ? 0.0 : rb;
            lk[T.slot[i]] = make_double4(A.ax[i] + dt / e * vxb, A.ay[i] + dt / e * vyb, A.drho[i] + dt * rbe, 0.0);
            vbx[i] = dt * xb;
            vby[i] = dt * yb;
            A.vx[i] = (2.0 / e - 1.0) * vxb;
            A.vy[i] = (2.0 / e - 1.0) * vyb;
            v[0] += xb * T.vxh[i] + yb * T.vyh[i] + (vxb * T.ax[i] + vyb * T.ay[i]) / e + rbe * T.drho[i];
            if (cc) A.rho[i] = 0.0;
        }
        // x and rho pass through with coefficient 1 (A.x, A.y, A.rho unchanged,
        // but for a clamped density)
    }
    block_partials<kAdjQ>(v, part, gridDim.x);
}

// Pair pass, reversed: gather over the step's cells, kAdjLanes lanes per
// particle (cell order of the step). Lane l of a particle's group takes the
// candidates l, l + kAdjLanes, ... of each row of three cells (consecutive
// bins: one run of slots, ~20 candidates, instead of nine cells of ~7 that
// left most lanes idle); the group then
// folds its sums with a fixed xor-shuffle tree, so the result does not depend
// on scheduling. (A thread per particle leaves most of the chip idle at the
// particle counts the adjoint is used for, and each pair costs two adjoint
// evaluations.)
constexpr int kAdjLanes = 8;

// The group's first lane then reverses the predictor for its particle (PEC;
// for Euler the pair inputs are the start state), which needs nothing of
// any other particle: the half-step adjoint goes straight into A, and the
// predictor's dt term into this workgroup's partial 0.
template <int KERNEL>
__global__ __launch_bounds__(kAdjBlock) void k_adj_pair(const TapeHdr *hdr, TapeSlot T, int n, Phys ph,
                                                         const double4 *lk, Adj A, const double *vbx,
                                                         const double *vby, double damping, int euler,
                                                         double *part) {
    double red[kAdjQ] = {0.0, 0.0, 0.0, 0.0};
    const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int t = (int)(gt / kAdjLanes);  // uniform within a lane group
    const int l = (int)(gt % kAdjLanes);
    const bool live = hdr->active && t < n;
    int k = -1;
    double4 ak = make_double4(0.0, 0.0, 0.0, -1.0);
    if (live) {
        k = T.id[t];
        ak = T.pa[t];
    }
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // xbar, ybar, vxbar, vybar, rhobar
    if (live && ak.w >= 0.0) {
        const Grid g = hdr->grid;
        const int ncx = (int)g.ncx, ncy = (int)g.ncy;
        const double4 bk = T.pb[t], sk = T.ps[t], lmk = lk[t];
        const double xk = ak.x, yk = ak.y, hk = ak.z, vxk = bk.x, vyk = bk.y;
        const Side K{hk, bk.z, bk.w, sk.x, sk.y, sk.z};
        double ga = 0.0, gc = 0.0;
        const int2 c = T.cxy[t];
        const double qmax = ph.qmax;
        const int lo = c.x - 1 > 0 ? c.x - 1 : 0;
        const int hi = c.x + 1 < ncx - 1 ? c.x + 1 : ncx - 1;
        for (int db = -1; db <= 1 && lo <= hi; ++db) {
            const int cy = c.y + db;
            if (!(ncy > cy && cy > -1)) continue;
            {
                const int row = ncx * cy;
                const int s0 = T.cellStart[row + lo], s1 = T.cellStart[row + hi + 1];
                for (int u = s0 + l; u < s1; u += kAdjLanes) {
                    if (u == t) continue;
                    const double4 aj = T.pa[u];
                    const double rx = xk - aj.x, ry = yk - aj.y;
                    const double d = sqrt(rx * rx + ry * ry);
                    const double hh = 0.5 * (hk + aj.z);
                    if (!(d / hh <= qmax)) continue;  // the forward's neighbour rule
                    if (aj.w < 0.0) {
                        wall_vjp(rx, ry, lmk.x, lmk.y, ph, acc[0], acc[1]);
                        continue;
                    }
                    double G, Gp;
                    if (!grad_factor<KERNEL>(d, hh, G, Gp)) continue;  // outside the support
                    const double4 bj = T.pb[u], sj = T.ps[u], lmj = lk[u];
                    const Side J{aj.z, bj.z, bj.w, sj.x, sj.y, sj.z};
                    pair_vjp2(rx, ry, d, G, Gp, vxk - bj.x, vyk - bj.y, K, J, lmk.x, lmk.y, lmk.z, lmj.x, lmj.y,
                              lmj.z, ph, acc, ga, gc);
                }
            }
        }
        red[1] = ga;
        red[2] = gc;
    }
    // fold the group (lanes of one particle are kAdjLanes consecutive lanes)
#pragma unroll
    for (int o = kAdjLanes / 2; o > 0; o >>= 1)
#pragma unroll
        for (int q = 0; q < 5; ++q) acc[q] += __shfl_xor(acc[q], o, 64);
    if (live && l == 0 && ak.w >= 0.0) {
        const double dt = hdr->dt, e = 1.0 + 0.5 * damping;
        // a predicted density the strict clamp set to 0 passes no adjoint back
        const double xhb = acc[0], yhb = acc[1], rhb = (T.clamp[k] & 1) ? 0.0 : acc[4];
        const double vhx = vbx[k] + acc[2], vhy = vby[k] + acc[3];
        if (euler) {
            A.x[k] += xhb;
            A.y[k] += yhb;
            A.vx[k] += vhx;
            A.vy[k] += vhy;
            A.rho[k] += rhb;
            A.ax[k] = 0.0;
            A.ay[k] = 0.0;
            A.drho[k] = 0.0;
        } else {
            A.x[k] += xhb;
            A.y[k] += yhb;
            A.vx[k] += 0.5 * dt * xhb + vhx / e;
            A.vy[k] += 0.5 * dt * yhb + vhy / e;
            A.rho[k] += rhb;
            A.ax[k] = 0.5 * dt * vhx / e;
            A.ay[k] = 0.5 * dt * vhy / e;
            A.drho[k] = 0.5 * dt * rhb;
            red[0] = 0.5 * (xhb * T.vx0[k] + yhb * T.vy0[k] + (vhx * T.axp[k] + vhy * T.ayp[k]) / e +
                            rhb * T.drhop[k]);
        }
    }
    block_partials<kAdjQ>(red, part, gridDim.x);
}

// Fold the partials of the two kernels (the corrector's dt term; the pair
// pass's dL/dalpha, dL/dc0 and the predictor's dt term); add this step's
// dL/dalpha and dL/dc0. One workgroup.
__global__ void k_adj_final(const TapeHdr *hdr, TapeSlot T, Adj A, const double *p1, int R1, const double *p2,
                            int R2, double co, double fixed_dt, double *g_alpha, double *g_co, int *force_steps) {
    if (!hdr->active) return;
    __shared__ double lds[3 * 16];
    double dtb = 0.0, ga = 0.0, gc = 0.0;
    for (int b = threadIdx.x; b < R1; b += blockDim.x) dtb += p1[b];
    for (int b = threadIdx.x; b < R2; b += blockDim.x) {
        dtb += p2[b];
        ga += p2[1 * R2 + b];
        gc += p2[2 * R2 + b];
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    dtb = wave_sum(dtb);
    ga = wave_sum(ga);
    gc = wave_sum(gc);
    if (lane == 0) {
        lds[w] = dtb;
        lds[16 + w] = ga;
        lds[32 + w] = gc;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    dtb = ga = gc = 0.0;
    for (int k = 0; k < nw; ++k) {
        dtb += lds[k];
        ga += lds[16 + k];
        gc += lds[32 + k];
    }
    // dt = min(courant, force) (src/Equations/TimeStep.py:40-45): the Courant
    // value gamma_c h_min / c0 is taken unless force < courant.
    if (!(fixed_dt > 0.0)) {
        if (!(hdr->dt_f < hdr->dt_c)) {
            gc += dtb * (-hdr->dt / co);
        } else if (hdr->amax2 >= 1e-12 && hdr->amax_id >= 0 && hdr->amax_id < 0x7fffffff) {
            // dt = gamma_f sqrt(h_min / a2max): d dt / d a_k = -dt a_k / a2max
            const int k = hdr->amax_id;
            const double coef = -dtb * hdr->dt / hdr->amax2;
            A.ax[k] += coef * T.axp[k];
            A.ay[k] += coef * T.ayp[k];
            atomicAdd(force_steps, 1);
        }
    }
    g_alpha[0] += ga;
    g_co[0] += gc;
}

inline int nblk_(long long n, int bs) {
    long long b = (n + bs - 1) / bs;
    return (int)(b < 1 ? 1 : b);
}
inline int redblk_(long long n) {
    int b = nblk_(n, kAdjBlock);
    return b > kRedBlocks ? kRedBlocks : b;
}

}  // namespace

// ---------------------------------------------------------------------------
// hooks for sph_step_record
// ---------------------------------------------------------------------------
int hipsph::tape_record_pair(sph_ctx *c, sph_tape *t, int slot, hipStream_t s, bool lean) {
    const int R = nblk_(c->n, kAdjBlock);
    k_tape_pair<<<R, kAdjBlock, 0, s>>>(c->ctrl, c->buf[c->cur], c->buf[1 - c->cur], c->perm, c->cxy_s,
                                        c->cellStart, c->n, t->slot[slot], t->hdr + slot, t->cells,
                                        make_phys(c->P), t->amax_v, t->amax_i, lean,
                                        c->P.strict && c->P.integrator == SPH_INTEGRATOR_PEC);
    k_tape_amax_final<<<1, 1024, 0, s>>>(t->amax_v, t->amax_i, R, t->hdr + slot);
    HIPCHK(hipGetLastError());
    return 0;
}

// The forms the adjoint covers; the parameters may change after the tape is
// made, so recording checks again.
int hipsph::tape_supported(const sph_ctx *c) {
    if (c->P.use_xsph || c->P.use_summation_density || c->P.dynamic_h || c->P.coupled_integrator)
        return fail(SPH_EUNSUPPORTED,
                    "the adjoint covers the plain WCSPH step (no XSPH, summation density, dynamic h or coupled integrator)");
    return 0;
}

// The fluid mask by reference index, from the current state: taken when a
// recording starts (slot 0), so a tape made before the state was uploaded,
// or kept across uploads, sees the labels of the steps it records.
int hipsph::tape_refresh_labels(sph_ctx *c, sph_tape *t, hipStream_t s) {
    k_tape_static<<<nblk_(c->n, kAdjBlock), kAdjBlock, 0, s>>>(c->buf[c->cur], c->n, t->label);
    HIPCHK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int sph_tape_create(sph_ctx_t *ctx, int32_t nslots, sph_tape_t **out) {
    if (!ctx || !out) return fail(SPH_EINVAL, "NULL argument");
    *out = nullptr;
    if (nslots < 1 || nslots > 100000) return fail(SPH_EINVAL, "nslots=%d out of range", nslots);
    int rc = tape_supported(ctx);
    if (rc) return rc;
    const size_t n = (size_t)ctx->n;
    sph_tape *t = new (std::nothrow) sph_tape;
    if (!t) return fail(SPH_ENOMEM, "host allocation failed");
    t->ctx = ctx;
    t->nslots = nslots;
    t->cells = ctx->cap_cells + 2;
    // per slot: 10 double arrays, 3 double4 arrays, slot, clamp, id, cxy, cellStart
    // (each carved array is rounded up to 256 bytes: 20 per slot, 20 fixed)
    const size_t per = n * (10 * sizeof(double) + 3 * sizeof(double4) + 2 * sizeof(int) + sizeof(int2) + 1) +
                       (size_t)t->cells * sizeof(int) + 20 * 256;
    const size_t npart = (size_t)kAdjQ * (kRedBlocks + nblk_((long long)ctx->n * kAdjLanes, kAdjBlock));
    const size_t fixed = n + n * 2 * sizeof(double) + n * sizeof(double4) + npart * sizeof(double) +
                         sizeof(TapeHdr) * nslots + nblk_(ctx->n, kAdjBlock) * (sizeof(double) + sizeof(int)) + 20 * 256 + 4096;
    if (hipMalloc(&t->arena, per * nslots + fixed) != hipSuccess) {
        delete t;
        return fail(SPH_ENOMEM, "tape allocation of %zu bytes failed", per * nslots + fixed);
    }
    char *p = (char *)t->arena;
    char *const end = p + per * nslots + fixed;
    auto take = [&](size_t bytes) {
        char *q = p;
        p += (bytes + 255) & ~(size_t)255;
        return (void *)q;
    };
    t->hdr = (TapeHdr *)take(sizeof(TapeHdr) * nslots);
    t->part = (double *)take(npart * sizeof(double));
    t->force_steps = (int *)take(sizeof(int));
    t->amax_v = (double *)take(nblk_(ctx->n, kAdjBlock) * sizeof(double));
    t->amax_i = (int *)take(nblk_(ctx->n, kAdjBlock) * sizeof(int));
    t->label = (int8_t *)take(n);
    double **scr[] = {&t->vbx, &t->vby};
    for (double **d : scr) *d = (double *)take(n * sizeof(double));
    t->lk = (double4 *)take(n * sizeof(double4));
    t->slot.resize(nslots);
    t->rec_exec.assign(2 * (size_t)nslots, nullptr);
    t->adj_exec.assign(nslots, nullptr);
    t->gen = ctx->gen;
    for (int k = 0; k < nslots; ++k) {
        TapeSlot &S = t->slot[k];
        double **f[] = {&S.vxh, &S.vyh, &S.ax, &S.ay, &S.drho, &S.vx0, &S.vy0, &S.axp, &S.ayp, &S.drhop};
        for (double **d : f) *d = (double *)take(n * sizeof(double));
        double4 **f4[] = {&S.pa, &S.pb, &S.ps};
        for (double4 **d : f4) *d = (double4 *)take(n * sizeof(double4));
        S.slot = (int *)take(n * sizeof(int));
        S.clamp = (int8_t *)take(n);
        S.id = (int *)take(n * sizeof(int));
        S.cxy = (int2 *)take(n * sizeof(int2));
        S.cellStart = (int *)take((size_t)t->cells * sizeof(int));
    }
    if (p > end) {  // sizing above must cover every carved array
        hipFree(t->arena);
        delete t;
        return fail(SPH_EINVAL, "internal: tape arena too small");
    }
    if (hipMemset(t->hdr, 0, sizeof(TapeHdr) * nslots) != hipSuccess ||
        hipMemset(t->force_steps, 0, sizeof(int)) != hipSuccess || hipMemset(t->label, 0, n) != hipSuccess) {
        hipFree(t->arena);
        delete t;
        return fail(SPH_EHIP, "tape init failed");
    }
    // the labels of the current state, if there is one (a recording from
    // slot 0 takes them again)
    if (ctx->loaded) k_tape_static<<<nblk_(ctx->n, kAdjBlock), kAdjBlock>>>(ctx->buf[ctx->cur], ctx->n, t->label);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        hipFree(t->arena);
        delete t;
        return fail(SPH_EHIP, "tape init kernel failed");
    }
    *out = t;
    return 0;
}

int sph_tape_destroy(sph_tape_t *tape) {
    if (!tape) return 0;
    hipDeviceSynchronize();
    tape_drop_graphs(tape);
    hipFree(tape->arena);
    delete tape;
    return 0;
}

int sph_tape_force_steps(sph_tape_t *tape, int64_t *out) {
    if (!tape || !out) return fail(SPH_EINVAL, "NULL argument");
    int v = 0;
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(&v, tape->force_steps, sizeof(int), hipMemcpyDeviceToHost));
    *out = v;
    return 0;
}

int sph_adjoint_step(sph_ctx_t *ctx, sph_tape_t *tape, int32_t slot, const sph_adjoint_t *adj, void *stream) {
    if (!ctx || !tape || !adj) return fail(SPH_EINVAL, "NULL argument");
    if (tape->ctx != ctx) return fail(SPH_EINVAL, "tape belongs to another context");
    if (slot < 0 || slot >= tape->nslots) return fail(SPH_EINVAL, "slot %d out of range", slot);
    if (!adj->x || !adj->y || !adj->vx || !adj->vy || !adj->rho || !adj->ax || !adj->ay || !adj->drho ||
        !adj->g_alpha || !adj->g_co)
        return fail(SPH_EINVAL, "every sph_adjoint_t pointer is required");
    return adjoint_launch(ctx, tape, slot, adj, (hipStream_t)stream);
}

}  // extern "C"

void hipsph::tape_drop_graphs(sph_tape *t) {
    for (auto &g : t->rec_exec)
        if (g) hipGraphExecDestroy(g), g = nullptr;
    for (auto &g : t->adj_exec)
        if (g) hipGraphExecDestroy(g), g = nullptr;
}

int hipsph::adjoint_launch(sph_ctx *ctx, sph_tape *tape, int slot, const sph_adjoint_t *adj, hipStream_t s) {
    const int n = ctx->n;
    const int R = redblk_(n);
    const int Rp = nblk_((long long)n * kAdjLanes, kAdjBlock);  // the pair kernel: one partial per workgroup
    double *p1 = tape->part, *p2 = tape->part + kAdjQ * kRedBlocks;
    const TapeSlot &T = tape->slot[slot];
    const TapeHdr *H = tape->hdr + slot;
    Adj A{adj->x, adj->y, adj->vx, adj->vy, adj->rho, adj->ax, adj->ay, adj->drho};
    const Phys ph = make_phys(ctx->P);
    const int euler = ctx->P.integrator == SPH_INTEGRATOR_EULER;
    k_adj_correct<<<R, kAdjBlock, 0, s>>>(H, T, tape->label, n, A, ctx->P.damping, euler, tape->lk, tape->vbx,
                                           tape->vby, p1);
    const double dmp = ctx->P.damping;
    switch (ctx->P.kernel) {
        case SPH_KERNEL_GAUSSIAN:
            k_adj_pair<SPH_KERNEL_GAUSSIAN><<<Rp, kAdjBlock, 0, s>>>(H, T, n, ph, tape->lk, A, tape->vbx, tape->vby,
                                                                     dmp, euler, p2);
            break;
        case SPH_KERNEL_WENDLAND:
            k_adj_pair<SPH_KERNEL_WENDLAND><<<Rp, kAdjBlock, 0, s>>>(H, T, n, ph, tape->lk, A, tape->vbx, tape->vby,
                                                                     dmp, euler, p2);
            break;
        default:
            k_adj_pair<SPH_KERNEL_CUBIC><<<Rp, kAdjBlock, 0, s>>>(H, T, n, ph, tape->lk, A, tape->vbx, tape->vby,
                                                                  dmp, euler, p2);
    }
    k_adj_final<<<1, 1024, 0, s>>>(H, T, A, p1, R, p2, Rp, ctx->P.co, ctx->P.fixed_dt, adj->g_alpha, adj->g_co,
                                   tape->force_steps);
    HIPCHK(hipGetLastError());
    return 0;
}
