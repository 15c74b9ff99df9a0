// sph_batch.hip — many independent WCSPH simulations advanced together.
//
// A batch holds K instances of uneven particle counts (an ensemble of dam
// breaks with different geometry, resolution and constants). All instances
// share one concatenated, cell-ordered state and one segmented cell grid, and
// every stage of the step is ONE launch covering all instances:
//   * instance k owns particle slots [off_k, off_k + n_k) for good (its
//     particles only ever move among its own slots) and the bins
//     [cellBase_k, cellBase_k + ncells_k] of the grid, rebased every step by
//     a scan over the instances of (ncells_k + 1);
//   * reductions (bounds, dt) run over fixed chunks of one instance each and a
//     workgroup per instance folds them;
//   * keys, the prefix scan, scatter, bin sort, reorder and the pair pass are
//     the single-simulation kernels with a per-particle instance lookup
//     (the scan/scatter/sort kernels are literally shared, over the union of
//     the instances' bins);
//   * every instance has its own control block (BInst): clock, dt, grid,
//     halt flags and constants (alpha, c0, B, r0, D ...).
// A step is captured in a hipGraph like the single simulation's. The
// arithmetic per particle is the single engine's, in the same order, so an
// instance evolves exactly as it would alone.
//
// Built with -ffp-contract=off (as sph_engine.hip).
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <vector>

#include "sph_ctx.h"
#include "sph_common.h"

using namespace hipsph;

namespace {

// Two segmentations of the slots, each segment inside one instance: chunks
// of kChunk for the bounds reduction, tiles of kTile for the dt reduction,
// the corrector and the uniform pair pass with the corrector fused in (one
// 256-thread workgroup per tile).
constexpr int kChunk = 2048;
constexpr int kTile = 256;
static_assert(kTile == 256, "the fused batch pair pass runs one 256-thread workgroup per tile");

struct BState {  // the concatenated cell-ordered state (as State)
    State s;
};

// ---------------------------------------------------------------------------
// per-instance dt (src/Solver.py:210-230, src/Equations/TimeStep.py:40-90)
// ---------------------------------------------------------------------------
// segment c (a tile) covers [cs[c], cs[c + 1]) inside one instance
__global__ void kb_dt_partial(const int *cs, const int *segInst, const BInst *inst, State s, double *part,
                              Ctrl *view) {
    __shared__ double lds[kDtQ * 16];
    const int c = blockIdx.x, NC = gridDim.x;
    if (c == 0 && threadIdx.x == 0) view->pec_parts = 0;
    const BInst &I = inst[segInst[c]];
    DtVars v;
    for (int t = cs[c] + threadIdx.x; t < cs[c + 1]; t += blockDim.x)
        if (s.label[t] == kFluid) v.add(I.uh_on ? I.uh : s.h[t], s.ax[t], s.ay[t]);
    v.store(part, NC, c, lds);
}

// One workgroup per instance: fold its chunks, open or close its step, and
// carry the instance's bounds of the border-cell path as k_dt_final does
// (view->pec_parts: the partials come from the fused PEC pass).
__global__ void kb_dt_final(BInst *inst, const int *chunk0, const double *part, int NC, const Ctrl *view) {
    __shared__ double lds[4 * 16];
    BInst &I = inst[blockIdx.x];
    const int c0 = chunk0[blockIdx.x], c1 = chunk0[blockIdx.x + 1];
    const int kind[4] = {0, 1, 2, 2};
    double v[4] = {INFINITY, -INFINITY, 0.0, 0.0};
    for (int c = c0 + threadIdx.x; c < c1; c += blockDim.x) {
        const double a = part[0 * NC + c], b = part[1 * NC + c];
        v[0] = a < v[0] ? a : v[0];
        v[1] = b > v[1] ? b : v[1];
        v[2] += part[2 * NC + c];
        v[3] += part[3 * NC + c];
    }
    block_fold<4>(v, kind, lds);
    if (threadIdx.x != 0) return;
    const int pec = view->pec_parts;
    const double dt0 = I.dt, V = I.vbound, A0 = I.abound;
    I.disp = I.vbound = I.abound = INFINITY;
    const double mn = v[0], mx = v[1], cnt = v[2], bad = v[3];
    const unsigned halt = SPH_FLAG_CAPACITY | SPH_FLAG_NONFINITE | SPH_FLAG_NO_FLUID | SPH_FLAG_ZERO_DENSITY;
    if ((I.flags & halt) || !(I.t < I.duration)) {
        I.active = 0;
        return;
    }
    if (cnt == 0) {
        I.flags |= SPH_FLAG_NO_FLUID;
        I.active = 0;
        return;
    }
    if (bad != 0) I.flags |= SPH_FLAG_NONFINITE;
    double dt, c, f;
    if (I.fixed_dt > 0.0) {
        dt = I.fixed_dt;
        c = 0.0;
        f = 0.0;
    } else {
        c = I.gamma_c * mn / I.co;
        f = (mx < 1e-12) ? 1e10 : I.gamma_f * sqrt(mn / mx);
        dt = (f < c) ? f : c;
        if (dt < 1e-6) I.flags |= SPH_FLAG_DT_SMALL;
    }
    I.dt = dt;
    I.dt_c = c;
    I.dt_f = f;
    if (I.advance_t) I.t = I.t + dt;
    I.step = I.step + 1;
    I.active = 1;
    if (bad == 0) {
        const double A1 = sqrt(mx);
        I.abound = A1;
        if (pec) {
            const double V1 = V + dt0 * A1;
            const double d = dt0 * (V + 0.5 * dt0 * A0) + 0.5 * dt0 * V + 0.5 * dt * V1;
            if (d == d) I.disp = d;
            I.vbound = V1;
        }
    }
}

// ---------------------------------------------------------------------------
// neighbour update over the segmented grid (src/Tools/NNLinkedList.py)
// ---------------------------------------------------------------------------
__device__ inline void bupd_pos(const State &a, int i, bool pred, double dt, double &x, double &y) {
    x = a.x[i];
    y = a.y[i];
    if (pred) {
        // the velocity is loaded with the label, not after it, and selected
        const double vx = a.vx[i], vy = a.vy[i];
        const bool f = a.label[i] == kFluid;
        x = f ? pec_half_pos(x, vx, dt) : x;
        y = f ? pec_half_pos(y, vy, dt) : y;
    }
}

// Per chunk of 2048 slots (one instance's): the bounds' partials. With PEC's
// predictor the full scan also finds the largest |vx|, |vy| (kb_grid keeps it
// as the instance's vbound). With cellStart (the fused PEC step) and an
// instance whose carried bound allows it (as the single engine's k_bounds),
// the instance's chunks share the walk of its border cells instead.
__global__ void kb_bounds(const int *cs, const int *chunkInst, const int *chunk0, BInst *inst, State a, int pec,
                          double *part, const int *cellStart, Ctrl *view) {
    __shared__ double lds[kBoundsVQ * 16];
    const int c = blockIdx.x, NC = gridDim.x;
    if (c == 0 && threadIdx.x == 0) view->pec_parts = 0;  // kb_dt_final has read it
    const int k = chunkInst[c];
    BInst &I = inst[k];
    const bool pred = pec && I.active;
    double xmn = INFINITY, xmx = -INFINITY, ymn = INFINITY, ymx = -INFINITY, hmn = INFINITY, hmx = -INFINITY;
    double bad = 0.0, vmx = -INFINITY;
    auto take = [&](int i) {
        double xi, yi;
        bupd_pos(a, i, pred, I.dt, xi, yi);
        const double hi = I.uh_on ? I.uh : a.h[i];
        if (!(xi - xi == 0.0) || !(yi - yi == 0.0)) bad += 1.0;
        xmn = xi < xmn ? xi : xmn;
        xmx = xi > xmx ? xi : xmx;
        ymn = yi < ymn ? yi : ymn;
        ymx = yi > ymx ? yi : ymx;
        hmn = hi < hmn ? hi : hmn;
        hmx = hi > hmx ? hi : hmx;
    };
    if (cellStart && pred && I.uh_on && border_margin_ok(I.grid, I.disp)) {
        const int j = c - chunk0[k], nck = chunk0[k + 1] - chunk0[k];
        if (j == 0 && threadIdx.x == 0) I.border_updates += 1;
        hmn = hmx = I.uh;
        border_walk(cellStart, I.cellBase, I.grid, j * (int)blockDim.x + (int)threadIdx.x, nck * (int)blockDim.x,
                    take);
    } else {
        for (int i = cs[c] + threadIdx.x; i < cs[c + 1]; i += blockDim.x) {
            take(i);
            if (pred) {
                const double ux = fabs(a.vx[i]), uy = fabs(a.vy[i]);
                const double w = (isfinite(ux) && isfinite(uy)) ? (ux > uy ? ux : uy) : INFINITY;
                vmx = w > vmx ? w : vmx;
            }
        }
    }
    const int kind[kBoundsVQ] = {0, 1, 0, 1, 0, 1, 2, 1};
    double v[kBoundsVQ] = {xmn, xmx, ymn, ymx, hmn, hmx, bad, vmx};
    block_partials<kBoundsVQ>(v, kind, lds, part, NC, c);
}

// One workgroup per instance: its grid (_get_cell_size :103-110,
// _get_number_of_cells :112-127). A finished or halted instance gets
// ncells = 0 (its particles park in its one bin).
__global__ void kb_grid(BInst *inst, const int *chunk0, const double *part, int NC) {
    __shared__ double lds[kBoundsVQ * 16];
    BInst &I = inst[blockIdx.x];
    const int c0 = chunk0[blockIdx.x], c1 = chunk0[blockIdx.x + 1];
    const int kind[kBoundsVQ] = {0, 1, 0, 1, 0, 1, 2, 1};
    double v[kBoundsVQ] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY, 0.0, -INFINITY};
    for (int c = c0 + threadIdx.x; c < c1; c += blockDim.x) {
#pragma unroll
        for (int q = 0; q < kBoundsVQ; ++q) {
            const double u = part[q * NC + c];
            v[q] = kind[q] == 0 ? (u < v[q] ? u : v[q]) : kind[q] == 1 ? (u > v[q] ? u : v[q]) : v[q] + u;
        }
    }
    block_fold<kBoundsVQ>(v, kind, lds);
    if (threadIdx.x != 0) return;
    if (I.active && v[7] >= 0.0) I.vbound = v[7];  // a full scan's velocity bound
    Grid g = I.grid;
    g.ncells = 0;
    if (I.active) {
        const double xmn = v[0], xmx = v[1], ymn = v[2], ymx = v[3], hmn = v[4], hmx = v[5], bad = v[6];
        double cell_size = hmn * I.scale;
        if (cell_size < 1e-6) cell_size = 1.0;
        const double cell_size1 = 1. / cell_size;
        const double fx = ceil(cell_size1 * (xmx - xmn));
        const double fy = ceil(cell_size1 * (ymx - ymn));
        if (bad != 0.0 || !(fx < 2.0e9) || !(fy < 2.0e9)) {
            I.flags |= SPH_FLAG_NONFINITE;
            I.active = 0;
        } else {
            long long ncx = (long long)fx, ncy = (long long)fy;
            if (ncx < 1) ncx = 1;
            if (ncy < 1) ncy = 1;
            if (ncx * ncy > I.cap) {
                I.flags |= SPH_FLAG_CAPACITY;
                I.active = 0;
            } else {
                g.xmin = xmn;
                g.xmax = xmx;
                g.ymin = ymn;
                g.ymax = ymx;
                g.cs = cell_size;
                g.uniform_h = hmn == hmx;
                g.ncx = ncx;
                g.ncy = ncy;
                g.ncells = ncx * ncy;
            }
        }
    }
    I.grid = g;
}

// Rebase the instances' bins (exclusive scan of ncells + 1) and publish the
// union as one grid for the shared scan / scatter / bin-sort kernels.
__global__ void kb_bases(BInst *inst, int K, Ctrl *view, long long capBins) {
    __shared__ long long lds[16];
    long long carry = 0;
    for (int b0 = 0; b0 < K; b0 += blockDim.x) {
        const int k = b0 + threadIdx.x;
        long long v = k < K ? inst[k].grid.ncells + 1 : 0;
        // inclusive scan in the wave, then across waves
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
        long long sc = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const long long u = __shfl_up(sc, o, 64);
            if (lane >= o) sc += u;
        }
        if (lane == 63) lds[w] = sc;
        __syncthreads();
        long long off = 0, tot = 0;
        for (int q = 0; q < nw; ++q) {
            if (q < w) off += lds[q];
            tot += lds[q];
        }
        __syncthreads();
        if (k < K) inst[k].cellBase = carry + off + sc - v;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        view->active = carry <= capBins;
        if (carry > capBins) {  // cannot happen with per-instance caps summed into capBins; kept as a guard
            for (int k = 0; k < K; ++k) {
                inst[k].flags |= SPH_FLAG_CAPACITY;
                inst[k].active = 0;
            }
        }
        view->grid.ncells = carry - 1;
    }
}

// _bin :129-141 with the instance's grid; rank by runs as k_keys does.
__global__ void kb_keys(BInst *inst, const int *instOf, State a, int n, int pec, int *counts, int *bin,
                        int *rank, int *binflag, const Ctrl *view) {
    if (!view->active) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int b = -1;
    if (t < n) {
        BInst &I = inst[instOf[t]];
        const Grid g = I.grid;
        if (I.active) {
            double x, y;
            bupd_pos(a, t, pec != 0, I.dt, x, y);
            const int2 c = cell_of(x, y, g);
            const long long key = (long long)c.x + g.ncx * (long long)c.y;
            if (key >= 0 && key < g.ncells) {
                b = (int)(I.cellBase + key);
            } else {  // as k_keys: the overflow bin, and the instance's flag
                b = (int)(I.cellBase + g.ncells);
                atomicOr(&I.flags, (unsigned)SPH_FLAG_OOB_KEY);
            }
        } else {
            b = (int)I.cellBase;  // parked: the instance's single bin
        }
    }
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
        if (base != 0) binflag[b] = 1;
    }
    base = __shfl(base, leader, 64);
    if (t < n) {
        bin[t] = b;
        rank[t] = base + (lane - leader);
    }
}

// Permute the state into the new order with the PEC predictor of active
// instances fused in, and write the pair records (as k_reorder). With PEC the
// reorder is lean as k_reorder's for the fluid of active instances: no
// x0 ... rho0 (kb_correct gathers them from the source buffer), no half-step
// x, y, vx, vy, rho (kb_correct forms the half-step velocity again and
// overwrites the rest), no rates of a particle inside its instance's grid
// (written by the pair pass).
__global__ void kb_reorder(BInst *inst, const int *instOf, State a, State b, const int *perm, int2 *cxy_s,
                           int n, int fmt, void *recA, void *recB, int pec) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int k = instOf[t];  // slots never change instance
    const BInst &I = inst[k];
    const int s = perm[t];
    const bool hs = I.uh_on;  // m and h stay in place (and m is not read: the uniform records have none)
    double x = a.x[s], y = a.y[s], vx = a.vx[s], vy = a.vy[s], rho = a.rho[s];
    const double m = hs ? 0.0 : a.m[s], h = hs ? I.uh : a.h[s];
    const double ax = a.ax[s], ay = a.ay[s], drho = a.drho[s];
    const int8_t lab = a.label[s];
    const bool skip = I.active && pec && lab == kFluid;
    if (skip) {
        const double dt = I.dt;
        x = pec_half_pos(x, vx, dt);
        y = pec_half_pos(y, vy, dt);
        vx = pec_half_vel(vx, ax, dt, I.damping);
        vy = pec_half_vel(vy, ay, dt, I.damping);
        rho = pec_half_rho(rho, drho, dt, I.strict);
    }
    // x0 ... rho0 are not kept: nothing in a batch step reads them
    // the raw cell the keys used (parked particles of an inactive instance: 0, 0)
    const Grid &g = I.grid;
    const int2 cc = I.active ? cell_of(x, y, g) : make_int2(0, 0);
    if (!skip) {
        b.x[t] = x;
        b.y[t] = y;
        b.vx[t] = vx;
        b.vy[t] = vy;
        b.rho[t] = rho;
    }
    if (!hs) {
        b.m[t] = m;
        b.h[t] = h;
    }
    b.label[t] = lab;
    b.id[t] = a.id[s];
    if (!skip || !(cc.x >= 0 && cc.x < g.ncx && cc.y >= 0 && cc.y < g.ncy && h > 0.0 && I.qmax >= 0.0)) {
        b.ax[t] = ax;
        b.ay[t] = ay;
        b.drho[t] = drho;
    }
    cxy_s[t] = cc;
    const double pr = write_pair_records(t, x, y, vx, vy, rho, m, h, lab, inst_phys(I), fmt, recA, recB);
    if (I.active && lab == kFluid && !isfinite(pr)) atomicOr(&inst[k].flags, (unsigned)SPH_FLAG_ZERO_DENSITY);
}

// Per instance (one workgroup each): min and max h over its particles and
// min and max m over its fluid, in out[4k .. 4k+3].
__global__ void kb_hm_range(const BInst *inst, State st, double *out) {
    __shared__ double lds[4 * 16];
    const int k = blockIdx.x, s0 = inst[k].off, s1 = s0 + inst[k].n;
    double v[4] = {INFINITY, -INFINITY, INFINITY, -INFINITY};
    for (int i = s0 + threadIdx.x; i < s1; i += blockDim.x) {
        const double h = st.h[i];
        v[0] = h < v[0] ? h : v[0];
        v[1] = h > v[1] ? h : v[1];
        if (st.label[i] == kFluid) {
            const double m = st.m[i];
            v[2] = m < v[2] ? m : v[2];
            v[3] = m > v[3] ? m : v[3];
        }
    }
    const int kind[4] = {0, 1, 0, 1};
    block_fold<4>(v, kind, lds);
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; ++q) out[4 * k + q] = v[q];
}

// PEC.correct / Euler (fluid of active instances) + the next dt partials.
// PEC's start-of-step values come from the buffer before the reorder, at the
// particle's previous slot perm[j] (the lean reorder did not copy them).
__global__ void kb_correct(const int *cs, const int *chunkInst, const BInst *inst, State s, State o,
                           const int *perm, int euler, double *part, Ctrl *view) {
    __shared__ double lds[kDtQ * 16];
    const int c = blockIdx.x, NC = gridDim.x;
    if (c == 0 && threadIdx.x == 0) view->pec_parts = 0;
    const BInst &I = inst[chunkInst[c]];
    const double dt = I.dt, damping = I.damping;
    DtVars v;
    for (int j = cs[c] + threadIdx.x; j < cs[c + 1]; j += blockDim.x) {
        if (s.label[j] != kFluid) continue;
        const double ax = s.ax[j], ay = s.ay[j];
        v.add(I.uh_on ? I.uh : s.h[j], ax, ay);
        if (!I.active) continue;
        if (!euler) {
            const int q = perm[j];
            const double x0 = o.x[q], y0 = o.y[q], vx0 = o.vx[q], vy0 = o.vy[q], rho0 = o.rho[q];
            const double vxh = pec_half_vel(vx0, o.ax[q], dt, damping);  // as kb_reorder formed it
            const double vyh = pec_half_vel(vy0, o.ay[q], dt, damping);
            double mid_x = x0 + 0.5 * dt * vxh;
            double mid_y = y0 + 0.5 * dt * vyh;
            double mid_vx = (vx0 + 0.5 * dt * ax) / (1 + 0.5 * damping);
            double mid_vy = (vy0 + 0.5 * dt * ay) / (1 + 0.5 * damping);
            s.x[j] = 2 * mid_x - x0;
            s.y[j] = 2 * mid_y - y0;
            s.vx[j] = 2 * mid_vx - vx0;
            s.vy[j] = 2 * mid_vy - vy0;
            double mid_rho = rho0 + 0.5 * dt * s.drho[j];
            double r = 2 * mid_rho - rho0;
            if (I.strict && r < 0.0) r = 0.0;
            s.rho[j] = r;
        } else {
            double vx = s.vx[j], vy = s.vy[j];
            s.x[j] = s.x[j] + dt * vx + 0.5 * dt * dt * ax;
            s.y[j] = s.y[j] + dt * vy + 0.5 * dt * dt * ay;
            s.vx[j] = vx + dt * ax;
            s.vy[j] = vy + dt * ay;
            s.rho[j] = s.rho[j] + dt * s.drho[j];
        }
    }
    v.store(part, NC, c, lds);
}

__global__ void kb_upload(State d, sph_state_t in, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
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
    d.id[i] = i;
}

// m0, h0: the upload's buffer, where an instance with uh_on keeps m and h at
// slot = id.
__global__ void kb_download(State s, sph_state_t out, const BInst *inst, const int *instOf, int n,
                            const double *m0, const double *h0) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = s.id[t];
    const BInst &I = inst[instOf[t]];
    if (out.x) out.x[i] = s.x[t];
    if (out.y) out.y[i] = s.y[t];
    if (out.vx) out.vx[i] = s.vx[t];
    if (out.vy) out.vy[i] = s.vy[t];
    if (out.rho) out.rho[i] = s.rho[t];
    if (out.ax) out.ax[i] = s.ax[t];
    if (out.ay) out.ay[i] = s.ay[t];
    if (out.drho) out.drho[i] = s.drho[t];
    if (out.m) out.m[i] = I.uh_on ? m0[i] : s.m[t];
    if (out.h) out.h[i] = I.uh_on ? h0[i] : s.h[t];
    if (out.label) out.label[i] = s.label[t];
    if (out.p) out.p[i] = tait_p(s.rho[t], s.label[t], I.gamma, I.B, I.rho0, I.Pb);
}

inline int nblk_b(long long n, int bs = kBlock) {
    long long b = (n + bs - 1) / bs;
    return (int)(b < 1 ? 1 : b);
}

}  // namespace

// ---------------------------------------------------------------------------
// the batch object
// ---------------------------------------------------------------------------
struct sph_batch {
    int K = 0, n = 0, NC = 0;
    int kernel = 0, precision = 0, integrator = 0;
    std::vector<sph_params_t> P;
    State buf[2];
    void *arena[2] = {nullptr, nullptr};
    int cur = 0;
    BInst *inst = nullptr;
    Ctrl *view = nullptr;
    int *instOf = nullptr, *cs = nullptr, *chunkInst = nullptr, *chunk0 = nullptr;
    int NT = 0;  // tiles: starts ts (NT + 1), instance of each, first tile of each instance (K + 1)
    int *ts = nullptr, *tileInst = nullptr, *tile0 = nullptr;
    int *bin = nullptr, *rank = nullptr, *perm = nullptr, *permKey = nullptr;
    int2 *cxy_s = nullptr;
    int *counts = nullptr, *cellStart = nullptr, *tileSums = nullptr, *binflag = nullptr;
    double *part_b = nullptr, *part_d = nullptr;
    void *recA = nullptr, *recB = nullptr;
    long long capBins = 0;
    bool dt_fresh = false;
    bool loaded = false;  // a state was uploaded: every id indexes its instance
    hipStream_t own = nullptr;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int pair_threads = 256;
    int list_len = 0;  // 0: the single-simulation pass's length
    int persistent = 0;    // the pair pass as a persistent work-queue kernel
    int border_bounds = 1;  // the fused PEC step's bounds from the border cells when the carried bound allows
    // every instance has one h for all its particles and one fluid mass and
    // allows the uniform pass (found at upload): records kRecUniform(64) and
    // the uniform pair pass, each instance's constants in uniK
    bool uni = false;
    void *uniK = nullptr;
    double *hm = nullptr;  // per-instance h and fluid-mass ranges (4 per instance)
    int *queue = nullptr;  // its counters (9 x 32 ints), zero between launches
    int cus = 0;           // compute units of the batch's device
    int queue_grid = 0;    // its grid (0: to be sized before the next launch)
    bool timing = false;
    std::vector<hipEvent_t> ev;
    size_t n_ev = 0;
    std::vector<int> offs;
};

namespace {

void free_batch(sph_batch *b) {
    if (!b) return;
    for (int k = 0; k < 2; ++k) {
        if (b->arena[k]) hipFree(b->arena[k]);
        if (b->gexec[k]) hipGraphExecDestroy(b->gexec[k]);
    }
    void *ptrs[] = {b->inst, b->view, b->instOf, b->cs, b->chunkInst, b->chunk0, b->ts, b->tileInst, b->tile0,
                    b->bin, b->rank, b->perm,
                    b->permKey, b->cxy_s, b->counts, b->cellStart, b->tileSums, b->binflag,
                    b->part_b, b->part_d, b->recA, b->recB, b->queue, b->uniK, b->hm};
    for (void *p : ptrs)
        if (p) hipFree(p);
    for (auto e : b->ev) hipEventDestroy(e);
    if (b->own) hipStreamDestroy(b->own);
    delete b;
}

// The persistent pair pass's grid: one resident occupancy of the device
// (hipOccupancy for the chosen kernel, list and workgroup size). Host only;
// sized before a step is launched or captured.
void size_queue(sph_batch *b) {
    if (!b->persistent || b->queue_grid > 0) return;
    const int per = b->uni ? (b->precision == SPH_PREC_FP32C ? pair_batch_uniform_occupancy_fp32(b->kernel)
                                                             : pair_batch_uniform_occupancy_fp64(b->kernel))
                    : b->precision == SPH_PREC_FP32C
                        ? pair_batch_queue_occupancy_fp32c(b->kernel, b->list_len, b->pair_threads)
                        : pair_batch_queue_occupancy_fp64(b->kernel, b->list_len, b->pair_threads);
    b->queue_grid = (b->cus > 0 ? b->cus : kXcds) * per;
}

int b_dt(sph_batch *b, hipStream_t s) {
    if (!b->dt_fresh)
        kb_dt_partial<<<b->NT, kBlock, 0, s>>>(b->ts, b->tileInst, b->inst, b->buf[b->cur], b->part_d, b->view);
    kb_dt_final<<<b->K, kBlock, 0, s>>>(b->inst, b->tile0, b->part_d, b->NT, b->view);
    HIPCHK(hipGetLastError());
    return 0;
}

int b_nn(sph_batch *b, hipStream_t s) {
    const int n = b->n, pec = b->integrator == SPH_INTEGRATOR_PEC;
    State a = b->buf[b->cur], c = b->buf[1 - b->cur];
    // the border-cell bounds only in the fused PEC step's update
    const bool fused = b->uni && !b->persistent && pec;
    kb_bounds<<<b->NC, kBlock, 0, s>>>(b->cs, b->chunkInst, b->chunk0, b->inst, a, pec, b->part_b,
                                       fused && b->border_bounds ? b->cellStart : nullptr, b->view);
    kb_grid<<<b->K, kBlock, 0, s>>>(b->inst, b->chunk0, b->part_b, b->NC);
    kb_bases<<<1, 1024, 0, s>>>(b->inst, b->K, b->view, b->capBins);
    kb_keys<<<nblk_b(n), kBlock, 0, s>>>(b->inst, b->instOf, a, n, pec, b->counts, b->bin, b->rank, b->binflag,
                                         b->view);
    long long tiles_cap = (b->capBins + kScanTile - 1) / kScanTile;
    int tiles = (int)(tiles_cap < 1024 ? tiles_cap : 1024);
    k_scan_reduce<<<tiles, kBlock, 0, s>>>(b->view, b->counts, b->tileSums);
    k_scan_apply<<<tiles, kBlock, 0, s>>>(b->view, b->counts, b->tileSums, b->cellStart);
    k_scatter<<<nblk_b(n), kBlock, 0, s>>>(b->view, b->cellStart, b->bin, b->rank, a.id, b->perm, b->permKey, n);
    long long cb = (b->capBins + kBlock - 1) / kBlock;
    k_cellsort<<<(int)(cb < kRedBlocks ? (cb < 1 ? 1 : cb) : kRedBlocks), kBlock, 0, s>>>(
        b->view, b->cellStart, b->perm, b->permKey, b->binflag, b->bin, b->rank);
    const int fmt = b->precision == SPH_PREC_FP32C ? (b->uni ? kRecUniform : kRecFp32)
                                                   : (b->uni ? kRecUniform64 : kRecFp64);
    kb_reorder<<<nblk_b(n), kBlock, 0, s>>>(b->inst, b->instOf, a, c, b->perm, b->cxy_s, n, fmt, b->recA, b->recB,
                                            pec);
    HIPCHK(hipGetLastError());
    b->cur = 1 - b->cur;
    return 0;
}

// An event around the pair pass (timing on).
int b_mark(sph_batch *b, hipStream_t s) {
    if (b->n_ev >= b->ev.size()) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        b->ev.push_back(e);
    }
    HIPCHK(hipEventRecord(b->ev[b->n_ev++], s));
    return 0;
}

int b_pair(sph_batch *b, hipStream_t s) {
    State st = b->buf[b->cur];
    PairIn in{b->recA, b->recB, st.h, b->cxy_s, b->cellStart};
    PairOut out{st.ax, st.ay, st.drho, nullptr, 0};
    int *q = b->persistent ? b->queue : nullptr;
    hipError_t e;
    if (b->uni)  // the uniform pass: its own workgroup size (256) and list length
        e = b->precision == SPH_PREC_FP32C
                ? launch_pair_batch_fp32u(b->inst, b->instOf, in, out, b->uniK, b->kernel, b->n, q, b->queue_grid, s)
                : launch_pair_batch_fp64u(b->inst, b->instOf, in, out, b->uniK, b->kernel, b->n, q, b->queue_grid, s);
    else
        e = b->precision == SPH_PREC_FP32C
                ? launch_pair_batch_fp32c(b->inst, b->instOf, in, out, b->kernel, b->n, b->pair_threads, b->list_len,
                                          q, b->queue_grid, s)
                : launch_pair_batch_fp64(b->inst, b->instOf, in, out, b->kernel, b->n, b->pair_threads, b->list_len,
                                         q, b->queue_grid, s);
    if (e != hipSuccess) return fail(SPH_EHIP, "batch pair launch: %s", hipGetErrorString(e));
    return 0;
}

// The uniform pair pass with PEC's corrector and the tiles' dt partials
// fused in: one workgroup per tile (one instance each).
int b_pair_pec(sph_batch *b, hipStream_t s) {
    State st = b->buf[b->cur];
    PairIn in{b->recA, b->recB, st.h, b->cxy_s, b->cellStart};
    PairOut out{st.ax, st.ay, st.drho, nullptr, 0};
    const PecArgs pec{st, b->buf[1 - b->cur], b->perm, b->part_d, HmConst{}, 0, &b->view->pec_parts};
    hipError_t e = b->precision == SPH_PREC_FP32C
                       ? launch_pair_batch_fp32u_pec(b->inst, b->ts, b->tileInst, b->NT, in, out, b->uniK, b->kernel,
                                                     b->n, pec, s)
                       : launch_pair_batch_fp64u_pec(b->inst, b->ts, b->tileInst, b->NT, in, out, b->uniK, b->kernel,
                                                     b->n, pec, s);
    if (e != hipSuccess) return fail(SPH_EHIP, "batch pair launch: %s", hipGetErrorString(e));
    return 0;
}

int b_step(sph_batch *b, hipStream_t s) {
    int rc;
    if ((rc = b_dt(b, s))) return rc;
    if ((rc = b_nn(b, s))) return rc;
    // timed: the pair pass and the corrector, fused or apart
    if (b->timing && (rc = b_mark(b, s))) return rc;
    if (b->uni && !b->persistent && b->integrator == SPH_INTEGRATOR_PEC) {
        if ((rc = b_pair_pec(b, s))) return rc;
    } else {
        if ((rc = b_pair(b, s))) return rc;
        kb_correct<<<b->NT, kBlock, 0, s>>>(b->ts, b->tileInst, b->inst, b->buf[b->cur], b->buf[1 - b->cur],
                                            b->perm, b->integrator == SPH_INTEGRATOR_EULER, b->part_d, b->view);
        HIPCHK(hipGetLastError());
    }
    if (b->timing && (rc = b_mark(b, s))) return rc;
    b->dt_fresh = true;
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int sph_batch_create(const sph_params_t *params, int32_t ninst, sph_batch_t **out) {
    if (!params || !out) return fail(SPH_EINVAL, "NULL argument");
    *out = nullptr;
    if (ninst < 1 || ninst > (1 << 20)) return fail(SPH_EINVAL, "ninst=%d out of range", ninst);
    long long ntot = 0, capTot = 0;
    for (int k = 0; k < ninst; ++k) {
        const sph_params_t &P = params[k];
        if (P.n <= 0) return fail(SPH_EINVAL, "instance %d: n=%lld", k, (long long)P.n);
        if (P.kernel != params[0].kernel || P.precision != params[0].precision ||
            P.integrator != params[0].integrator)
            return fail(SPH_EINVAL, "instance %d: kernel, precision and integrator must agree across the batch", k);
        if (P.kernel < 0 || P.kernel > 2) return fail(SPH_EINVAL, "unknown kernel id %d", P.kernel);
        if (P.precision != SPH_PREC_FP64 && P.precision != SPH_PREC_FP32C)
            return fail(SPH_EINVAL, "unknown precision %d", P.precision);
        if (P.integrator != SPH_INTEGRATOR_PEC && P.integrator != SPH_INTEGRATOR_EULER)
            return fail(SPH_EINVAL, "unknown integrator %d", P.integrator);
        if (P.use_xsph || P.use_summation_density || P.dynamic_h || P.coupled_integrator)
            return fail(SPH_EUNSUPPORTED, "a batch runs the plain WCSPH step (no XSPH, summation density, dynamic h or coupled integrator)");
        if (!(P.scale > 0) || !(P.qmax > 0) || !(P.rho0 > 0))
            return fail(SPH_EINVAL, "instance %d: scale, qmax and rho0 must be positive", k);
        ntot += P.n;
        const long long cap = P.cell_capacity > 0 ? P.cell_capacity : P.n * 8 + 4096;
        capTot += cap + 1;
    }
    if (ntot > (1LL << 30)) return fail(SPH_EINVAL, "%lld particles in the batch (max 2^30)", ntot);
    if (capTot > (1LL << 31) - 16) return fail(SPH_EINVAL, "too many cells in the batch");
    sph_batch *b = new (std::nothrow) sph_batch;
    if (!b) return fail(SPH_ENOMEM, "host allocation failed");
    b->K = ninst;
    b->n = (int)ntot;
    b->kernel = params[0].kernel;
    b->precision = params[0].precision;
    b->integrator = params[0].integrator;
    b->P.assign(params, params + ninst);
    b->capBins = capTot;
    const size_t n = (size_t)ntot;
    // chunk and tile tables
    std::vector<int> cs, chunkInst, chunk0, ts, tileInst, tile0;
    std::vector<BInst> hI(ninst);
    long long off = 0;
    for (int k = 0; k < ninst; ++k) {
        const sph_params_t &P = params[k];
        chunk0.push_back((int)cs.size());
        for (long long c = 0; c < P.n; c += kChunk) {
            cs.push_back((int)(off + c));
            chunkInst.push_back(k);
        }
        tile0.push_back((int)ts.size());
        for (long long c = 0; c < P.n; c += kTile) {
            ts.push_back((int)(off + c));
            tileInst.push_back(k);
        }
        BInst &I = hI[k];
        std::memset(&I, 0, sizeof I);
        I.disp = I.vbound = I.abound = INFINITY;
        I.off = (int)off;
        I.n = (int)P.n;
        I.active = 1;
        I.advance_t = 1;
        I.cap = P.cell_capacity > 0 ? P.cell_capacity : P.n * 8 + 4096;
        I.duration = P.duration;
        I.fixed_dt = P.fixed_dt;
        I.damping = P.damping;
        I.rho0 = P.rho0;
        I.gamma = P.gamma;
        I.co = P.co;
        I.B = P.B;
        I.Pb = P.Pb;
        I.alpha = P.alpha;
        I.beta = P.beta;
        I.r0 = P.r0;
        I.D = P.D;
        I.p1 = P.p1;
        I.p2 = P.p2;
        I.g = P.gravity;
        I.qmax = P.qmax;
        I.scale = P.scale;
        I.gamma_c = P.gamma_c;
        I.gamma_f = P.gamma_f;
        I.strict = P.strict;
        I.kernel = P.kernel;
        b->offs.push_back((int)off);
        off += P.n;
    }
    b->offs.push_back((int)off);
    cs.push_back((int)off);
    chunk0.push_back((int)chunkInst.size());
    b->NC = (int)chunkInst.size();
    ts.push_back((int)off);
    tile0.push_back((int)tileInst.size());
    b->NT = (int)tileInst.size();
    std::vector<int> instOf(n);
    for (int k = 0; k < ninst; ++k)
        for (int i = b->offs[k]; i < b->offs[k + 1]; ++i) instOf[i] = k;
    const size_t nd = 15;
    const size_t bytes = n * (nd * sizeof(double) + sizeof(int) + 8) + 1024;
    bool ok = true;
    for (int k = 0; k < 2 && ok; ++k) {
        ok = hipMalloc(&b->arena[k], bytes) == hipSuccess;
        if (!ok) break;
        char *p = (char *)b->arena[k];
        double **fields[] = {&b->buf[k].x,  &b->buf[k].y,   &b->buf[k].vx,   &b->buf[k].vy, &b->buf[k].rho,
                             &b->buf[k].ax, &b->buf[k].ay,  &b->buf[k].drho, &b->buf[k].x0, &b->buf[k].y0,
                             &b->buf[k].vx0, &b->buf[k].vy0, &b->buf[k].rho0, &b->buf[k].m, &b->buf[k].h};
        for (double **f : fields) {
            *f = (double *)p;
            p += n * sizeof(double);
        }
        b->buf[k].id = (int *)p;
        p += n * sizeof(int);
        b->buf[k].label = (int8_t *)p;
    }
    const long long tiles = (capTot + kScanTile - 1) / kScanTile;
    const size_t ra = b->precision == SPH_PREC_FP32C ? sizeof(Rec<float>::A) : sizeof(Rec<double>::A);
    // the uniform forms' records (16-B A, float4 / double4 B) fit in these
    const size_t rb = b->precision == SPH_PREC_FP32C ? sizeof(float4) : sizeof(Rec<double>::B);
    static_assert(sizeof(Rec<float>::A) >= sizeof(double2) && sizeof(Rec<double>::A) >= sizeof(double2) &&
                      sizeof(Rec<double>::B) >= sizeof(double4) && sizeof(float4) >= sizeof(Rec<float>::B),
                  "record buffers hold every format");
    const size_t uk = b->precision == SPH_PREC_FP32C ? uniform_const_size_fp32() : uniform_const_size_fp64();
    ok = ok && hipMalloc(&b->inst, sizeof(BInst) * ninst) == hipSuccess &&
         hipMalloc(&b->view, sizeof(Ctrl)) == hipSuccess && hipMalloc(&b->instOf, sizeof(int) * n) == hipSuccess &&
         hipMalloc(&b->cs, sizeof(int) * cs.size()) == hipSuccess &&
         hipMalloc(&b->chunkInst, sizeof(int) * chunkInst.size()) == hipSuccess &&
         hipMalloc(&b->chunk0, sizeof(int) * chunk0.size()) == hipSuccess &&
         hipMalloc(&b->ts, sizeof(int) * ts.size()) == hipSuccess &&
         hipMalloc(&b->tileInst, sizeof(int) * tileInst.size()) == hipSuccess &&
         hipMalloc(&b->tile0, sizeof(int) * tile0.size()) == hipSuccess &&
         hipMalloc(&b->bin, sizeof(int) * n) == hipSuccess && hipMalloc(&b->rank, sizeof(int) * n) == hipSuccess &&
         hipMalloc(&b->perm, sizeof(int) * n) == hipSuccess && hipMalloc(&b->permKey, sizeof(int) * n) == hipSuccess &&
         hipMalloc(&b->cxy_s, sizeof(int2) * n) == hipSuccess &&
         hipMalloc(&b->counts, sizeof(int) * (capTot + 1)) == hipSuccess &&
         hipMalloc(&b->cellStart, sizeof(int) * (capTot + 2)) == hipSuccess &&
         hipMalloc(&b->tileSums, sizeof(int) * (tiles + 1)) == hipSuccess &&
         hipMalloc(&b->binflag, sizeof(int) * (capTot + 1)) == hipSuccess &&
         hipMalloc(&b->part_b, sizeof(double) * kBoundsVQ * b->NC) == hipSuccess &&
         hipMalloc(&b->part_d, sizeof(double) * kDtQ * b->NT) == hipSuccess &&
         hipMalloc(&b->recA, ra * n) == hipSuccess && hipMalloc(&b->recB, rb * n) == hipSuccess &&
         hipMalloc(&b->queue, sizeof(int) * kQueueInts) == hipSuccess &&
         hipMalloc(&b->uniK, uk * ninst) == hipSuccess && hipMalloc(&b->hm, sizeof(double) * 4 * ninst) == hipSuccess;
    if (!ok) {
        free_batch(b);
        return fail(SPH_ENOMEM, "device allocation for a batch of %lld particles failed", ntot);
    }
    Ctrl v;
    std::memset(&v, 0, sizeof v);
    v.active = 1;
    ok = hipMemcpy(b->inst, hI.data(), sizeof(BInst) * ninst, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(b->view, &v, sizeof v, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(b->instOf, instOf.data(), sizeof(int) * n, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(b->cs, cs.data(), sizeof(int) * cs.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(b->chunkInst, chunkInst.data(), sizeof(int) * chunkInst.size(), hipMemcpyHostToDevice) ==
             hipSuccess &&
         hipMemcpy(b->chunk0, chunk0.data(), sizeof(int) * chunk0.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(b->ts, ts.data(), sizeof(int) * ts.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(b->tileInst, tileInst.data(), sizeof(int) * tileInst.size(), hipMemcpyHostToDevice) ==
             hipSuccess &&
         hipMemcpy(b->tile0, tile0.data(), sizeof(int) * tile0.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemset(b->counts, 0, sizeof(int) * (capTot + 1)) == hipSuccess &&
         hipMemset(b->cellStart, 0, sizeof(int) * (capTot + 2)) == hipSuccess &&
         hipMemset(b->binflag, 0, sizeof(int) * (capTot + 1)) == hipSuccess &&
         hipMemset(b->queue, 0, sizeof(int) * kQueueInts) == hipSuccess &&
         hipStreamCreateWithFlags(&b->own, hipStreamNonBlocking) == hipSuccess;
    int dev = 0;
    if (ok && hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        b->cus = 0;
    if (!ok) {
        free_batch(b);
        return fail(SPH_EHIP, "batch initialisation failed");
    }
    *out = b;
    return 0;
}

int sph_batch_destroy(sph_batch_t *b) {
    if (!b) return 0;
    hipDeviceSynchronize();
    free_batch(b);
    return 0;
}

int sph_batch_upload(sph_batch_t *b, const sph_state_t *st, void *stream) {
    if (!b || !st || !st->x || !st->y || !st->vx || !st->vy || !st->rho || !st->m || !st->h || !st->label)
        return fail(SPH_EINVAL, "upload needs x, y, vx, vy, rho, m, h and label");
    hipStream_t s = (hipStream_t)stream;
    b->cur = 0;
    b->dt_fresh = false;
    kb_upload<<<nblk_b(b->n), kBlock, 0, s>>>(b->buf[0], *st, b->n);
    HIPCHK(hipGetLastError());
    b->loaded = true;
    HIPCHK(hipStreamSynchronize(s));
    std::vector<BInst> hI(b->K);
    HIPCHK(hipMemcpy(hI.data(), b->inst, sizeof(BInst) * b->K, hipMemcpyDeviceToHost));
    for (auto &I : hI) {
        I.t = 0.0;
        I.step = 0;
        I.dt = I.dt_c = I.dt_f = 0.0;
        I.flags = 0;
        I.active = 1;
        I.disp = I.vbound = I.abound = INFINITY;
    }
    HIPCHK(hipMemcpy(b->inst, hI.data(), sizeof(BInst) * b->K, hipMemcpyHostToDevice));
    // the uniform pass when every instance has one h and one fluid mass and
    // its parameters allow it; a batch has no dynamic h, so this holds until
    // the next upload
    kb_hm_range<<<b->K, kBlock, 0, s>>>(b->inst, b->buf[0], b->hm);
    HIPCHK(hipGetLastError());
    std::vector<double> hm(4 * (size_t)b->K);
    HIPCHK(hipMemcpyAsync(hm.data(), b->hm, sizeof(double) * hm.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    bool uni = b->n <= kUniformMaxN;  // the records of all instances, addressed with 32-bit byte offsets
    for (int k = 0; k < b->K && uni; ++k) {
        const double *v = &hm[4 * (size_t)k];
        uni = v[0] == v[1] && v[2] == v[3] && uniform_eligible(b->P[k], v[0]);
    }
    if (uni) {
        const bool f32 = b->precision == SPH_PREC_FP32C;
        const size_t uk = f32 ? uniform_const_size_fp32() : uniform_const_size_fp64();
        std::vector<unsigned char> blk(uk * (size_t)b->K);
        for (int k = 0; k < b->K; ++k) {
            const UniformParams u = make_uniform_params(b->P[k], hm[4 * (size_t)k], hm[4 * (size_t)k + 2]);
            if (f32) uniform_const_fp32(u, blk.data() + uk * k);
            else uniform_const_fp64(u, blk.data() + uk * k);
        }
        HIPCHK(hipMemcpy(b->uniK, blk.data(), blk.size(), hipMemcpyHostToDevice));
    }
    // m and h stay in place for the uniform pass (the upload's buffer, slot =
    // id): every instance's one h in its control block
    for (int k = 0; k < b->K; ++k) {
        hI[k].uh_on = uni ? 1 : 0;
        hI[k].uh = uni ? hm[4 * (size_t)k] : 0.0;
    }
    HIPCHK(hipMemcpy(b->inst, hI.data(), sizeof(BInst) * b->K, hipMemcpyHostToDevice));
    if (uni != b->uni) {  // another record format and pair kernel: recapture
        b->uni = uni;
        b->queue_grid = 0;
        for (auto &g : b->gexec)
            if (g) hipGraphExecDestroy(g), g = nullptr;
    }  // same form: the graphs read the constants from uniK, rewritten above
    return 0;
}

int sph_batch_pair_form(sph_batch_t *b, int32_t *form) {
    if (!b || !form) return fail(SPH_EINVAL, "NULL argument");
    if (!b->loaded) return fail(SPH_ESTATE, "no state: call sph_batch_upload first");
    *form = b->precision == SPH_PREC_FP32C ? (b->uni ? SPH_PAIR_FP32C_UNIFORM : SPH_PAIR_FP32C)
                                           : (b->uni ? SPH_PAIR_FP64_UNIFORM : SPH_PAIR_FP64);
    return 0;
}

int sph_batch_download(sph_batch_t *b, const sph_state_t *st, void *stream) {
    if (!b || !st) return fail(SPH_EINVAL, "NULL argument");
    if (!b->loaded) return fail(SPH_ESTATE, "no state: call sph_batch_upload first");
    kb_download<<<nblk_b(b->n), kBlock, 0, (hipStream_t)stream>>>(b->buf[b->cur], *st, b->inst, b->instOf, b->n,
                                                                  b->buf[0].m, b->buf[0].h);
    HIPCHK(hipGetLastError());
    return 0;
}

int sph_batch_step(sph_batch_t *b, int64_t nsteps, int32_t use_graph, void *stream) {
    if (!b) return fail(SPH_EINVAL, "batch is NULL");
    if (nsteps < 0) return fail(SPH_EINVAL, "nsteps < 0");
    if (!b->loaded) return fail(SPH_ESTATE, "no state: call sph_batch_upload first");
    size_queue(b);
    hipStream_t s = (hipStream_t)stream;
    if (!use_graph || b->timing) {
        for (int64_t k = 0; k < nsteps; ++k) {
            int rc = b_step(b, s);
            if (rc) return rc;
        }
        return 0;
    }
    hipStream_t g = b->own;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreateWithFlags(&e0, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&e1, hipEventDisableTiming));
    HIPCHK(hipEventRecord(e0, s));
    HIPCHK(hipStreamWaitEvent(g, e0, 0));
    int rc = 0;
    if (!b->dt_fresh && nsteps > 0) {
        rc = b_step(b, g);
        --nsteps;
    }
    for (int64_t k = 0; k < nsteps && !rc; ++k) {
        const int start = b->cur;
        if (!b->gexec[start]) {
            // the capture ends, and the graph is released, whatever fails
            hipGraph_t graph = nullptr;
            hipError_t ce = hipStreamBeginCapture(g, hipStreamCaptureModeThreadLocal);
            if (ce != hipSuccess) {
                rc = fail(SPH_EHIP, "begin capture: %s", hipGetErrorString(ce));
                break;
            }
            rc = b_step(b, g);
            ce = hipStreamEndCapture(g, &graph);
            b->cur = start;  // capture flipped it without running anything
            if (!rc && ce != hipSuccess) rc = fail(SPH_EHIP, "end capture: %s", hipGetErrorString(ce));
            if (!rc && (ce = hipGraphInstantiate(&b->gexec[start], graph, nullptr, nullptr, 0)) != hipSuccess) {
                b->gexec[start] = nullptr;
                rc = fail(SPH_EHIP, "graph instantiate: %s", hipGetErrorString(ce));
            }
            if (graph) hipGraphDestroy(graph);
            if (rc) break;
        }
        hipError_t le = hipGraphLaunch(b->gexec[start], g);
        if (le != hipSuccess) rc = fail(SPH_EHIP, "graph launch: %s", hipGetErrorString(le));
        else b->cur = 1 - start;
    }
    hipEventRecord(e1, g);
    hipStreamWaitEvent(s, e1, 0);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc;
}

int sph_batch_status(sph_batch_t *b, double *out_d, int64_t *out_i, void *stream) {
    if (!b) return fail(SPH_EINVAL, "batch is NULL");
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    std::vector<BInst> hI(b->K);
    HIPCHK(hipMemcpy(hI.data(), b->inst, sizeof(BInst) * b->K, hipMemcpyDeviceToHost));
    for (int k = 0; k < b->K; ++k) {
        if (out_d) {
            out_d[4 * k + 0] = hI[k].t;
            out_d[4 * k + 1] = hI[k].dt;
            out_d[4 * k + 2] = hI[k].dt_c;
            out_d[4 * k + 3] = hI[k].dt_f;
        }
        if (out_i) {
            out_i[2 * k + 0] = hI[k].step;
            out_i[2 * k + 1] = hI[k].flags;
        }
    }
    return 0;
}

int sph_batch_set_tuning(sph_batch_t *b, int32_t pair_threads) {
    if (!b) return fail(SPH_EINVAL, "batch is NULL");
    if (pair_threads != 64 && pair_threads != 128 && pair_threads != 256)
        return fail(SPH_EINVAL, "pair_threads must be 64, 128 or 256");
    if (pair_threads != b->pair_threads) {
        b->pair_threads = pair_threads;
        b->queue_grid = 0;
        for (auto &g : b->gexec)
            if (g) hipGraphExecDestroy(g), g = nullptr;
    }
    return 0;
}

int sph_batch_set_list_len(sph_batch_t *b, int32_t list_len) {
    if (!b) return fail(SPH_EINVAL, "batch is NULL");
    if (list_len != 0 && list_len != 16 && list_len != 24 && list_len != 32)
        return fail(SPH_EINVAL, "list_len must be 0, 16, 24 or 32");
    if (list_len != b->list_len) {
        b->list_len = list_len;
        b->queue_grid = 0;
        for (auto &g : b->gexec)
            if (g) hipGraphExecDestroy(g), g = nullptr;
    }
    return 0;
}

int sph_batch_set_persistent(sph_batch_t *b, int32_t on) {
    if (!b) return fail(SPH_EINVAL, "batch is NULL");
    if (on != 0 && on != 1) return fail(SPH_EINVAL, "persistent must be 0 or 1");
    if (on != b->persistent) {
        b->persistent = on;
        b->queue_grid = 0;
        for (auto &g : b->gexec)
            if (g) hipGraphExecDestroy(g), g = nullptr;
    }
    return 0;
}

int sph_batch_set_border_bounds(sph_batch_t *b, int32_t on) {
    if (!b) return fail(SPH_EINVAL, "batch is NULL");
    b->border_bounds = on ? 1 : 0;
    for (auto &g : b->gexec)
        if (g) hipGraphExecDestroy(g), g = nullptr;
    return 0;
}

int sph_batch_border_count(sph_batch_t *b, int64_t *counts, void *stream) {
    if (!b || !counts) return fail(SPH_EINVAL, "NULL argument");
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    std::vector<BInst> hI(b->K);
    HIPCHK(hipMemcpy(hI.data(), b->inst, sizeof(BInst) * b->K, hipMemcpyDeviceToHost));
    for (int k = 0; k < b->K; ++k) counts[k] = hI[k].border_updates;
    return 0;
}

int sph_batch_timing_enable(sph_batch_t *b, int32_t on) {
    if (!b) return fail(SPH_EINVAL, "batch is NULL");
    b->timing = on != 0;
    b->n_ev = 0;
    return 0;
}

int sph_batch_timing_read(sph_batch_t *b, double *avg_ms_pair, int64_t *count, int32_t reset) {
    if (!b) return fail(SPH_EINVAL, "batch is NULL");
    HIPCHK(hipDeviceSynchronize());
    double tot = 0.0;
    int64_t k = 0;
    for (size_t i = 0; i + 1 < b->n_ev; i += 2) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, b->ev[i], b->ev[i + 1]));
        tot += ms;
        ++k;
    }
    if (avg_ms_pair) *avg_ms_pair = k ? tot / k : 0.0;
    if (count) *count = k;
    if (reset) b->n_ev = 0;
    return 0;
}

}  // extern "C"
