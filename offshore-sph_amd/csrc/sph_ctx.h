// sph_ctx.h — the context and tape objects behind the C ABI, shared by the
// engine and the adjoint translation units.
#pragma once
#include <string>
#include <vector>

#include "sph_internal.h"

// One simulation on one GPU (opaque `sph_ctx_t` of the C ABI).
struct sph_ctx {
    sph_params_t P;
    int n = 0;
    int dev = 0;
    hipsph::State buf[2];
    void *arena[2] = {nullptr, nullptr};
    int cur = 0;
    hipsph::Ctrl *ctrl = nullptr;
    double *dt_hist = nullptr;
    int *bin = nullptr, *rank = nullptr, *perm = nullptr, *permKey = nullptr;
    int2 *cxy_s = nullptr;  // raw cell (cx, cy) of each slot in cell order
    int *counts = nullptr, *cellStart = nullptr, *tileSums = nullptr, *binflag = nullptr;
    double *part_b = nullptr, *part_d = nullptr;  // per-workgroup partials: bounds, dt
    double *part_k = nullptr;                     // kinetic-energy partials + result
    long long cap_cells = 0;
    bool dt_fresh = false;  // part_d holds the dt partials of the current state (left by the corrector)
    bool loaded = false;    // a state was uploaded: every id indexes 0 .. n-1
    bool mid = false;       // a split step is open (sph_step_begin ran, sph_step_end not yet)
    bool xs_fluid_zero = false;  // XSPH off and a reorder ran: the fluid's XSPH velocity reads as 0
    void *recA = nullptr, *recB = nullptr;  // packed pair-pass records (Rec<double> / Rec<float>)
    int *slot_of = nullptr;
    hipStream_t own = nullptr;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    // bumped whenever values that captured graphs bake into kernel arguments
    // (constants, launch shape) change; graphs captured under another value
    // are stale
    unsigned long long gen = 0;
    int pair_threads = 256;
    int pair_staging = 1;  // the uniform float pair passes stage each workgroup's records in LDS
    int border_bounds = 1;  // the fused PEC step's bounds from the border cells when the displacement allows
    // one h for every particle and one mass for every fluid particle in the
    // current state (found at upload, restore and scatter): the fused float
    // pass then takes its uniform form, with kRecUniform records
    bool uni_hm = false;
    double uni_h = 0.0, uni_m = 0.0;
    // one h and one mass for every particle (walls included) and no dynamic
    // h: both buffers hold them in every slot and nothing permutes them
    bool static_hm = false;
    int rec_fmt = 0;  // format of the records the last reorder wrote
    // timing of the pair pass and of the neighbour update
    bool timing = false;
    std::vector<hipEvent_t> ev_pair, ev_nn;
    size_t n_pair = 0, n_nn = 0;
};

namespace hipsph {

// Per recorded step: what the reverse pass reads besides the arrays.
struct TapeHdr {
    Grid grid;
    double dt, dt_c, dt_f;
    double amax2;  // max |a_prev|^2 over the fluid (the force criterion's argument)
    int amax_id;   // a fluid particle attaining it (smallest id)
    int active;
};

// One recorded step. Per-particle values are stored in reference index
// order (id); the walk structures in the cell order of the step's pair pass.
struct TapeSlot {
    // id order
    double *vxh, *vyh;          // pair inputs: the predicted velocity
    double *ax, *ay, *drho;     // pair outputs
    double *vx0, *vy0;          // velocity at the start of the step
    double *axp, *ayp, *drhop;  // rates carried into the step (previous pair pass)
    int *slot;                  // id -> cell slot
    int8_t *clamp;              // PEC's strict clamp took the predicted (bit 0) / corrected (bit 1) density
    // cell-slot order (the step's walk order, read by the pair adjoint)
    double4 *pa;     // (x, y, h, w) of the predicted state; w = m for fluid, -label otherwise
    double4 *pb;     // (vx, vy, rho, m) of the predicted state
    double4 *ps;     // (p / rho^2, its rho- and c0-derivatives, 0) for fluid, 0 otherwise
    int *id;         // cell slot -> id
    int2 *cxy;       // cell of each slot
    int *cellStart;  // cell offsets [0 .. ncells + 1]
};

}  // namespace hipsph

// A step tape for the adjoint (opaque `sph_tape_t` of the C ABI).
struct sph_tape {
    sph_ctx *ctx = nullptr;
    int nslots = 0;
    long long cells = 0;  // cellStart entries per slot
    void *arena = nullptr;
    hipsph::TapeHdr *hdr = nullptr;  // device, one per slot
    std::vector<hipsph::TapeSlot> slot;
    // static per-particle data in id order
    int8_t *label = nullptr;
    // reverse-pass scratch: id order, and (lam_x, lam_y, kap, 0) in the cell
    // slot order of the step being reversed
    double *vbx = nullptr, *vby = nullptr;
    double4 *lk = nullptr;
    double *part = nullptr;  // reduction partials
    double *amax_v = nullptr;  // argmax partials (value, id) per workgroup
    int *amax_i = nullptr;
    int *force_steps = nullptr;  // device counter: reversed steps whose dt came from the force criterion
    // captured steps of sph_adjoint_segment: recording per (slot, buffer
    // parity), reversal per slot for the adjoint buffers in adj_key
    std::vector<hipGraphExec_t> rec_exec, adj_exec;
    sph_adjoint_t adj_key{};
    unsigned long long gen = 0;  // the context's gen when the cached graphs were captured
};

namespace hipsph {

// Error text for sph_last_error(); returns `code`.
int fail(int code, const char *fmt, ...);
Phys make_phys(const sph_params_t &P);

// Tape hooks called by sph_step_record (implemented with the adjoint).
int tape_record_pair(sph_ctx *c, sph_tape *t, int slot, hipStream_t s, bool lean);
// The reverse of one recorded step (the kernels of sph_adjoint_step).
int adjoint_launch(sph_ctx *c, sph_tape *t, int slot, const sph_adjoint_t *adj, hipStream_t s);
void tape_drop_graphs(sph_tape *t);
int tape_supported(const sph_ctx *c);
int tape_refresh_labels(sph_ctx *c, sph_tape *t, hipStream_t s);

}  // namespace hipsph

#define HIPCHK(expr)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess) return hipsph::fail(SPH_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)
