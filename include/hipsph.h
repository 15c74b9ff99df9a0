/*
 * hipsph.h — C ABI of the MI355X (gfx950) WCSPH step engine.
 *
 * The reference (KoningJasper/Offshore-SPH) has no FFI: its hot path is a set
 * of numba dispatchers called from Python. Each entry point below replaces one
 * of those dispatchers and cites it. Host code binds this header through
 * ctypes (offshore-sph_amd/hipsph/_lib.py).
 *
 * Conventions
 *   - Every function returns 0 on success and a negative SPH_E* code on error;
 *     sph_last_error() gives a message for the calling thread. No C++
 *     exception crosses this boundary.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *     Every call is stream-ordered; calls that return host values synchronise
 *     that stream and say so.
 *   - Device pointers are borrowed: the library never frees memory it did not
 *     allocate. A context owns its workspace (cell-ordered state, keys,
 *     permutation, cell offsets); step calls allocate nothing.
 *   - A context is not thread-safe: use one per host thread.
 *   - All arrays are structure-of-arrays; indices are positions in the
 *     active-particle array, the reference's `pA[indexes]` order.
 */
#ifndef HIPSPH_H
#define HIPSPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---- */
#define SPH_OK 0
#define SPH_EINVAL (-1)      /* bad argument */
#define SPH_EHIP (-2)        /* HIP runtime error */
#define SPH_ENOMEM (-3)      /* allocation failed */
#define SPH_ECAPACITY (-4)   /* cell grid larger than the context's capacity */
#define SPH_ESTATE (-5)      /* call out of order (e.g. compute before nn_update, a step or
                                download before the first state upload) */
#define SPH_ENUMERIC (-6)    /* non-finite value met (position, dt) */
#define SPH_EUNSUPPORTED (-7)

/* ---- device status flags (sph_status) ---- */
#define SPH_FLAG_CAPACITY 1u     /* cell grid exceeded capacity; step halted */
#define SPH_FLAG_NONFINITE 2u    /* NaN/Inf position or bound; step halted */
#define SPH_FLAG_OOB_KEY 4u      /* a particle's flat cell key fell outside [0, n_cells) */
#define SPH_FLAG_DT_SMALL 8u     /* dt < 1e-6 (src/Solver.py:220-224) */
#define SPH_FLAG_NO_FLUID 16u    /* dt reduction over zero fluid particles */
#define SPH_FLAG_ZERO_DENSITY 32u /* a fluid particle reached the pair pass with p / rho^2 not finite
                                    (rho = 0, e.g. after PEC's strict clamp): the reference's momentum
                                    sums turn NaN from there on, because its zero-gradient terms
                                    multiply p / rho^2 by 0 (src/Equations/Momentum.py:36-56); the
                                    step completes and the next one halts */

/* ---- enums ---- */
enum { SPH_KERNEL_CUBIC = 0, SPH_KERNEL_GAUSSIAN = 1, SPH_KERNEL_WENDLAND = 2 };
/* FP64: pair arithmetic in double, neighbours in the reference's order, no FMA
 * contraction (bit-reproducible against the CPU oracle for the cubic spline).
 * FP32C: pair arithmetic in float (separations rounded once from double
 * positions), TwoSum-compensated sums; integrator, time and dt stay double. */
enum { SPH_PREC_FP64 = 0, SPH_PREC_FP32C = 1 };
enum { SPH_INTEGRATOR_PEC = 0, SPH_INTEGRATOR_EULER = 1 };
enum { SPH_FLUID = 0, SPH_BOUNDARY = 1, SPH_TEMPBOUNDARY = 2, SPH_COUPLED = 3 };
/* Integrator of the particles labelled SPH_COUPLED: none (they stay where
 * they are, as the reference leaves them without a coupling hook), or
 * Newmark-beta (src/Integrators/NewmarkBeta.py:41-66). */
enum { SPH_COUPLED_NONE = 0, SPH_COUPLED_NEWMARK = 1 };

typedef struct sph_ctx sph_ctx_t;

/* Constants of WCSPH (src/Methods/WCSPH.py:35-79), of the integrator
 * (src/Integrators/PEC.py:13-27), of the neighbour grid
 * (src/Tools/NNLinkedList.py:23-35, src/Solver.py:109) and of the time step
 * (src/Solver.py:210-230). */
typedef struct {
    int64_t n;               /* active particles */
    int32_t kernel;          /* SPH_KERNEL_* */
    int32_t precision;       /* SPH_PREC_* */
    int32_t integrator;      /* SPH_INTEGRATOR_* */
    int32_t strict;          /* PEC(strict=...) : clamp rho at 0 */
    int32_t use_xsph;        /* 1: WCSPH.compute_velocity's XSPH velocity in the loop
                              * (src/Methods/WCSPH.py:171-189, src/Equations/XSPH.py),
                              * used by PEC for the positions (src/Integrators/PEC.py:
                              * 43-48,66-71; the reference requires it on in both) */
    int32_t use_summation_density; /* 1: density by summation before the loop's
                              * pass (src/Tools/SolverTools.py:129-140), no continuity */
    double rho0, gamma, co, B, Pb, alpha, beta;
    double r0, D, p1, p2;    /* Lennard-Jones wall force */
    double gravity;          /* 9.81, subtracted from a_y */
    double scale;            /* cell size = scale * min h (2.0) */
    double qmax;             /* neighbour cut-off q <= qmax (3.0) */
    double gamma_c, gamma_f; /* Courant and force factors (0.25, 0.25) */
    double damping;          /* PEC damping */
    double fixed_dt;         /* > 0: use this dt (Solver(timeStep=...)); else dynamic */
    double duration;         /* device-side stepping stops once t >= duration */
    int64_t cell_capacity;   /* max cells of the grid; 0 = automatic */
    double epsilon;          /* XSPH factor (WCSPH.epsilon, 0.5; src/Methods/WCSPH.py:65) */
    int32_t dynamic_h;       /* 1: Solver(h=None): every step, after the neighbour update,
                              * fluid h = 1.3 (m / rho)^(1/2) of the predicted density
                              * (src/Tools/SolverTools.py:106-118, src/Solver.py:242-246) */
    int32_t coupled_integrator; /* SPH_COUPLED_*: Solver(couplingIntegrator=NewmarkBeta(...))
                              * predicts and corrects the coupled particles with the
                              * fluid (src/Solver.py:381-382,397-398) */
    double newmark_beta, newmark_gamma; /* NewmarkBeta(beta, gamma); 2 beta >= gamma */
} sph_params_t;

/* Structure-of-arrays state in the active-particle order. Device pointers. */
typedef struct {
    double *x, *y, *vx, *vy, *rho, *ax, *ay, *drho, *m, *h;
    double *p;     /* download only: Tait pressure + Pb (src/Tools/SolverTools.py:122-127); may be NULL */
    int8_t *label;
    double *xsphx, *xsphy; /* XSPH velocity (particle_dtype xsphx, xsphy); may be NULL (upload: 0) */
} sph_state_t;

/* ---- context ---- */
/* Replaces Solver.__init__ + setup's device side (src/Solver.py:46-141,173-208). */
int sph_ctx_create(const sph_params_t *params, sph_ctx_t **out);
int sph_ctx_destroy(sph_ctx_t *ctx);
/* Update constants that may change between steps (damping, alpha, beta, co,
 * B, duration, fixed_dt). n, kernel, precision and use_xsph are fixed at creation.
 * Synchronises the device first (the control block is rewritten). */
int sph_ctx_set_params(sph_ctx_t *ctx, const sph_params_t *params);

/* Copy the state in (reference index order) / out (un-permuted back to that
 * order). Replaces the masked copies `pA[indexes]` / scatter-back of
 * src/Solver.py:238,254. upload also resets t and the step count when
 * reset_clock != 0. */
int sph_state_upload(sph_ctx_t *ctx, const sph_state_t *st, int32_t reset_clock, void *stream);
int sph_state_download(sph_ctx_t *ctx, const sph_state_t *st, void *stream);

/* ---- neighbour search: src/Tools/NNLinkedList.py ---- */
/* NNLinkedList.update (:37-39 -> _init :86-101, _get_cell_size :103-110,
 * _get_number_of_cells :112-127, _bin :129-141): bounds, grid, counting sort
 * of cell keys with a prefix scan, reorder of the state into cell order. */
int sph_nn_update(sph_ctx_t *ctx, void *stream);
/* Grid of the last update: out_d = {xmin, xmax, ymin, ymax, cell_size},
 * out_i = {ncx, ncy, n_cells}. Synchronises `stream`. */
int sph_nn_grid(sph_ctx_t *ctx, double out_d[5], int64_t out_i[3], void *stream);
/* For parity: NNLinkedList.near(i) (:82-84 -> nearPos :41-80) for every
 * particle i, as CSR in the reference's order (cells x-shift outer, y-shift
 * inner; descending index inside a cell), plus each particle's flat key
 * cx + ncx*cy (:136-137). Device outputs: keys[n], off[n+1], idx[cap],
 * q[cap], r[cap] (q, r may be NULL). *total receives the entry count; when it
 * exceeds cap nothing past cap is written and SPH_ECAPACITY is returned.
 * Synchronises `stream`. */
int sph_nn_dump(sph_ctx_t *ctx, int64_t *keys, int64_t *off, int32_t *idx, double *q, double *r,
                int64_t cap, int64_t *total, void *stream);

/* ---- interaction loop: src/Tools/SolverTools.py:120-175 ---- */
/* One fused pass over the fluid particles: Tait pressure, continuity
 * (src/Equations/Continuity.py:4-16), momentum with artificial viscosity and
 * gravity (src/Equations/Momentum.py:5-57, src/Methods/WCSPH.py:151-169) and
 * the Lennard-Jones wall force (src/Equations/BoundaryForce.py:6-42).
 * Requires a prior sph_nn_update. */
int sph_compute(sph_ctx_t *ctx, void *stream);

/* The same pass split by equation, for timing the stages apart (and for
 * callers that want one of them): sph_continuity writes drho only
 * (src/Equations/Continuity.py:4-16 via WCSPH.compute_density_change,
 * src/Methods/WCSPH.py:191-203); sph_momentum writes ax, ay only (momentum
 * + AV + gravity + Lennard-Jones wall force, src/Equations/Momentum.py:5-57,
 * BoundaryForce.py:6-42, WCSPH.py:151-169). Each evaluates the same
 * expressions in the same order as sph_compute. Requires a prior
 * sph_nn_update. */
int sph_continuity(sph_ctx_t *ctx, void *stream);
int sph_momentum(sph_ctx_t *ctx, void *stream);

/* For parity: one pair pass (as sph_compute) that also reports, per particle
 * in reference order, the number of neighbours the pass itself found
 * (len(near(i)) of src/Tools/SolverTools.py:163; -1 for non-fluid).
 * counts is a device int32[n]. Synchronises `stream`. */
int sph_compute_counts(sph_ctx_t *ctx, int32_t *counts, void *stream);

/* ---- time step and integrators ---- */
/* Solver._minTimeStep + TimeStep.compute (src/Solver.py:210-230,
 * src/Equations/TimeStep.py:11-90), evaluated on the device into the
 * context's control block; also starts a step (t += dt, step += 1) unless
 * t >= duration, in which case the step's later calls are no-ops. */
int sph_dt(sph_ctx_t *ctx, void *stream);
/* PEC.predict (src/Integrators/PEC.py:32-59) on the fluid; no-op for Euler
 * (src/Integrators/Euler.py:13-15). */
int sph_predict(sph_ctx_t *ctx, void *stream);
/* PEC.correct (src/Integrators/PEC.py:61-89) / Euler.correct
 * (src/Integrators/Euler.py:17-26) on the fluid. */
int sph_correct(sph_ctx_t *ctx, void *stream);
/* nsteps iterations of the Solver.run loop body (src/Solver.py:366-410):
 * dt -> predict -> nn_update -> compute -> correct, all on the device.
 * use_graph != 0 captures the step in a hipGraph and replays it. */
int sph_step(sph_ctx_t *ctx, int64_t nsteps, int32_t use_graph, void *stream);

/* One loop body split at the coupling hook (src/Solver.py:389-392):
 * sph_step_begin runs dt -> predict (fluid and, with coupled_integrator,
 * the coupled particles) -> nn_update -> compute and leaves the state of the
 * hook's call: the predicted positions, velocities and density and this
 * step's rates. The caller may then read and write records
 * (sph_state_gather / sph_state_scatter, sph_state_download), query the
 * step's grid (sph_near_pos, sph_probe_pressure) and must call
 * sph_step_end, which runs the correctors with the values it finds and
 * advances the clock. begin + end give sph_step's result bit for bit.
 * Other step calls fail with SPH_ESTATE while a split step is open. */
int sph_step_begin(sph_ctx_t *ctx, void *stream);
int sph_step_end(sph_ctx_t *ctx, void *stream);

/* ---- records by reference index: the coupling hook's view of the state ---- */
/* The hook reads and writes particleArray[cInd] (examples/IceBreak.py:110-178,
 * src/Solver.py:391). `ids` (host, k entries) are indices into the uploaded
 * active array; the non-NULL fields of `st` are device arrays of k entries in
 * the order of `ids`. Gather fills them (p is the Tait pressure + Pb of the
 * record's density); scatter writes x, y, vx, vy, rho, ax, ay, drho, m, h,
 * xsphx, xsphy into the current state (label and p must be NULL; ids must be
 * distinct). Synchronise `stream`. */
int sph_state_gather(sph_ctx_t *ctx, const int64_t *ids, int64_t k, const sph_state_t *st, void *stream);
int sph_state_scatter(sph_ctx_t *ctx, const int64_t *ids, int64_t k, const sph_state_t *st, void *stream);

/* NNLinkedList.nearPos (src/Tools/NNLinkedList.py:41-80) for npts points
 * (x, y, h: device arrays) against the grid of the last neighbour update and
 * the current positions: as sph_nn_dump, off (npts + 1 entries) and, when
 * the total fits in cap, idx (reference indices), q and r in the reference's
 * order. Returns SPH_ECAPACITY with *total set when it does not fit. */
int sph_near_pos(sph_ctx_t *ctx, int64_t npts, const double *x, const double *y, const double *h, int64_t *off,
                 int32_t *idx, double *q, double *r, int64_t cap, int64_t *total, void *stream);
/* The SPH pressure probe of the coupled example (pressure_SPH,
 * examples/IceBreak.py:253-285) at npts points, one lane each: nearPos of
 * (x, y) with the probe's h, W(r, h) of the context's kernel, the Shepard
 * normalisation (src/Equations/Shepard.py:5-12), the summation density over
 * the fluid (src/Equations/SummationDensity.py:5-13) and
 * p = ((rho / rho0)^gamma - 1) B. rho, p: device arrays of npts. */
int sph_probe_pressure(sph_ctx_t *ctx, int64_t npts, const double *x, const double *y, const double *h,
                       double *rho, double *p, void *stream);

/* Device clock and flags: out_d = {t, dt, dt_c, dt_f}, out_i = {steps, flags}.
 * Synchronises `stream`. */
int sph_status(sph_ctx_t *ctx, double out_d[4], int64_t out_i[2], void *stream);
/* Settling control of Solver.run (src/Solver.py:366-410): sets the device
 * clock to `t` and whether a step advances it (`advance_time` = 0 while the
 * reference is settling, src/Solver.py:404-406). Synchronises `stream`. */
int sph_set_clock(sph_ctx_t *ctx, double t, int32_t advance_time, void *stream);
/* KineticEnergy (src/Equations/KineticEnergy.py:5-12) of the fluid of the
 * current state: sum of 0.5 m (vx^2 + vy^2), reduced on the device (order of
 * the sum differs from the serial loop). Synchronises `stream`. */
int sph_kinetic_energy(sph_ctx_t *ctx, double *out, void *stream);
/* The dt series of the last `count` steps (dt_a, dt_c, dt_f interleaved) from
 * the device ring of the most recent SPH_DT_HISTORY steps. Synchronises. */
#define SPH_DT_HISTORY 65536
int sph_dt_history(sph_ctx_t *ctx, double *out3, int64_t count, void *stream);

/* ---- reverse mode (no counterpart in the reference) ----
 * For gradients of a loss on a later state with respect to the initial
 * state, the artificial-viscosity alpha and the sound speed c0 (the caller
 * keeps B = rho0 c0^2 / gamma): a tape holds `nslots` recorded steps;
 * sph_step_record performs one ordinary step (as sph_step) and records it in
 * `slot`; sph_adjoint_step maps the adjoint of that step's outputs to the
 * adjoint of its inputs, in place, and adds the step's contributions to
 * *g_alpha and *g_co. Steps are reversed in the opposite order of recording;
 * checkpointing (re-recording a segment from a saved state) is up to the
 * caller. All arrays are device double[n] in reference index order; entries
 * of non-fluid particles are ignored and set to 0. PEC and Euler; with PEC's
 * strict clamp a density the clamp set to 0 passes no adjoint back (the
 * derivative of max(rho, 0) on the clamped side).
 * dt is differentiated on both branches of min(courant, force)
 * (src/Equations/TimeStep.py:40-56): the Courant value through c0, the force
 * value through the carried acceleration of the particle attaining max |a|^2
 * (recorded; smallest index on ties). sph_tape_force_steps counts the
 * reversed steps whose dt came from the force criterion. */
typedef struct sph_tape sph_tape_t;
typedef struct {
    double *x, *y, *vx, *vy, *rho; /* adjoint of the state: in after the step, out before it */
    double *ax, *ay, *drho;        /* adjoint of the rates carried into the next step's predictor */
    double *g_alpha, *g_co;        /* device scalars, accumulated */
} sph_adjoint_t;
int sph_tape_create(sph_ctx_t *ctx, int32_t nslots, sph_tape_t **out);
int sph_tape_destroy(sph_tape_t *tape);
int sph_tape_force_steps(sph_tape_t *tape, int64_t *out);
int sph_step_record(sph_ctx_t *ctx, sph_tape_t *tape, int32_t slot, void *stream);
int sph_adjoint_step(sph_ctx_t *ctx, sph_tape_t *tape, int32_t slot, const sph_adjoint_t *adj, void *stream);
/* One checkpoint segment in a single call: record `nsteps` steps from the
 * current state into slots 0..nsteps-1, then reverse them (as
 * sph_step_record and sph_adjoint_step, in that order). Steps after the first
 * are replayed from hipGraphs cached in the tape (per slot and buffer parity
 * for the recording; per slot for the reversal, for the adjoint buffers of
 * the last call). */
int sph_adjoint_segment(sph_ctx_t *ctx, sph_tape_t *tape, int32_t nsteps, const sph_adjoint_t *adj, void *stream);
/* A saved state back into the context without a host round trip (for
 * checkpoint restores): as sph_state_upload without resetting the step count,
 * with the clock set to t and advancing, flags cleared. */
int sph_state_restore(sph_ctx_t *ctx, const sph_state_t *st, double t, void *stream);

/* ---- batched ensemble (no counterpart in the reference) ----
 * K independent simulations of uneven sizes advanced together: one launch per
 * stage covers every instance (segmented cell grid, per-instance clock, dt,
 * grid and constants), and a step replays as one hipGraph. Each instance
 * evolves exactly as sph_step would advance it alone. params[k] describes
 * instance k (n, constants, duration, cell_capacity); kernel, precision and
 * integrator must be the same for all. State arrays are the instances'
 * states concatenated in instance order (instance k at offset sum of n_j for
 * j < k), in each instance's reference index order. */
typedef struct sph_batch sph_batch_t;
int sph_batch_create(const sph_params_t *params, int32_t ninst, sph_batch_t **out);
int sph_batch_destroy(sph_batch_t *batch);
int sph_batch_upload(sph_batch_t *batch, const sph_state_t *st, void *stream);
int sph_batch_download(sph_batch_t *batch, const sph_state_t *st, void *stream);
int sph_batch_step(sph_batch_t *batch, int64_t nsteps, int32_t use_graph, void *stream);
/* Per instance k: out_d[4k..4k+3] = {t, dt, dt_c, dt_f}, out_i[2k..2k+1] =
 * {steps, flags}. Synchronises `stream`. */
int sph_batch_status(sph_batch_t *batch, double *out_d, int64_t *out_i, void *stream);
/* Pair-pass workgroup size (64, 128, 256), as sph_ctx_set_tuning. */
int sph_batch_set_tuning(sph_batch_t *batch, int32_t pair_threads);
/* Pair-pass contributor list length per thread (the LDS tile: list_len x 4 B
 * x the workgroup): 16, 24 or 32, or 0 for the single-simulation pass's
 * (24 for the float cubic/Wendland pass, else 32). Chosen by the autotuner. */
int sph_batch_set_list_len(sph_batch_t *batch, int32_t list_len);
/* Pair-pass schedule: 0 = one thread per slot (one workgroup per tile of
 * slots), 1 = a persistent kernel of about one resident occupancy whose
 * workgroups take tiles from a per-XCD work queue until it is drained.
 * Results are identical either way. Chosen by the autotuner. */
int sph_batch_set_persistent(sph_batch_t *batch, int32_t on);
/* No reference counterpart: the border-cell bounds of sph_ctx_set_border_bounds
 * for every instance of the batch (its fused PEC step), 1 (the default) or 0;
 * and how many neighbour updates of each instance took them (counts: one per
 * instance). Synchronises `stream`. */
int sph_batch_set_border_bounds(sph_batch_t *batch, int32_t on);
int sph_batch_border_count(sph_batch_t *batch, int64_t *counts, void *stream);
/* Eager steps with timing on record HIP events around each step's pair pass
 * and corrector (one kernel when the corrector is fused into the uniform
 * pass); timing_read returns their mean in ms. */
int sph_batch_timing_enable(sph_batch_t *batch, int32_t on);
int sph_batch_timing_read(sph_batch_t *batch, double *avg_ms_pair, int64_t *count, int32_t reset);

/* ---- element-wise plugin functions on device arrays ---- */
/* Kernel.evaluate / Kernel.gradient (src/Kernels/Kernel.py:6-12) for
 * CubicSpline (src/Kernels/CubicSpline.py:10-70), Gaussian
 * (src/Kernels/Gaussian.py:16-59, incl. derivative :27-37 via
 * sph_kernel_derivative) and Wendland (src/Kernels/Wendland.py:9-64). */
int sph_kernel_evaluate(int32_t kernel, int64_t n, const double *r, const double *h, double *out, void *stream);
int sph_kernel_gradient(int32_t kernel, int64_t n, const double *x, const double *r, const double *h, double *out,
                        void *stream);
int sph_kernel_derivative(int32_t kernel, int64_t n, const double *r, const double *h, double *out, void *stream);
/* TaitEOS (src/Equations/TaitEOS.py:6-31) + Pb. */
int sph_tait_eos(int64_t n, double gamma, double B, double rho0, double Pb, const double *rho, const int8_t *label,
                 double *p_out, void *stream);

/* ---- launch shape of the pair pass ---- */
/* No reference counterpart (numba picks no launch shape): the workgroup size
 * of the pair pass (64, 128 or 256 threads; each thread walks one particle's
 * 3x3 cells and keeps a 32-entry contributor list in LDS). Set by the
 * autotuner (offshore-sph_amd/hipsph/autotune.py). */
int sph_ctx_set_tuning(sph_ctx_t *ctx, int32_t pair_threads);
int sph_ctx_get_tuning(sph_ctx_t *ctx, int32_t *pair_threads);
/* No reference counterpart: whether the uniform float pair passes (FP32C,
 * one h, one fluid mass; sph_compute, the split passes and the step's fused
 * pass) stage each workgroup's neighbour records in LDS before the walk (1,
 * the default) or read them from global memory (0). Results are
 * bit-identical either way; for tests and tuning. */
int sph_ctx_set_pair_staging(sph_ctx_t *ctx, int32_t on);
/* No reference counterpart (NNLinkedList._init, src/Tools/NNLinkedList.py:
 * 86-101, scans every particle): whether the neighbour update of a fused PEC
 * step (one h) takes the grid's bounds from the particles of the border cells
 * alone when the previous corrector's displacement bound shows that no other
 * particle can hold an extreme (1, the default), or scans every particle (0).
 * The grid is the same either way; for tests. */
int sph_ctx_set_border_bounds(sph_ctx_t *ctx, int32_t on);
/* No reference counterpart: how many neighbour updates of this context took
 * their bounds from the border cells. Synchronises `stream`. */
int sph_nn_border_count(sph_ctx_t *ctx, int64_t *count, void *stream);
/* No reference counterpart: which form of the fused pair pass the context
 * runs, chosen from the state at each upload, restore and scatter:
 * SPH_PAIR_FP64 / SPH_PAIR_FP32C (general), or the forms for one h for every
 * particle and one mass for every fluid particle (the cubic spline or
 * Wendland, no XSPH, summation density or dynamic h), SPH_PAIR_FP32C_UNIFORM
 * (float terms summed in double) and SPH_PAIR_FP64_UNIFORM (bit-identical to
 * SPH_PAIR_FP64). */
enum { SPH_PAIR_FP64 = 0, SPH_PAIR_FP32C = 1, SPH_PAIR_FP32C_UNIFORM = 2, SPH_PAIR_FP64_UNIFORM = 3 };
int sph_pair_form(sph_ctx_t *ctx, int32_t *form);
/* The same for a batch, chosen at each sph_batch_upload: a uniform form when
 * every instance has one h and one fluid mass (each its own) and allows it. */
int sph_batch_pair_form(sph_batch_t *batch, int32_t *form);

/* ---- nn.update + _loop with no allocation ----
 * No counterpart in the reference (whose _loop allocates per particle). For
 * callers that cannot allocate: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc. One call does what sph_nn_update
 * then sph_compute do on a context holding the same state
 * (src/Tools/NNLinkedList.py:37-39, src/Tools/SolverTools.py:120-175), with
 * the same kernels and the same results, on arrays the caller passes, and
 * with every piece of workspace in one static device array of the library
 * (about 6 MB per device, held until the process ends): at most
 * SPH_NOALLOC_MAX_PARTICLES particles (more is refused with SPH_EINVAL) and
 * SPH_NOALLOC_MAX_CELLS grid cells. params: n, kernel, precision and the
 * WCSPH constants (no XSPH, summation density or dynamic h). st: device x,
 * y, vx, vy, rho, m, h, label and, optionally, ax, ay, drho (the rates a
 * target keeps when it finds no neighbour; NULL reads as 0), n entries in
 * the reference's active order. drho, ax, ay: device outputs of n entries.
 * status: a device int32 that receives the SPH_FLAG_* bits of the call; with
 * SPH_FLAG_CAPACITY (a grid of more than SPH_NOALLOC_MAX_CELLS cells) or
 * SPH_FLAG_NONFINITE the outputs are not written. The call allocates
 * nothing and does not synchronise; the first call on a device resolves the
 * array's address (hipGetSymbolAddress), so make that one outside a graph
 * capture. Calls on one device share the array: calls that overlap on one
 * device race. */
#define SPH_NOALLOC_MAX_PARTICLES 32768
#define SPH_NOALLOC_MAX_CELLS 131072
int sph_loop_noalloc(const sph_params_t *params, const sph_state_t *st, double *drho, double *ax, double *ay,
                     int32_t *status, void *stream);

/* ---- diagnostics ---- */
const char *sph_last_error(void);
/* Library build string (compile flags, target). */
const char *sph_version(void);
/* Average duration (ms) of the launches of the dominant kernel (the fused
 * pair pass) since the last reset, measured with HIP events on the stream it
 * ran on; *count receives the number of timed launches. Timing is enabled by
 * sph_timing_enable(ctx, 1). */
int sph_timing_enable(sph_ctx_t *ctx, int32_t on);
int sph_timing_read(sph_ctx_t *ctx, double *avg_ms_pair, double *avg_ms_nn, int64_t *count, int32_t reset);

#ifdef __cplusplus
}
#endif
#endif
