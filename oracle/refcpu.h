/*
 * refcpu — CPU oracle for the WCSPH step of KoningJasper/Offshore-SPH.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP engine
 * (tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it);
 * the product never links, loads or calls it.
 *
 * It is a plain-C, FP64, serial restatement of the reference algorithm. It
 * keeps the reference's data layout (the 154-byte `particle_dtype` record and
 * the 113-byte `computed_dtype` neighbour record, both packed exactly like
 * numpy's unaligned structured dtypes) and its cost structure (a head/next
 * linked-list grid rebuilt every step, one neighbour vector per query, a
 * gathered record per neighbour, per-axis kernel gradients, a serial loop).
 * Every function cites the reference file:line it restates.
 *
 * Floating point: compiled with -ffp-contract=off and without fast-math, so
 * every + - * / sqrt is one IEEE-754 double operation in the order the
 * reference writes it. Integer powers that the reference evaluates with pow()
 * (Tait exponent 7, Lennard-Jones exponents 4 and 2, (2-q)^2) are evaluated as
 * left-to-right products when the exponent is a small integer (ref_ipow);
 * the numba reference runs those call sites with fastmath, so they are only
 * reproducible to a few ulp there anyway.
 *
 * What pins each function (tests/test_oracle_golden.py: the reference's own
 * known answers; tests/test_oracle_invariants.py: mathematics):
 *   cubic W, grad W       cubic KAT (test/test_kernels_cubic.py), 2D
 *                         normalisation, W(0), grad W = dW/dr x/r
 *   gaussian W, dW/dq, grad W   PySPH restatement fixture, normalisation of
 *                         the 3h-truncated kernel (1 - e^-9), W(0), dW/dr
 *   wendland W, grad W    normalisation, W(0), dW/dr
 *   tait, co, B, height   test/test_numba_taiteos.py KAT + closed forms
 *   continuity            test/test_numba_continuity.py KAT, uniform flow
 *                         gives 0, two-particle closed form
 *   momentum              test/test_numba_momentum.py KAT (pressure term),
 *                         two-particle closed forms (pressure; Monaghan AV
 *                         with its branch, sign and alpha-linearity),
 *                         momentum conservation, mirror/rotation symmetry
 *   boundary force        test/test_eq_boundary.py KAT (588.6)
 *   timestep              test/test_eq_courant.py KAT
 *   pec / euler           test/test_integrators_{pec,euler}.py KATs
 *   nn update / near      test/test_linked_list.py KAT (scale 3), brute force
 *                         over the 3x3-cell rule incl. order, cubic-support
 *                         completeness, edge aliasing, IEEE boundary pairs
 *   loop                  dense-pair model (no neighbour list) for all three
 *                         kernels + the invariants above
 *   solver step / run     the pieces above + mask test (walls fixed, fluid
 *                         only integrated)
 *   XSPH (ref_xsph, the   dense-pair model in the all-pairs form of the
 *   loop's velocity)      reference's momentum test (test/test_numba_momentum
 *                         .py:117-123), walls included; Galilean invariance,
 *                         momentum conservation, linearity in epsilon; zero
 *                         without XSPH
 *   summation density     dense-pair model (fluid neighbours, self included),
 *                         lattice normalisation of all three kernels, the
 *                         pressure of the loop taken from the old density,
 *                         zero continuity, momentum with the new density
 *   pec XSPH branch       reduces bit for bit to the plain branch when the
 *                         XSPH velocity equals the one that branch uses
 *   newmark predict /   the textbook Newmark-beta update with a_n and
 *   correct               a_n+1 (closed form); constant acceleration exact;
 *                         an undamped multi-node oscillator (average
 *                         acceleration, beta 1/4, gamma 1/2) conserves its
 *                         discrete energy
 *   solver begin / end    begin + end = step, bit for bit; coupled particles
 *                         integrated only with the coupling integrator
 *   probe pressure        Shepard + summation density of a uniform density
 *                         field returns it (Tait 0 at rho0); brute force over
 *                         the 3x3-cell rule with random densities; no fluid
 *                         neighbour gives rho 0, p = -B
 *   dynamic h (solver   h = 1.3 (m/rho)^(1/2) of the predicted density
 *   step, h_fixed = 0)    after a step, walls unchanged; 1.3 dx on a lattice
 *                         at rho = m / dx^2; 0 for a zero density. The square
 *                         root is the correctly rounded sqrt (the reference's
 *                         fastmath pow(x, 0.5), see ref_solver_step).
 */
#ifndef REFCPU_H
#define REFCPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#pragma pack(push, 1)
/* src/Common.py:26-57 particle_dtype (154 bytes, field order preserved). */
typedef struct {
    uint8_t deleted;
    int8_t label;
    double m, rho, p, c, drho, h, x, y, vx, vy, ax, ay, xsphx, xsphy, x0, y0, vx0, vy0, rho0;
} ref_particle_t;

/* src/Common.py:60-98 computed_dtype (113 bytes, field order preserved). */
typedef struct {
    int8_t label;
    double m, p, rho, h, q, c, r, w, dw_x, dw_y, x, y, vx, vy;
} ref_computed_t;
#pragma pack(pop)

enum { REF_FLUID = 0, REF_BOUNDARY = 1, REF_TEMPBOUNDARY = 2, REF_COUPLED = 3 };
enum { REF_KERNEL_CUBIC = 0, REF_KERNEL_GAUSSIAN = 1 };
enum { REF_INTEGRATOR_PEC = 0, REF_INTEGRATOR_EULER = 1 };

/* src/Methods/WCSPH.py:8-80 — the WCSPH jitclass fields. */
typedef struct {
    int32_t useXSPH, useSummationDensity;
    double height, rho0, r0, D, p1, p2, epsilon, gamma, co, B, Pb, alpha, beta;
} ref_wcsph_t;

/* src/Tools/NNLinkedList.py:7-20 — the NNLinkedList jitclass fields. */
typedef struct {
    double scale;
    double xmin, xmax, ymin, ymax, cell_size;
    int64_t ncx, ncy, n_cells;
    int64_t *heads, *nexts;
    int64_t cap_cells, cap_n;
    int64_t oob_binned; /* particles whose flat key fell outside [0, n_cells) */
} ref_nn_t;

/* Solver-level loop state (src/Solver.py:29-44,366-410) for the parity harness. */
typedef struct {
    int32_t kernel;      /* REF_KERNEL_* */
    int32_t integrator;  /* REF_INTEGRATOR_* */
    int32_t pec_useXSPH, pec_strict;
    double h_fixed;      /* Solver(h=...) ; <= 0 means dynamic h (computeH) */
    double damping;
    double gamma_c, gamma_f;
    double t;
    int64_t t_step;
    int32_t nthreads;    /* 1 = serial like the reference; >1 = OpenMP over fluid */
    int32_t coupled_integrator; /* 1: Newmark-beta on the coupled particles (Solver(coupling=...,
                                 * couplingIntegrator=NewmarkBeta(...)), src/Solver.py:381-398) */
    double nm_beta, nm_gamma;
    double dt;           /* the dt of the current step (set by the first half) */
} ref_solver_t;

/* --- L1: kernels (src/Kernels/CubicSpline.py, src/Kernels/Gaussian.py) --- */
double ref_cubic_evaluate(double r, double h);
double ref_cubic_gradient(double x, double r, double h);
double ref_gaussian_evaluate(double r, double h);
double ref_gaussian_derivative(double r, double h);
double ref_gaussian_gradient(double x, double r, double h);
void ref_kernel_evaluate_n(int32_t kernel, int64_t n, const double *r, const double *h, double *out);
void ref_kernel_gradient_n(int32_t kernel, int64_t n, const double *x, const double *r, const double *h, double *out);
double ref_ipow(double x, double p);

/* --- L1: equations --- */
double ref_tait_eos(double gamma, double B, double rho0, double rho, int8_t label);
double ref_tait_co(double H);
double ref_tait_B(double co, double rho0, double gamma);
double ref_tait_height(double rho0, double H, double B, double gamma, double y);
double ref_continuity(const ref_computed_t *comp, int64_t J);
void ref_momentum(double alpha, double beta, const ref_particle_t *p, const ref_computed_t *comp, int64_t J, double out[2]);
void ref_boundary_force(double r0, double D, double p1, double p2, const ref_computed_t *comp, int64_t J, double out[2]);
void ref_timestep(int64_t J, const ref_particle_t *pA, double gamma_c, double gamma_f, double out[3]);

/* --- L4: method --- */
void ref_wcsph_init(ref_wcsph_t *m, double height, double r0, double rho0, int32_t useXSPH, double Pb, int32_t useSD);

/* --- L5: integrators (src/Integrators/PEC.py, Euler.py) --- */
void ref_pec_predict(int32_t useXSPH, int32_t strict, double dt, ref_particle_t *pA, int64_t J, double damping);
void ref_pec_correct(int32_t useXSPH, int32_t strict, double dt, ref_particle_t *pA, int64_t J, double damping);
void ref_euler_correct(double dt, ref_particle_t *pA, int64_t J, double damping);

/* --- L2: neighbour search (src/Tools/NNLinkedList.py) --- */
void ref_nn_init(ref_nn_t *nn, double scale);
void ref_nn_free(ref_nn_t *nn);
int ref_nn_update(ref_nn_t *nn, const ref_particle_t *pA, int64_t n);
int64_t ref_nn_near_pos(const ref_nn_t *nn, double x, double y, double h, const ref_particle_t *pA,
                        double *h_out, double *q_out, double *r_out, int64_t *idx_out, int64_t cap);
int64_t ref_nn_near(const ref_nn_t *nn, int64_t i, const ref_particle_t *pA,
                    double *h_out, double *q_out, double *r_out, int64_t *idx_out, int64_t cap);
int64_t ref_nn_key(const ref_nn_t *nn, double x, double y);
/* CSR dump of near(i) for every i, for parity. Returns total entries or -needed when cap too small. */
int64_t ref_nn_dump(const ref_nn_t *nn, const ref_particle_t *pA, int64_t n, int64_t *keys,
                    int64_t *off, int64_t *idx, double *q, double *r, int64_t cap);

/* --- L3: interaction loop (src/Tools/SolverTools.py:120-175) --- */
int ref_loop(ref_particle_t *pA, int64_t n, int32_t kernel, const ref_wcsph_t *m, const ref_nn_t *nn, int32_t nthreads);

/* --- L6: one Solver.run loop body (src/Solver.py:366-410, damping given, no settling) --- */
int ref_solver_step(ref_particle_t *pA, int64_t n, ref_solver_t *s, const ref_wcsph_t *m, ref_nn_t *nn, double dt_out[3]);
int ref_solver_run(ref_particle_t *pA, int64_t n, ref_solver_t *s, const ref_wcsph_t *m, ref_nn_t *nn,
                   int64_t nsteps, double *dt_hist);
/* The same body split at the coupling hook (src/Solver.py:389-392): begin = dt,
 * predictors, _compute; end = correctors and the clock. */
int ref_solver_begin(ref_particle_t *pA, int64_t n, ref_solver_t *s, const ref_wcsph_t *m, ref_nn_t *nn,
                     double dt_out[3]);
int ref_solver_end(ref_particle_t *pA, int64_t n, ref_solver_t *s);

/* --- coupling: src/Integrators/NewmarkBeta.py, examples/IceBreak.py:253-285 --- */
void ref_newmark_predict(double beta, double gamma, double dt, ref_particle_t *pA, int64_t J);
void ref_newmark_correct(double beta, double gamma, double dt, ref_particle_t *pA, int64_t J);
int64_t ref_probe_pressure(const ref_nn_t *nn, const ref_particle_t *pA, int64_t n, int32_t kernel,
                           const ref_wcsph_t *m, double x, double y, double h, double out[2]);

#ifdef __cplusplus
}
#endif
#endif
