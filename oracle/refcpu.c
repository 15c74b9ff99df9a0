/*
 * refcpu.c — CPU oracle (test infrastructure only; see refcpu.h).
 *
 * Plain C restatement of the KoningJasper/Offshore-SPH WCSPH step. Each
 * function cites the reference file:line it follows. Build: oracle/Makefile
 * (gcc -O2 -ffp-contract=off, no fast-math).
 */
#include "refcpu.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define REF_PI 3.141592653589793
#define REF_G 9.81

/* Integer powers as left-to-right products; pow() otherwise. */
double ref_ipow(double x, double p) {
    if (p >= 1.0 && p <= 16.0 && p == floor(p)) {
        int n = (int)p;
        double r = x;
        for (int k = 1; k < n; ++k) r = r * x;
        return r;
    }
    if (p == 0.0) return 1.0;
    return pow(x, p);
}

/* ------------------------------------------------------------------------ */
/* Kernels                                                                   */
/* ------------------------------------------------------------------------ */

/* src/Kernels/CubicSpline.py:10-36 */
double ref_cubic_evaluate(double r, double h) {
    const double fac = 10.0 / (7.0 * REF_PI);
    double alpha = fac / (h * h);
    double q = r / h;
    if (q > 2.0) return 0.0;
    if (q > 1.0) return alpha * 0.25 * ref_ipow(2.0 - q, 3.0);
    return alpha * (1.0 - 1.5 * ref_ipow(q, 2.0) * (1.0 - 0.5 * q));
}

/* src/Kernels/CubicSpline.py:38-70 */
double ref_cubic_gradient(double x, double r, double h) {
    const double fac = 10.0 / (7.0 * REF_PI);
    double alpha = fac / (h * h);
    double q = r / h;
    double grad;
    if ((q > 2.0) || (r < 1e-10)) return 0.0;
    if (q > 1.0)
        grad = -0.75 * ref_ipow(2.0 - q, 2.0);
    else
        grad = -3.0 * q * (1.0 - 0.75 * q);
    double norm = alpha * grad / (h * r);
    return norm * x;
}

/* src/Kernels/Gaussian.py:16-25 */
double ref_gaussian_evaluate(double r, double h) {
    const double alpha = 1.0 / REF_PI;
    double q = r / h;
    if (q <= 3.0) {
        double alpha_c = alpha / (h * h);
        return alpha_c * exp(-q * q);
    }
    return 0.0;
}

/* src/Kernels/Gaussian.py:27-37 */
double ref_gaussian_derivative(double r, double h) {
    const double alpha = 1.0 / REF_PI;
    double q = r / h;
    if (q <= 3.0) {
        double alpha_c = alpha / (h * h);
        return -2.0 * q * alpha_c * exp(-q * q);
    }
    return 0.0;
}

/* src/Kernels/Gaussian.py:39-60 */
double ref_gaussian_gradient(double x, double r, double h) {
    const double alpha = 1.0 / REF_PI;
    double tmp = r * h;
    if (tmp > 1e-12) {
        double q = r / h;
        if (q <= 3.0) {
            double alpha_c = alpha / (h * h);
            double dwdq = -2.0 * q * alpha_c * exp(-q * q);
            return dwdq / tmp * x;
        }
        return 0.0;
    }
    return 0.0;
}

/* src/Kernels/Wendland.py:10-30 */
static double ref_wendland_evaluate(double r, double h) {
    const double fac = 9.0 / (4.0 * REF_PI);
    double alpha = fac / (h * h);
    double q = r / h;
    if (q >= 2.0) return 0.0;
    double inner = 1.0 - 0.5 * q;
    return alpha * (ref_ipow(inner, 6.0) * (35.0 / 12.0 * q * q + 3.0 * q + 1.0));
}

/* src/Kernels/Wendland.py:32-64 */
static double ref_wendland_gradient(double x, double r, double h) {
    const double fac = 9.0 / (4.0 * REF_PI);
    double alpha = fac / (h * h);
    double q = r / h;
    if ((q >= 2.0) || (r < 1e-10)) return 0.0;
    double inner = 1.0 - 0.5 * q;
    double grad = ref_ipow(inner, 5.0) * (-14.0 / 3.0) * q * (1 + 2.5 * q);
    double norm = alpha * grad / (h * r);
    return norm * x;
}

static double kernel_w(int32_t kernel, double r, double h) {
    if (kernel == REF_KERNEL_GAUSSIAN) return ref_gaussian_evaluate(r, h);
    if (kernel == 2) return ref_wendland_evaluate(r, h);
    return ref_cubic_evaluate(r, h);
}

static double kernel_g(int32_t kernel, double x, double r, double h) {
    if (kernel == REF_KERNEL_GAUSSIAN) return ref_gaussian_gradient(x, r, h);
    if (kernel == 2) return ref_wendland_gradient(x, r, h);
    return ref_cubic_gradient(x, r, h);
}

void ref_kernel_evaluate_n(int32_t kernel, int64_t n, const double *r, const double *h, double *out) {
    for (int64_t j = 0; j < n; ++j) out[j] = kernel_w(kernel, r[j], h[j]);
}

void ref_kernel_gradient_n(int32_t kernel, int64_t n, const double *x, const double *r, const double *h, double *out) {
    for (int64_t j = 0; j < n; ++j) out[j] = kernel_g(kernel, x[j], r[j], h[j]);
}

/* ------------------------------------------------------------------------ */
/* Equations                                                                 */
/* ------------------------------------------------------------------------ */

/* src/Equations/TaitEOS.py:6-31 */
double ref_tait_eos(double gamma, double B, double rho0, double rho, int8_t label) {
    if (label != REF_FLUID) return 0.0;
    double ratio = ref_ipow(rho / rho0, gamma);
    return (ratio - 1.0) * B;
}

/* src/Equations/TaitEOS.py:40-44 */
double ref_tait_co(double H) { return 10.0 * sqrt(2.0 * REF_G * H); }

/* src/Equations/TaitEOS.py:33-38 */
double ref_tait_B(double co, double rho0, double gamma) { return co * co * rho0 / gamma; }

/* src/Equations/TaitEOS.py:46-65 */
double ref_tait_height(double rho0, double H, double B, double gamma, double y) {
    double frac = rho0 * REF_G * (H - y) / B;
    return rho0 * pow(1.0 + frac, 1.0 / gamma);
}

/* src/Equations/Continuity.py:4-16 */
double ref_continuity(const ref_computed_t *comp, int64_t J) {
    double arho = 0.0;
    for (int64_t j = 0; j < J; ++j) {
        if (comp[j].label != REF_FLUID) continue;
        double dot = comp[j].vx * comp[j].dw_x + comp[j].vy * comp[j].dw_y;
        arho += comp[j].m * dot;
    }
    return arho;
}

/* src/Equations/Momentum.py:5-57 */
void ref_momentum(double alpha, double beta, const ref_particle_t *p, const ref_computed_t *comp, int64_t J,
                  double out[2]) {
    double slf = p->p / (p->rho * p->rho);
    double a0 = 0.0, a1 = 0.0;
    for (int64_t j = 0; j < J; ++j) {
        if (comp[j].label != REF_FLUID) continue;
        double othr = comp[j].p / (comp[j].rho * comp[j].rho);
        double dot = comp[j].vx * comp[j].x + comp[j].vy * comp[j].y;
        double PI_ij = 0.0;
        if (dot < 0) {
            double hij = 0.5 * (p->h + comp[j].h);
            double cij = 0.5 * (p->c + comp[j].c);
            double rhoij = 0.5 * (p->rho + comp[j].rho);
            double muij = hij * dot / (comp[j].r * comp[j].r + 0.01 * hij * hij);
            double ppij = muij * (beta * muij - alpha * cij);
            PI_ij = ppij / rhoij;
        }
        double factor = slf + othr + PI_ij;
        a0 += -comp[j].m * factor * comp[j].dw_x;
        a1 += -comp[j].m * factor * comp[j].dw_y;
    }
    out[0] = a0;
    out[1] = a1;
}

/* src/Equations/BoundaryForce.py:6-42 */
void ref_boundary_force(double r0, double D, double p1, double p2, const ref_computed_t *comp, int64_t J,
                        double out[2]) {
    double f0 = 0.0, f1 = 0.0;
    for (int64_t j = 0; j < J; ++j) {
        if ((comp[j].label == REF_FLUID) || (comp[j].r > r0)) continue;
        if (comp[j].r > 1e-12) {
            double frac = r0 / comp[j].r;
            double tmp = ref_ipow(frac, p1) - ref_ipow(frac, p2);
            double fac = D * tmp;
            f0 += fac * comp[j].x / ref_ipow(comp[j].r, 2.0);
            f1 += fac * comp[j].y / ref_ipow(comp[j].r, 2.0);
        }
    }
    out[0] = f0;
    out[1] = f1;
}

/* src/Equations/XSPH.py:5-31 */
static void ref_xsph(double epsilon, const ref_particle_t *p, const ref_computed_t *comp, int64_t J, double out[2]) {
    double x0 = 0.0, x1 = 0.0;
    for (int64_t j = 0; j < J; ++j) {
        double rho_ij = 0.5 * (p->rho + comp[j].rho);
        double fac = -epsilon * comp[j].m * comp[j].w / rho_ij;
        x0 += fac * comp[j].vx;
        x1 += fac * comp[j].vy;
    }
    out[0] = x0;
    out[1] = x1;
}

/* src/Equations/TimeStep.py:11-90 (jitclass: IEEE, no fastmath) */
void ref_timestep(int64_t J, const ref_particle_t *pA, double gamma_c, double gamma_f, double out[3]) {
    double min_h = INFINITY, max_c = -INFINITY, max_a2 = -INFINITY;
    int64_t cnt = 0;
    for (int64_t j = 0; j < J; ++j) {
        if (pA[j].label != REF_FLUID) continue;
        double a2 = pA[j].ax * pA[j].ax + pA[j].ay * pA[j].ay;
        if (pA[j].h < min_h) min_h = pA[j].h;
        if (pA[j].c > max_c) max_c = pA[j].c;
        if (a2 > max_a2) max_a2 = a2;
        ++cnt;
    }
    if (cnt == 0) {
        out[0] = out[1] = out[2] = NAN; /* np.min of an empty array raises in the reference */
        return;
    }
    double c = gamma_c * min_h / max_c;                              /* courant :47-49 */
    double f = (max_a2 < 1e-12) ? 1e10 : gamma_f * sqrt(min_h / max_a2); /* force :51-56 */
    out[0] = (f < c) ? f : c;                                        /* min(c, f) :45 */
    out[1] = c;
    out[2] = f;
}

/* ------------------------------------------------------------------------ */
/* Method                                                                    */
/* ------------------------------------------------------------------------ */

/* src/Methods/WCSPH.py:35-79 */
void ref_wcsph_init(ref_wcsph_t *m, double height, double r0, double rho0, int32_t useXSPH, double Pb, int32_t useSD) {
    memset(m, 0, sizeof(*m));
    m->height = height;
    m->rho0 = rho0;
    m->useXSPH = useXSPH;
    m->useSummationDensity = useSD;
    m->epsilon = 0.5;
    m->gamma = 7.0;
    m->co = ref_tait_co(m->height);
    m->B = ref_tait_B(m->co, m->rho0, m->gamma);
    m->Pb = Pb;
    m->alpha = 0.01;
    m->beta = 0.0;
    m->r0 = r0;
    m->D = 5 * REF_G * m->height;
    m->p1 = 4;
    m->p2 = 2;
}

/* ------------------------------------------------------------------------ */
/* Integrators (jitclasses: IEEE, no fastmath)                               */
/* ------------------------------------------------------------------------ */

/* src/Integrators/PEC.py:32-59 */
void ref_pec_predict(int32_t useXSPH, int32_t strict, double dt, ref_particle_t *pA, int64_t J, double damping) {
    for (int64_t j = 0; j < J; ++j) {
        ref_particle_t *p = &pA[j];
        p->x0 = p->x;
        p->y0 = p->y;
        p->vx0 = p->vx;
        p->vy0 = p->vy;
        if (useXSPH && p->label == REF_FLUID) {
            p->x = p->x + 0.5 * dt * p->xsphx;
            p->y = p->y + 0.5 * dt * p->xsphy;
        } else {
            p->x = p->x + 0.5 * dt * p->vx;
            p->y = p->y + 0.5 * dt * p->vy;
        }
        p->vx = (p->vx + 0.5 * dt * p->ax) / (1 + 0.5 * damping);
        p->vy = (p->vy + 0.5 * dt * p->ay) / (1 + 0.5 * damping);
        p->rho0 = p->rho;
        p->rho = p->rho + 0.5 * dt * p->drho;
        if (strict && p->rho < 0.0) p->rho = 0.0;
    }
}

/* src/Integrators/PEC.py:61-89 */
void ref_pec_correct(int32_t useXSPH, int32_t strict, double dt, ref_particle_t *pA, int64_t J, double damping) {
    for (int64_t j = 0; j < J; ++j) {
        ref_particle_t *p = &pA[j];
        double mid_x, mid_y;
        if (useXSPH && p->label == REF_FLUID) {
            mid_x = p->x0 + 0.5 * dt * p->xsphx;
            mid_y = p->y0 + 0.5 * dt * p->xsphy;
        } else {
            mid_x = p->x0 + 0.5 * dt * p->vx;
            mid_y = p->y0 + 0.5 * dt * p->vy;
        }
        double mid_vx = (p->vx0 + 0.5 * dt * p->ax) / (1 + 0.5 * damping);
        double mid_vy = (p->vy0 + 0.5 * dt * p->ay) / (1 + 0.5 * damping);
        p->x = 2 * mid_x - p->x0;
        p->y = 2 * mid_y - p->y0;
        p->vx = 2 * mid_vx - p->vx0;
        p->vy = 2 * mid_vy - p->vy0;
        double mid_rho = p->rho0 + 0.5 * dt * p->drho;
        p->rho = 2 * mid_rho - p->rho0;
        if (strict && p->rho < 0.0) p->rho = 0.0;
    }
}

/* src/Integrators/Euler.py:17-26 (predict at :13-15 is a no-op) */
void ref_euler_correct(double dt, ref_particle_t *pA, int64_t J, double damping) {
    (void)damping;
    for (int64_t j = 0; j < J; ++j) {
        ref_particle_t *p = &pA[j];
        p->x = p->x + dt * p->vx + 0.5 * dt * dt * p->ax;
        p->y = p->y + dt * p->vy + 0.5 * dt * dt * p->ay;
        p->vx = p->vx + dt * p->ax;
        p->vy = p->vy + dt * p->ay;
        p->rho = p->rho + dt * p->drho;
    }
}

/* ------------------------------------------------------------------------ */
/* Neighbour search: src/Tools/NNLinkedList.py                               */
/* ------------------------------------------------------------------------ */

void ref_nn_init(ref_nn_t *nn, double scale) {
    memset(nn, 0, sizeof(*nn));
    nn->scale = scale; /* :23-35 */
}

void ref_nn_free(ref_nn_t *nn) {
    free(nn->heads);
    free(nn->nexts);
    nn->heads = nn->nexts = NULL;
    nn->cap_cells = nn->cap_n = 0;
}

/* :164-176 _real_to_int */
static int64_t real_to_int(double val, double step) { return (int64_t)floor(val / step); }

int64_t ref_nn_key(const ref_nn_t *nn, double x, double y) {
    int64_t cx = real_to_int(x - nn->xmin, nn->cell_size);
    int64_t cy = real_to_int(y - nn->ymin, nn->cell_size);
    return cx + nn->ncx * cy; /* :200-208 _flatten_raw */
}

/* :37-39 update -> :86-101 _init, :103-110 _get_cell_size, :112-127
 * _get_number_of_cells, :129-141 _bin. A flat key outside [0, n_cells) is an
 * out-of-bounds write in the reference (no check at :136-141); here such a
 * particle is left out of every cell list and counted in oob_binned. */
int ref_nn_update(ref_nn_t *nn, const ref_particle_t *pA, int64_t n) {
    if (n <= 0) return -1;
    double xmin = pA[0].x, xmax = pA[0].x, ymin = pA[0].y, ymax = pA[0].y, hmin = pA[0].h;
    for (int64_t i = 1; i < n; ++i) {
        if (pA[i].x < xmin) xmin = pA[i].x;
        if (pA[i].x > xmax) xmax = pA[i].x;
        if (pA[i].y < ymin) ymin = pA[i].y;
        if (pA[i].y > ymax) ymax = pA[i].y;
        if (pA[i].h < hmin) hmin = pA[i].h;
    }
    nn->xmin = xmin;
    nn->xmax = xmax;
    nn->ymin = ymin;
    nn->ymax = ymax;
    double cell_size = hmin * nn->scale;
    if (cell_size < 1e-6) cell_size = 1.0;
    nn->cell_size = cell_size;
    double cell_size1 = 1. / nn->cell_size;
    int64_t ncx = (int64_t)ceil(cell_size1 * (nn->xmax - nn->xmin));
    int64_t ncy = (int64_t)ceil(cell_size1 * (nn->ymax - nn->ymin));
    if (ncx < 1) ncx = 1;
    if (ncy < 1) ncy = 1;
    nn->ncx = ncx;
    nn->ncy = ncy;
    nn->n_cells = ncx * ncy;
    if (nn->n_cells > nn->cap_cells) {
        free(nn->heads);
        nn->heads = (int64_t *)malloc(sizeof(int64_t) * (size_t)nn->n_cells);
        nn->cap_cells = nn->n_cells;
    }
    if (n > nn->cap_n) {
        free(nn->nexts);
        nn->nexts = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
        nn->cap_n = n;
    }
    if (!nn->heads || !nn->nexts) return -2;
    for (int64_t c = 0; c < nn->n_cells; ++c) nn->heads[c] = -1;
    for (int64_t i = 0; i < n; ++i) nn->nexts[i] = -1;
    nn->oob_binned = 0;
    for (int64_t i = 0; i < n; ++i) {
        int64_t cid = ref_nn_key(nn, pA[i].x, pA[i].y);
        if (cid < 0 || cid >= nn->n_cells) {
            nn->oob_binned++;
            continue;
        }
        nn->nexts[i] = nn->heads[cid];
        nn->heads[cid] = i;
    }
    return 0;
}

/* :178-198 _get_valid_cell_index */
static int64_t valid_cell_index(const ref_nn_t *nn, int64_t cid_x, int64_t cid_y) {
    int64_t cell_index = -1;
    if ((nn->ncx > cid_x && cid_x > -1) && (nn->ncy > cid_y && cid_y > -1)) {
        cell_index = cid_x + nn->ncx * cid_y;
        if (!(-1 < cell_index && cell_index < nn->n_cells)) cell_index = -1;
    }
    return cell_index;
}

/* :41-80 nearPos. Returns the number of neighbours (all of them, even past
 * cap; only the first cap are written). */
int64_t ref_nn_near_pos(const ref_nn_t *nn, double x, double y, double h, const ref_particle_t *pA,
                        double *h_out, double *q_out, double *r_out, int64_t *idx_out, int64_t cap) {
    static const int64_t shifts[3] = {-1, 0, 1};
    int64_t _cid_x = real_to_int(x - nn->xmin, nn->cell_size);
    int64_t _cid_y = real_to_int(y - nn->ymin, nn->cell_size);
    int64_t k = 0;
    for (int ix = 0; ix < 3; ++ix) {
        for (int iy = 0; iy < 3; ++iy) {
            int64_t cell_index = valid_cell_index(nn, _cid_x + shifts[ix], _cid_y + shifts[iy]);
            if (cell_index > -1) {
                int64_t _next = nn->heads[cell_index];
                while (_next != -1) {
                    double dx = x - pA[_next].x;
                    double dy = y - pA[_next].y;
                    double xij2 = dx * dx + dy * dy; /* :143-144 _norm2 */
                    double r = sqrt(xij2);
                    double h_ij = 0.5 * (h + pA[_next].h);
                    double q_ij = r / h_ij;
                    if (q_ij <= 3.0) {
                        if (k < cap) {
                            if (idx_out) idx_out[k] = _next;
                            if (r_out) r_out[k] = r;
                            if (q_out) q_out[k] = q_ij;
                            if (h_out) h_out[k] = h_ij;
                        }
                        ++k;
                    }
                    _next = nn->nexts[_next];
                }
            }
        }
    }
    return k;
}

/* :82-84 near */
int64_t ref_nn_near(const ref_nn_t *nn, int64_t i, const ref_particle_t *pA,
                    double *h_out, double *q_out, double *r_out, int64_t *idx_out, int64_t cap) {
    return ref_nn_near_pos(nn, pA[i].x, pA[i].y, pA[i].h, pA, h_out, q_out, r_out, idx_out, cap);
}

int64_t ref_nn_dump(const ref_nn_t *nn, const ref_particle_t *pA, int64_t n, int64_t *keys,
                    int64_t *off, int64_t *idx, double *q, double *r, int64_t cap) {
    int64_t total = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (keys) keys[i] = ref_nn_key(nn, pA[i].x, pA[i].y);
        if (off) off[i] = total;
        int64_t room = cap - total;
        if (room < 0) room = 0;
        int64_t k = ref_nn_near(nn, i, pA, NULL, q ? q + total : NULL, r ? r + total : NULL,
                                idx ? idx + total : NULL, room);
        total += k;
    }
    if (off) off[n] = total;
    return total <= cap ? total : -total;
}

/* ------------------------------------------------------------------------ */
/* Interaction loop: src/Tools/SolverTools.py                                */
/* ------------------------------------------------------------------------ */

typedef struct {
    double *h, *q, *r;
    int64_t *idx;
    ref_computed_t *comp;
    double *w, *dwx, *dwy;
    int64_t cap;
} ref_scratch_t;

static int scratch_reserve(ref_scratch_t *s, int64_t k) {
    if (k <= s->cap) return 0;
    int64_t c = k + 64;
    s->h = (double *)realloc(s->h, sizeof(double) * c);
    s->q = (double *)realloc(s->q, sizeof(double) * c);
    s->r = (double *)realloc(s->r, sizeof(double) * c);
    s->w = (double *)realloc(s->w, sizeof(double) * c);
    s->dwx = (double *)realloc(s->dwx, sizeof(double) * c);
    s->dwy = (double *)realloc(s->dwy, sizeof(double) * c);
    s->idx = (int64_t *)realloc(s->idx, sizeof(int64_t) * c);
    s->comp = (ref_computed_t *)realloc(s->comp, sizeof(ref_computed_t) * c);
    s->cap = c;
    return (s->h && s->q && s->r && s->w && s->dwx && s->dwy && s->idx && s->comp) ? 0 : -1;
}

static void scratch_free(ref_scratch_t *s) {
    free(s->h); free(s->q); free(s->r); free(s->w); free(s->dwx); free(s->dwy); free(s->idx); free(s->comp);
    memset(s, 0, sizeof(*s));
}

/* :74-104 _assignProps, then :8-44 computeProps (kernel per array, gradient per axis). */
static void compute_props(int64_t i, const ref_particle_t *pA, int64_t J, ref_scratch_t *s, int32_t kernel) {
    for (int64_t j = 0; j < J; ++j) {
        const ref_particle_t *o = &pA[s->idx[j]];
        ref_computed_t *c = &s->comp[j];
        memset(c, 0, sizeof(*c));
        c->label = o->label;
        c->p = o->p;
        c->m = o->m;
        c->rho = o->rho;
        c->h = s->h[j];
        c->q = s->q[j];
        c->r = s->r[j];
        c->x = pA[i].x - o->x;
        c->y = pA[i].y - o->y;
        c->vx = pA[i].vx - o->vx;
        c->vy = pA[i].vy - o->vy;
    }
    for (int64_t j = 0; j < J; ++j) s->w[j] = kernel_w(kernel, s->comp[j].r, s->comp[j].h);
    for (int64_t j = 0; j < J; ++j) s->dwx[j] = kernel_g(kernel, s->comp[j].x, s->comp[j].r, s->comp[j].h);
    for (int64_t j = 0; j < J; ++j) s->dwy[j] = kernel_g(kernel, s->comp[j].y, s->comp[j].r, s->comp[j].h);
    for (int64_t j = 0; j < J; ++j) {
        s->comp[j].w = s->w[j];
        s->comp[j].dw_x = s->dwx[j];
        s->comp[j].dw_y = s->dwy[j];
    }
}

static int64_t near_into(const ref_nn_t *nn, int64_t i, const ref_particle_t *pA, ref_scratch_t *s) {
    int64_t k = ref_nn_near(nn, i, pA, s->h, s->q, s->r, s->idx, s->cap);
    if (k > s->cap) {
        if (scratch_reserve(s, k)) return -1;
        k = ref_nn_near(nn, i, pA, s->h, s->q, s->r, s->idx, s->cap);
    }
    return k;
}

/* Per-fluid-particle body of :160-173 (_loop regular loop). */
static int loop_one(ref_particle_t *pA, int64_t i, int32_t kernel, const ref_wcsph_t *m, const ref_nn_t *nn,
                    ref_scratch_t *s) {
    int64_t J = near_into(nn, i, pA, s);
    if (J < 0) return -1;
    if (J == 0) return 0; /* :166-169 early exit keeps previous properties */
    compute_props(i, pA, J, s, kernel);
    ref_particle_t *p = &pA[i];
    /* WCSPH.compute_density_change (src/Methods/WCSPH.py:191-203) */
    double drho = m->useSummationDensity ? 0.0 : ref_continuity(s->comp, J);
    /* WCSPH.compute_acceleration (src/Methods/WCSPH.py:151-169) */
    double a[2];
    ref_momentum(m->alpha, m->beta, p, s->comp, J, a);
    a[1] = a[1] - 9.81;
    double b[2];
    ref_boundary_force(m->r0, m->D, m->p1, m->p2, s->comp, J, b);
    p->drho = drho;
    p->ax = a[0] + b[0];
    p->ay = a[1] + b[1];
    /* WCSPH.compute_velocity (src/Methods/WCSPH.py:171-189) */
    if (m->useXSPH) {
        double xs[2];
        ref_xsph(m->epsilon, p, s->comp, J, xs);
        double vx = p->vx, vy = p->vy;
        p->xsphx = vx + xs[0];
        p->xsphy = vy + xs[1];
    } else {
        p->xsphx = 0.0;
        p->xsphy = 0.0;
    }
    return 0;
}

/* :120-175 _loop */
int ref_loop(ref_particle_t *pA, int64_t n, int32_t kernel, const ref_wcsph_t *m, const ref_nn_t *nn, int32_t nthreads) {
    /* :122-127 pressure (Tait + Pb) and speed of sound for every active particle */
    for (int64_t i = 0; i < n; ++i) {
        pA[i].p = ref_tait_eos(m->gamma, m->B, m->rho0, pA[i].rho, pA[i].label) + m->Pb;
        pA[i].c = m->co;
    }
    int err = 0;
    if (m->useSummationDensity) {
        /* :129-140 summation density (pressure above is already computed: p lags rho) */
        ref_scratch_t s;
        memset(&s, 0, sizeof(s));
        scratch_reserve(&s, 128);
        for (int64_t i = 0; i < n; ++i) {
            if (pA[i].label != REF_FLUID) continue;
            int64_t J = near_into(nn, i, pA, &s);
            if (J < 0) { err = -1; break; }
            double rho = 0.0;
            for (int64_t j = 0; j < J; ++j) {
                const ref_particle_t *o = &pA[s.idx[j]];
                double w = kernel_w(kernel, s.r[j], s.h[j]);
                if (o->label != REF_FLUID) continue;
                rho += o->m * w;
            }
            pA[i].rho = rho;
        }
        scratch_free(&s);
        if (err) return err;
    }
    if (nthreads <= 1) {
        ref_scratch_t s;
        memset(&s, 0, sizeof(s));
        scratch_reserve(&s, 128);
        for (int64_t i = 0; i < n; ++i) {
            if (pA[i].label != REF_FLUID) continue;
            if (loop_one(pA, i, kernel, m, nn, &s)) { err = -1; break; }
        }
        scratch_free(&s);
    } else {
        /* Each fluid particle writes only its own record and reads p, rho, x,
         * y, vx, vy, m, h, label of others, none of which the loop writes, so
         * the parallel loop gives the serial results bit for bit. */
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
        {
            ref_scratch_t s;
            memset(&s, 0, sizeof(s));
            scratch_reserve(&s, 128);
#pragma omp for schedule(dynamic, 256)
            for (int64_t i = 0; i < n; ++i) {
                if (pA[i].label != REF_FLUID) continue;
                if (loop_one(pA, i, kernel, m, nn, &s)) err = -1;
            }
            scratch_free(&s);
        }
#else
        return ref_loop(pA, n, kernel, m, nn, 1);
#endif
    }
    return err;
}

/* ------------------------------------------------------------------------ */
/* Solver loop body: src/Solver.py:366-410 with the masked copies of        */
/* :217,238,254,380,396 (fluid copy for dt/predict/correct, active copy for */
/* the neighbour update and the loop).                                      */
/* ------------------------------------------------------------------------ */

typedef struct {
    ref_particle_t *fluid, *active, *coupled;
    int64_t *fidx, *aidx, *cidx;
    int64_t nf, na, nc, cap;
} ref_masks_t;

static void masks_free(ref_masks_t *mk) {
    free(mk->fluid); free(mk->active); free(mk->coupled);
    free(mk->fidx); free(mk->aidx); free(mk->cidx);
}

/* f_indexes, c_indexes and indexes (src/Solver.py:178-181) */
static int masks_build(ref_masks_t *mk, const ref_particle_t *pA, int64_t n) {
    if (n > mk->cap) {
        masks_free(mk);
        mk->fluid = (ref_particle_t *)malloc(sizeof(ref_particle_t) * n);
        mk->active = (ref_particle_t *)malloc(sizeof(ref_particle_t) * n);
        mk->coupled = (ref_particle_t *)malloc(sizeof(ref_particle_t) * n);
        mk->fidx = (int64_t *)malloc(sizeof(int64_t) * n);
        mk->aidx = (int64_t *)malloc(sizeof(int64_t) * n);
        mk->cidx = (int64_t *)malloc(sizeof(int64_t) * n);
        mk->cap = n;
        if (!mk->fluid || !mk->active || !mk->coupled || !mk->fidx || !mk->aidx || !mk->cidx) return -1;
    }
    mk->nf = mk->na = mk->nc = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (pA[i].deleted) continue;
        mk->aidx[mk->na++] = i;
        if (pA[i].label == REF_FLUID) mk->fidx[mk->nf++] = i;
        if (pA[i].label == REF_COUPLED) mk->cidx[mk->nc++] = i;
    }
    return 0;
}

static void gather(ref_particle_t *dst, const ref_particle_t *src, const int64_t *idx, int64_t k) {
    for (int64_t j = 0; j < k; ++j) dst[j] = src[idx[j]];
}

static void scatter(ref_particle_t *dst, const ref_particle_t *src, const int64_t *idx, int64_t k) {
    for (int64_t j = 0; j < k; ++j) dst[idx[j]] = src[j];
}

/* The loop body up to the coupling hook (src/Solver.py:366-392): dt, the
 * predictors, _compute. The dt is kept in s->dt for the second half. */
static int solver_begin_masked(ref_particle_t *pA, ref_solver_t *s, const ref_wcsph_t *m, ref_nn_t *nn,
                               ref_masks_t *mk, double dt_out[3]) {
    double ts[3];
    /* :210-230 _minTimeStep on the fluid copy */
    gather(mk->fluid, pA, mk->fidx, mk->nf);
    ref_timestep(mk->nf, mk->fluid, s->gamma_c, s->gamma_f, ts);
    double dt = ts[0];
    s->dt = dt;
    if (dt_out) { dt_out[0] = ts[0]; dt_out[1] = ts[1]; dt_out[2] = ts[2]; }
    /* :380 predict on the fluid copy */
    if (s->integrator == REF_INTEGRATOR_PEC) {
        ref_pec_predict(s->pec_useXSPH, s->pec_strict, dt, mk->fluid, mk->nf, s->damping);
        scatter(pA, mk->fluid, mk->fidx, mk->nf);
    }
    /* :381-382 the coupling integrator's predict on the coupled copy */
    if (s->coupled_integrator) {
        gather(mk->coupled, pA, mk->cidx, mk->nc);
        ref_newmark_predict(s->nm_beta, s->nm_gamma, dt, mk->coupled, mk->nc);
        scatter(pA, mk->coupled, mk->cidx, mk->nc);
    }
    /* :232-256 _compute: nn.update on the active copy, h for fluid, _loop */
    gather(mk->active, pA, mk->aidx, mk->na);
    if (ref_nn_update(nn, mk->active, mk->na)) return -1;
    if (s->h_fixed > 0) {
        for (int64_t j = 0; j < mk->nf; ++j) pA[mk->fidx[j]].h = s->h_fixed;
    } else {
        /* SolverTools.computeH :106-118 with sigma 1.3, d = 2: (m/rho) ** (1/d)
         * in a fastmath=True numba function, where the exponent folds to 0.5
         * and LLVM's fast-math library-call simplification replaces
         * pow(x, 0.5) by sqrt(x). The correctly rounded sqrt is used here;
         * glibc's pow(x, 0.5) differs from it in ~0.08 % of arguments. */
        for (int64_t j = 0; j < mk->nf; ++j) {
            ref_particle_t *p = &pA[mk->fidx[j]];
            p->h = (p->rho > 1e-12) ? 1.3 * sqrt(p->m / p->rho) : 0.0;
        }
    }
    gather(mk->active, pA, mk->aidx, mk->na);
    if (ref_loop(mk->active, mk->na, s->kernel, m, nn, s->nthreads)) return -1;
    scatter(pA, mk->active, mk->aidx, mk->na);
    return 0;
}

/* The loop body after the coupling hook (src/Solver.py:394-410): the
 * correctors with the dt of the first half, then the clock. */
static void solver_end_masked(ref_particle_t *pA, ref_solver_t *s, ref_masks_t *mk) {
    const double dt = s->dt;
    /* :396 correct on the fluid copy */
    gather(mk->fluid, pA, mk->fidx, mk->nf);
    if (s->integrator == REF_INTEGRATOR_PEC)
        ref_pec_correct(s->pec_useXSPH, s->pec_strict, dt, mk->fluid, mk->nf, s->damping);
    else
        ref_euler_correct(dt, mk->fluid, mk->nf, s->damping);
    scatter(pA, mk->fluid, mk->fidx, mk->nf);
    /* :397-398 the coupling integrator's correct on the coupled copy */
    if (s->coupled_integrator) {
        gather(mk->coupled, pA, mk->cidx, mk->nc);
        ref_newmark_correct(s->nm_beta, s->nm_gamma, dt, mk->coupled, mk->nc);
        scatter(pA, mk->coupled, mk->cidx, mk->nc);
    }
    s->t += dt;
    s->t_step += 1;
}

int ref_solver_step(ref_particle_t *pA, int64_t n, ref_solver_t *s, const ref_wcsph_t *m, ref_nn_t *nn, double dt_out[3]) {
    ref_masks_t mk;
    memset(&mk, 0, sizeof(mk));
    int rc = masks_build(&mk, pA, n);
    if (!rc) rc = solver_begin_masked(pA, s, m, nn, &mk, dt_out);
    if (!rc) solver_end_masked(pA, s, &mk);
    masks_free(&mk);
    return rc;
}

int ref_solver_run(ref_particle_t *pA, int64_t n, ref_solver_t *s, const ref_wcsph_t *m, ref_nn_t *nn,
                   int64_t nsteps, double *dt_hist) {
    ref_masks_t mk;
    memset(&mk, 0, sizeof(mk));
    int rc = masks_build(&mk, pA, n);
    for (int64_t k = 0; k < nsteps && !rc; ++k) {
        rc = solver_begin_masked(pA, s, m, nn, &mk, dt_hist ? dt_hist + 3 * k : NULL);
        if (!rc) solver_end_masked(pA, s, &mk);
    }
    masks_free(&mk);
    return rc;
}

int ref_solver_begin(ref_particle_t *pA, int64_t n, ref_solver_t *s, const ref_wcsph_t *m, ref_nn_t *nn,
                     double dt_out[3]) {
    ref_masks_t mk;
    memset(&mk, 0, sizeof(mk));
    int rc = masks_build(&mk, pA, n);
    if (!rc) rc = solver_begin_masked(pA, s, m, nn, &mk, dt_out);
    masks_free(&mk);
    return rc;
}

int ref_solver_end(ref_particle_t *pA, int64_t n, ref_solver_t *s) {
    ref_masks_t mk;
    memset(&mk, 0, sizeof(mk));
    int rc = masks_build(&mk, pA, n);
    if (!rc) solver_end_masked(pA, s, &mk);
    masks_free(&mk);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Coupling: Newmark-beta (src/Integrators/NewmarkBeta.py) and the SPH      */
/* pressure probe of the coupled IceBreak example (examples/IceBreak.py).   */
/* ------------------------------------------------------------------------ */

/* NewmarkBeta.predict :41-52 (damping is not used there) */
void ref_newmark_predict(double beta, double gamma, double dt, ref_particle_t *pA, int64_t J) {
    for (int64_t j = 0; j < J; ++j) {
        ref_particle_t *p = &pA[j];
        p->x = p->x + p->vx * dt + p->ax * (0.5 - beta) * (dt * dt);
        p->y = p->y + p->vy * dt + p->ay * (0.5 - beta) * (dt * dt);
        p->vx = p->vx + p->ax * (1 - gamma) * dt;
        p->vy = p->vy + p->ay * (1 - gamma) * dt;
    }
}

/* NewmarkBeta.correct :54-66 */
void ref_newmark_correct(double beta, double gamma, double dt, ref_particle_t *pA, int64_t J) {
    for (int64_t j = 0; j < J; ++j) {
        ref_particle_t *p = &pA[j];
        p->x = p->x + p->ax * beta * (dt * dt);
        p->y = p->y + p->ay * beta * (dt * dt);
        p->vx = p->vx + p->ax * gamma * dt;
        p->vy = p->vy + p->ay * gamma * dt;
    }
}

/* pressure_SPH (examples/IceBreak.py:253-285) at the point (x, y): the
 * neighbours of nearPos (src/Tools/NNLinkedList.py:41-80) against the last
 * update, W(r, h) with the probe's own h (:274), the Shepard-normalised
 * weights (src/Equations/Shepard.py:5-12: fluid with rho >= 1e-3 in the
 * normaliser), the summation density over the fluid
 * (src/Equations/SummationDensity.py:5-13) and the Tait pressure
 * ((rho / rho0)^gamma - 1) B (:282-284). out = (rho, p); returns the
 * neighbour count or -1. */
int64_t ref_probe_pressure(const ref_nn_t *nn, const ref_particle_t *pA, int64_t n, int32_t kernel,
                           const ref_wcsph_t *m, double x, double y, double h, double out[2]) {
    int64_t J = ref_nn_near_pos(nn, x, y, h, pA, NULL, NULL, NULL, NULL, 0);
    double *r = (double *)malloc(sizeof(double) * (J > 0 ? J : 1));
    int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (J > 0 ? J : 1));
    if (!r || !idx) { free(r); free(idx); return -1; }
    ref_nn_near_pos(nn, x, y, h, pA, NULL, NULL, r, idx, J);
    (void)n;
    double norm = 0.0;
    for (int64_t j = 0; j < J; ++j) {
        const ref_particle_t *b = &pA[idx[j]];
        if (b->rho < 1e-3 || b->label != REF_FLUID) continue;
        norm += kernel_w(kernel, r[j], h) * b->m / b->rho;
    }
    double rho = 0.0;
    for (int64_t j = 0; j < J; ++j) {
        const ref_particle_t *b = &pA[idx[j]];
        if (b->label != REF_FLUID) continue;
        rho += b->m * (kernel_w(kernel, r[j], h) / norm);
    }
    out[0] = rho;
    out[1] = ref_tait_eos(m->gamma, m->B, m->rho0, rho, REF_FLUID);
    free(r);
    free(idx);
    return J;
}
