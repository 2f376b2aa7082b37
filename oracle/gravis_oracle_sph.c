/*
 * gravis_oracle_sph.c -- CPU oracle for sim-core's SPH step (test
 * infrastructure only; see gravis_oracle.c for the rules).
 *
 * Restates, op for op in IEEE f64 (-ffp-contract=off):
 *   kernel     crates/sim-core/src/sph/kernel.rs:17-108 (Wendland C2, q<2)
 *   neighbours crates/sim-core/src/sph/neighbors.rs:56-119 and the ball
 *              query of octree.rs:245-308 (order = reference tree order)
 *   density    crates/sim-core/src/sph/density.rs:42-157 (fixed-point h,
 *              <=5 iterations, tol 1e-3, re-query when 2h_new > 2.2h)
 *   forces     crates/sim-core/src/sph/forces.rs:69-232 (grad-h pressure,
 *              Monaghan AV, Price conductivity, Morris-Monaghan switch)
 *   solver     crates/sim-core/src/sph/solver.rs:61-212 (compute,
 *              compute_timestep, step_with_sph)
 *   ICs        crates/sim-core/src/scenarios/evrard_collapse.rs:45-71,
 *              scenarios/mod.rs:17-51
 *
 * The same functions carry a build-defined physics variant (no reference
 * counterpart), selected by the params' knobs; the defaults reproduce
 * sim-core exactly:
 *   kernel 1      cubic-spline M4, W = (1/pi h^3) f(q), f = 1 - 1.5q^2 +
 *                 0.75q^3 (q<1), 0.25(2-q)^3 (1<=q<2)
 *   h_solver 1    Newton-Raphson on rho(h) - m (eta/h)^3 = 0 to |dh|/h <
 *                 h_tol, falling back to the fixed-point step when the
 *                 derivative is not positive, steps clamped to [h/2, 2h];
 *                 re-query when 2 h_new exceeds the list's radius
 *   alpha_switch 0  alpha fixed (no Morris-Monaghan switch)
 *   soft_h 1      gravity softening eps_i = h_i of the target
 *   ICs           orc_evrard_1r: rho ~ 1/r (r = R sqrt(U)), h from the
 *                 analytic density
 * pair_fp32 only concerns the device (FP32 pair loops); the oracle is FP64.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SPH_PI 3.14159265358979323846
#define NORM_3D (21.0 / (16.0 * SPH_PI))
#define DENSITY_FLOOR 1e-10
#define ENERGY_FLOOR 1e-10
#define DT_FLOOR 1e-6
#define MAX_H_ITER 5
#define H_TOL 1e-3
#define GAS 4

#include "gravis_oracle.h"

void orc_bf_accel_eps(size_t n, const double* eps_i, const double* x, const double* y,
                      const double* z, const double* m, double* ax, double* ay, double* az);

typedef struct {
    size_t n;
    double *x, *y, *z, *vx, *vy, *vz, *ax, *ay, *az, *m;
    double *rho, *p, *u, *dudt, *h, *c, *alpha;
    uint8_t* type;
} orc_sph_state;

typedef struct {
    double gamma, eta, h_max, cfl, force_factor;
    double beta_visc, alpha_min, alpha_max, alpha_cond, sigma_visc;
    double h_tol;
    int32_t kernel, h_solver, max_h_iter, alpha_switch, pair_fp32, soft_h;
} orc_sph_params;

void orc_sph_default_params(orc_sph_params* p) {
    p->gamma = 5.0 / 3.0;
    p->eta = 1.2;
    p->h_max = INFINITY;
    p->cfl = 0.3;
    p->force_factor = 0.25;
    p->beta_visc = 2.0;
    p->alpha_min = 0.1;
    p->alpha_max = 2.0;
    p->alpha_cond = 1.0;
    p->sigma_visc = 0.1;
    p->h_tol = H_TOL;
    p->kernel = 0;
    p->h_solver = 0;
    p->max_h_iter = MAX_H_ITER;
    p->alpha_switch = 1;
    p->pair_fp32 = 0;
    p->soft_h = 0;
}

/* The physics variant: cubic spline, Newton h, fixed alpha, no conduction,
 * eps = h (see the header). */
void orc_sph_variant_params(orc_sph_params* p) {
    orc_sph_default_params(p);
    p->kernel = 1;
    p->h_solver = 1;
    p->h_tol = 1e-4;
    p->max_h_iter = 30;
    p->alpha_switch = 0;
    p->alpha_cond = 0.0;
    p->soft_h = 1;
}

/* ---- kernel.rs ---------------------------------------------------------- */
double orc_w(double r, double h) {
    double q = r / h;
    if (q >= 2.0) return 0.0;
    double h3 = h * h * h;
    double s = 1.0 - 0.5 * q;
    double s4 = s * s * s * s;
    return NORM_3D / h3 * s4 * (1.0 + 2.0 * q);
}

static void grad_w(double dx, double dy, double dz, double r, double h, double g[3]) {
    g[0] = g[1] = g[2] = 0.0;
    if (r < 1e-30) return;
    double q = r / h;
    if (q >= 2.0) return;
    double h3 = h * h * h;
    double s = 1.0 - 0.5 * q;
    double s3 = s * s * s;
    double dwdq = NORM_3D / h3 * (-5.0 * q) * s3;
    double f = dwdq / (h * r);
    g[0] = f * dx; g[1] = f * dy; g[2] = f * dz;
}

double orc_dw_dh(double r, double h) {
    double q = r / h;
    if (q >= 2.0) return 0.0;
    double h3 = h * h * h;
    double s = 1.0 - 0.5 * q;
    double s3 = s * s * s;
    double s4 = s3 * s;
    double w = NORM_3D / h3 * s4 * (1.0 + 2.0 * q);
    double dwdq = NORM_3D / h3 * (-5.0 * q) * s3;
    return -3.0 / h * w - q / h * dwdq;
}

static double abs_grad_w(double r, double h) {
    if (r < 1e-30) return 0.0;
    double q = r / h;
    if (q >= 2.0) return 0.0;
    double h3 = h * h * h;
    double s = 1.0 - 0.5 * q;
    double s3 = s * s * s;
    double dwdq = NORM_3D / h3 * (-5.0 * q) * s3;
    return -dwdq / h;
}

void orc_grad_w(double dx, double dy, double dz, double r, double h, double* g) {
    grad_w(dx, dy, dz, r, h, g);
}

double orc_abs_grad_w(double r, double h) { return abs_grad_w(r, h); }

/* ---- cubic spline M4 (variant), support 2h, sigma = 1/pi --------------- */
#define CS_NORM (1.0 / SPH_PI)
static double cs_f(double q) {
    if (q < 1.0) return 1.0 - 1.5 * q * q + 0.75 * q * q * q;
    if (q < 2.0) { double t = 2.0 - q; return 0.25 * t * t * t; }
    return 0.0;
}
static double cs_df(double q) {
    if (q < 1.0) return -3.0 * q + 2.25 * q * q;
    if (q < 2.0) { double t = 2.0 - q; return -0.75 * t * t; }
    return 0.0;
}
double orc_cs_w(double r, double h) { return CS_NORM / (h * h * h) * cs_f(r / h); }
double orc_cs_dw_dh(double r, double h) {
    double q = r / h;
    return -CS_NORM / (h * h * h * h) * (3.0 * cs_f(q) + q * cs_df(q));
}
static void cs_grad(double dx, double dy, double dz, double r, double h, double g[3]) {
    g[0] = g[1] = g[2] = 0.0;
    if (r < 1e-30) return;
    double q = r / h;
    if (q >= 2.0) return;
    double f = CS_NORM / (h * h * h * h) * cs_df(q) / r;
    g[0] = f * dx; g[1] = f * dy; g[2] = f * dz;
}
static double cs_abs_grad(double r, double h) {
    if (r < 1e-30) return 0.0;
    double q = r / h;
    if (q >= 2.0) return 0.0;
    return -CS_NORM / (h * h * h * h) * cs_df(q);
}
void orc_cs_grad_w(double dx, double dy, double dz, double r, double h, double* g) {
    cs_grad(dx, dy, dz, r, h, g);
}

static double k_w(int k, double r, double h) { return k ? orc_cs_w(r, h) : orc_w(r, h); }
static double k_dwdh(int k, double r, double h) { return k ? orc_cs_dw_dh(r, h) : orc_dw_dh(r, h); }
static void k_grad(int k, double dx, double dy, double dz, double r, double h, double g[3]) {
    if (k) cs_grad(dx, dy, dz, r, h, g);
    else grad_w(dx, dy, dz, r, h, g);
}
static double k_absgrad(int k, double r, double h) {
    return k ? cs_abs_grad(r, h) : abs_grad_w(r, h);
}

/* ---- reference octree ball query (octree.rs:245-308) ------------------- */
#define SNONE 0xFFFFFFFFu
typedef struct {
    double c[3], hw, mass;
    double com[3];
    uint32_t child[8], pidx;
} snode;
typedef struct {
    snode* v;
    size_t len, cap;
} stree;

static uint32_t s_alloc(stree* t, const double c[3], double hw) {
    if (t->len == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 256;
        t->v = (snode*)realloc(t->v, t->cap * sizeof(snode));
    }
    snode* nd = &t->v[t->len];
    memcpy(nd->c, c, sizeof nd->c);
    nd->hw = hw;
    nd->mass = 0.0;
    nd->com[0] = nd->com[1] = nd->com[2] = 0.0;
    for (int k = 0; k < 8; ++k) nd->child[k] = SNONE;
    nd->pidx = SNONE;
    return (uint32_t)t->len++;
}

static int s_oct(const snode* nd, double x, double y, double z) {
    return (x >= nd->c[0] ? 1 : 0) | (y >= nd->c[1] ? 2 : 0) | (z >= nd->c[2] ? 4 : 0);
}

static void s_cb(const snode* nd, int o, double c[3], double* hw) {
    double q = nd->hw * 0.5;
    c[0] = nd->c[0] + ((o & 1) ? q : -q);
    c[1] = nd->c[1] + ((o & 2) ? q : -q);
    c[2] = nd->c[2] + ((o & 4) ? q : -q);
    *hw = q;
}

static void s_com(snode* nd, double px, double py, double pz, double m) {
    double nm = nd->mass + m;
    if (nm > 0.0) {
        double inv = 1.0 / nm;
        nd->com[0] = (nd->mass * nd->com[0] + m * px) * inv;
        nd->com[1] = (nd->mass * nd->com[1] + m * py) * inv;
        nd->com[2] = (nd->mass * nd->com[2] + m * pz) * inv;
    }
    nd->mass = nm;
}

static void s_insert(stree* t, uint32_t ni, uint32_t pi, double px, double py, double pz,
                     double m, int depth) {
    if (t->v[ni].mass == 0.0) {
        snode* nd = &t->v[ni];
        nd->pidx = pi; nd->mass = m;
        nd->com[0] = px; nd->com[1] = py; nd->com[2] = pz;
        return;
    }
    if (depth >= 64) { s_com(&t->v[ni], px, py, pz, m); return; }
    uint32_t ex = t->v[ni].pidx;
    double c[3], hw;
    if (ex != SNONE) {
        double oc[3] = {t->v[ni].com[0], t->v[ni].com[1], t->v[ni].com[2]};
        double om = t->v[ni].mass;
        t->v[ni].pidx = SNONE;
        int oo = s_oct(&t->v[ni], oc[0], oc[1], oc[2]);
        s_cb(&t->v[ni], oo, c, &hw);
        uint32_t ci = s_alloc(t, c, hw);
        t->v[ni].child[oo] = ci;
        s_insert(t, ci, ex, oc[0], oc[1], oc[2], om, depth + 1);
        int no = s_oct(&t->v[ni], px, py, pz);
        if (t->v[ni].child[no] == SNONE) {
            s_cb(&t->v[ni], no, c, &hw);
            uint32_t c2 = s_alloc(t, c, hw);
            t->v[ni].child[no] = c2;
        }
        uint32_t c3 = t->v[ni].child[no];
        s_com(&t->v[ni], px, py, pz, m);
        s_insert(t, c3, pi, px, py, pz, m, depth + 1);
    } else {
        s_com(&t->v[ni], px, py, pz, m);
        int o = s_oct(&t->v[ni], px, py, pz);
        if (t->v[ni].child[o] == SNONE) {
            s_cb(&t->v[ni], o, c, &hw);
            uint32_t ci = s_alloc(t, c, hw);
            t->v[ni].child[o] = ci;
        }
        s_insert(t, t->v[ni].child[o], pi, px, py, pz, m, depth + 1);
    }
}

static void s_build(stree* t, const orc_sph_state* s) {
    t->len = 0;
    double box[4];
    orc_root_box(s->n, s->x, s->y, s->z, box);
    s_alloc(t, box, box[3]);
    for (size_t i = 0; i < s->n; ++i) s_insert(t, 0, (uint32_t)i, s->x[i], s->y[i], s->z[i], s->m[i], 0);
}

typedef struct {
    uint32_t* v;
    size_t len, cap;
} ulist;

static void ul_push(ulist* l, uint32_t x) {
    if (l->len == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 64;
        l->v = (uint32_t*)realloc(l->v, l->cap * 4);
    }
    l->v[l->len++] = x;
}

static int sphere_box(const double p[3], double r, const snode* nd) {
    double d2 = 0.0;
    for (int d = 0; d < 3; ++d) {
        double lo = nd->c[d] - nd->hw, hi = nd->c[d] + nd->hw;
        if (p[d] < lo) { double t = lo - p[d]; d2 += t * t; }
        else if (p[d] > hi) { double t = p[d] - hi; d2 += t * t; }
    }
    return d2 <= r * r;
}

static void s_query(const stree* t, uint32_t ni, const double p[3], double r, double r2,
                    const orc_sph_state* s, ulist* out) {
    const snode* nd = &t->v[ni];
    if (nd->mass == 0.0) return;
    if (!sphere_box(p, r, nd)) return;
    if (nd->pidx != SNONE) {
        uint32_t pi = nd->pidx;
        double dx = s->x[pi] - p[0], dy = s->y[pi] - p[1], dz = s->z[pi] - p[2];
        if (dx * dx + dy * dy + dz * dz <= r2) ul_push(out, pi);
        return;
    }
    for (int k = 0; k < 8; ++k)
        if (nd->child[k] != SNONE) s_query(t, nd->child[k], p, r, r2, s, out);
}

static void ball(const stree* t, const orc_sph_state* s, size_t i, double radius, ulist* out) {
    out->len = 0;
    double p[3] = {s->x[i], s->y[i], s->z[i]};
    s_query(t, 0, p, radius, radius * radius, s, out);
}

/* Octree::query_ball from particle i's position over a tree of the whole
 * state; writes at most cap indices, returns the count found. */
size_t orc_ball_query(const orc_sph_state* s, size_t i, double radius, uint32_t* out, size_t cap) {
    stree t = {0};
    s_build(&t, s);
    ulist l = {0};
    ball(&t, s, i, radius, &l);
    size_t k = l.len;
    for (size_t q = 0; q < k && q < cap; ++q) out[q] = l.v[q];
    free(l.v);
    free(t.v);
    return k;
}

/* ---- density.rs:42-157 ------------------------------------------------- */
static void density_dwdh(const orc_sph_state* s, const ulist* nb, size_t i, double h, int kern,
                         double* rho, double* dsum) {
    double xi = s->x[i], yi = s->y[i], zi = s->z[i];
    double r_ = 0.0, d_ = 0.0;
    for (size_t k = 0; k < nb->len; ++k) {
        uint32_t j = nb->v[k];
        double dx = xi - s->x[j], dy = yi - s->y[j], dz = zi - s->z[j];
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        if (r < 2.0 * h) {
            double mj = s->m[j];
            r_ += mj * k_w(kern, r, h);
            d_ += mj * k_dwdh(kern, r, h);
        }
    }
    *rho = r_;
    *dsum = d_;
}

/* One gas particle's list, density / smoothing-length iteration
 * (density.rs:42-104) and equation of state (:111-119). It reads the
 * positions, the masses and its own h only, so particles can be done in
 * any order. lr_i is the radius the list nl was built with. */
static void density_one(orc_sph_state* s, const stree* t, const orc_sph_params* pp, size_t i,
                        double* omega_i, ulist* nl) {
    *omega_i = 1.0;
    if (s->type[i] != GAS) return;
    double lr = 2.0 * s->h[i] * 1.1;
    ball(t, s, i, lr, nl);
    double h = s->h[i], mi = s->m[i];
    for (int it = 0; it < pp->max_h_iter; ++it) {
        double rho, dsum;
        density_dwdh(s, nl, i, h, pp->kernel, &rho, &dsum);
        double h_new;
        if (pp->h_solver == 0) {
            h_new = rho > 0.0 ? fmin(pp->eta * cbrt(mi / rho), pp->h_max) : fmin(h, pp->h_max);
        } else {
            double e = pp->eta / h;
            double rt = mi * e * e * e;
            double gp = dsum + 3.0 * rt / h;
            double hn = h - (rho - rt) / gp;
            if (!(gp > 0.0) || !(rho > 0.0)) hn = rho > 0.0 ? pp->eta * cbrt(mi / rho) : 2.0 * h;
            if (hn < 0.5 * h) hn = 0.5 * h;
            if (hn > 2.0 * h) hn = 2.0 * h;
            h_new = fmin(hn, pp->h_max);
        }
        if (rho > 0.0) {
            double dh = -h_new / (3.0 * rho);
            *omega_i = 1.0 - dh * dsum;
            if (*omega_i < 0.1) *omega_i = 0.1;
        }
        s->rho[i] = rho;
        s->h[i] = h_new;
        if (fabs(h_new - h) / fmax(h, 1e-30) < pp->h_tol) break;
        /* sim-core compares with the current h (density.rs:96-97); the
         * variant with the radius the list was built with */
        double sr = pp->h_solver == 0 ? 2.0 * h * 1.1 : lr;
        if (2.0 * h_new > sr) {
            lr = 2.0 * h_new * 1.1;
            ball(t, s, i, lr, nl);
        }
        h = h_new;
    }
    double rho = s->rho[i], u = s->u[i];
    s->p[i] = (pp->gamma - 1.0) * rho * u;
    s->c[i] = sqrt(pp->gamma * (pp->gamma - 1.0) * fmax(u, 0.0));
}

/* Computes density/h/omega/EOS; returns per-particle neighbour lists. */
static ulist* compute_density(orc_sph_state* s, const stree* t, const orc_sph_params* pp,
                              double* omega) {
    ulist* nl = (ulist*)calloc(s->n, sizeof(ulist));
    for (size_t i = 0; i < s->n; ++i) density_one(s, t, pp, i, &omega[i], &nl[i]);
    return nl;
}

/* ---- forces.rs:69-232 -------------------------------------------------- */
/* One gas particle's force sums over its final list; ADDS to a and du/dt. */
static void force_one(orc_sph_state* s, const ulist* nli, const double* omega,
                      const orc_sph_params* pp, size_t i, double* div_v, double* max_mu,
                      double* dalpha) {
    *div_v = 0.0; *max_mu = 0.0; *dalpha = 0.0;
    if (s->type[i] != GAS) return;
    double xi = s->x[i], yi = s->y[i], zi = s->z[i];
    double vxi = s->vx[i], vyi = s->vy[i], vzi = s->vz[i];
    double hi = s->h[i];
    double rho_i = fmax(s->rho[i], DENSITY_FLOOR);
    double p_i = s->p[i], u_i = s->u[i], c_i = s->c[i], a_i = s->alpha[i];
    double om_i = omega[i];
    double pr_i = rho_i > 0.0 ? p_i / (om_i * rho_i * rho_i) : 0.0;
    double ax = 0, ay = 0, az = 0, du = 0, dv = 0, mmu = 0;
    for (size_t k = 0; k < nli->len; ++k) {
        uint32_t j = nli->v[k];
        if (j == i) continue;
        double dx = xi - s->x[j], dy = yi - s->y[j], dz = zi - s->z[j];
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        double hj = s->h[j];
        double havg = 0.5 * (hi + hj);
        if (r >= 2.0 * hi && r >= 2.0 * hj) continue;
        double mj = s->m[j];
        double rho_j = fmax(s->rho[j], DENSITY_FLOOR);
        double p_j = s->p[j], c_j = s->c[j], a_j = s->alpha[j], om_j = omega[j];
        double dvx = vxi - s->vx[j], dvy = vyi - s->vy[j], dvz = vzi - s->vz[j];
        double vr = dvx * dx + dvy * dy + dvz * dz;
        double gi[3], gj[3], ga[3];
        k_grad(pp->kernel, dx, dy, dz, r, hi, gi);
        k_grad(pp->kernel, dx, dy, dz, r, hj, gj);
        k_grad(pp->kernel, dx, dy, dz, r, havg, ga);
        double pr_j = rho_j > 0.0 ? p_j / (om_j * rho_j * rho_j) : 0.0;
        double pij = 0.0;
        if (vr < 0.0) {
            double aavg = 0.5 * (a_i + a_j), cavg = 0.5 * (c_i + c_j),
                   ravg = 0.5 * (rho_i + rho_j);
            double eta2 = 0.01 * havg * havg;
            double mu = havg * vr / (r * r + eta2);
            mmu = fmax(mmu, fabs(mu));
            pij = ravg > 0.0 ? (-aavg * cavg * mu + pp->beta_visc * mu * mu) / ravg : 0.0;
        }
        ax -= mj * (pr_i * gi[0] + pr_j * gj[0] + pij * ga[0]);
        ay -= mj * (pr_i * gi[1] + pr_j * gj[1] + pij * ga[1]);
        az -= mj * (pr_i * gi[2] + pr_j * gj[2] + pij * ga[2]);
        double vgi = dvx * gi[0] + dvy * gi[1] + dvz * gi[2];
        double vga = dvx * ga[0] + dvy * ga[1] + dvz * ga[2];
        du += mj * (pr_i * vgi + 0.5 * pij * vga);
        dv -= mj * vgi;
        if (pp->alpha_cond > 0.0) {
            double u_j = s->u[j];
            double rs = rho_i + rho_j;
            if (rs > 0.0) {
                double ravg = 0.5 * rs;
                double dp = fabs(p_i - p_j);
                double vsig = sqrt(dp / ravg);
                double agw = k_absgrad(pp->kernel, r, havg);
                du += mj * pp->alpha_cond * vsig * (u_i - u_j) / rs * agw;
            }
        }
    }
    s->ax[i] += ax; s->ay[i] += ay; s->az[i] += az;
    s->dudt[i] += du;
    if (rho_i > 0.0) *div_v = dv / rho_i;
    *max_mu = mmu;
    double tau = c_i > 0.0 ? hi / (2.0 * pp->sigma_visc * c_i) : 1e10;
    double src = fmax(-*div_v, 0.0) * (pp->alpha_max - a_i);
    *dalpha = pp->alpha_switch ? -(a_i - pp->alpha_min) / tau + src : 0.0;
}

static void compute_forces(orc_sph_state* s, const ulist* nl, const double* omega,
                           const orc_sph_params* pp, double* div_v, double* max_mu,
                           double* dalpha) {
    for (size_t i = 0; i < s->n; ++i)
        force_one(s, &nl[i], omega, pp, i, &div_v[i], &max_mu[i], &dalpha[i]);
}

/* SphSolver::compute (solver.rs:61-80): density + forces; intermediates out. */
void orc_sph_compute(orc_sph_state* s, const orc_sph_params* pp, double* div_v, double* max_mu,
                     double* dalpha, double* omega_out, uint32_t* nneigh_out) {
    size_t n = s->n;
    stree t = {0};
    s_build(&t, s);
    double* omega = (double*)malloc(n * sizeof(double));
    ulist* nl = compute_density(s, &t, pp, omega);
    compute_forces(s, nl, omega, pp, div_v, max_mu, dalpha);
    if (omega_out) memcpy(omega_out, omega, n * sizeof(double));
    for (size_t i = 0; i < n; ++i) {
        if (nneigh_out) nneigh_out[i] = (uint32_t)nl[i].len;
        free(nl[i].v);
    }
    free(nl);
    free(omega);
    free(t.v);
}

/* SphSolver::compute restricted to nt target particles (a checker for
 * large states): the density iteration of each target, then of every
 * particle on a target's final list, then the forces of the targets only.
 * Each particle's density depends on the positions, the masses and its own
 * h alone, so every value written equals what orc_sph_compute writes for
 * that particle. Writes rho, h, p, c of the particles it visits, ADDS the
 * targets' a and du/dt, and fills the per-target arrays (k-th entry for
 * targets[k]; any may be NULL). Parallel over particles (OpenMP). */
void orc_sph_compute_targets(orc_sph_state* s, const orc_sph_params* pp, size_t nt,
                             const int64_t* targets, double* div_v, double* max_mu,
                             double* dalpha, double* omega_out, uint32_t* nneigh_out,
                             int nthreads) {
    size_t n = s->n;
    stree t = {0};
    s_build(&t, s);
    double* omega = (double*)malloc(n * sizeof(double));
    ulist* nl = (ulist*)calloc(n, sizeof(ulist));
    uint8_t* done = (uint8_t*)calloc(n, 1);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    for (size_t k = 0; k < nt; ++k) done[targets[k]] = 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (long long k = 0; k < (long long)nt; ++k) {
        size_t i = (size_t)targets[k];
        density_one(s, &t, pp, i, &omega[i], &nl[i]);
    }
    /* the particles the targets' forces read */
    size_t cap = 1024, cnt = 0;
    uint32_t* todo = (uint32_t*)malloc(cap * sizeof(uint32_t));
    for (size_t k = 0; k < nt; ++k) {
        const ulist* l = &nl[targets[k]];
        for (size_t q = 0; q < l->len; ++q) {
            uint32_t j = l->v[q];
            if (done[j]) continue;
            done[j] = 1;
            if (cnt == cap) { cap *= 2; todo = (uint32_t*)realloc(todo, cap * sizeof(uint32_t)); }
            todo[cnt++] = j;
        }
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (long long k = 0; k < (long long)cnt; ++k)
        density_one(s, &t, pp, todo[k], &omega[todo[k]], &nl[todo[k]]);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (long long k = 0; k < (long long)nt; ++k) {
        size_t i = (size_t)targets[k];
        double dv, mm, da;
        force_one(s, &nl[i], omega, pp, i, &dv, &mm, &da);
        if (div_v) div_v[k] = dv;
        if (max_mu) max_mu[k] = mm;
        if (dalpha) dalpha[k] = da;
        if (omega_out) omega_out[k] = omega[i];
        if (nneigh_out) nneigh_out[k] = (uint32_t)nl[i].len;
    }
    for (size_t i = 0; i < n; ++i) free(nl[i].v);
    free(todo);
    free(done);
    free(nl);
    free(omega);
    free(t.v);
}

/* Brute force at nt targets with the softening eps_i[k] of target k
 * (the variant's eps = h): a_k = sum_{j != t_k} m_j d / (r^2 + eps_k^2)^{3/2},
 * sources in index order, as orc_bf_accel_eps. Overwrites ax/ay/az[k]. */
void orc_bf_targets_eps(size_t n, const double* eps_t, const double* x, const double* y,
                        const double* z, const double* m, size_t nt, const int64_t* targets,
                        double* ax, double* ay, double* az, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (long long k = 0; k < (long long)nt; ++k) {
        size_t i = (size_t)targets[k];
        double e2 = eps_t[k] * eps_t[k];
        double a0 = 0, a1 = 0, a2 = 0;
        for (size_t j = 0; j < n; ++j) {
            if (j == i) continue;
            double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
            double r2 = dx * dx + dy * dy + dz * dz + e2;
            double w = 1.0 / (r2 * sqrt(r2));
            double f = m[j] * w;
            a0 += f * dx; a1 += f * dy; a2 += f * dz;
        }
        ax[k] = a0; ay[k] = a1; az[k] = a2;
    }
}

/* SphSolver::compute_timestep (solver.rs:90-131). */
double orc_sph_timestep(const orc_sph_state* s, const orc_sph_params* pp, const double* div_v,
                        const double* max_mu) {
    double dt = INFINITY;
    for (size_t i = 0; i < s->n; ++i) {
        if (s->type[i] != GAS) continue;
        double h = s->h[i];
        double am = sqrt(s->ax[i] * s->ax[i] + s->ay[i] * s->ay[i] + s->az[i] * s->az[i]);
        if (am > 0.0) dt = fmin(dt, pp->force_factor * sqrt(h / am));
        double c = s->c[i], a = s->alpha[i];
        double vsig = c + h * fabs(div_v[i]) + 1.2 * (a * c + pp->beta_visc * max_mu[i]);
        if (vsig > 0.0) dt = fmin(dt, pp->cfl * h / vsig);
    }
    return dt < DT_FLOOR ? DT_FLOOR : dt;
}

/* step_with_sph (solver.rs:145-212). gravity: 0 none, 1 brute force,
 * 2 reference Barnes-Hut, 3 the linear-octree group walk (leaf 16, group
 * 64). Returns the next dt. */
double orc_sph_step(orc_sph_state* s, const orc_sph_params* pp, int gravity, double eps,
                    double theta, double dt) {
    size_t n = s->n;
    double h = 0.5 * dt;
    for (size_t i = 0; i < n; ++i) { s->vx[i] += s->ax[i] * h; s->vy[i] += s->ay[i] * h; s->vz[i] += s->az[i] * h; }
    for (size_t i = 0; i < n; ++i)
        if (s->type[i] == GAS) {
            s->u[i] += s->dudt[i] * h;
            if (s->u[i] < ENERGY_FLOOR) s->u[i] = ENERGY_FLOOR;
        }
    for (size_t i = 0; i < n; ++i) { s->x[i] += s->vx[i] * dt; s->y[i] += s->vy[i] * dt; s->z[i] += s->vz[i] * dt; }
    for (size_t i = 0; i < n; ++i) { s->ax[i] = 0.0; s->ay[i] = 0.0; s->az[i] = 0.0; s->dudt[i] = 0.0; }
    if (pp->soft_h && gravity == 1) orc_bf_accel_eps(n, s->h, s->x, s->y, s->z, s->m, s->ax, s->ay, s->az);
    else if (pp->soft_h && gravity == 3)
        orc_bh_group_accel_eps(n, s->h, theta, 16, 64, 4096, s->x, s->y, s->z, s->m, s->ax,
                               s->ay, s->az, 0);
    else if (gravity == 1) orc_bf_accel(n, eps, s->x, s->y, s->z, s->m, s->ax, s->ay, s->az);
    else if (gravity == 2) orc_bh_accel(n, eps, theta, s->x, s->y, s->z, s->m, s->ax, s->ay, s->az, 0);
    else if (gravity == 3)
        orc_bh_group_accel(n, eps, theta, 16, 64, 4096, 0, s->x, s->y, s->z, s->m, s->ax, s->ay,
                           s->az, 0);
    double* tmp = (double*)malloc(3 * n * sizeof(double));
    orc_sph_compute(s, pp, tmp, tmp + n, tmp + 2 * n, NULL, NULL);
    for (size_t i = 0; i < n; ++i) { s->vx[i] += s->ax[i] * h; s->vy[i] += s->ay[i] * h; s->vz[i] += s->az[i] * h; }
    for (size_t i = 0; i < n; ++i)
        if (s->type[i] == GAS) {
            s->u[i] += s->dudt[i] * h;
            if (s->u[i] < ENERGY_FLOOR) s->u[i] = ENERGY_FLOOR;
            if (pp->alpha_switch) {
                s->alpha[i] += tmp[2 * n + i] * dt;
                /* f64::clamp (solver.rs:205-207): NaN stays NaN */
                double al = s->alpha[i];
                s->alpha[i] = al < pp->alpha_min ? pp->alpha_min
                                                 : (al > pp->alpha_max ? pp->alpha_max : al);
            }
        }
    double dtn = orc_sph_timestep(s, pp, tmp, tmp + n);
    free(tmp);
    return dtn;
}

/* EvrardCollapse::generate (evrard_collapse.rs:45-57) with
 * fill_uniform_gas_sphere (scenarios/mod.rs:33-51) and sphere_spacing
 * (mod.rs:17-22). State arrays must hold n entries; everything not set by
 * add_gas (particle.rs:136-154) is zero. */
void orc_evrard(size_t n, double total_mass, double radius, double u0, uint64_t seed,
                orc_sph_state* s) {
    orc_rng rng;
    orc_rng_seed_u64(&rng, seed);
    double mass = total_mass / (double)n;
    double vol = 4.0 / 3.0 * SPH_PI * (radius * radius * radius);
    double spacing = cbrt(vol / (double)n);
    double h = 1.5 * spacing;
    size_t cnt = 0;
    while (cnt < n) {
        double x = orc_rng_f64(&rng) * 2.0 * radius - radius;
        double y = orc_rng_f64(&rng) * 2.0 * radius - radius;
        double z = orc_rng_f64(&rng) * 2.0 * radius - radius;
        if (x * x + y * y + z * z <= radius * radius) {
            s->x[cnt] = x; s->y[cnt] = y; s->z[cnt] = z;
            s->vx[cnt] = s->vy[cnt] = s->vz[cnt] = 0.0;
            s->ax[cnt] = s->ay[cnt] = s->az[cnt] = 0.0;
            s->m[cnt] = mass;
            s->type[cnt] = GAS;
            s->rho[cnt] = s->p[cnt] = s->dudt[cnt] = s->c[cnt] = 0.0;
            s->u[cnt] = u0;
            s->h[cnt] = h;
            s->alpha[cnt] = 1.0;
            cnt++;
        }
    }
}

/* Brute force with Plummer softening eps_i of each TARGET (the variant's
 * eps = h): a_i += sum_{j != i} m_j d / (r^2 + eps_i^2)^{3/2}, gather form,
 * sources in index order. */
void orc_bf_accel_eps(size_t n, const double* eps_i, const double* x, const double* y,
                      const double* z, const double* m, double* ax, double* ay, double* az) {
    for (size_t i = 0; i < n; ++i) {
        double e2 = eps_i[i] * eps_i[i];
        double a0 = 0, a1 = 0, a2 = 0;
        for (size_t j = 0; j < n; ++j) {
            if (j == i) continue;
            double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
            double r2 = dx * dx + dy * dy + dz * dz + e2;
            double w = 1.0 / (r2 * sqrt(r2));
            double f = m[j] * w;
            a0 += f * dx; a1 += f * dy; a2 += f * dz;
        }
        ax[i] += a0; ay[i] += a1; az[i] += a2;
    }
}

/* Variant initial conditions: n gas particles in a sphere of radius R and
 * mass M with rho(r) = M / (2 pi R^2 r), i.e. M(<r) = M (r/R)^2, sampled as
 * r = R sqrt(U) with U = random::<f64>(), cos(theta) in U[-1, 1),
 * phi in U[0, 2 pi) (rand's random_range), drawn in that order per particle
 * from ChaCha20Rng::seed_from_u64(seed); at rest, u = u0, alpha 1, m = M/n,
 * h = eta (m / rho(r))^(1/3) from the analytic density (r floored at
 * 1e-6 R). */
void orc_evrard_1r(size_t n, double total_mass, double radius, double u0, double eta,
                   uint64_t seed, orc_sph_state* s) {
    orc_rng rng;
    orc_rng_seed_u64(&rng, seed);
    double mass = total_mass / (double)n;
    for (size_t i = 0; i < n; ++i) {
        double r = radius * sqrt(orc_rng_f64(&rng));
        double ct = orc_rng_range(&rng, -1.0, 1.0);
        double ph = orc_rng_range(&rng, 0.0, 2.0 * SPH_PI);
        double st = sqrt(fmax(0.0, 1.0 - ct * ct));
        s->x[i] = r * st * cos(ph);
        s->y[i] = r * st * sin(ph);
        s->z[i] = r * ct;
        s->vx[i] = s->vy[i] = s->vz[i] = 0.0;
        s->ax[i] = s->ay[i] = s->az[i] = 0.0;
        s->m[i] = mass;
        s->type[i] = GAS;
        s->rho[i] = s->p[i] = s->dudt[i] = s->c[i] = 0.0;
        s->u[i] = u0;
        double rr = fmax(r, 1e-6 * radius);
        double rho = total_mass / (2.0 * SPH_PI * radius * radius * rr);
        s->h[i] = eta * cbrt(mass / rho);
        s->alpha[i] = 1.0;
    }
}

/* ColdCollapse::generate (cold_collapse.rs:56-102): gas = 0 rejection with
 * random_range(-1..1) per axis scaled by R (type 0); gas = 1
 * fill_uniform_gas_sphere with u = 0.05, h = 1.5 spacing (type 4); then
 * shift_to_com_frame (skipped when com_shift = 0, so tests can check the
 * samples themselves). Every other field 0. */
void orc_cold_collapse_ex(size_t n, double radius, double total_mass, uint64_t seed, int gas,
                          int com_shift, orc_sph_state* s) {
    orc_rng rng;
    orc_rng_seed_u64(&rng, seed);
    double m = total_mass / (double)n;
    double* f[17] = {s->x, s->y, s->z, s->vx, s->vy, s->vz, s->ax, s->ay, s->az,
                     s->m, s->rho, s->p, s->u, s->dudt, s->h, s->c, s->alpha};
    for (int k = 0; k < 17; ++k) memset(f[k], 0, n * sizeof(double));
    if (gas) {
        double vol = 4.0 / 3.0 * SPH_PI * (radius * radius * radius);
        double h = 1.5 * cbrt(vol / (double)n);
        size_t cnt = 0;
        while (cnt < n) {
            double x = orc_rng_f64(&rng) * 2.0 * radius - radius;
            double y = orc_rng_f64(&rng) * 2.0 * radius - radius;
            double z = orc_rng_f64(&rng) * 2.0 * radius - radius;
            if (x * x + y * y + z * z <= radius * radius) {
                s->x[cnt] = x; s->y[cnt] = y; s->z[cnt] = z;
                s->m[cnt] = m; s->u[cnt] = 0.05; s->h[cnt] = h; s->alpha[cnt] = 1.0;
                s->type[cnt] = GAS;
                cnt++;
            }
        }
    } else {
        for (size_t i = 0; i < n; ++i) {
            double x, y, z;
            do {
                x = orc_rng_range(&rng, -1.0, 1.0);
                y = orc_rng_range(&rng, -1.0, 1.0);
                z = orc_rng_range(&rng, -1.0, 1.0);
            } while (!(x * x + y * y + z * z <= 1.0));
            s->x[i] = x * radius; s->y[i] = y * radius; s->z[i] = z * radius;
            s->m[i] = m;
            s->type[i] = 0;
        }
    }
    if (com_shift) orc_shift_to_com(n, s->x, s->y, s->z, s->vx, s->vy, s->vz, s->m);
}

void orc_cold_collapse(size_t n, double radius, double total_mass, uint64_t seed, int gas,
                       orc_sph_state* s) {
    orc_cold_collapse_ex(n, radius, total_mass, seed, gas, 1, s);
}

double orc_evrard_softening(size_t n, double radius) {
    double vol = 4.0 / 3.0 * SPH_PI * (radius * radius * radius);
    return 0.3 * cbrt(vol / (double)n);
}
