/*
 * gravis_oracle_adjoint.c -- CPU oracle for the batched ensemble and its
 * discrete adjoint (test infrastructure only; see gravis_oracle.c).
 *
 * The reference has no ensemble and no reverse mode; these are the build's
 * definitions, pinned by central finite differences of the forward map
 * (tests/test_oracle_adjoint.py) rather than by reference numbers.
 *
 * Forward (per system, independent): sim-core's LeapfrogKDK
 * (crates/sim-core/src/integrator.rs:41-77) with softened brute force in the
 * gather form a_i = sum_{j != i} m_j d_ij (|d_ij|^2 + eps^2)^{-3/2},
 * d_ij = x_j - x_i, sources in index order; a(x_0) computed first as
 * headless does (crates/headless/src/main.rs:151-152).
 *
 * Adjoint of one step  vh = v + h a(x);  x' = x + dt vh;  v' = vh + h a(x')
 * with h = dt/2, cotangents (lx, lv) of (x', v'):
 *   lvh = lv' ; lx' += J(x')^T (h lv')          [K2]
 *   lvh += dt lx' ; lx = lx'                     [D]
 *   lv = lvh ; lx += J(x)^T (h lvh)              [K1]
 * The J^T products at a shared time level are merged into one pair sweep
 * with weight 2h (h at t = 0 and at t = T). For cotangent u of the
 * accelerations, with d = x_j - x_k, s = |d|^2 + eps^2, w = s^{-3/2}:
 *   (J_x^T u)_k = sum_{j != k} [ w v - 3 s^{-5/2} d (d . v) ],
 *                 v = m_k u_j - m_j u_k
 *   (J_m^T u)_k = - sum_{j != k} w (u_j . d)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "gravis_oracle.h"

/* Threads for the per-particle sums (1 = serial, the timed CPU baseline).
 * Each particle's sum runs in index order on one thread, so results do not
 * depend on the count. */
static int ens_threads = 1;
void orc_ens_set_threads(int t) { ens_threads = t > 0 ? t : 1; }

static void accel_gather(size_t n, double eps2, const double* x, const double* y,
                         const double* z, const double* m, double* ax, double* ay, double* az) {
#pragma omp parallel for schedule(static) num_threads(ens_threads) if (ens_threads > 1 && n >= 256)
    for (size_t i = 0; i < n; ++i) {
        double a0 = 0, a1 = 0, a2 = 0;
        for (size_t j = 0; j < n; ++j) {
            if (j == i) continue;
            double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
            double s = dx * dx + dy * dy + dz * dz + eps2;
            double w = 1.0 / (s * sqrt(s));
            double f = m[j] * w;
            a0 += f * dx; a1 += f * dy; a2 += f * dz;
        }
        ax[i] = a0; ay[i] = a1; az[i] = a2;
    }
}

/* lx += J_x^T u, lm += J_m^T u at positions (x, y, z). */
static void vjp(size_t n, double eps2, const double* x, const double* y, const double* z,
                const double* m, const double* ux, const double* uy, const double* uz,
                double* lx, double* ly, double* lz, double* lm) {
#pragma omp parallel for schedule(static) num_threads(ens_threads) if (ens_threads > 1 && n >= 256)
    for (size_t k = 0; k < n; ++k) {
        double g0 = 0, g1 = 0, g2 = 0, gm = 0;
        for (size_t j = 0; j < n; ++j) {
            if (j == k) continue;
            double dx = x[j] - x[k], dy = y[j] - y[k], dz = z[j] - z[k];
            double s = dx * dx + dy * dy + dz * dz + eps2;
            double ri = 1.0 / sqrt(s);
            double w = ri * ri * ri;
            double q = 3.0 * w * ri * ri;
            double v0 = m[k] * ux[j] - m[j] * ux[k];
            double v1 = m[k] * uy[j] - m[j] * uy[k];
            double v2 = m[k] * uz[j] - m[j] * uz[k];
            double dv = dx * v0 + dy * v1 + dz * v2;
            g0 += w * v0 - q * dv * dx;
            g1 += w * v1 - q * dv * dy;
            g2 += w * v2 - q * dv * dz;
            gm -= w * (ux[j] * dx + uy[j] * dy + uz[j] * dz);
        }
        lx[k] += g0; ly[k] += g1; lz[k] += g2;
        lm[k] += gm;
    }
}

/* Forward: nsteps of KDK from (x, v); state updated in place; a(x) output. */
void orc_ens_forward(size_t n, double eps, double dt, uint64_t nsteps, double* x, double* y,
                     double* z, double* vx, double* vy, double* vz, const double* m) {
    double eps2 = eps * eps, h = 0.5 * dt;
    double* a = (double*)malloc(3 * n * sizeof(double));
    double *ax = a, *ay = a + n, *az = a + 2 * n;
    accel_gather(n, eps2, x, y, z, m, ax, ay, az);
    for (uint64_t t = 0; t < nsteps; ++t) {
        for (size_t i = 0; i < n; ++i) {
            vx[i] = vx[i] + ax[i] * h; vy[i] = vy[i] + ay[i] * h; vz[i] = vz[i] + az[i] * h;
            x[i] = x[i] + vx[i] * dt; y[i] = y[i] + vy[i] * dt; z[i] = z[i] + vz[i] * dt;
        }
        accel_gather(n, eps2, x, y, z, m, ax, ay, az);
        for (size_t i = 0; i < n; ++i) {
            vx[i] = vx[i] + ax[i] * h; vy[i] = vy[i] + ay[i] * h; vz[i] = vz[i] + az[i] * h;
        }
    }
    free(a);
}

/* Reverse mode of orc_ens_forward for one system. Inputs: initial state and
 * the cotangents (gx, gy, gz) of the final positions and (gvx, gvy, gvz) of
 * the final velocities (may be NULL = 0). Outputs: cotangents of the initial
 * positions, velocities and masses. Stores the whole trajectory (small n). */
void orc_ens_adjoint(size_t n, double eps, double dt, uint64_t nsteps, const double* x0,
                     const double* y0, const double* z0, const double* vx0, const double* vy0,
                     const double* vz0, const double* m, const double* gx, const double* gy,
                     const double* gz, const double* gvx, const double* gvy, const double* gvz,
                     double* ox, double* oy, double* oz, double* ovx, double* ovy, double* ovz,
                     double* om) {
    double eps2 = eps * eps, h = 0.5 * dt;
    size_t T = (size_t)nsteps;
    double* traj = (double*)malloc((T + 1) * 3 * n * sizeof(double));
    double* v = (double*)malloc(3 * n * sizeof(double));
    double* a = (double*)malloc(3 * n * sizeof(double));
    memcpy(traj, x0, n * 8); memcpy(traj + n, y0, n * 8); memcpy(traj + 2 * n, z0, n * 8);
    memcpy(v, vx0, n * 8); memcpy(v + n, vy0, n * 8); memcpy(v + 2 * n, vz0, n * 8);
    accel_gather(n, eps2, traj, traj + n, traj + 2 * n, m, a, a + n, a + 2 * n);
    for (size_t t = 0; t < T; ++t) {
        double* xs = traj + t * 3 * n;
        double* xn = traj + (t + 1) * 3 * n;
        for (size_t c = 0; c < 3; ++c)
            for (size_t i = 0; i < n; ++i) {
                v[c * n + i] = v[c * n + i] + a[c * n + i] * h;
                xn[c * n + i] = xs[c * n + i] + v[c * n + i] * dt;
            }
        accel_gather(n, eps2, xn, xn + n, xn + 2 * n, m, a, a + n, a + 2 * n);
        for (size_t c = 0; c < 3; ++c)
            for (size_t i = 0; i < n; ++i) v[c * n + i] = v[c * n + i] + a[c * n + i] * h;
    }
    double* lx = (double*)calloc(3 * n, sizeof(double));
    double* lv = (double*)calloc(3 * n, sizeof(double));
    double* u = (double*)calloc(3 * n, sizeof(double));
    memcpy(lx, gx, n * 8); memcpy(lx + n, gy, n * 8); memcpy(lx + 2 * n, gz, n * 8);
    if (gvx) { memcpy(lv, gvx, n * 8); memcpy(lv + n, gvy, n * 8); memcpy(lv + 2 * n, gvz, n * 8); }
    for (size_t i = 0; i < n; ++i) om[i] = 0.0;
    /* top: K2 of the last step at x_T with weight h */
    {
        double* xT = traj + T * 3 * n;
        for (size_t i = 0; i < 3 * n; ++i) u[i] = h * lv[i];
        if (T > 0)
            vjp(n, eps2, xT, xT + n, xT + 2 * n, m, u, u + n, u + 2 * n, lx, lx + n, lx + 2 * n, om);
    }
    /* lv now holds lvh of step T-1 */
    for (size_t tt = T; tt-- > 0;) {
        double* xt = traj + tt * 3 * n;
        for (size_t i = 0; i < 3 * n; ++i) lv[i] = lv[i] + dt * lx[i];  /* D */
        double wgt = tt > 0 ? 2.0 * h : h;                                 /* K1 (+K2 of tt-1) */
        for (size_t i = 0; i < 3 * n; ++i) u[i] = wgt * lv[i];
        vjp(n, eps2, xt, xt + n, xt + 2 * n, m, u, u + n, u + 2 * n, lx, lx + n, lx + 2 * n, om);
    }
    memcpy(ox, lx, n * 8); memcpy(oy, lx + n, n * 8); memcpy(oz, lx + 2 * n, n * 8);
    memcpy(ovx, lv, n * 8); memcpy(ovy, lv + n, n * 8); memcpy(ovz, lv + 2 * n, n * 8);
    free(traj); free(v); free(a); free(lx); free(lv); free(u);
}

/* Seeded ensemble layout: system s has N_s = lo + (u64 draw of a ChaCha20
 * stream seeded with `seed`) mod (hi - lo + 1) particles, drawn in system
 * order. Writes counts[nsys]. */
void orc_ens_counts(uint64_t seed, size_t nsys, uint32_t lo, uint32_t hi, uint32_t* counts) {
    orc_rng r;
    orc_rng_seed_u64(&r, seed);
    uint64_t span = (uint64_t)(hi - lo) + 1;
    for (size_t s = 0; s < nsys; ++s) counts[s] = lo + (uint32_t)(orc_rng_next_u64(&r) % span);
}

/* Seeded loss targets U[-1, 1)^3, n per call, drawn x, y, z per particle. */
void orc_ens_targets(uint64_t seed, size_t n, double* tx, double* ty, double* tz) {
    orc_rng r;
    orc_rng_seed_u64(&r, seed);
    for (size_t i = 0; i < n; ++i) {
        tx[i] = orc_rng_range(&r, -1.0, 1.0);
        ty[i] = orc_rng_range(&r, -1.0, 1.0);
        tz[i] = orc_rng_range(&r, -1.0, 1.0);
    }
}
