/*
 * gravis_oracle.c -- CPU oracle for the gravis force-and-integration path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it. The product path (gravis_amd/csrc, libgravis_hip.so) never links,
 * loads or calls anything here.
 *
 * It restates, op for op in IEEE f64 with no FMA contraction
 * (-ffp-contract=off, no fast-math), the reference sim-core hot path:
 *
 *   rng        rand_core 0.9.5 seed_from_u64 (PCG32 expansion) + rand_chacha
 *              0.9.0 ChaCha20Rng + rand 0.9.2 float conversions; call sites
 *              crates/sim-core/src/scenarios/plummer_sphere.rs:21-49,119
 *   plummer    crates/sim-core/src/scenarios/plummer_sphere.rs:13-152
 *   two-body   crates/sim-core/src/scenarios/two_body.rs:49-105
 *   galaxy     crates/sim-core/src/scenarios/galaxy_collision.rs:20-450,
 *              scenarios/mod.rs:25-30 (box_muller)
 *   com shift  crates/sim-core/src/particle.rs:168-226
 *   brute      crates/sim-core/src/gravity.rs:47-80
 *   kdk        crates/sim-core/src/integrator.rs:41-77
 *   diag       crates/sim-core/src/diagnostics.rs:50-148
 *   octree     crates/sim-core/src/octree.rs:14-86,145-239,312-324
 *   bh walk    crates/sim-core/src/barnes_hut.rs:44-192
 *
 * and it defines, for the parts the reference does not have (pinned by the
 * tests against brute force, closed forms and invariants, not by reference
 * numbers):
 *
 *   morton63   21 bits per axis over the reference root cube (octree.rs:14-44),
 *              axis bit order x:1 y:2 z:4 as octree.rs:67-79
 *   sort       stable sort of the 63-bit keys, ties by original index
 *   ltree      leaf-16 linear octree, breadth-first node order, FP64 COM sums
 *   gwalk      group walk: the particles of one group (consecutive sibling
 *              cells packed to at most 64) share one interaction list built
 *              breadth first with a box-distance opening test
 *              s^2 < theta^2 * d_min^2
 *
 * Parity status of each function is stated in tests/test_oracle_*.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "gravis_oracle.h"

#define ORC_PI 3.14159265358979323846

/* ------------------------------------------------------------------------ */
/* ChaCha20 (RFC 8439 block function; djb layout: 64-bit counter in words    */
/* 12-13, 64-bit stream id in words 14-15, as rand_chacha uses it).          */
/* ------------------------------------------------------------------------ */

static inline uint32_t rotl32(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }

#define QROUND(a, b, c, d)                 \
    a += b; d ^= a; d = rotl32(d, 16);     \
    c += d; b ^= c; b = rotl32(b, 12);     \
    a += b; d ^= a; d = rotl32(d, 8);      \
    c += d; b ^= c; b = rotl32(b, 7);

void orc_chacha20_block(const uint32_t key[8], uint64_t counter, uint64_t stream,
                        uint32_t out[16]) {
    uint32_t s[16];
    s[0] = 0x61707865u; s[1] = 0x3320646eu; s[2] = 0x79622d32u; s[3] = 0x6b206574u;
    for (int i = 0; i < 8; ++i) s[4 + i] = key[i];
    s[12] = (uint32_t)counter; s[13] = (uint32_t)(counter >> 32);
    s[14] = (uint32_t)stream;  s[15] = (uint32_t)(stream >> 32);
    uint32_t w[16];
    memcpy(w, s, sizeof w);
    for (int r = 0; r < 10; ++r) {
        QROUND(w[0], w[4], w[8], w[12]);
        QROUND(w[1], w[5], w[9], w[13]);
        QROUND(w[2], w[6], w[10], w[14]);
        QROUND(w[3], w[7], w[11], w[15]);
        QROUND(w[0], w[5], w[10], w[15]);
        QROUND(w[1], w[6], w[11], w[12]);
        QROUND(w[2], w[7], w[8], w[13]);
        QROUND(w[3], w[4], w[9], w[14]);
    }
    for (int i = 0; i < 16; ++i) out[i] = w[i] + s[i];
}

/* orc_rng: see gravis_oracle.h */

/* rand_core 0.9 SeedableRng::seed_from_u64: PCG32 (XSH-RR) words fill the
 * 32-byte seed, little endian. */
void orc_rng_seed_u64(orc_rng* r, uint64_t state) {
    const uint64_t MUL = 6364136223846793005ull;
    const uint64_t INC = 11634580027462260723ull;
    for (int i = 0; i < 8; ++i) {
        state = state * MUL + INC;
        uint32_t xorshifted = (uint32_t)(((state >> 18) ^ state) >> 27);
        uint32_t rot = (uint32_t)(state >> 59);
        r->key[i] = (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
    }
    r->counter = 0;
    r->idx = 64;
}

static void rng_refill(orc_rng* r) {
    for (int b = 0; b < 4; ++b) orc_chacha20_block(r->key, r->counter + (uint64_t)b, 0, r->buf + 16 * b);
    r->counter += 4;
    r->idx = 0;
}

static uint32_t rng_next_u32(orc_rng* r) {
    if (r->idx >= 64) rng_refill(r);
    return r->buf[r->idx++];
}

/* rand_core BlockRng::next_u64: low word first. */
uint64_t orc_rng_next_u64(orc_rng* r) {
    uint64_t lo = rng_next_u32(r);
    uint64_t hi = rng_next_u32(r);
    return (hi << 32) | lo;
}

/* rand 0.9 StandardUniform for f64: 53 high bits times 2^-53. */
double orc_rng_f64(orc_rng* r) {
    uint64_t v = orc_rng_next_u64(r);
    return (double)(v >> 11) * (1.0 / 9007199254740992.0);
}

/* rand 0.9 UniformFloat<f64>::sample_single(lo, hi): 52 mantissa bits give a
 * value in [1,2); (v - 1) * scale + lo, retried with a one-ulp smaller scale
 * if rounding lands on hi. */
double orc_rng_range(orc_rng* r, double lo, double hi) {
    double scale = hi - lo;
    for (;;) {
        uint64_t v = orc_rng_next_u64(r);
        uint64_t bits = (v >> 12) | 0x3FF0000000000000ull;
        double v12;
        memcpy(&v12, &bits, 8);
        double v01 = v12 - 1.0;
        double res = v01 * scale + lo;
        if (res < hi) return res;
        uint64_t sb;
        memcpy(&sb, &scale, 8);
        sb -= 1;
        memcpy(&scale, &sb, 8);
    }
}

/* Test hooks: raw stream and converted draws. */
void orc_rng_stream_u64(uint64_t seed, size_t n, uint64_t* out) {
    orc_rng r;
    orc_rng_seed_u64(&r, seed);
    for (size_t i = 0; i < n; ++i) out[i] = orc_rng_next_u64(&r);
}

void orc_rng_seed_key(uint64_t seed, uint32_t key[8]) {
    orc_rng r;
    orc_rng_seed_u64(&r, seed);
    memcpy(key, r.key, sizeof r.key);
}

/* ------------------------------------------------------------------------ */
/* Initial conditions                                                         */
/* ------------------------------------------------------------------------ */

/* particle.rs:189-226 shift_to_com_frame (total_mass by a sequential sum). */
void orc_shift_to_com(size_t n, double* x, double* y, double* z, double* vx, double* vy,
                      double* vz, const double* m) {
    double mt = 0.0;
    for (size_t i = 0; i < n; ++i) mt += m[i];
    if (mt == 0.0) return;
    double cx = 0, cy = 0, cz = 0, cvx = 0, cvy = 0, cvz = 0;
    for (size_t i = 0; i < n; ++i) {
        cx += m[i] * x[i]; cy += m[i] * y[i]; cz += m[i] * z[i];
        cvx += m[i] * vx[i]; cvy += m[i] * vy[i]; cvz += m[i] * vz[i];
    }
    cx /= mt; cy /= mt; cz /= mt; cvx /= mt; cvy /= mt; cvz /= mt;
    for (size_t i = 0; i < n; ++i) {
        x[i] -= cx; y[i] -= cy; z[i] -= cz;
        vx[i] -= cvx; vy[i] -= cvy; vz[i] -= cvz;
    }
}

/* plummer_sphere.rs:13-58 sample_plummer_particle. */
static void plummer_particle(orc_rng* rng, double total_mass, double a, double pos[3],
                             double vel[3]) {
    double x_mass = orc_rng_range(rng, 0.001, 1.0);
    double r = a / sqrt(pow(x_mass, -2.0 / 3.0) - 1.0);
    double cos_t = orc_rng_range(rng, -1.0, 1.0);
    double sin_t = sqrt(1.0 - cos_t * cos_t);
    double phi = orc_rng_range(rng, 0.0, 2.0 * ORC_PI);
    pos[0] = r * sin_t * cos(phi);
    pos[1] = r * sin_t * sin(phi);
    pos[2] = r * cos_t;
    double v_esc = sqrt(2.0 * 1.0 * total_mass / sqrt(r * r + a * a));
    double q;
    for (;;) {
        q = orc_rng_range(rng, 0.0, 1.0);
        double g = q * q * pow(1.0 - q * q, 3.5);
        double u = orc_rng_f64(rng);
        if (u < g / 0.1) break;
    }
    double v = q * v_esc;
    double cos_v = orc_rng_range(rng, -1.0, 1.0);
    double sin_v = sqrt(1.0 - cos_v * cos_v);
    double phi_v = orc_rng_range(rng, 0.0, 2.0 * ORC_PI);
    vel[0] = v * sin_v * cos(phi_v);
    vel[1] = v * sin_v * sin(phi_v);
    vel[2] = v * cos_v;
}

/* plummer_sphere.rs:106-133 PlummerSphere::generate. */
void orc_plummer(size_t n, double a, double total_mass, uint64_t seed, double* x, double* y,
                 double* z, double* vx, double* vy, double* vz, double* m) {
    orc_rng rng;
    orc_rng_seed_u64(&rng, seed);
    double mp = total_mass / (double)n;
    for (size_t i = 0; i < n; ++i) {
        double p[3], v[3];
        plummer_particle(&rng, total_mass, a, p, v);
        x[i] = p[0]; y[i] = p[1]; z[i] = p[2];
        vx[i] = v[0]; vy[i] = v[1]; vz[i] = v[2];
        m[i] = mp;
    }
    orc_shift_to_com(n, x, y, z, vx, vy, vz, m);
}

/* plummer_sphere.rs:135-152. powi(3) is x*x*x. */
double orc_plummer_suggested_dt(double a, double total_mass) {
    return sqrt(a * a * a / (1.0 * total_mass)) / 100.0;
}
double orc_plummer_suggested_softening(size_t n, double a) {
    double rh = 1.305 * a;
    double ms = pow(4.0 / 3.0 * ORC_PI * (rh * rh * rh) / (double)n, 1.0 / 3.0);
    return ms * 0.5;
}

/* two_body.rs:49-105; m_total = 1. */
void orc_two_body(double e, double mass_ratio, double a, double* x, double* y, double* z,
                  double* vx, double* vy, double* vz, double* m) {
    double mt = 1.0;
    double m1 = mt * mass_ratio, m2 = mt * (1.0 - mass_ratio);
    double rp = a * (1.0 - e);
    double vp = sqrt(1.0 * mt * (2.0 / rp - 1.0 / a));
    x[0] = -m2 / mt * rp; x[1] = m1 / mt * rp;
    y[0] = y[1] = z[0] = z[1] = 0.0;
    vx[0] = vx[1] = vz[0] = vz[1] = 0.0;
    vy[0] = -m2 / mt * vp; vy[1] = m1 / mt * vp;
    m[0] = m1; m[1] = m2;
}
double orc_two_body_period(double a) { return 2.0 * ORC_PI * sqrt(a * a * a / (1.0 * 1.0)); }

/* ------------------------------------------------------------------------ */
/* GalaxyCollision (galaxy_collision.rs:20-450). Parameter block in the
 * field order of the product's gravis_galaxy_params. */
typedef struct {
    uint64_t n_per_galaxy, seed;
    double disk_fraction, halo_mass_ratio, baryon_mass, disk_scale_radius,
        disk_scale_height, bulge_scale_radius, halo_scale_radius, disk_max_radius,
        halo_max_radius, separation, impact_parameter, approach_velocity;
} orc_galaxy_params;

typedef struct {
    double *x, *y, *z, *vx, *vy, *vz, *m;
    uint8_t* type;
    size_t count;
} gal_out;

static void gal_push(gal_out* o, double x, double y, double z, double vx, double vy, double vz,
                     double m, uint8_t t) {
    size_t i = o->count++;
    o->x[i] = x; o->y[i] = y; o->z[i] = z;
    o->vx[i] = vx; o->vy[i] = vy; o->vz[i] = vz;
    o->m[i] = m; o->type[i] = t;
}

/* galaxy_collision.rs:193-199 */
static double gal_disk_mass(double r, double m_disk, double rd) {
    if (r <= 0.0) return 0.0;
    double x = r / rd;
    return m_disk * (1.0 - (1.0 + x) * exp(-x));
}

/* galaxy_collision.rs:236-257 (G = 1) */
static double gal_vc(double r, double m_disk, double rd, double m_bulge, double b_a,
                     double m_halo, double h_a) {
    if (r <= 0.0) return 0.0;
    double v2_disk = 1.0 * gal_disk_mass(r, m_disk, rd) / r;
    double v2_bulge = 1.0 * m_bulge * r * r / pow(r * r + b_a * b_a, 1.5);
    double v2_halo = 1.0 * m_halo * r / ((r + h_a) * (r + h_a));
    return sqrt(v2_disk + v2_bulge + v2_halo);
}

/* galaxy_collision.rs:259-346 generate_disk */
static void gal_disk(const orc_galaxy_params* p, orc_rng* rng, size_t n_disk, double m_disk,
                     double m_bulge, double m_halo, gal_out* o) {
    double mpp = m_disk / (double)n_disk;
    double r_d = p->disk_scale_radius;
    double max_x = p->disk_max_radius / r_d;
    for (size_t k = 0; k < n_disk; ++k) {
        double r;
        for (;;) {
            double u = orc_rng_f64(rng);
            double r_try = u * p->disk_max_radius;
            double x = r_try / r_d;
            double f = x * exp(-x);
            double f_max = exp(-1.0);
            double accept = orc_rng_f64(rng);
            if (accept < f / f_max && x < max_x) { r = r_try; break; }
        }
        double phi = orc_rng_range(rng, 0.0, 2.0 * ORC_PI);
        double x = r * cos(phi);
        double y = r * sin(phi);
        double u = orc_rng_range(rng, 0.01, 0.99);
        double z = p->disk_scale_height * atanh(2.0 * u - 1.0);
        double v_c = gal_vc(r, m_disk, r_d, m_bulge, p->bulge_scale_radius, m_halo,
                            p->halo_scale_radius);
        double sigma_r = 0.15 * v_c;
        double v_r = orc_rng_range(rng, -1.0, 1.0) * sigma_r;
        double v_phi = v_c + orc_rng_range(rng, -1.0, 1.0) * sigma_r * 0.5;
        double sigma_z = 0.1 * v_c;
        double v_z = orc_rng_range(rng, -1.0, 1.0) * sigma_z;
        double vx = v_r * cos(phi) - v_phi * sin(phi);
        double vy = v_r * sin(phi) + v_phi * cos(phi);
        gal_push(o, x, y, z, vx, vy, v_z, mpp, 1);
    }
}

/* galaxy_collision.rs:348-366 generate_bulge */
static void gal_bulge(const orc_galaxy_params* p, orc_rng* rng, size_t n_bulge, double m_bulge,
                      gal_out* o) {
    double mpp = m_bulge / (double)n_bulge;
    for (size_t k = 0; k < n_bulge; ++k) {
        double pos[3], vel[3];
        plummer_particle(rng, m_bulge, p->bulge_scale_radius, pos, vel);
        gal_push(o, pos[0], pos[1], pos[2], vel[0], vel[1], vel[2], mpp, 2);
    }
}

/* scenarios/mod.rs:25-30 box_muller */
static void gal_box_muller(orc_rng* rng, double* g1, double* g2) {
    double u1 = orc_rng_range(rng, 0.001, 1.0);
    double u2 = orc_rng_range(rng, 0.0, 2.0 * ORC_PI);
    double r = sqrt(-2.0 * log(u1));
    *g1 = r * cos(u2);
    *g2 = r * sin(u2);
}

/* galaxy_collision.rs:368-450 generate_halo */
static void gal_halo(const orc_galaxy_params* p, orc_rng* rng, size_t n_halo, double m_disk,
                     double m_bulge, double m_halo, gal_out* o) {
    double mpp = m_halo / (double)n_halo;
    double a = p->halo_scale_radius;
    double xt = p->halo_max_radius / (p->halo_max_radius + a);
    double u_max = xt * xt;
    for (size_t k = 0; k < n_halo; ++k) {
        double u = orc_rng_range(rng, 0.001, u_max);
        double sqrt_u = sqrt(u);
        double r = a * sqrt_u / (1.0 - sqrt_u);
        double cos_t = orc_rng_range(rng, -1.0, 1.0);
        double sin_t = sqrt(1.0 - cos_t * cos_t);
        double phi = orc_rng_range(rng, 0.0, 2.0 * ORC_PI);
        double px = r * sin_t * cos(phi);
        double py = r * sin_t * sin(phi);
        double pz = r * cos_t;
        double v_c = gal_vc(r, m_disk, p->disk_scale_radius, m_bulge, p->bulge_scale_radius,
                            m_halo, a);
        double sigma = v_c / sqrt(2.0);
        double g1, g2, g3, unused;
        gal_box_muller(rng, &g1, &g2);
        gal_box_muller(rng, &g3, &unused);
        gal_push(o, px, py, pz, sigma * g1, sigma * g2, sigma * g3, mpp, 3);
    }
}

/* galaxy_collision.rs:75-157 generate; n = 2 * n_per_galaxy. Every array
 * must hold n entries. com_shift = 0 skips the final shift_to_com_frame
 * (a test hook: the raw samples are checked against the profiles). */
void orc_galaxy_collision(const orc_galaxy_params* p, int com_shift, double* x, double* y,
                          double* z, double* vx, double* vy, double* vz, double* m,
                          uint8_t* type) {
    orc_rng rng;
    orc_rng_seed_u64(&rng, p->seed);
    double m_disk = p->baryon_mass * p->disk_fraction;
    double m_bulge = p->baryon_mass * (1.0 - p->disk_fraction);
    double m_halo = p->baryon_mass * p->halo_mass_ratio;
    double m_total = m_disk + m_bulge + m_halo;
    size_t n_disk = (size_t)round((double)p->n_per_galaxy * m_disk / m_total);
    size_t n_bulge = (size_t)round((double)p->n_per_galaxy * m_bulge / m_total);
    size_t n_halo = p->n_per_galaxy - n_disk - n_bulge;
    gal_out o = {x, y, z, vx, vy, vz, m, type, 0};
    size_t bounds[3] = {0, 0, 0};
    for (int g = 0; g < 2; ++g) {
        gal_disk(p, &rng, n_disk, m_disk, m_bulge, m_halo, &o);
        gal_bulge(p, &rng, n_bulge, m_bulge, &o);
        gal_halo(p, &rng, n_halo, m_disk, m_bulge, m_halo, &o);
        bounds[g + 1] = o.count;
    }
    double half_sep = p->separation / 2.0;
    double half_imp = p->impact_parameter / 2.0;
    double half_vel = p->approach_velocity / 2.0;
    for (size_t i = bounds[0]; i < bounds[1]; ++i) {
        x[i] -= half_sep; y[i] -= half_imp; vx[i] += half_vel;
    }
    for (size_t i = bounds[1]; i < bounds[2]; ++i) {
        x[i] += half_sep; y[i] += half_imp; vx[i] -= half_vel;
    }
    if (com_shift) orc_shift_to_com(o.count, x, y, z, vx, vy, vz, m);
}

/* galaxy_collision.rs:159-185 */
double orc_galaxy_suggested_dt(const orc_galaxy_params* p) {
    double m_enc_disk = 0.26 * p->baryon_mass * p->disk_fraction;
    double r = p->disk_scale_radius;
    double a_b = p->bulge_scale_radius;
    double m_bulge = p->baryon_mass * (1.0 - p->disk_fraction);
    double m_enc_bulge = m_bulge * (r * r * r) / pow(r * r + a_b * a_b, 1.5);
    double a_h = p->halo_scale_radius;
    double m_halo = p->baryon_mass * p->halo_mass_ratio;
    double x = r / a_h;
    double m_enc_halo = m_halo * x * x / ((1.0 + x) * (1.0 + x));
    double m_enc = m_enc_disk + m_enc_bulge + m_enc_halo;
    double v_c = sqrt(1.0 * m_enc / r);
    double t_orb = 2.0 * ORC_PI * r / v_c;
    return t_orb / 500.0;
}

/* ------------------------------------------------------------------------ */
/* Brute force (gravity.rs:47-80): i ascending, j>i ascending, a_i then a_j. */
/* ------------------------------------------------------------------------ */
void orc_bf_accel(size_t n, double eps, const double* x, const double* y, const double* z,
                  const double* m, double* ax, double* ay, double* az) {
    double eps2 = eps * eps;
    for (size_t i = 0; i < n; ++i) {
        for (size_t j = i + 1; j < n; ++j) {
            double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
            double r2 = dx * dx + dy * dy + dz * dz + eps2;
            double inv_r3 = 1.0 / (r2 * sqrt(r2));
            double fi = 1.0 * m[j] * inv_r3;
            double fj = 1.0 * m[i] * inv_r3;
            ax[i] += fi * dx; ay[i] += fi * dy; az[i] += fi * dz;
            ax[j] -= fj * dx; ay[j] -= fj * dy; az[j] -= fj * dz;
        }
    }
}

/* Per-target brute force restricted to a subset of targets (sampling oracle
 * for large N): a_t = sum over j != t, j ascending, of the pair term in the
 * same arithmetic as barnes_hut.rs:114-135. Also returns sum_j |a_tj|, the
 * condition number used to normalise force errors. */
void orc_bf_targets(size_t n, double eps, const double* x, const double* y, const double* z,
                    const double* m, size_t nt, const int64_t* targets, double* ax, double* ay,
                    double* az, double* absum, int nthreads) {
    double eps2 = eps * eps;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (long long k = 0; k < (long long)nt; ++k) {
        size_t t = (size_t)targets[k];
        double a0 = 0, a1 = 0, a2 = 0, s = 0;
        double px = x[t], py = y[t], pz = z[t];
        for (size_t j = 0; j < n; ++j) {
            if (j == t) continue;
            double dx = x[j] - px, dy = y[j] - py, dz = z[j] - pz;
            double r2 = dx * dx + dy * dy + dz * dz + eps2;
            double inv_r3 = 1.0 / (r2 * sqrt(r2));
            double f = 1.0 * m[j] * inv_r3;
            a0 += f * dx; a1 += f * dy; a2 += f * dz;
            s += fabs(f) * sqrt(dx * dx + dy * dy + dz * dz);
        }
        ax[k] = a0; ay[k] = a1; az[k] = a2;
        if (absum) absum[k] = s;
    }
}

/* ------------------------------------------------------------------------ */
/* Reference octree + Barnes-Hut (octree.rs, barnes_hut.rs).                 */
/* ------------------------------------------------------------------------ */
#define ONONE 0xFFFFFFFFu
#define OMAX_DEPTH 64

typedef struct {
    double c[3], hw;
    double mass, com[3];
    uint32_t child[8];
    uint32_t pidx;
} onode;

typedef struct {
    onode* nodes;
    size_t len, cap;
} otree;

/* octree.rs:14-44 BoundingBox::containing */
void orc_root_box(size_t n, const double* x, const double* y, const double* z, double out[4]) {
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < n; ++i) {
        double p[3] = {x[i], y[i], z[i]};
        for (int d = 0; d < 3; ++d) {
            mn[d] = fmin(mn[d], p[d]);
            mx[d] = fmax(mx[d], p[d]);
        }
    }
    out[0] = 0.5 * (mn[0] + mx[0]);
    out[1] = 0.5 * (mn[1] + mx[1]);
    out[2] = 0.5 * (mn[2] + mx[2]);
    double ext = fmax(fmax(mx[0] - mn[0], mx[1] - mn[1]), mx[2] - mn[2]);
    out[3] = 0.5 * ext * 1.001 + 1e-10;
}

static uint32_t oalloc(otree* t, const double c[3], double hw) {
    if (t->len == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 64;
        t->nodes = (onode*)realloc(t->nodes, t->cap * sizeof(onode));
    }
    onode* nd = &t->nodes[t->len];
    nd->c[0] = c[0]; nd->c[1] = c[1]; nd->c[2] = c[2]; nd->hw = hw;
    nd->mass = 0.0; nd->com[0] = nd->com[1] = nd->com[2] = 0.0;
    for (int k = 0; k < 8; ++k) nd->child[k] = ONONE;
    nd->pidx = ONONE;
    return (uint32_t)t->len++;
}

static int ooctant(const onode* nd, double x, double y, double z) {
    int o = 0;
    if (x >= nd->c[0]) o |= 1;
    if (y >= nd->c[1]) o |= 2;
    if (z >= nd->c[2]) o |= 4;
    return o;
}

static void ochild_bounds(const onode* nd, int oct, double c[3], double* hw) {
    double q = nd->hw * 0.5;
    c[0] = nd->c[0] + ((oct & 1) ? q : -q);
    c[1] = nd->c[1] + ((oct & 2) ? q : -q);
    c[2] = nd->c[2] + ((oct & 4) ? q : -q);
    *hw = q;
}

static void oupdate_com(onode* nd, double px, double py, double pz, double mass) {
    double nm = nd->mass + mass;
    if (nm > 0.0) {
        double inv = 1.0 / nm;
        nd->com[0] = (nd->mass * nd->com[0] + mass * px) * inv;
        nd->com[1] = (nd->mass * nd->com[1] + mass * py) * inv;
        nd->com[2] = (nd->mass * nd->com[2] + mass * pz) * inv;
    }
    nd->mass = nm;
}

/* octree.rs:188-239 insert */
static void oinsert(otree* t, uint32_t ni, uint32_t pi, double px, double py, double pz,
                    double mass, int depth) {
    onode* nd = &t->nodes[ni];
    if (nd->mass == 0.0) {
        nd->pidx = pi; nd->mass = mass;
        nd->com[0] = px; nd->com[1] = py; nd->com[2] = pz;
        return;
    }
    if (depth >= OMAX_DEPTH) {
        oupdate_com(nd, px, py, pz, mass);
        return;
    }
    uint32_t existing = nd->pidx;
    if (existing != ONONE) {
        double oc[3] = {nd->com[0], nd->com[1], nd->com[2]};
        double om = nd->mass;
        t->nodes[ni].pidx = ONONE;
        int ooct = ooctant(&t->nodes[ni], oc[0], oc[1], oc[2]);
        double cc[3], chw;
        ochild_bounds(&t->nodes[ni], ooct, cc, &chw);
        uint32_t ci = oalloc(t, cc, chw);
        t->nodes[ni].child[ooct] = ci;
        oinsert(t, ci, existing, oc[0], oc[1], oc[2], om, depth + 1);
        int noct = ooctant(&t->nodes[ni], px, py, pz);
        if (t->nodes[ni].child[noct] == ONONE) {
            ochild_bounds(&t->nodes[ni], noct, cc, &chw);
            uint32_t c2 = oalloc(t, cc, chw);
            t->nodes[ni].child[noct] = c2;
        }
        uint32_t c3 = t->nodes[ni].child[noct];
        oupdate_com(&t->nodes[ni], px, py, pz, mass);
        oinsert(t, c3, pi, px, py, pz, mass, depth + 1);
    } else {
        oupdate_com(nd, px, py, pz, mass);
        int oct = ooctant(&t->nodes[ni], px, py, pz);
        if (t->nodes[ni].child[oct] == ONONE) {
            double cc[3], chw;
            ochild_bounds(&t->nodes[ni], oct, cc, &chw);
            uint32_t ci = oalloc(t, cc, chw);
            t->nodes[ni].child[oct] = ci;
        }
        uint32_t ci = t->nodes[ni].child[oct];
        oinsert(t, ci, pi, px, py, pz, mass, depth + 1);
    }
}

static void obuild(otree* t, size_t n, const double* x, const double* y, const double* z,
                   const double* m) {
    t->len = 0;
    double box[4];
    orc_root_box(n, x, y, z, box);
    oalloc(t, box, box[3]);
    for (size_t i = 0; i < n; ++i) oinsert(t, 0, (uint32_t)i, x[i], y[i], z[i], m[i], 0);
}

/* Test hook: build the reference octree and report node count plus root
 * mass/COM (octree.rs tests). */
size_t orc_ref_octree_summary(size_t n, const double* x, const double* y, const double* z,
                              const double* m, double root[4]) {
    otree t = {0};
    obuild(&t, n, x, y, z, m);
    root[0] = t.nodes[0].mass;
    root[1] = t.nodes[0].com[0]; root[2] = t.nodes[0].com[1]; root[3] = t.nodes[0].com[2];
    size_t len = t.len;
    free(t.nodes);
    return len;
}

/* barnes_hut.rs:44-108 tree_walk_accel */
static void owalk(const otree* t, double theta2, double px, double py, double pz, uint32_t self,
                  double eps2, double acc[3]) {
    uint32_t stack[512];
    int sl = 1;
    stack[0] = 0;
    acc[0] = acc[1] = acc[2] = 0.0;
    while (sl > 0) {
        const onode* nd = &t->nodes[stack[--sl]];
        if (nd->mass == 0.0) continue;
        int accept;
        if (nd->pidx != ONONE) {
            if (nd->pidx == self) continue;
            accept = 1;
        } else {
            double dx = nd->com[0] - px, dy = nd->com[1] - py, dz = nd->com[2] - pz;
            double d2 = dx * dx + dy * dy + dz * dz;
            double s = nd->hw * 2.0;
            accept = (s * s < theta2 * d2);
        }
        if (accept) {
            double dx = nd->com[0] - px, dy = nd->com[1] - py, dz = nd->com[2] - pz;
            double r2 = dx * dx + dy * dy + dz * dz + eps2;
            double inv_r3 = 1.0 / (r2 * sqrt(r2));
            double f = 1.0 * nd->mass * inv_r3;
            acc[0] += f * dx; acc[1] += f * dy; acc[2] += f * dz;
            continue;
        }
        for (int k = 0; k < 8; ++k)
            if (nd->child[k] != ONONE) stack[sl++] = nd->child[k];
    }
}

/* barnes_hut.rs:139-192 compute_accelerations (accumulate into a). */
void orc_bh_accel(size_t n, double eps, double theta, const double* x, const double* y,
                  const double* z, const double* m, double* ax, double* ay, double* az,
                  int nthreads) {
    if (n == 0) return;
    otree t = {0};
    obuild(&t, n, x, y, z, m);
    double eps2 = eps * eps, theta2 = theta * theta;
    double* tmp = (double*)malloc(3 * n * sizeof(double));
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64) if (n >= 1000)
#endif
    for (long long i = 0; i < (long long)n; ++i)
        owalk(&t, theta2, x[i], y[i], z[i], (uint32_t)i, eps2, tmp + 3 * i);
    for (size_t i = 0; i < n; ++i) {
        ax[i] += tmp[3 * i]; ay[i] += tmp[3 * i + 1]; az[i] += tmp[3 * i + 2];
    }
    free(tmp);
    free(t.nodes);
}

/* ------------------------------------------------------------------------ */
/* Linear octree (the HIP backend's Barnes-Hut definition).                  */
/* ------------------------------------------------------------------------ */
#define MORTON_BITS 21
#define MORTON_MAXQ 2097151u

static inline uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x1FFFFFu;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

static inline uint32_t quantise(double v, double lo, double inv) {
    double q = (v - lo) * inv;
    if (!(q >= 0.0)) return 0u;
    if (q >= 2097151.0) return MORTON_MAXQ;
    return (uint32_t)q;
}

/* 63-bit key over the root cube box = (cx, cy, cz, hw): per axis
 * q = floor((p - (c - hw)) * (2^21 / (2 hw))) clamped to [0, 2^21 - 1]. */
void orc_morton63(size_t n, const double* x, const double* y, const double* z,
                  const double box[4], uint64_t* keys) {
    double inv = 2097152.0 / (2.0 * box[3]);
    double lx = box[0] - box[3], ly = box[1] - box[3], lz = box[2] - box[3];
    for (size_t i = 0; i < n; ++i) {
        uint32_t qx = quantise(x[i], lx, inv), qy = quantise(y[i], ly, inv),
                 qz = quantise(z[i], lz, inv);
        keys[i] = spread3(qx) | (spread3(qy) << 1) | (spread3(qz) << 2);
    }
}

/* Stable LSD radix sort, 8-bit digits; perm[k] = original index. */
void orc_sort_keys(size_t n, const uint64_t* keys, uint64_t* skeys, uint32_t* perm) {
    uint64_t* k0 = (uint64_t*)malloc(n * sizeof(uint64_t));
    uint64_t* k1 = (uint64_t*)malloc(n * sizeof(uint64_t));
    uint32_t* v0 = (uint32_t*)malloc(n * sizeof(uint32_t));
    uint32_t* v1 = (uint32_t*)malloc(n * sizeof(uint32_t));
    for (size_t i = 0; i < n; ++i) { k0[i] = keys[i]; v0[i] = (uint32_t)i; }
    for (int pass = 0; pass < 8; ++pass) {
        int sh = 8 * pass;
        size_t cnt[257] = {0};
        for (size_t i = 0; i < n; ++i) cnt[((k0[i] >> sh) & 255) + 1]++;
        for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
        for (size_t i = 0; i < n; ++i) {
            size_t d = (k0[i] >> sh) & 255;
            k1[cnt[d]] = k0[i]; v1[cnt[d]] = v0[i]; cnt[d]++;
        }
        uint64_t* tk = k0; k0 = k1; k1 = tk;
        uint32_t* tv = v0; v0 = v1; v1 = tv;
    }
    memcpy(skeys, k0, n * sizeof(uint64_t));
    memcpy(perm, v0, n * sizeof(uint32_t));
    free(k0); free(k1); free(v0); free(v1);
}

/* Linear octree node (structure of arrays owned by the caller).
 *   level[v]       depth, root = 0
 *   first[v],cnt[v] range in sorted order
 *   child[v]       index of first child (children contiguous), nchild[v]
 *   mass[v], mx/my/mz[v]  FP64 sums of m and m*x over the range
 *   com = mx / mass (first particle position if mass == 0)
 * Node order: breadth first; children of a node in increasing octant digit.
 * A node is a leaf when cnt <= leaf or level == 21.
 * Leaf sums: sequential over the range (ascending). Internal sums:
 * sequential over children (ascending). No FMA. */
typedef struct {
    uint32_t *level, *first, *cnt, *child, *nchild;
    double *mass, *comx, *comy, *comz;
    double *blo, *bhi; /* particle bounding box, 3 per node */
    size_t len, cap;
} ltree;

static void lt_reserve(ltree* t, size_t need) {
    if (need <= t->cap) return;
    size_t c = t->cap ? t->cap : 1024;
    while (c < need) c *= 2;
    t->level = (uint32_t*)realloc(t->level, c * 4);
    t->first = (uint32_t*)realloc(t->first, c * 4);
    t->cnt = (uint32_t*)realloc(t->cnt, c * 4);
    t->child = (uint32_t*)realloc(t->child, c * 4);
    t->nchild = (uint32_t*)realloc(t->nchild, c * 4);
    t->mass = (double*)realloc(t->mass, c * 8);
    t->comx = (double*)realloc(t->comx, c * 8);
    t->comy = (double*)realloc(t->comy, c * 8);
    t->comz = (double*)realloc(t->comz, c * 8);
    t->blo = (double*)realloc(t->blo, 3 * c * 8);
    t->bhi = (double*)realloc(t->bhi, 3 * c * 8);
    t->cap = c;
}

static void lt_free(ltree* t) {
    free(t->level); free(t->first); free(t->cnt); free(t->child); free(t->nchild);
    free(t->mass); free(t->comx); free(t->comy); free(t->comz);
    free(t->blo); free(t->bhi);
}

static inline uint32_t digit_at(uint64_t key, uint32_t child_level) {
    return (uint32_t)((key >> (3 * (MORTON_BITS - child_level))) & 7u);
}

static void lt_build(ltree* t, size_t n, const uint64_t* sk, uint32_t leaf, const double* sx,
                     const double* sy, const double* sz, const double* sm) {
    t->len = 0;
    lt_reserve(t, 1);
    t->level[0] = 0; t->first[0] = 0; t->cnt[0] = (uint32_t)n;
    t->len = 1;
    size_t lb = 0, le = 1;
    while (lb < le) {
        for (size_t v = lb; v < le; ++v) {
            uint32_t L = t->level[v];
            if (t->cnt[v] <= leaf || L == MORTON_BITS) {
                t->child[v] = 0; t->nchild[v] = 0;
                continue;
            }
            uint32_t b = t->first[v], e = b + t->cnt[v];
            t->child[v] = (uint32_t)t->len;
            uint32_t nc = 0;
            uint32_t s = b;
            while (s < e) {
                uint32_t d = digit_at(sk[s], L + 1);
                uint32_t q = s + 1;
                while (q < e && digit_at(sk[q], L + 1) == d) ++q;
                lt_reserve(t, t->len + 1);
                t->level[t->len] = L + 1;
                t->first[t->len] = s;
                t->cnt[t->len] = q - s;
                t->len++;
                nc++;
                s = q;
            }
            t->nchild[v] = nc;
        }
        lb = le;
        le = t->len;
    }
    /* bottom-up sums: nodes in reverse BFS order have children done first */
    for (size_t vi = t->len; vi-- > 0;) {
        double M = 0, X = 0, Y = 0, Z = 0;
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        if (t->nchild[vi] == 0) {
            for (uint32_t k = t->first[vi]; k < t->first[vi] + t->cnt[vi]; ++k) {
                M = M + sm[k];
                X = X + sm[k] * sx[k]; Y = Y + sm[k] * sy[k]; Z = Z + sm[k] * sz[k];
                lo[0] = fmin(lo[0], sx[k]); hi[0] = fmax(hi[0], sx[k]);
                lo[1] = fmin(lo[1], sy[k]); hi[1] = fmax(hi[1], sy[k]);
                lo[2] = fmin(lo[2], sz[k]); hi[2] = fmax(hi[2], sz[k]);
            }
        } else {
            for (uint32_t c = t->child[vi]; c < t->child[vi] + t->nchild[vi]; ++c) {
                M = M + t->mass[c];
                X = X + t->comx[c]; Y = Y + t->comy[c]; Z = Z + t->comz[c];
                for (int d = 0; d < 3; ++d) {
                    lo[d] = fmin(lo[d], t->blo[3 * c + d]);
                    hi[d] = fmax(hi[d], t->bhi[3 * c + d]);
                }
            }
        }
        t->mass[vi] = M; t->comx[vi] = X; t->comy[vi] = Y; t->comz[vi] = Z;
        for (int d = 0; d < 3; ++d) { t->blo[3 * vi + d] = lo[d]; t->bhi[3 * vi + d] = hi[d]; }
    }
}

/* After lt_build, com arrays hold m*x sums; convert to positions in place
 * (kept separate so internal sums above use the raw m*x sums). */
static void lt_finish_com(ltree* t, const double* sx, const double* sy, const double* sz) {
    for (size_t v = 0; v < t->len; ++v) {
        double M = t->mass[v];
        if (M != 0.0) {
            t->comx[v] = t->comx[v] / M; t->comy[v] = t->comy[v] / M; t->comz[v] = t->comz[v] / M;
        } else {
            uint32_t f = t->first[v];
            t->comx[v] = sx[f]; t->comy[v] = sy[f]; t->comz[v] = sz[f];
        }
    }
}

/* Test hook: full tree export. Arrays must hold cap entries; returns the node
 * count, or 0 if cap is too small. com_out holds the FP64 com (x,y,z) and
 * mass as 4 doubles per node. */
size_t orc_linear_tree(size_t n, const double* x, const double* y, const double* z,
                       const double* m, uint32_t leaf, size_t cap, uint64_t* skeys_out,
                       uint32_t* perm_out, uint32_t* level, uint32_t* first, uint32_t* cnt,
                       uint32_t* child, uint32_t* nchild, double* com_out, double box_out[4]) {
    double box[4];
    orc_root_box(n, x, y, z, box);
    if (box_out) memcpy(box_out, box, sizeof box);
    uint64_t* keys = (uint64_t*)malloc(n * 8);
    uint64_t* sk = (uint64_t*)malloc(n * 8);
    uint32_t* perm = (uint32_t*)malloc(n * 4);
    orc_morton63(n, x, y, z, box, keys);
    orc_sort_keys(n, keys, sk, perm);
    double* s = (double*)malloc(4 * n * 8);
    for (size_t k = 0; k < n; ++k) {
        s[k] = x[perm[k]]; s[n + k] = y[perm[k]]; s[2 * n + k] = z[perm[k]]; s[3 * n + k] = m[perm[k]];
    }
    ltree t = {0};
    lt_build(&t, n, sk, leaf, s, s + n, s + 2 * n, s + 3 * n);
    lt_finish_com(&t, s, s + n, s + 2 * n);
    size_t len = t.len;
    if (skeys_out) memcpy(skeys_out, sk, n * 8);
    if (perm_out) memcpy(perm_out, perm, n * 4);
    if (len <= cap) {
        for (size_t v = 0; v < len; ++v) {
            level[v] = t.level[v]; first[v] = t.first[v]; cnt[v] = t.cnt[v];
            child[v] = t.child[v]; nchild[v] = t.nchild[v];
            com_out[4 * v] = t.comx[v]; com_out[4 * v + 1] = t.comy[v];
            com_out[4 * v + 2] = t.comz[v]; com_out[4 * v + 3] = t.mass[v];
        }
    } else {
        len = 0;
    }
    lt_free(&t);
    free(keys); free(sk); free(perm); free(s);
    return len;
}

/* Group walk for one group of sorted particles [g0, g1).
 * Breadth-first frontier; for each node in frontier order:
 *   accept (s^2 < theta^2 d_min^2, d_min = distance from COM to the group
 *   box)                      -> monopole source appended to the list
 *   leaf, not accepted        -> the leaf's particles appended in order
 *   internal, not accepted    -> children appended to the next frontier
 * Accelerations are accumulated in list order with the reference pair
 * arithmetic (barnes_hut.rs:114-135); a source at r2 == 0 is skipped. */
static size_t g_max_frontier = 0;

/* The device walk's interaction-list bookkeeping, restated so the sums are
 * formed in the same order: frontier nodes are taken in chunks of 64; a
 * chunk's sources (1 per accepted node, the particles of an opened leaf)
 * get consecutive positions in a list of LIST_CAP entries, which restarts
 * at 0 (a flush) when the chunk would not fit, or, for a chunk larger than
 * the list, whenever it is full. A group of at most 32 (16) targets is held
 * by 2 (4) lanes per target; lane copy r sums the sources at positions
 * p with p % R == r, and the copies are added as (r0 + r1) [+ (r2 + r3)]. */
#define LIST_CAP 1024u
#define WAVE 64u

static int group_replicas(size_t cnt) { return cnt <= 16 ? 4 : (cnt <= 32 ? 2 : 1); }

typedef struct {
    double* p;   /* (g1 - g0) * R * 3 partial sums: target, copy, component */
    int R;
    uint32_t nl; /* list position */
} rsum;

static void apply_source(const double* sx, const double* sy, const double* sz, double eps2,
                         const double* e2s, size_t g0, size_t g1, double srcx, double srcy,
                         double srcz, double srcm, rsum* acc, int r) {
    for (size_t i = g0; i < g1; ++i) {
        double dx = srcx - sx[i], dy = srcy - sy[i], dz = srcz - sz[i];
        double r2 = dx * dx + dy * dy + dz * dz + (e2s ? e2s[i] : eps2);
        if (r2 == 0.0) continue;
        double inv_r3 = 1.0 / (r2 * sqrt(r2));
        double fm = 1.0 * srcm * inv_r3;
        double* q = acc->p + ((i - g0) * acc->R + r) * 3;
        q[0] = q[0] + fm * dx; q[1] = q[1] + fm * dy; q[2] = q[2] + fm * dz;
    }
}

/* One breadth-first walk for targets [g0, g1) with opening box (lo, hi).
 * Returns 1 if a next frontier exceeds fcap (the sums are then discarded). */
static int walk_box(const ltree* t, double side0, double theta2, double eps2, const double* e2s,
                    const double* sx, const double* sy, const double* sz, const double* sm,
                    size_t g0, size_t g1, const double lo[3], const double hi[3], rsum* acc,
                    uint32_t* fa, uint32_t* fb, uint32_t fcap, uint64_t* ninter) {
    size_t nf = 1;
    fa[0] = 0;
    acc->nl = 0;
    uint64_t cnt_inter = 0;
    while (nf > 0) {
        size_t nn = 0;
        for (size_t base = 0; base < nf; base += WAVE) {
            size_t end = base + WAVE < nf ? base + WAVE : nf;
            int acc_flag[WAVE];
            uint32_t total = 0;
            for (size_t f = base; f < end; ++f) {
                uint32_t v = fa[f];
                double cx = t->comx[v], cy = t->comy[v], cz = t->comz[v];
                double ex = fmax(fmax(lo[0] - cx, cx - hi[0]), 0.0);
                double ey = fmax(fmax(lo[1] - cy, cy - hi[1]), 0.0);
                double ez = fmax(fmax(lo[2] - cz, cz - hi[2]), 0.0);
                double d2 = ex * ex + ey * ey + ez * ez;
                double s = ldexp(side0, -(int)t->level[v]);
                int accept = (s * s < theta2 * d2);
                acc_flag[f - base] = accept;
                if (accept) total += 1;
                else if (t->nchild[v] == 0) total += t->cnt[v];
            }
            if (total > 0 && acc->nl + total > LIST_CAP) acc->nl = 0;
            for (size_t f = base; f < end; ++f) {
                uint32_t v = fa[f];
                int accept = acc_flag[f - base];
                if (!accept && t->nchild[v] != 0) continue;
                uint32_t kb, ke;
                int single = accept;
                if (single) { kb = 0; ke = 1; } else { kb = t->first[v]; ke = kb + t->cnt[v]; }
                for (uint32_t k = kb; k < ke; ++k) {
                    if (total > LIST_CAP && acc->nl == LIST_CAP) acc->nl = 0;
                    int r = (int)(acc->nl % (uint32_t)acc->R);
                    acc->nl++;
                    if (single)
                        apply_source(sx, sy, sz, eps2, e2s, g0, g1, t->comx[v], t->comy[v],
                                     t->comz[v], t->mass[v], acc, r);
                    else
                        apply_source(sx, sy, sz, eps2, e2s, g0, g1, sx[k], sy[k], sz[k], sm[k],
                                     acc, r);
                    cnt_inter += (uint64_t)(g1 - g0);
                }
            }
            for (size_t f = base; f < end; ++f) {
                uint32_t v = fa[f];
                if (!acc_flag[f - base] && t->nchild[v] != 0)
                    for (uint32_t c = 0; c < t->nchild[v]; ++c) fb[nn++] = t->child[v] + c;
            }
        }
        uint32_t* tmp = fa; fa = fb; fb = tmp;
        nf = nn;
        if (nf > g_max_frontier) {
#ifdef _OPENMP
#pragma omp critical
#endif
            if (nf > g_max_frontier) g_max_frontier = nf;
        }
        if (nf > fcap) {
            if (ninter) *ninter += cnt_inter;
            return 1;
        }
    }
    if (ninter) *ninter += cnt_inter;
    return 0;
}

/* copies combined as (r0 + r1) [+ (r2 + r3)] */
static double combine(const double* q, int R, int c) {
    if (R == 1) return q[c];
    double a = q[c] + q[3 + c];
    if (R == 2) return a;
    return a + (q[6 + c] + q[9 + c]);
}

static void gwalk(const ltree* t, double side0, double theta2, double eps2, const double* e2s,
                  const double* sx,
                  const double* sy, const double* sz, const double* sm, size_t g0, size_t g1,
                  double* ax, double* ay, double* az, uint32_t* fa, uint32_t* fb, uint32_t fcap,
                  uint64_t* ninter, double* scratch) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = g0; i < g1; ++i) {
        lo[0] = fmin(lo[0], sx[i]); hi[0] = fmax(hi[0], sx[i]);
        lo[1] = fmin(lo[1], sy[i]); hi[1] = fmax(hi[1], sy[i]);
        lo[2] = fmin(lo[2], sz[i]); hi[2] = fmax(hi[2], sz[i]);
    }
    rsum acc = {scratch, group_replicas(g1 - g0), 0};
    const size_t np = (g1 - g0) * (size_t)acc.R * 3;
    for (size_t k = 0; k < np; ++k) scratch[k] = 0.0;
    if (walk_box(t, side0, theta2, eps2, e2s, sx, sy, sz, sm, g0, g1, lo, hi, &acc, fa, fb,
                 fcap, ninter)) {
        for (size_t k = 0; k < np; ++k) scratch[k] = 0.0;
        for (size_t i = g0; i < g1; ++i) {
            double p[3] = {sx[i], sy[i], sz[i]};
            rsum one = {scratch + (i - g0) * (size_t)acc.R * 3, acc.R, 0};
            walk_box(t, side0, theta2, eps2, e2s, sx, sy, sz, sm, i, i + 1, p, p, &one, fa, fb,
                     0xFFFFFFFFu, ninter);
        }
    }
    for (size_t i = g0; i < g1; ++i) {
        const double* q = scratch + (i - g0) * (size_t)acc.R * 3;
        ax[i] = combine(q, acc.R, 0);
        ay[i] = combine(q, acc.R, 1);
        az[i] = combine(q, acc.R, 2);
    }
}

/* Largest breadth-first frontier seen by the group walk since the last call
 * (test hook for the device walk's frontier capacity). */
size_t orc_bh_max_frontier(void) {
    size_t v = g_max_frontier;
    g_max_frontier = 0;
    return v;
}

/* Target groups (at most `group` particles each, contiguous in sorted order):
 * for every internal node holding more than `group` particles, its children
 * are visited in order and consecutive children holding at most `group`
 * particles are packed greedily while the total stays within `group` or,
 * with `split`, each maximal run of such consecutive children is cut into
 * runs of `group` particles; a child leaf holding more than `group`
 * particles (coincident points at level 21) is cut into runs of `group`; a
 * root holding at most `group` particles is one group. Every particle lies
 * in exactly one group. Group order does not affect any result. Returns the
 * group count; gf/gc receive (first, count). */
typedef struct {
    uint32_t *f, *c;
    size_t n, cap;
} glist;

static void gl_push(glist* g, uint32_t f, uint32_t c) {
    if (g->n == g->cap) {
        g->cap = g->cap ? 2 * g->cap : 64;
        g->f = (uint32_t*)realloc(g->f, g->cap * 4);
        g->c = (uint32_t*)realloc(g->c, g->cap * 4);
    }
    g->f[g->n] = f;
    g->c[g->n] = c;
    g->n++;
}

static void gl_runs(glist* g, uint32_t first, uint32_t cnt, uint32_t group) {
    for (uint32_t b = 0; b < cnt; b += group) gl_push(g, first + b, cnt - b < group ? cnt - b : group);
}

static size_t lt_groups(const ltree* t, uint32_t group, int split, uint32_t** gf,
                        uint32_t** gc) {
    glist g = {0};
    if (t->cnt[0] <= group || t->nchild[0] == 0) {
        gl_runs(&g, t->first[0], t->cnt[0], group);
    } else {
        for (size_t p = 0; p < t->len; ++p) {
            if (t->cnt[p] <= group || t->nchild[p] == 0) continue;
            uint32_t rf = 0, rc = 0;
            for (uint32_t q = 0; q < t->nchild[p]; ++q) {
                uint32_t c = t->child[p] + q;
                if (t->cnt[c] <= group) {
                    if (split) {
                        if (rc == 0) rf = t->first[c];
                        rc += t->cnt[c];
                        continue;
                    }
                    if (rc > 0) {
                        if (rc + t->cnt[c] <= group) {
                            rc += t->cnt[c];
                            continue;
                        }
                        gl_push(&g, rf, rc);
                    }
                    rf = t->first[c];
                    rc = t->cnt[c];
                } else {
                    if (rc > 0) gl_runs(&g, rf, rc, group);
                    rc = 0;
                    if (t->nchild[c] == 0) gl_runs(&g, t->first[c], t->cnt[c], group);
                }
            }
            if (rc > 0) gl_runs(&g, rf, rc, group);
        }
    }
    *gf = g.f;
    *gc = g.c;
    return g.n;
}

/* Full linear-octree Barnes-Hut: box, keys, sort, tree, group walk.
 * Accumulates into ax/ay/az in the caller's order. Returns the total number
 * of pair interactions evaluated (targets x sources). */
static uint64_t bh_group(size_t n, double eps, const double* eps_i, double theta, uint32_t leaf,
                         uint32_t group, uint32_t fcap, int split, const double* x,
                         const double* y, const double* z, const double* m, double* ax,
                         double* ay, double* az, int nthreads) {
    if (n == 0) return 0;
    double box[4];
    orc_root_box(n, x, y, z, box);
    uint64_t* keys = (uint64_t*)malloc(n * 8);
    uint64_t* sk = (uint64_t*)malloc(n * 8);
    uint32_t* perm = (uint32_t*)malloc(n * 4);
    orc_morton63(n, x, y, z, box, keys);
    orc_sort_keys(n, keys, sk, perm);
    double* s = (double*)malloc(8 * n * 8);
    double *sx = s, *sy = s + n, *sz = s + 2 * n, *sm = s + 3 * n;
    double *bx = s + 4 * n, *by = s + 5 * n, *bz = s + 6 * n, *e2 = s + 7 * n;
    for (size_t k = 0; k < n; ++k) {
        sx[k] = x[perm[k]]; sy[k] = y[perm[k]]; sz[k] = z[perm[k]]; sm[k] = m[perm[k]];
        if (eps_i) e2[k] = eps_i[perm[k]] * eps_i[perm[k]];
    }
    ltree t = {0};
    lt_build(&t, n, sk, leaf, sx, sy, sz, sm);
    lt_finish_com(&t, sx, sy, sz);
    double side0 = 2.0 * box[3];
    double eps2 = eps * eps, theta2 = theta * theta;
    uint32_t *gfirst, *gcnt;
    size_t ng = lt_groups(&t, group, split, &gfirst, &gcnt);
    uint64_t total = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel reduction(+ : total)
#endif
    {
        uint32_t* fa = (uint32_t*)malloc(t.len * 4 + 4);
        uint32_t* fb = (uint32_t*)malloc(t.len * 4 + 4);
        double* scratch = (double*)malloc(64 * 4 * 3 * sizeof(double));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (long long g = 0; g < (long long)ng; ++g) {
            size_t g0 = gfirst[g], g1 = g0 + gcnt[g];
            gwalk(&t, side0, theta2, eps2, eps_i ? e2 : NULL, sx, sy, sz, sm, g0, g1, bx, by, bz,
                  fa, fb, fcap, &total, scratch);
        }
        free(fa); free(fb); free(scratch);
    }
    for (size_t k = 0; k < n; ++k) {
        ax[perm[k]] += bx[k]; ay[perm[k]] += by[k]; az[perm[k]] += bz[k];
    }
    free(gfirst);
    free(gcnt);
    lt_free(&t);
    free(keys); free(sk); free(perm); free(s);
    return total;
}

/* split = 1: the target groups of the device's Barnes-Hut gravity path,
 * 0: those of its SPH step (see lt_groups). */
uint64_t orc_bh_group_accel(size_t n, double eps, double theta, uint32_t leaf, uint32_t group,
                            uint32_t fcap, int split, const double* x, const double* y,
                            const double* z, const double* m, double* ax, double* ay, double* az,
                            int nthreads) {
    return bh_group(n, eps, NULL, theta, leaf, group, fcap, split, x, y, z, m, ax, ay, az,
                    nthreads);
}

/* The same walk with Plummer softening eps_i[i] of each TARGET i
 * (a_i = sum m_j d / (r^2 + eps_i^2)^{3/2}, sources as above). */
uint64_t orc_bh_group_accel_eps(size_t n, const double* eps_i, double theta, uint32_t leaf,
                                uint32_t group, uint32_t fcap, const double* x, const double* y,
                                const double* z, const double* m, double* ax, double* ay,
                                double* az, int nthreads) {
    return bh_group(n, 0.0, eps_i, theta, leaf, group, fcap, 0, x, y, z, m, ax, ay, az,
                    nthreads);
}

/* ------------------------------------------------------------------------ */
/* KDK leapfrog (integrator.rs:41-77)                                         */
/* ------------------------------------------------------------------------ */
enum { ORC_BF = 0, ORC_BH_REF = 1, ORC_BH_LINEAR = 2 };

static void orc_force(int algo, size_t n, double eps, double theta, const double* x,
                      const double* y, const double* z, const double* m, double* ax, double* ay,
                      double* az, int nthreads) {
    for (size_t i = 0; i < n; ++i) { ax[i] = 0.0; ay[i] = 0.0; az[i] = 0.0; }
    if (algo == ORC_BF) orc_bf_accel(n, eps, x, y, z, m, ax, ay, az);
    else if (algo == ORC_BH_REF) orc_bh_accel(n, eps, theta, x, y, z, m, ax, ay, az, nthreads);
    else orc_bh_group_accel(n, eps, theta, 16, 64, 16384, 1, x, y, z, m, ax, ay, az, nthreads);
}

void orc_kdk_steps(size_t n, int algo, double eps, double theta, double* x, double* y,
                   double* z, double* vx, double* vy, double* vz, const double* m, double* ax,
                   double* ay, double* az, double dt, uint64_t nsteps, int nthreads) {
    double h = 0.5 * dt;
    for (uint64_t s = 0; s < nsteps; ++s) {
        for (size_t i = 0; i < n; ++i) { vx[i] += ax[i] * h; vy[i] += ay[i] * h; vz[i] += az[i] * h; }
        for (size_t i = 0; i < n; ++i) { x[i] += vx[i] * dt; y[i] += vy[i] * dt; z[i] += vz[i] * dt; }
        orc_force(algo, n, eps, theta, x, y, z, m, ax, ay, az, nthreads);
        for (size_t i = 0; i < n; ++i) { vx[i] += ax[i] * h; vy[i] += ay[i] * h; vz[i] += az[i] * h; }
    }
}

/* headless/src/main.rs:150-216 run loop for gravity-only scenarios:
 * a(t0) from a fresh force call, then nsteps KDK steps; diagnostics rows
 * (12 doubles each, layout of orc_diagnostics) are written at step 0 and
 * every diag_interval steps. Returns the number of rows written. */
size_t orc_run(size_t n, int algo, double eps, double theta, double* x, double* y, double* z,
               double* vx, double* vy, double* vz, const double* m, double dt, uint64_t nsteps,
               uint64_t diag_interval, double* rows, size_t max_rows, int nthreads);

/* ------------------------------------------------------------------------ */
/* Diagnostics (diagnostics.rs:50-148).                                      */
/* out: [0] K [1] U [2] E_th [3] E [4..6] p [7..9] L [10..12] com            */
/*      [13..15] v_com [16] virial [17] time [18] step                       */
/* ------------------------------------------------------------------------ */
void orc_diagnostics(size_t n, double eps, int with_potential, const double* x, const double* y,
                     const double* z, const double* vx, const double* vy, const double* vz,
                     const double* m, const double* u, const uint8_t* ptype, double time,
                     uint64_t step, double* out) {
    double eps2 = eps * eps;
    double K = 0.0;
    for (size_t i = 0; i < n; ++i) {
        double v2 = vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i];
        K += 0.5 * m[i] * v2;
    }
    double U = 0.0;
    if (with_potential) {
        for (size_t i = 0; i < n; ++i)
            for (size_t j = i + 1; j < n; ++j) {
                double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
                double r2 = dx * dx + dy * dy + dz * dz + eps2;
                U -= 1.0 * m[i] * m[j] / sqrt(r2);
            }
    }
    double Eth = 0.0;
    if (u && ptype)
        for (size_t i = 0; i < n; ++i)
            if (ptype[i] == 4) Eth += m[i] * u[i];
    double E = K + U + Eth;
    double p[3] = {0, 0, 0};
    for (size_t i = 0; i < n; ++i) { p[0] += m[i] * vx[i]; p[1] += m[i] * vy[i]; p[2] += m[i] * vz[i]; }
    double L[3] = {0, 0, 0};
    for (size_t i = 0; i < n; ++i) {
        L[0] += m[i] * (y[i] * vz[i] - z[i] * vy[i]);
        L[1] += m[i] * (z[i] * vx[i] - x[i] * vz[i]);
        L[2] += m[i] * (x[i] * vy[i] - y[i] * vx[i]);
    }
    double mt = 0.0;
    for (size_t i = 0; i < n; ++i) mt += m[i];
    double c[3] = {0, 0, 0}, cv[3] = {0, 0, 0};
    if (mt > 0.0) {
        for (size_t i = 0; i < n; ++i) {
            c[0] += m[i] * x[i]; c[1] += m[i] * y[i]; c[2] += m[i] * z[i];
            cv[0] += m[i] * vx[i]; cv[1] += m[i] * vy[i]; cv[2] += m[i] * vz[i];
        }
        for (int d = 0; d < 3; ++d) { c[d] /= mt; cv[d] /= mt; }
    }
    double vir = fabs(U) > 0.0 ? 2.0 * K / fabs(U) : 0.0;
    out[0] = K; out[1] = U; out[2] = Eth; out[3] = E;
    out[4] = p[0]; out[5] = p[1]; out[6] = p[2];
    out[7] = L[0]; out[8] = L[1]; out[9] = L[2];
    out[10] = c[0]; out[11] = c[1]; out[12] = c[2];
    out[13] = cv[0]; out[14] = cv[1]; out[15] = cv[2];
    out[16] = vir; out[17] = time; out[18] = (double)step;
}

size_t orc_run(size_t n, int algo, double eps, double theta, double* x, double* y, double* z,
               double* vx, double* vy, double* vz, const double* m, double dt, uint64_t nsteps,
               uint64_t diag_interval, double* rows, size_t max_rows, int nthreads) {
    double* a = (double*)calloc(3 * n, sizeof(double));
    double *ax = a, *ay = a + n, *az = a + 2 * n;
    orc_force(algo, n, eps, theta, x, y, z, m, ax, ay, az, nthreads);
    size_t nr = 0;
    if (nr < max_rows) orc_diagnostics(n, eps, 1, x, y, z, vx, vy, vz, m, NULL, NULL, 0.0, 0, rows + 19 * nr++);
    double t = 0.0;
    for (uint64_t s = 1; s <= nsteps; ++s) {
        orc_kdk_steps(n, algo, eps, theta, x, y, z, vx, vy, vz, m, ax, ay, az, dt, 1, nthreads);
        t += dt;
        if (diag_interval > 0 && s % diag_interval == 0 && nr < max_rows)
            orc_diagnostics(n, eps, 1, x, y, z, vx, vy, vz, m, NULL, NULL, t, s, rows + 19 * nr++);
    }
    free(a);
    return nr;
}
