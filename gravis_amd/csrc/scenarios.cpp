// scenarios.cpp -- initial-condition generators of sim-core, host side.
//
// Random streams reproduce ChaCha20Rng::seed_from_u64 (rand_chacha 0.9.0 /
// rand_core 0.9.5) and rand 0.9.2's f64 conversions, so a given seed gives
// the reference's particles:
//   - seed expansion: PCG32 (XSH-RR) words fill the 256-bit key;
//   - ChaCha20 block function (20 rounds), 64-bit block counter from 0,
//     stream id 0, output consumed as consecutive little-endian u32 words,
//     u64 = low word | high word << 32;
//   - random::<f64>()      = (u64 >> 11) * 2^-53;
//   - random_range(lo..hi) = ((u64 >> 12 | 1.0 bits) - 1) * (hi - lo) + lo,
//     redrawn with a one-ulp smaller scale if the result reaches hi.
// Scenario code: scenarios/plummer_sphere.rs:13-152, scenarios/two_body.rs:
// 49-105, particle.rs:189-226 (COM shift). Compiled without contraction so
// the arithmetic rounds as the Rust code does.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/gravis.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

class ChaChaStream {
  public:
    explicit ChaChaStream(uint64_t seed) {
        // rand_core seed_from_u64
        uint64_t st = seed;
        for (int i = 0; i < 8; ++i) {
            st = st * 6364136223846793005ull + 11634580027462260723ull;
            const uint32_t xs = static_cast<uint32_t>(((st >> 18) ^ st) >> 27);
            const uint32_t rot = static_cast<uint32_t>(st >> 59);
            key_[i] = (xs >> rot) | (xs << ((32u - rot) & 31u));
        }
    }

    uint64_t next_u64() {
        const uint64_t lo = next_u32();
        const uint64_t hi = next_u32();
        return lo | (hi << 32);
    }

    double uniform01() { return static_cast<double>(next_u64() >> 11) * 0x1.0p-53; }

    double range(double lo, double hi) {
        double scale = hi - lo;
        for (;;) {
            const uint64_t bits = (next_u64() >> 12) | 0x3FF0000000000000ull;
            double one_two;
            std::memcpy(&one_two, &bits, sizeof one_two);
            const double res = (one_two - 1.0) * scale + lo;
            if (res < hi) return res;
            uint64_t sb;
            std::memcpy(&sb, &scale, sizeof sb);
            --sb;
            std::memcpy(&scale, &sb, sizeof sb);
        }
    }

  private:
    static uint32_t rotl(uint32_t v, int s) { return (v << s) | (v >> (32 - s)); }

    static void quarter(uint32_t* w, int a, int b, int c, int d) {
        w[a] += w[b]; w[d] = rotl(w[d] ^ w[a], 16);
        w[c] += w[d]; w[b] = rotl(w[b] ^ w[c], 12);
        w[a] += w[b]; w[d] = rotl(w[d] ^ w[a], 8);
        w[c] += w[d]; w[b] = rotl(w[b] ^ w[c], 7);
    }

    void block(uint64_t counter, uint32_t* out) const {
        uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
        for (int i = 0; i < 8; ++i) in[4 + i] = key_[i];
        in[12] = static_cast<uint32_t>(counter);
        in[13] = static_cast<uint32_t>(counter >> 32);
        in[14] = 0u;
        in[15] = 0u;
        uint32_t w[16];
        std::memcpy(w, in, sizeof w);
        for (int r = 0; r < 10; ++r) {
            quarter(w, 0, 4, 8, 12); quarter(w, 1, 5, 9, 13);
            quarter(w, 2, 6, 10, 14); quarter(w, 3, 7, 11, 15);
            quarter(w, 0, 5, 10, 15); quarter(w, 1, 6, 11, 12);
            quarter(w, 2, 7, 8, 13); quarter(w, 3, 4, 9, 14);
        }
        for (int i = 0; i < 16; ++i) out[i] = w[i] + in[i];
    }

    uint32_t next_u32() {
        if (pos_ == 16) {
            block(counter_++, words_);
            pos_ = 0;
        }
        return words_[pos_++];
    }

    uint32_t key_[8];
    uint32_t words_[16] = {};
    uint64_t counter_ = 0;
    int pos_ = 16;
};

void com_shift(size_t n, double* x, double* y, double* z, double* vx, double* vy, double* vz,
               const double* m) {
    double mt = 0.0;
    for (size_t i = 0; i < n; ++i) mt += m[i];
    if (mt == 0.0) return;
    double c[6] = {0, 0, 0, 0, 0, 0};
    for (size_t i = 0; i < n; ++i) {
        c[0] += m[i] * x[i]; c[1] += m[i] * y[i]; c[2] += m[i] * z[i];
        c[3] += m[i] * vx[i]; c[4] += m[i] * vy[i]; c[5] += m[i] * vz[i];
    }
    for (double& v : c) v /= mt;
    for (size_t i = 0; i < n; ++i) {
        x[i] -= c[0]; y[i] -= c[1]; z[i] -= c[2];
        vx[i] -= c[3]; vy[i] -= c[4]; vz[i] -= c[5];
    }
}

// sample_plummer_particle (plummer_sphere.rs:13-58): one position and
// velocity of a Plummer sphere of mass M and scale radius a.
void plummer_particle(ChaChaStream& rng, double total_mass, double a, double pos[3],
                      double vel[3]) {
    // position: inverse CDF of M(<r) = r^3 / (r^2 + a^2)^(3/2)
    const double xm = rng.range(0.001, 1.0);
    const double r = a / std::sqrt(std::pow(xm, -2.0 / 3.0) - 1.0);
    const double ct = rng.range(-1.0, 1.0);
    const double st = std::sqrt(1.0 - ct * ct);
    const double ph = rng.range(0.0, 2.0 * kPi);
    pos[0] = r * st * std::cos(ph);
    pos[1] = r * st * std::sin(ph);
    pos[2] = r * ct;
    // speed: rejection from q^2 (1 - q^2)^(7/2) with bound 0.1
    const double vesc = std::sqrt(2.0 * 1.0 * total_mass / std::sqrt(r * r + a * a));
    double q;
    for (;;) {
        q = rng.range(0.0, 1.0);
        const double gq = q * q * std::pow(1.0 - q * q, 3.5);
        if (rng.uniform01() < gq / 0.1) break;
    }
    const double v = q * vesc;
    const double cv = rng.range(-1.0, 1.0);
    const double sv = std::sqrt(1.0 - cv * cv);
    const double pv = rng.range(0.0, 2.0 * kPi);
    vel[0] = v * sv * std::cos(pv);
    vel[1] = v * sv * std::sin(pv);
    vel[2] = v * cv;
}

}  // namespace

extern "C" {

gravis_status gravis_scenario_plummer(size_t n, double scale_radius, double total_mass,
                                      uint64_t seed, double* x, double* y, double* z,
                                      double* vx, double* vy, double* vz, double* m) {
    if (n == 0 || !(scale_radius > 0.0) || !(total_mass > 0.0)) return GRAVIS_INVALID_ARG;
    if (!x || !y || !z || !vx || !vy || !vz || !m) return GRAVIS_INVALID_ARG;
    ChaChaStream rng(seed);
    const double mp = total_mass / static_cast<double>(n);
    for (size_t i = 0; i < n; ++i) {
        double p[3], v[3];
        plummer_particle(rng, total_mass, scale_radius, p, v);
        x[i] = p[0]; y[i] = p[1]; z[i] = p[2];
        vx[i] = v[0]; vy[i] = v[1]; vz[i] = v[2];
        m[i] = mp;
    }
    com_shift(n, x, y, z, vx, vy, vz, m);
    return GRAVIS_OK;
}

double gravis_plummer_suggested_dt(double scale_radius, double total_mass) {
    const double a3 = scale_radius * scale_radius * scale_radius;
    return std::sqrt(a3 / (1.0 * total_mass)) / 100.0;
}

double gravis_plummer_suggested_softening(size_t n, double scale_radius) {
    const double rh = 1.305 * scale_radius;
    const double spacing = std::pow(4.0 / 3.0 * kPi * (rh * rh * rh) / static_cast<double>(n),
                                    1.0 / 3.0);
    return spacing * 0.5;
}

gravis_status gravis_scenario_two_body(double e, double mass_ratio, double a, double* x,
                                       double* y, double* z, double* vx, double* vy,
                                       double* vz, double* m) {
    if (!(e >= 0.0 && e < 1.0) || !(a > 0.0) || !(mass_ratio > 0.0 && mass_ratio < 1.0))
        return GRAVIS_INVALID_ARG;
    if (!x || !y || !z || !vx || !vy || !vz || !m) return GRAVIS_INVALID_ARG;
    const double mt = 1.0;
    const double m1 = mt * mass_ratio, m2 = mt * (1.0 - mass_ratio);
    const double rp = a * (1.0 - e);
    const double vp = std::sqrt(1.0 * mt * (2.0 / rp - 1.0 / a));
    x[0] = -m2 / mt * rp;
    x[1] = m1 / mt * rp;
    y[0] = y[1] = z[0] = z[1] = 0.0;
    vx[0] = vx[1] = vz[0] = vz[1] = 0.0;
    vy[0] = -m2 / mt * vp;
    vy[1] = m1 / mt * vp;
    m[0] = m1;
    m[1] = m2;
    return GRAVIS_OK;
}

double gravis_two_body_period(double a) {
    return 2.0 * kPi * std::sqrt(a * a * a / (1.0 * 1.0));
}

}  // extern "C"

extern "C" gravis_status gravis_scenario_plummer_pair(size_t n_each, double scale_radius,
                                                      double total_mass, uint64_t seed_a,
                                                      uint64_t seed_b, double separation,
                                                      double impact, double* x, double* y,
                                                      double* z, double* vx, double* vy,
                                                      double* vz, double* m) {
    if (!(separation > 0.0) || !(impact >= 0.0)) return GRAVIS_INVALID_ARG;
    gravis_status s = gravis_scenario_plummer(n_each, scale_radius, total_mass, seed_a, x, y, z,
                                              vx, vy, vz, m);
    if (s != GRAVIS_OK) return s;
    const size_t o = n_each;
    s = gravis_scenario_plummer(n_each, scale_radius, total_mass, seed_b, x + o, y + o, z + o,
                                vx + o, vy + o, vz + o, m + o);
    if (s != GRAVIS_OK) return s;
    // parabolic approach: E_orb = 0 at the initial separation
    const double r0 = std::sqrt(separation * separation + impact * impact);
    const double vrel = std::sqrt(2.0 * 1.0 * (2.0 * total_mass) / r0);
    const double hs = separation / 2.0, hb = impact / 2.0, hv = vrel / 2.0;
    for (size_t i = 0; i < n_each; ++i) {
        x[i] -= hs; y[i] -= hb; vx[i] += hv;
        x[o + i] += hs; y[o + i] += hb; vx[o + i] -= hv;
    }
    com_shift(2 * n_each, x, y, z, vx, vy, vz, m);
    return GRAVIS_OK;
}

// Ensemble layout (no reference counterpart): counts[s] = lo + u64 % (hi -
// lo + 1) from one ChaCha20 stream seeded with `seed`, in system order.
extern "C" gravis_status gravis_scenario_ensemble_counts(uint64_t seed, int nsys, uint32_t lo,
                                                         uint32_t hi, uint32_t* counts) {
    if (nsys <= 0 || hi < lo || !counts) return GRAVIS_INVALID_ARG;
    ChaChaStream rng(seed);
    const uint64_t span = (uint64_t)(hi - lo) + 1;
    for (int s = 0; s < nsys; ++s) counts[s] = lo + (uint32_t)(rng.next_u64() % span);
    return GRAVIS_OK;
}

// Loss targets U[-1, 1)^3, drawn x, y, z per particle from a stream seeded
// with `seed`.
extern "C" gravis_status gravis_scenario_targets(uint64_t seed, size_t n, double* tx, double* ty,
                                                 double* tz) {
    if (!tx || !ty || !tz) return GRAVIS_INVALID_ARG;
    ChaChaStream rng(seed);
    for (size_t i = 0; i < n; ++i) {
        tx[i] = rng.range(-1.0, 1.0);
        ty[i] = rng.range(-1.0, 1.0);
        tz[i] = rng.range(-1.0, 1.0);
    }
    return GRAVIS_OK;
}

// EvrardCollapse::generate (crates/sim-core/src/scenarios/evrard_collapse.rs:
// 45-57): sphere_spacing (scenarios/mod.rs:17-22) then
// fill_uniform_gas_sphere (mod.rs:33-51), one rejection draw of x, y, z per
// trial; Particles::add_gas (particle.rs:136-154) sets alpha 1 and zeroes
// the rest.
extern "C" gravis_status gravis_scenario_evrard(size_t n, double total_mass, double radius,
                                                double u0, uint64_t seed,
                                                gravis_sph_state* s) {
    if (!s || s->n != n) return GRAVIS_INVALID_ARG;
    if (n == 0) return GRAVIS_OK;
    double* f[17] = {s->x, s->y, s->z, s->vx, s->vy, s->vz, s->ax, s->ay, s->az,
                     s->m, s->rho, s->p, s->u, s->dudt, s->h, s->c, s->alpha};
    for (double* a : f)
        if (!a) return GRAVIS_INVALID_ARG;
    if (!s->type || !(radius > 0.0)) return GRAVIS_INVALID_ARG;
    ChaChaStream rng(seed);
    const double mass = total_mass / static_cast<double>(n);
    const double vol = 4.0 / 3.0 * kPi * (radius * radius * radius);
    const double h = 1.5 * std::cbrt(vol / static_cast<double>(n));
    size_t cnt = 0;
    while (cnt < n) {
        const double x = rng.uniform01() * 2.0 * radius - radius;
        const double y = rng.uniform01() * 2.0 * radius - radius;
        const double z = rng.uniform01() * 2.0 * radius - radius;
        if (x * x + y * y + z * z <= radius * radius) {
            for (double* a : f) a[cnt] = 0.0;
            s->x[cnt] = x;
            s->y[cnt] = y;
            s->z[cnt] = z;
            s->m[cnt] = mass;
            s->u[cnt] = u0;
            s->h[cnt] = h;
            s->alpha[cnt] = 1.0;
            s->type[cnt] = 4;
            ++cnt;
        }
    }
    return GRAVIS_OK;
}

// EvrardCollapse::suggested_softening (evrard_collapse.rs:67-70).
extern "C" double gravis_evrard_softening(size_t n, double radius) {
    const double vol = 4.0 / 3.0 * kPi * (radius * radius * radius);
    return 0.3 * std::cbrt(vol / static_cast<double>(n));
}

// Variant initial conditions (no reference counterpart): rho(r) =
// M / (2 pi R^2 r), so M(<r) = M (r/R)^2 and r = R sqrt(U), U =
// random::<f64>(); cos(theta) in U[-1, 1) and phi in U[0, 2 pi) by
// random_range, drawn per particle in that order; at rest, u = u0,
// alpha 1, m = M/n, h = eta (m / rho(r))^(1/3) (r floored at 1e-6 R).
extern "C" gravis_status gravis_scenario_evrard_1r(size_t n, double total_mass, double radius,
                                                   double u0, double eta, uint64_t seed,
                                                   gravis_sph_state* s) {
    if (!s || s->n != n) return GRAVIS_INVALID_ARG;
    if (n == 0) return GRAVIS_OK;
    double* f[17] = {s->x, s->y, s->z, s->vx, s->vy, s->vz, s->ax, s->ay, s->az,
                     s->m, s->rho, s->p, s->u, s->dudt, s->h, s->c, s->alpha};
    for (double* a : f)
        if (!a) return GRAVIS_INVALID_ARG;
    if (!s->type || !(radius > 0.0) || !(eta > 0.0)) return GRAVIS_INVALID_ARG;
    ChaChaStream rng(seed);
    const double mass = total_mass / static_cast<double>(n);
    for (size_t i = 0; i < n; ++i) {
        const double r = radius * std::sqrt(rng.uniform01());
        const double ct = rng.range(-1.0, 1.0);
        const double ph = rng.range(0.0, 2.0 * kPi);
        const double st = std::sqrt(std::fmax(0.0, 1.0 - ct * ct));
        for (double* a : f) a[i] = 0.0;
        s->x[i] = r * st * std::cos(ph);
        s->y[i] = r * st * std::sin(ph);
        s->z[i] = r * ct;
        s->m[i] = mass;
        s->u[i] = u0;
        const double rr = std::fmax(r, 1e-6 * radius);
        const double rho = total_mass / (2.0 * kPi * radius * radius * rr);
        s->h[i] = eta * std::cbrt(mass / rho);
        s->alpha[i] = 1.0;
        s->type[i] = 4;
    }
    return GRAVIS_OK;
}

// ColdCollapse::generate (crates/sim-core/src/scenarios/cold_collapse.rs:
// 56-102) with fill_uniform_gas_sphere (scenarios/mod.rs:33-51) for the gas
// form, then shift_to_com_frame (particle.rs:189-226).
extern "C" gravis_status gravis_scenario_cold_collapse(size_t n, double radius,
                                                       double total_mass, uint64_t seed,
                                                       int gas, gravis_sph_state* s) {
    if (!s || s->n != n || n == 0) return GRAVIS_INVALID_ARG;
    double* f[17] = {s->x, s->y, s->z, s->vx, s->vy, s->vz, s->ax, s->ay, s->az,
                     s->m, s->rho, s->p, s->u, s->dudt, s->h, s->c, s->alpha};
    for (double* a : f)
        if (!a) return GRAVIS_INVALID_ARG;
    if (!s->type || !(radius > 0.0) || !(total_mass > 0.0)) return GRAVIS_INVALID_ARG;
    ChaChaStream rng(seed);
    const double m = total_mass / static_cast<double>(n);
    for (double* a : f)
        for (size_t i = 0; i < n; ++i) a[i] = 0.0;
    if (gas) {
        const double vol = 4.0 / 3.0 * kPi * (radius * radius * radius);
        const double h = 1.5 * std::cbrt(vol / static_cast<double>(n));
        size_t cnt = 0;
        while (cnt < n) {
            const double x = rng.uniform01() * 2.0 * radius - radius;
            const double y = rng.uniform01() * 2.0 * radius - radius;
            const double z = rng.uniform01() * 2.0 * radius - radius;
            if (x * x + y * y + z * z <= radius * radius) {
                s->x[cnt] = x;
                s->y[cnt] = y;
                s->z[cnt] = z;
                s->m[cnt] = m;
                s->u[cnt] = 0.05;
                s->h[cnt] = h;
                s->alpha[cnt] = 1.0;
                s->type[cnt] = 4;
                ++cnt;
            }
        }
    } else {
        for (size_t i = 0; i < n; ++i) {
            double x, y, z;
            for (;;) {
                x = rng.range(-1.0, 1.0);
                y = rng.range(-1.0, 1.0);
                z = rng.range(-1.0, 1.0);
                if (x * x + y * y + z * z <= 1.0) break;
            }
            s->x[i] = x * radius;
            s->y[i] = y * radius;
            s->z[i] = z * radius;
            s->m[i] = m;
            s->type[i] = 0;
        }
    }
    com_shift(n, s->x, s->y, s->z, s->vx, s->vy, s->vz, s->m);
    return GRAVIS_OK;
}

// GalaxyCollision (crates/sim-core/src/scenarios/galaxy_collision.rs:20-450).
// Each galaxy: exponential disk (rejection against U[0, R_max]), Plummer
// bulge, truncated Hernquist halo with isotropic Gaussian velocities; both
// galaxies are drawn from one stream, then offset along x/y with opposite
// x velocities, then moved to the COM frame.
namespace {

constexpr uint8_t kDiskStar = 1, kBulgeStar = 2, kDarkMatter = 3;

// disk_enclosed_mass + circular_velocity (galaxy_collision.rs:193-257).
double disk_enclosed(double r, double m_disk, double rd) {
    if (r <= 0.0) return 0.0;
    const double x = r / rd;
    return m_disk * (1.0 - (1.0 + x) * std::exp(-x));
}

double circular_velocity(double r, double m_disk, double rd, double m_bulge, double ab,
                         double m_halo, double ah) {
    if (r <= 0.0) return 0.0;
    const double v2_disk = 1.0 * disk_enclosed(r, m_disk, rd) / r;
    const double v2_bulge = 1.0 * m_bulge * r * r / std::pow(r * r + ab * ab, 1.5);
    const double v2_halo = 1.0 * m_halo * r / ((r + ah) * (r + ah));
    return std::sqrt(v2_disk + v2_bulge + v2_halo);
}

struct Sink {
    gravis_sph_state* s;
    size_t k;
    void add(double x, double y, double z, double vx, double vy, double vz, double m,
             uint8_t type) {
        s->x[k] = x; s->y[k] = y; s->z[k] = z;
        s->vx[k] = vx; s->vy[k] = vy; s->vz[k] = vz;
        s->m[k] = m;
        s->type[k] = type;
        ++k;
    }
};

// generate_galaxy = generate_disk, generate_bulge, generate_halo
// (galaxy_collision.rs:204-450), in that order from one stream.
void galaxy(const gravis_galaxy_params& g, ChaChaStream& rng, size_t n_disk, size_t n_bulge,
            size_t n_halo, double m_disk, double m_bulge, double m_halo, Sink& out) {
    const double rd = g.disk_scale_radius, ab = g.bulge_scale_radius, ah = g.halo_scale_radius;
    // disk
    {
        const double mp = m_disk / static_cast<double>(n_disk);
        const double max_x = g.disk_max_radius / rd;
        for (size_t i = 0; i < n_disk; ++i) {
            double r;
            for (;;) {
                const double u = rng.uniform01();
                const double r_try = u * g.disk_max_radius;
                const double x = r_try / rd;
                const double f = x * std::exp(-x);
                const double f_max = std::exp(-1.0);
                const double accept = rng.uniform01();
                if (accept < f / f_max && x < max_x) {
                    r = r_try;
                    break;
                }
            }
            const double phi = rng.range(0.0, 2.0 * kPi);
            const double c = std::cos(phi), s = std::sin(phi);
            const double px = r * c, py = r * s;
            const double uz = rng.range(0.01, 0.99);
            const double pz = g.disk_scale_height * std::atanh(2.0 * uz - 1.0);
            const double vc = circular_velocity(r, m_disk, rd, m_bulge, ab, m_halo, ah);
            const double sigma_r = 0.15 * vc;
            const double v_r = rng.range(-1.0, 1.0) * sigma_r;
            const double v_phi = vc + rng.range(-1.0, 1.0) * sigma_r * 0.5;
            const double sigma_z = 0.1 * vc;
            const double v_z = rng.range(-1.0, 1.0) * sigma_z;
            out.add(px, py, pz, v_r * c - v_phi * s, v_r * s + v_phi * c, v_z, mp, kDiskStar);
        }
    }
    // bulge
    {
        const double mp = m_bulge / static_cast<double>(n_bulge);
        for (size_t i = 0; i < n_bulge; ++i) {
            double p[3], v[3];
            plummer_particle(rng, m_bulge, ab, p, v);
            out.add(p[0], p[1], p[2], v[0], v[1], v[2], mp, kBulgeStar);
        }
    }
    // halo: r = a sqrt(U) / (1 - sqrt(U)), U in [0.001, (R/(R+a))^2)
    {
        const double mp = m_halo / static_cast<double>(n_halo);
        const double xm = g.halo_max_radius / (g.halo_max_radius + ah);
        const double u_max = xm * xm;
        const double sqrt2 = std::sqrt(2.0);
        for (size_t i = 0; i < n_halo; ++i) {
            const double u = rng.range(0.001, u_max);
            const double su = std::sqrt(u);
            const double r = ah * su / (1.0 - su);
            const double ct = rng.range(-1.0, 1.0);
            const double st = std::sqrt(1.0 - ct * ct);
            const double phi = rng.range(0.0, 2.0 * kPi);
            const double px = r * st * std::cos(phi);
            const double py = r * st * std::sin(phi);
            const double pz = r * ct;
            const double vc = circular_velocity(r, m_disk, rd, m_bulge, ab, m_halo, ah);
            const double sigma = vc / sqrt2;
            // box_muller (scenarios/mod.rs:25-30), twice
            double gs[4];
            for (int b = 0; b < 2; ++b) {
                const double u1 = rng.range(0.001, 1.0);
                const double u2 = rng.range(0.0, 2.0 * kPi);
                const double rr = std::sqrt(-2.0 * std::log(u1));
                gs[2 * b] = rr * std::cos(u2);
                gs[2 * b + 1] = rr * std::sin(u2);
            }
            out.add(px, py, pz, sigma * gs[0], sigma * gs[1], sigma * gs[2], mp, kDarkMatter);
        }
    }
}

// Particle counts proportional to mass (galaxy_collision.rs:80-88).
void galaxy_counts(const gravis_galaxy_params& g, double* md, double* mb, double* mh,
                   size_t* nd, size_t* nb, size_t* nh) {
    *md = g.baryon_mass * g.disk_fraction;
    *mb = g.baryon_mass * (1.0 - g.disk_fraction);
    *mh = g.baryon_mass * g.halo_mass_ratio;
    const double mt = *md + *mb + *mh;
    const double n = static_cast<double>(g.n_per_galaxy);
    *nd = static_cast<size_t>(std::round(n * *md / mt));
    *nb = static_cast<size_t>(std::round(n * *mb / mt));
    *nh = g.n_per_galaxy - *nd - *nb;
}

}  // namespace

extern "C" void gravis_galaxy_default_params(gravis_galaxy_params* g) {
    if (!g) return;
    g->n_per_galaxy = 50000;
    g->seed = 42;
    g->disk_fraction = 0.7;
    g->halo_mass_ratio = 10.0;
    g->baryon_mass = 1.0;
    g->disk_scale_radius = 1.0;
    g->disk_scale_height = 0.1;
    g->bulge_scale_radius = 0.3;
    g->halo_scale_radius = 5.0;
    g->disk_max_radius = 6.0;
    g->halo_max_radius = 30.0;
    g->separation = 15.0;
    g->impact_parameter = 2.0;
    g->approach_velocity = 0.5;
}

extern "C" gravis_status gravis_scenario_galaxy_collision(const gravis_galaxy_params* g,
                                                          gravis_sph_state* s) {
    if (!g || !s || g->n_per_galaxy == 0 || s->n != 2 * g->n_per_galaxy)
        return GRAVIS_INVALID_ARG;
    if (!(g->disk_scale_radius > 0.0) || !(g->halo_scale_radius > 0.0) ||
        !(g->disk_max_radius > 0.0) || !(g->halo_max_radius > 0.0) ||
        !(g->disk_fraction >= 0.0 && g->disk_fraction <= 1.0) || !(g->baryon_mass > 0.0) ||
        !(g->halo_mass_ratio >= 0.0))
        return GRAVIS_INVALID_ARG;
    double* f[17] = {s->x, s->y, s->z, s->vx, s->vy, s->vz, s->ax, s->ay, s->az,
                     s->m, s->rho, s->p, s->u, s->dudt, s->h, s->c, s->alpha};
    for (double* a : f)
        if (!a) return GRAVIS_INVALID_ARG;
    if (!s->type) return GRAVIS_INVALID_ARG;
    double md, mb, mh;
    size_t nd, nb, nh;
    galaxy_counts(*g, &md, &mb, &mh, &nd, &nb, &nh);
    if (nd + nb > g->n_per_galaxy) return GRAVIS_INVALID_ARG;
    const size_t n = s->n;
    for (double* a : f)
        for (size_t i = 0; i < n; ++i) a[i] = 0.0;
    ChaChaStream rng(g->seed);
    Sink out{s, 0};
    galaxy(*g, rng, nd, nb, nh, md, mb, mh, out);
    const size_t half = out.k;
    galaxy(*g, rng, nd, nb, nh, md, mb, mh, out);
    const double hs = g->separation / 2.0, hi = g->impact_parameter / 2.0,
                 hv = g->approach_velocity / 2.0;
    for (size_t i = 0; i < half; ++i) {
        s->x[i] -= hs;
        s->y[i] -= hi;
        s->vx[i] += hv;
    }
    for (size_t i = half; i < n; ++i) {
        s->x[i] += hs;
        s->y[i] += hi;
        s->vx[i] -= hv;
    }
    com_shift(n, s->x, s->y, s->z, s->vx, s->vy, s->vz, s->m);
    return GRAVIS_OK;
}

// suggested_dt / suggested_softening (galaxy_collision.rs:159-185): t_orb at
// the disk scale radius over 500; softening 0.05 R_d.
extern "C" double gravis_galaxy_suggested_dt(const gravis_galaxy_params* g) {
    const double m_enc_disk = 0.26 * g->baryon_mass * g->disk_fraction;
    const double r = g->disk_scale_radius;
    const double ab = g->bulge_scale_radius;
    const double m_bulge = g->baryon_mass * (1.0 - g->disk_fraction);
    const double m_enc_bulge = m_bulge * (r * r * r) / std::pow(r * r + ab * ab, 1.5);
    const double ah = g->halo_scale_radius;
    const double m_halo = g->baryon_mass * g->halo_mass_ratio;
    const double x = r / ah;
    const double m_enc_halo = m_halo * x * x / ((1.0 + x) * (1.0 + x));
    const double m_enc = m_enc_disk + m_enc_bulge + m_enc_halo;
    const double vc = std::sqrt(1.0 * m_enc / r);
    const double t_orb = 2.0 * kPi * r / vc;
    return t_orb / 500.0;
}

extern "C" double gravis_galaxy_suggested_softening(const gravis_galaxy_params* g) {
    return g->disk_scale_radius * 0.05;
}
