// headless.cpp -- gravis_headless: sim-core's headless runner
// (crates/headless/src/main.rs) on the gfx950 engine.
//
// Same flags (main.rs:16-68), scenario table (scenario_builder.rs:37-90),
// step loop (main.rs:150-215), stderr lines (main.rs:125-237) and
// diagnostics CSV (diagnostics.rs:170-189), with Rust's number formatting
// reproduced. Additions: --backend (only "hip": the CPU backend is sim-core's
// own headless), --precision {fp64,mixed}, --leaf-size, and --restart FILE
// (continue from a snapshot: Snapshot::read_from then initialize).
//
// The loop runs on the device between output events: the steps up to the
// next diagnostics or snapshot step are one gravis_kdk / gravis_sph_run call
// (state stays in HBM in between), so stdout/stderr/CSV/snapshots appear at
// the same steps as the reference's.
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gravis.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

// ---- Rust std::fmt for f64 ------------------------------------------------
// {:.Pe}: mantissa with P decimals, 'e', exponent without padding or '+'.
std::string fexp(double v, int prec, bool plus = false) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? (plus ? "+inf" : "inf") : "-inf";
    char buf[64];
    std::snprintf(buf, sizeof buf, plus ? "%+.*e" : "%.*e", prec, v);
    std::string s(buf);
    const size_t e = s.find('e');
    std::string mant = s.substr(0, e);
    int ex = std::atoi(s.c_str() + e + 1);
    return mant + "e" + std::to_string(ex);
}

// {:.P}
std::string ffix(double v, int prec) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    char buf[512];
    std::snprintf(buf, sizeof buf, "%.*f", prec, v);
    return buf;
}

// {} (Display): the shortest decimal that reads back to v, without exponent.
std::string fdisp(double v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    char buf[64];
    int p = 1;
    for (; p <= 17; ++p) {
        std::snprintf(buf, sizeof buf, "%.*e", p - 1, v);
        if (std::strtod(buf, nullptr) == v) break;
    }
    // digits and exponent of the shortest form
    std::string s(buf);
    const bool neg = s[0] == '-';
    if (neg) s = s.substr(1);
    const size_t e = s.find('e');
    const int ex = std::atoi(s.c_str() + e + 1);
    std::string dig;
    for (size_t i = 0; i < e; ++i)
        if (s[i] != '.') dig += s[i];
    std::string out;
    if (ex >= 0) {
        if ((int)dig.size() <= ex + 1) {
            out = dig + std::string(ex + 1 - dig.size(), '0');
        } else {
            out = dig.substr(0, ex + 1) + "." + dig.substr(ex + 1);
        }
    } else {
        out = "0." + std::string(-ex - 1, '0') + dig;
    }
    return (neg ? "-" : "") + out;
}

// ---- command line ---------------------------------------------------------
struct Cli {
    std::string scenario = "plummer";
    long long particles = -1;
    uint64_t steps = 1000;
    double dt = NAN;
    double softening = NAN;
    uint64_t snapshot_interval = 0;
    std::string snapshot_dir = "artifacts/snapshots";
    uint64_t diag_interval = 100;
    std::string diag_csv;
    uint64_t seed = 42;
    double eccentricity = 0.5;
    std::string algorithm = "brute-force";
    double theta = 0.5;
    std::string backend = "hip";
    std::string precision = "fp64";
    uint32_t leaf_size = 16;
    std::string restart;
};

[[noreturn]] void usage_error(const std::string& msg) {
    std::fprintf(stderr, "error: %s\n\nUsage: gravis_headless [OPTIONS]\n", msg.c_str());
    std::exit(2);
}

double to_f64(const std::string& flag, const std::string& v) {
    char* end = nullptr;
    const double d = std::strtod(v.c_str(), &end);
    if (v.empty() || *end) usage_error("invalid value '" + v + "' for '" + flag + "'");
    return d;
}

uint64_t to_u64(const std::string& flag, const std::string& v) {
    char* end = nullptr;
    if (v.empty() || v[0] == '-') usage_error("invalid value '" + v + "' for '" + flag + "'");
    const unsigned long long d = std::strtoull(v.c_str(), &end, 10);
    if (*end) usage_error("invalid value '" + v + "' for '" + flag + "'");
    return d;
}

Cli parse(int argc, char** argv) {
    Cli c;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i], v;
        const size_t eq = a.find('=');
        bool has_v = false;
        if (a.rfind("--", 0) == 0 && eq != std::string::npos) {
            v = a.substr(eq + 1);
            a = a.substr(0, eq);
            has_v = true;
        }
        auto val = [&]() -> std::string {
            if (has_v) return v;
            if (i + 1 >= argc) usage_error("a value is required for '" + a + "'");
            return argv[++i];
        };
        if (a == "-s" || a == "--scenario") c.scenario = val();
        else if (a == "-n" || a == "--particles") c.particles = (long long)to_u64(a, val());
        else if (a == "--steps") c.steps = to_u64(a, val());
        else if (a == "--dt") c.dt = to_f64(a, val());
        else if (a == "--softening") c.softening = to_f64(a, val());
        else if (a == "--snapshot-interval") c.snapshot_interval = to_u64(a, val());
        else if (a == "--snapshot-dir") c.snapshot_dir = val();
        else if (a == "--diag-interval") c.diag_interval = to_u64(a, val());
        else if (a == "--diag-csv") c.diag_csv = val();
        else if (a == "--seed") c.seed = to_u64(a, val());
        else if (a == "--eccentricity") c.eccentricity = to_f64(a, val());
        else if (a == "--algorithm") {
            c.algorithm = val();
            if (c.algorithm != "brute-force" && c.algorithm != "barnes-hut")
                usage_error("invalid value '" + c.algorithm +
                            "' for '--algorithm <ALGORITHM>' [possible values: brute-force, "
                            "barnes-hut]");
        } else if (a == "--theta") c.theta = to_f64(a, val());
        else if (a == "--backend") c.backend = val();
        else if (a == "--precision") {
            c.precision = val();
            if (c.precision != "fp64" && c.precision != "mixed")
                usage_error("--precision must be fp64 or mixed");
        } else if (a == "--leaf-size") c.leaf_size = (uint32_t)to_u64(a, val());
        else if (a == "--restart") c.restart = val();
        else if (a == "-h" || a == "--help") {
            std::printf(
                "Headless N-body simulation runner for benchmarks and artifact generation\n"
                "(sim-core's headless on the gfx950 engine)\n\n"
                "Options: -s/--scenario, -n/--particles, --steps, --dt, --softening,\n"
                "  --snapshot-interval, --snapshot-dir, --diag-interval, --diag-csv, --seed,\n"
                "  --eccentricity, --algorithm {brute-force,barnes-hut}, --theta,\n"
                "  --backend hip, --precision {fp64,mixed}, --leaf-size, --restart FILE\n");
            std::exit(0);
        } else usage_error("unexpected argument '" + a + "' found");
    }
    return c;
}

// ---- particle storage -------------------------------------------------------
struct State {
    size_t n = 0;
    std::vector<double> f[17];
    std::vector<uint8_t> type;
    gravis_sph_state view{};
    void resize(size_t count) {
        n = count;
        for (auto& a : f) a.assign(n, 0.0);
        type.assign(n, 0);
        refresh();
    }
    void refresh() {
        view.n = n;
        double** p[17] = {&view.x, &view.y, &view.z, &view.vx, &view.vy, &view.vz,
                          &view.ax, &view.ay, &view.az, &view.m, &view.rho, &view.p,
                          &view.u, &view.dudt, &view.h, &view.c, &view.alpha};
        for (int k = 0; k < 17; ++k) *p[k] = f[k].data();
        view.type = type.data();
    }
    bool has_gas() const {
        for (uint8_t t : type)
            if (t == 4) return true;
        return false;
    }
};

void check(gravis_ctx* ctx, gravis_status st, const char* what) {
    if (st == GRAVIS_OK) return;
    const char* msg = ctx ? gravis_last_error(ctx) : "";
    std::fprintf(stderr, "Error: %s failed: %s (%s)\n", what, gravis_status_string(st),
                 msg ? msg : "");
    std::exit(1);
}

struct Built {
    double dt = 0, softening = 0;
    std::string name;
};

// scenario_builder::build (scenario_builder.rs:37-90) for the scenarios this
// engine generates.
bool build(const Cli& c, State* s, Built* b) {
    const std::string& sc = c.scenario;
    if (sc == "plummer") {
        const size_t n = c.particles > 0 ? (size_t)c.particles : 1000;
        s->resize(n);
        check(nullptr,
              gravis_scenario_plummer(n, 1.0, 1.0, c.seed, s->view.x, s->view.y, s->view.z,
                                      s->view.vx, s->view.vy, s->view.vz, s->view.m),
              "plummer");
        b->name = "Plummer Sphere";
        b->dt = gravis_plummer_suggested_dt(1.0, 1.0);
        b->softening = gravis_plummer_suggested_softening(n, 1.0);
    } else if (sc == "two-body" || sc == "kepler") {
        s->resize(2);
        check(nullptr,
              gravis_scenario_two_body(c.eccentricity, 0.5, 1.0, s->view.x, s->view.y,
                                       s->view.z, s->view.vx, s->view.vy, s->view.vz, s->view.m),
              "two-body");
        b->name = "Two-Body Kepler Orbit";
        b->dt = gravis_two_body_period(1.0) / 10000.0;  // two_body.rs:95-98
        b->softening = 1.0 * (1.0 - c.eccentricity) * 0.001;  // two_body.rs:100-105
    } else if (sc == "cold-collapse" || sc == "cold-collapse-gas") {
        const bool gas = sc == "cold-collapse-gas";
        const size_t n = c.particles > 0 ? (size_t)c.particles : 5000;
        s->resize(n);
        check(nullptr, gravis_scenario_cold_collapse(n, 1.0, 1.0, c.seed, gas ? 1 : 0, &s->view),
              "cold-collapse");
        b->name = gas ? "Cold Collapse (Gas)" : "Cold Collapse";
        const double t_ff = kPi / 2.0 * std::sqrt((1.0 * 1.0 * 1.0) / (2.0 * 1.0 * 1.0));
        b->dt = gas ? t_ff / 500.0 : t_ff / 200.0;  // cold_collapse.rs:104-115
        b->softening =
            std::pow(4.0 / 3.0 * kPi * (1.0 * 1.0 * 1.0) / (double)n, 1.0 / 3.0) * 0.3;
    } else if (sc == "galaxy-collision") {
        gravis_galaxy_params g;
        gravis_galaxy_default_params(&g);
        g.n_per_galaxy = (c.particles > 0 ? (uint64_t)c.particles : 10000) / 2;
        g.seed = c.seed;
        s->resize(2 * g.n_per_galaxy);
        check(nullptr, gravis_scenario_galaxy_collision(&g, &s->view), "galaxy-collision");
        b->name = "Galaxy Collision";
        b->dt = gravis_galaxy_suggested_dt(&g);
        b->softening = gravis_galaxy_suggested_softening(&g);
    } else if (sc == "evrard-collapse") {
        const size_t n = c.particles > 0 ? (size_t)c.particles : 5000;
        s->resize(n);
        check(nullptr, gravis_scenario_evrard(n, 1.0, 1.0, 0.05, c.seed, &s->view), "evrard");
        b->name = "Evrard Collapse";
        const double t_ff = kPi / 2.0 * std::sqrt((1.0 * 1.0 * 1.0) / (2.0 * 1.0));
        b->dt = t_ff / 500.0;  // evrard_collapse.rs:60-66
        b->softening = gravis_evrard_softening(n, 1.0);
    } else {
        std::fprintf(stderr,
                     "Unknown scenario: %s. Available on this backend: plummer, two-body, "
                     "cold-collapse, cold-collapse-gas, galaxy-collision, evrard-collapse (the "
                     "reference's pure-hydro and protoplanetary scenarios run on sim-core's "
                     "CPU path)\n",
                     sc.c_str());
        return false;
    }
    if (!std::isnan(c.dt)) b->dt = c.dt;
    if (!std::isnan(c.softening)) b->softening = c.softening;
    return true;
}

struct Diag {
    gravis_diag d;
    double mom() const {
        return std::sqrt(d.momentum[0] * d.momentum[0] + d.momentum[1] * d.momentum[1] +
                         d.momentum[2] * d.momentum[2]);
    }
    double angmom() const {
        const double* l = d.angular_momentum;
        return std::sqrt(l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
    }
    double com_drift() const {
        const double* c = d.center_of_mass;
        return std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    }
    // diagnostics.rs:176-189
    std::string csv() const {
        return std::to_string(d.step) + "," + fexp(d.time, 6) + "," +
               fexp(d.kinetic_energy, 10) + "," + fexp(d.potential_energy, 10) + "," +
               fexp(d.thermal_energy, 10) + "," + fexp(d.total_energy, 10) + "," +
               fexp(mom(), 10) + "," + fexp(angmom(), 10) + "," + fexp(com_drift(), 10) + "," +
               ffix(d.virial_ratio, 6);
    }
};

Diag diagnose(gravis_ctx* ctx, State& s, double eps, bool full, double time, uint64_t step) {
    Diag r;
    const bool gas = s.has_gas();
    check(ctx,
          gravis_diagnostics(ctx, s.n, eps, full ? 1 : 0, s.view.x, s.view.y, s.view.z,
                             s.view.vx, s.view.vy, s.view.vz, s.view.m, gas ? s.view.u : nullptr,
                             gas ? s.view.type : nullptr, time, step, &r.d),
          "diagnostics");
    return r;
}

void write_snapshot(const Cli& c, State& s, double t, uint64_t step, double eps, double dt) {
    char name[64];
    std::snprintf(name, sizeof name, "/snap_%08llu.bin", (unsigned long long)step);
    const std::string path = c.snapshot_dir + name;
    if (gravis_snapshot_write(path.c_str(), t, step, eps, dt, &s.view) != GRAVIS_OK) {
        std::fprintf(stderr, "Failed to write snapshot %s\n", path.c_str());
        std::exit(1);
    }
}

void mkdirs(const std::string& dir) {
    std::string cur;
    for (size_t i = 0; i <= dir.size(); ++i) {
        if (i == dir.size() || dir[i] == '/') {
            if (!cur.empty()) mkdir(cur.c_str(), 0755);
        }
        if (i < dir.size()) cur += dir[i];
    }
}

}  // namespace

int main(int argc, char** argv) {
    const Cli cli = parse(argc, argv);
    // main.rs:73-95
    if (cli.particles == 0) {
        std::fprintf(stderr, "Error: --particles must be > 0\n");
        return 1;
    }
    if (cli.eccentricity < 0.0 || cli.eccentricity >= 1.0) {
        std::fprintf(stderr, "Error: --eccentricity must be in [0, 1), got %s\n",
                     fdisp(cli.eccentricity).c_str());
        return 1;
    }
    if (!std::isnan(cli.dt) && cli.dt <= 0.0) {
        std::fprintf(stderr, "Error: --dt must be positive, got %s\n", fdisp(cli.dt).c_str());
        return 1;
    }
    if (!std::isnan(cli.softening) && cli.softening < 0.0) {
        std::fprintf(stderr, "Error: --softening must be non-negative, got %s\n",
                     fdisp(cli.softening).c_str());
        return 1;
    }
    if (cli.backend != "hip") {
        std::fprintf(stderr,
                     "Error: --backend %s: this runner drives the gfx950 engine only "
                     "(--backend hip); the CPU backend is sim-core's own headless\n",
                     cli.backend.c_str());
        return 1;
    }

    State s;
    Built b;
    uint64_t step0 = 0;
    double sim_time = 0.0;
    if (!cli.restart.empty()) {
        // Snapshot::read_from (snapshot.rs:74-151)
        int ver = 0;
        uint64_t cnt = 0, st = 0;
        double t = 0, so = 0, dt = 0;
        if (gravis_snapshot_info(cli.restart.c_str(), &ver, &t, &st, &so, &dt, &cnt) !=
            GRAVIS_OK) {
            std::fprintf(stderr, "Error: cannot read snapshot %s\n", cli.restart.c_str());
            return 1;
        }
        s.resize(cnt);
        if (gravis_snapshot_read(cli.restart.c_str(), &s.view) != GRAVIS_OK) {
            std::fprintf(stderr, "Error: cannot read snapshot %s\n", cli.restart.c_str());
            return 1;
        }
        b.name = "Snapshot " + cli.restart;
        b.dt = std::isnan(cli.dt) ? dt : cli.dt;
        b.softening = std::isnan(cli.softening) ? so : cli.softening;
        step0 = st;
        sim_time = t;
    } else if (!build(cli, &s, &b)) {
        return 1;
    }
    const double dt = b.dt, softening = b.softening;
    const bool use_sph = s.has_gas();

    gravis_ctx* ctx = nullptr;
    check(nullptr, gravis_create(0, &ctx), "gravis_create");
    gravis_solver solver{};
    solver.algorithm = cli.algorithm == "barnes-hut" ? GRAVIS_BARNES_HUT : GRAVIS_BRUTE_FORCE;
    solver.precision = cli.precision == "mixed" ? GRAVIS_MIXED : GRAVIS_FP64;
    solver.softening = softening;
    solver.theta = cli.theta;
    solver.leaf_size = cli.leaf_size;
    solver.group_size = 64;
    gravis_sph_params sp{};
    gravis_sph_default_params(&sp);

    // main.rs:123-131
    std::fprintf(stderr, "Scenario:   %s\n", b.name.c_str());
    std::fprintf(stderr, "Particles:  %zu\n", s.n);
    std::fprintf(stderr, "Algorithm:  %s\n", cli.algorithm.c_str());
    if (cli.algorithm == "barnes-hut") std::fprintf(stderr, "Theta:      %s\n", fdisp(cli.theta).c_str());
    std::fprintf(stderr, "Steps:      %llu\n", (unsigned long long)cli.steps);
    std::fprintf(stderr, "dt:         %s\n", fexp(dt, 6).c_str());
    std::fprintf(stderr, "Softening:  %s\n", fexp(softening, 6).c_str());
    std::fprintf(stderr, "\n");

    FILE* csv = nullptr;
    if (!cli.diag_csv.empty()) {
        csv = std::fopen(cli.diag_csv.c_str(), "w");
        if (!csv) {
            std::fprintf(stderr, "Failed to create diagnostics CSV\n");
            return 1;
        }
        std::fprintf(csv, "step,time,kinetic_energy,potential_energy,thermal_energy,"
                          "total_energy,momentum_mag,angular_momentum_mag,com_drift,"
                          "virial_ratio\n");
    }
    if (cli.snapshot_interval > 0) mkdirs(cli.snapshot_dir);

    // main.rs:150-155: accelerations before the first step
    for (int k = 6; k <= 8; ++k) std::fill(s.f[k].begin(), s.f[k].end(), 0.0);
    check(ctx,
          gravis_accel(ctx, &solver, s.n, s.view.x, s.view.y, s.view.z, s.view.m, s.view.ax,
                       s.view.ay, s.view.az),
          "gravity");
    std::vector<double> iv(3 * s.n);
    if (use_sph)
        check(ctx,
              gravis_sph_compute(ctx, &sp, &s.view, iv.data(), iv.data() + s.n,
                                 iv.data() + 2 * s.n, nullptr, nullptr),
              "SPH compute");

    const bool use_full_diag = cli.diag_interval > 0 || !cli.diag_csv.empty();
    const Diag d0 = diagnose(ctx, s, softening, use_full_diag, sim_time, step0);
    std::fprintf(stderr, "Initial:    E=%s  2K/|U|=%s\n", fexp(d0.d.total_energy, 10).c_str(),
                 ffix(d0.d.virial_ratio, 4).c_str());
    if (csv) std::fprintf(csv, "%s\n", d0.csv().c_str());

    const auto wall0 = std::chrono::steady_clock::now();
    double current_dt = dt;
    const uint64_t last = step0 + cli.steps;
    uint64_t step = step0;
    while (step < last) {
        // next step with output
        uint64_t next = last;
        if (cli.diag_interval > 0) next = std::min(next, (step / cli.diag_interval + 1) * cli.diag_interval);
        if (cli.snapshot_interval > 0)
            next = std::min(next, (step / cli.snapshot_interval + 1) * cli.snapshot_interval);
        const uint64_t k = next - step;
        if (use_sph) {
            // step_with_sph + dt = min(dt_next, 1.5 dt), sim_time += dt (main.rs:181-187)
            check(ctx, gravis_sph_run(ctx, &sp, &solver, &s.view, k, &current_dt, &sim_time),
                  "SPH steps");
        } else {
            check(ctx,
                  gravis_kdk(ctx, &solver, s.n, s.view.x, s.view.y, s.view.z, s.view.vx,
                             s.view.vy, s.view.vz, s.view.m, s.view.ax, s.view.ay, s.view.az,
                             current_dt, k),
                  "KDK steps");
            for (uint64_t q = 0; q < k; ++q) sim_time += current_dt;
        }
        step = next;
        if (cli.diag_interval > 0 && step % cli.diag_interval == 0) {
            const Diag d = diagnose(ctx, s, softening, true, sim_time, step);
            const double de = (d.d.total_energy - d0.d.total_energy) / std::fabs(d0.d.total_energy);
            std::fprintf(stderr, "Step %8llu  t=%s  E=%s  dE/E=%s  2K/|U|=%s\n",
                         (unsigned long long)step, fexp(sim_time, 4).c_str(),
                         fexp(d.d.total_energy, 10).c_str(), fexp(de, 4, true).c_str(),
                         ffix(d.d.virial_ratio, 4).c_str());
            if (csv) std::fprintf(csv, "%s\n", d.csv().c_str());
        }
        if (cli.snapshot_interval > 0 && step % cli.snapshot_interval == 0)
            write_snapshot(cli, s, sim_time, step, softening, current_dt);
    }
    const double wall =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();

    // main.rs:218-246
    const Diag df = diagnose(ctx, s, softening, use_full_diag, sim_time, last);
    const double de = (df.d.total_energy - d0.d.total_energy) / std::fabs(d0.d.total_energy);
    std::fprintf(stderr, "\n=== Final ===\n");
    std::fprintf(stderr, "Sim time:   %s\n", fexp(sim_time, 6).c_str());
    std::fprintf(stderr, "Wall time:  %ss\n", ffix(wall, 3).c_str());
    std::fprintf(stderr, "Steps/sec:  %s\n", ffix((double)cli.steps / wall, 0).c_str());
    std::fprintf(stderr, "Energy:     %s  (dE/E = %s)\n", fexp(df.d.total_energy, 10).c_str(),
                 fexp(de, 6, true).c_str());
    std::fprintf(stderr, "Momentum:   %s\n", fexp(df.mom(), 6).c_str());
    std::fprintf(stderr, "Ang. Mom:   %s\n", fexp(df.angmom(), 6).c_str());
    std::fprintf(stderr, "COM drift:  %s\n", fexp(df.com_drift(), 6).c_str());
    std::fprintf(stderr, "Virial:     %s\n", ffix(df.d.virial_ratio, 4).c_str());
    const bool already = cli.diag_interval > 0 && last % cli.diag_interval == 0;
    if (!already && csv) std::fprintf(csv, "%s\n", df.csv().c_str());
    if (csv) std::fclose(csv);
    gravis_destroy(ctx);
    return 0;
}
