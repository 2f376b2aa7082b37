// api.hip -- the extern "C" boundary (include/gravis.h) over the gfx950
// kernels. Validates arguments the way the reference's CLI and asserts do
// (crates/headless/src/main.rs:73-95, octree.rs:16), converts every failure
// into a gravis_status plus a message, and never throws across the ABI.
#include <cmath>
#include <cstdarg>
#include <algorithm>
#include <cstring>
#include <new>

#include "internal.h"

namespace gravis {

gravis_status fail(Ctx* c, gravis_status s, const char* fmt, ...) {
    if (c) {
        char msg[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(msg, sizeof msg, fmt, ap);
        va_end(ap);
        c->err = msg;
    }
    return s;
}

gravis_status buffer(Ctx* c, const char* name, size_t bytes, void** out) {
    if (bytes == 0) bytes = 256;
    if (c->arena) {
        for (int k = 0; k < c->arena_n; ++k) {
            const Ctx::ArenaEntry& e = c->arena_tab[k];
            if (std::strcmp(e.name, name) != 0) continue;
            if (e.bytes < bytes)
                return fail(c, GRAVIS_INTERNAL, "scratch range %s: %zu bytes asked, %zu held",
                            name, bytes, e.bytes);
            *out = c->arena + e.off;
            return GRAVIS_OK;
        }
        const size_t need = (bytes + 255) / 256 * 256;
        if (c->arena_n == Ctx::kArenaEntries || c->arena_used + need > c->arena_cap)
            return fail(c, GRAVIS_INVALID_ARG,
                        "the no-allocation scratch (%zu bytes) cannot hold %s (%zu more bytes)",
                        c->arena_cap, name, need);
        c->arena_tab[c->arena_n++] = {name, c->arena_used, need};
        *out = c->arena + c->arena_used;
        c->arena_used += need;
        return GRAVIS_OK;
    }
    DevBuf& b = c->bufs[name];
    if (b.bytes < bytes) {
        if (b.p) {
            // The old buffer may still be read by queued work.
            hipError_t e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess)
                return fail(c, GRAVIS_HIP_ERROR, "sync before realloc: %s", hipGetErrorString(e));
            (void)hipFree(b.p);
            b.p = nullptr;
            b.bytes = 0;
        }
        size_t want = bytes + bytes / 8;  // headroom for growing N
        hipError_t e = hipMalloc(&b.p, want);
        if (e != hipSuccess) {
            b.p = nullptr;
            (void)hipGetLastError();
            return fail(c, GRAVIS_OUT_OF_MEMORY, "hipMalloc(%zu) for %s: %s", want, name,
                        hipGetErrorString(e));
        }
        b.bytes = want;
    }
    *out = b.p;
    return GRAVIS_OK;
}

gravis_status check_launch(Ctx* c, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(c, GRAVIS_HIP_ERROR, "%s launch failed: %s", what, hipGetErrorString(e));
    return GRAVIS_OK;
}

void timer_begin(Ctx* c) {
    if (c->profiling) (void)hipEventRecord(c->ev0, c->stream);
}

void timer_end(Ctx* c, const char* name) {
    if (!c->profiling) return;
    (void)hipEventRecord(c->ev1, c->stream);
    if (hipEventSynchronize(c->ev1) == hipSuccess) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess) {
            KernelTimer& t = c->timers[name];
            t.ms += ms;
            t.launches += 1;
        }
    }
}

}  // namespace gravis

using gravis::Ctx;
using gravis::fail;

namespace {

gravis_status check_solver(Ctx* c, const gravis_solver* s) {
    if (!s) return fail(c, GRAVIS_INVALID_ARG, "solver is NULL");
    if (!(s->softening >= 0.0))
        return fail(c, GRAVIS_INVALID_ARG, "softening must be non-negative, got %g", s->softening);
    if (s->algorithm != GRAVIS_BRUTE_FORCE && s->algorithm != GRAVIS_BARNES_HUT &&
        s->algorithm != GRAVIS_NO_GRAVITY)
        return fail(c, GRAVIS_INVALID_ARG, "unknown algorithm %d", s->algorithm);
    if (s->precision != GRAVIS_FP64 && s->precision != GRAVIS_MIXED)
        return fail(c, GRAVIS_INVALID_ARG, "unknown precision %d", s->precision);
    if (s->algorithm == GRAVIS_BARNES_HUT && !(s->theta >= 0.0))
        return fail(c, GRAVIS_INVALID_ARG, "theta must be non-negative, got %g", s->theta);
    return GRAVIS_OK;
}

bool any_null(std::initializer_list<const void*> ps) {
    for (const void* p : ps)
        if (!p) return true;
    return false;
}

gravis_status force_device(Ctx* c, const gravis_solver* s, size_t n, const double* x,
                           const double* y, const double* z, const double* m,
                           const gravis::FinishArgs& fa) {
    switch (s->algorithm) {
        case GRAVIS_BRUTE_FORCE:
            return gravis::bf_force(c, s, n, x, y, z, m, fa);
        case GRAVIS_BARNES_HUT:
            return gravis::bh_force(c, s, n, x, y, z, m, fa);
        default:
            if (fa.mode == gravis::kAssignKick) {
                GRAVIS_HIP(c, hipMemsetAsync(fa.ax, 0, n * sizeof(double), c->stream));
                GRAVIS_HIP(c, hipMemsetAsync(fa.ay, 0, n * sizeof(double), c->stream));
                GRAVIS_HIP(c, hipMemsetAsync(fa.az, 0, n * sizeof(double), c->stream));
            }
            return GRAVIS_OK;
    }
}

struct HostState {
    double *x, *y, *z, *vx, *vy, *vz, *m, *ax, *ay, *az;
};

gravis_status stage(Ctx* c, size_t n, HostState* d) {
    double* base;
    GRAVIS_TRY(gravis::buf(c, "host_state", 10 * n, &base));
    d->x = base; d->y = base + n; d->z = base + 2 * n;
    d->vx = base + 3 * n; d->vy = base + 4 * n; d->vz = base + 5 * n;
    d->m = base + 6 * n;
    d->ax = base + 7 * n; d->ay = base + 8 * n; d->az = base + 9 * n;
    return GRAVIS_OK;
}

gravis_status h2d(Ctx* c, double* d, const double* h, size_t n) {
    GRAVIS_HIP(c, hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    return GRAVIS_OK;
}

gravis_status d2h(Ctx* c, double* h, const double* d, size_t n) {
    GRAVIS_HIP(c, hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return GRAVIS_OK;
}

void fill_diag(const double r[19], double time, uint64_t step, gravis_diag* out) {
    out->time = time;
    out->step = step;
    out->kinetic_energy = r[0];
    out->potential_energy = r[1];
    out->thermal_energy = r[2];
    out->total_energy = r[3];
    for (int d = 0; d < 3; ++d) {
        out->momentum[d] = r[4 + d];
        out->angular_momentum[d] = r[7 + d];
        out->center_of_mass[d] = r[10 + d];
        out->center_of_mass_velocity[d] = r[13 + d];
    }
    out->virial_ratio = r[16];
}

gravis_status check_sph(Ctx* c, const gravis_sph_params* P, const gravis_sph_state* S) {
    if (!P) return fail(c, GRAVIS_INVALID_ARG, "SPH params is NULL");
    if (!S) return fail(c, GRAVIS_INVALID_ARG, "SPH state is NULL");
    if (S->n > 0 && any_null({S->x, S->y, S->z, S->vx, S->vy, S->vz, S->ax, S->ay, S->az, S->m,
                              S->rho, S->p, S->u, S->dudt, S->h, S->c, S->alpha, S->type}))
        return fail(c, GRAVIS_INVALID_ARG, "null array in SPH state");
    if (S->n > 0x7FFFFFFFull) return fail(c, GRAVIS_INVALID_ARG, "n too large");
    return GRAVIS_OK;
}

// The 17 f64 arrays of a gravis_sph_state in declaration order.
inline double** sph_fields(gravis_sph_state* S, int k) {
    double** f[17] = {&S->x, &S->y, &S->z, &S->vx, &S->vy, &S->vz, &S->ax, &S->ay, &S->az,
                      &S->m, &S->rho, &S->p, &S->u, &S->dudt, &S->h, &S->c, &S->alpha};
    return f[k];
}

// Device copy of a host SPH state (context buffer); copies it in.
gravis_status sph_stage_in(Ctx* c, const gravis_sph_state* H, gravis_sph_state* D) {
    const size_t n = H->n;
    double* base;
    uint8_t* ty;
    GRAVIS_TRY(gravis::buf(c, "sph_host_state", 17 * n, &base));
    GRAVIS_TRY(gravis::buf(c, "sph_host_type", n, &ty));
    gravis_sph_state h = *H;
    D->n = n;
    for (int k = 0; k < 17; ++k) {
        *sph_fields(D, k) = base + (size_t)k * n;
        GRAVIS_TRY(h2d(c, *sph_fields(D, k), *sph_fields(&h, k), n));
    }
    D->type = ty;
    GRAVIS_HIP(c, hipMemcpyAsync(ty, H->type, n, hipMemcpyHostToDevice, c->stream));
    return GRAVIS_OK;
}

// Copies the fields an SPH call changes back to the host state.
gravis_status sph_stage_out(Ctx* c, gravis_sph_state* D, gravis_sph_state* H, bool all) {
    const size_t n = H->n;
    for (int k = 0; k < 17; ++k) {
        const bool changed = all || k >= 10 || (k >= 6 && k <= 8);
        if (k == 9 || !changed) continue;  // masses never change
        GRAVIS_TRY(d2h(c, *sph_fields(H, k), *sph_fields(D, k), n));
    }
    return GRAVIS_OK;
}

}  // namespace

extern "C" {

const char* gravis_version(void) { return "gravis-mi355x 0.1 (gfx950)"; }

const char* gravis_status_string(gravis_status s) {
    switch (s) {
        case GRAVIS_OK: return "ok";
        case GRAVIS_INVALID_ARG: return "invalid argument";
        case GRAVIS_HIP_ERROR: return "HIP error";
        case GRAVIS_OUT_OF_MEMORY: return "out of device memory";
        case GRAVIS_UNSUPPORTED: return "unsupported";
        case GRAVIS_INTERNAL: return "internal capacity exceeded";
    }
    return "unknown status";
}

gravis_status gravis_create(int device, gravis_ctx** out) {
    if (!out) return GRAVIS_INVALID_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return GRAVIS_HIP_ERROR;
    }
    if (device < 0 || device >= count) return GRAVIS_INVALID_ARG;
    if (hipSetDevice(device) != hipSuccess) return GRAVIS_HIP_ERROR;
    gravis_ctx* g = new (std::nothrow) gravis_ctx();
    if (!g) return GRAVIS_OUT_OF_MEMORY;
    Ctx* c = &g->c;
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        c->num_cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete g;
        return GRAVIS_HIP_ERROR;
    }
    c->stream = c->own_stream;
    gravis::noalloc_setup(c);
    *out = g;
    return GRAVIS_OK;
}

void gravis_destroy(gravis_ctx* g) {
    if (!g) return;
    Ctx* c = &g->c;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->stream != c->own_stream) (void)hipStreamSynchronize(c->own_stream);
    for (auto& kv : c->bufs)
        if (kv.second.p) (void)hipFree(kv.second.p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete g;
}

const char* gravis_last_error(const gravis_ctx* g) { return g ? g->c.err.c_str() : "no context"; }

gravis_status gravis_set_stream(gravis_ctx* g, void* stream) {
    if (!g) return GRAVIS_INVALID_ARG;
    g->c.stream = stream ? static_cast<hipStream_t>(stream) : g->c.own_stream;
    return GRAVIS_OK;
}

gravis_status gravis_synchronize(gravis_ctx* g) {
    if (!g) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    GRAVIS_HIP(c, hipStreamSynchronize(c->stream));
    return GRAVIS_OK;
}

gravis_status gravis_set_profiling(gravis_ctx* g, int enable) {
    if (!g) return GRAVIS_INVALID_ARG;
    g->c.profiling = enable != 0;
    return GRAVIS_OK;
}

gravis_status gravis_kernel_time(gravis_ctx* g, const char* name, double* ms_total,
                                 uint64_t* launches) {
    if (!g || !name || !ms_total || !launches) return GRAVIS_INVALID_ARG;
    auto it = g->c.timers.find(name);
    if (it == g->c.timers.end()) {
        *ms_total = 0.0;
        *launches = 0;
        return GRAVIS_OK;
    }
    *ms_total = it->second.ms;
    *launches = it->second.launches;
    g->c.timers.erase(it);
    return GRAVIS_OK;
}

gravis_status gravis_accel_device(gravis_ctx* g, const gravis_solver* s, size_t n,
                                  const double* x, const double* y, const double* z,
                                  const double* m, double* ax, double* ay, double* az) {
    if (!g) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    GRAVIS_TRY(check_solver(c, s));
    if (n == 0) return GRAVIS_OK;  // barnes_hut.rs:141-143: empty system, nothing to do
    if (any_null({x, y, z, m, ax, ay, az})) return fail(c, GRAVIS_INVALID_ARG, "null array");
    GRAVIS_HIP(c, hipSetDevice(c->device));
    gravis::FinishArgs fa;
    fa.mode = gravis::kAccumulate;
    fa.ax = ax; fa.ay = ay; fa.az = az;
    return force_device(c, s, n, x, y, z, m, fa);
}

gravis_status gravis_accel(gravis_ctx* g, const gravis_solver* s, size_t n, const double* x,
                           const double* y, const double* z, const double* m, double* ax,
                           double* ay, double* az) {
    if (!g) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    GRAVIS_TRY(check_solver(c, s));
    if (n == 0) return GRAVIS_OK;
    if (any_null({x, y, z, m, ax, ay, az})) return fail(c, GRAVIS_INVALID_ARG, "null array");
    GRAVIS_HIP(c, hipSetDevice(c->device));
    HostState d;
    GRAVIS_TRY(stage(c, n, &d));
    GRAVIS_TRY(h2d(c, d.x, x, n)); GRAVIS_TRY(h2d(c, d.y, y, n)); GRAVIS_TRY(h2d(c, d.z, z, n));
    GRAVIS_TRY(h2d(c, d.m, m, n));
    GRAVIS_TRY(h2d(c, d.ax, ax, n)); GRAVIS_TRY(h2d(c, d.ay, ay, n)); GRAVIS_TRY(h2d(c, d.az, az, n));
    gravis::FinishArgs fa;
    fa.mode = gravis::kAccumulate;
    fa.ax = d.ax; fa.ay = d.ay; fa.az = d.az;
    GRAVIS_TRY(force_device(c, s, n, d.x, d.y, d.z, d.m, fa));
    GRAVIS_TRY(d2h(c, ax, d.ax, n)); GRAVIS_TRY(d2h(c, ay, d.ay, n)); GRAVIS_TRY(d2h(c, az, d.az, n));
    GRAVIS_HIP(c, hipStreamSynchronize(c->stream));
    return GRAVIS_OK;
}

gravis_status gravis_kdk_device(gravis_ctx* g, const gravis_solver* s, size_t n, double* x,
                                double* y, double* z, double* vx, double* vy, double* vz,
                                const double* m, double* ax, double* ay, double* az, double dt,
                                uint64_t nsteps) {
    if (!g) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    GRAVIS_TRY(check_solver(c, s));
    if (!(dt > 0.0)) return fail(c, GRAVIS_INVALID_ARG, "dt must be positive, got %g", dt);
    if (n == 0 || nsteps == 0) return GRAVIS_OK;
    if (any_null({x, y, z, vx, vy, vz, m, ax, ay, az}))
        return fail(c, GRAVIS_INVALID_ARG, "null array");
    GRAVIS_HIP(c, hipSetDevice(c->device));
    const double h = 0.5 * dt;
    gravis::FinishArgs fa;
    fa.mode = gravis::kAssignKick;
    fa.ax = ax; fa.ay = ay; fa.az = az;
    fa.vx = vx; fa.vy = vy; fa.vz = vz;
    fa.half_dt = h;
    if (s->algorithm == GRAVIS_BRUTE_FORCE && s->precision == GRAVIS_FP64 &&
        n <= c->kdk_fused_max_n)
        return gravis::kdk_small(c, n, x, y, z, vx, vy, vz, m, ax, ay, az, s->softening, dt,
                                 nsteps);
    if (s->algorithm == GRAVIS_BRUTE_FORCE && s->precision == GRAVIS_FP64 &&
        n > (size_t)gravis::kSmallN && n <= c->kdk_coop_max_n && nsteps >= 2) {
        // one cooperative launch, unless the stream is being captured (the
        // call synchronises it)
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        GRAVIS_HIP(c, hipStreamIsCapturing(c->stream, &cs));
        if (cs == hipStreamCaptureStatusNone) {
            // steps the runtime would not run cooperatively continue below
            GRAVIS_TRY(gravis::kdk_coop(c, n, x, y, z, vx, vy, vz, m, ax, ay, az, s->softening,
                                        dt, &nsteps));
            if (nsteps == 0) return GRAVIS_OK;
        }
    }
    for (uint64_t step = 0; step < nsteps; ++step) {
        GRAVIS_TRY(gravis::kick_drift(c, n, x, y, z, vx, vy, vz, ax, ay, az, h, dt));
        GRAVIS_TRY(force_device(c, s, n, x, y, z, m, fa));
    }
    return GRAVIS_OK;
}

gravis_status gravis_kdk(gravis_ctx* g, const gravis_solver* s, size_t n, double* x, double* y,
                         double* z, double* vx, double* vy, double* vz, const double* m,
                         double* ax, double* ay, double* az, double dt, uint64_t nsteps) {
    if (!g) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    GRAVIS_TRY(check_solver(c, s));
    if (!(dt > 0.0)) return fail(c, GRAVIS_INVALID_ARG, "dt must be positive, got %g", dt);
    if (n == 0 || nsteps == 0) return GRAVIS_OK;
    if (any_null({x, y, z, vx, vy, vz, m, ax, ay, az}))
        return fail(c, GRAVIS_INVALID_ARG, "null array");
    GRAVIS_HIP(c, hipSetDevice(c->device));
    HostState d;
    GRAVIS_TRY(stage(c, n, &d));
    const double* hin[10] = {x, y, z, vx, vy, vz, m, ax, ay, az};
    double* din[10] = {d.x, d.y, d.z, d.vx, d.vy, d.vz, d.m, d.ax, d.ay, d.az};
    for (int k = 0; k < 10; ++k) GRAVIS_TRY(h2d(c, din[k], hin[k], n));
    GRAVIS_TRY(gravis_kdk_device(g, s, n, d.x, d.y, d.z, d.vx, d.vy, d.vz, d.m, d.ax, d.ay,
                                 d.az, dt, nsteps));
    double* hout[9] = {x, y, z, vx, vy, vz, ax, ay, az};
    double* dout[9] = {d.x, d.y, d.z, d.vx, d.vy, d.vz, d.ax, d.ay, d.az};
    for (int k = 0; k < 9; ++k) GRAVIS_TRY(d2h(c, hout[k], dout[k], n));
    GRAVIS_HIP(c, hipStreamSynchronize(c->stream));
    return GRAVIS_OK;
}

gravis_status gravis_diagnostics_device(gravis_ctx* g, size_t n, double softening,
                                        int compute_potential, const double* x,
                                        const double* y, const double* z, const double* vx,
                                        const double* vy, const double* vz, const double* m,
                                        const double* u, const uint8_t* ptype, double time,
                                        uint64_t step, gravis_diag* out) {
    if (!g || !out) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    if (!(softening >= 0.0)) return fail(c, GRAVIS_INVALID_ARG, "softening must be >= 0");
    double r[19] = {0};
    if (n > 0) {
        if (any_null({x, y, z, vx, vy, vz, m})) return fail(c, GRAVIS_INVALID_ARG, "null array");
        GRAVIS_HIP(c, hipSetDevice(c->device));
        GRAVIS_TRY(gravis::diag_device(c, n, softening, compute_potential, x, y, z, vx, vy, vz,
                                       m, u, ptype, r));
    }
    fill_diag(r, time, step, out);
    return GRAVIS_OK;
}

gravis_status gravis_render_pack_device(gravis_ctx* g, size_t n, const double* x,
                                       const double* y, const double* z, const double* m,
                                       const double* u, const uint8_t* ptype, void* out) {
    if (!g) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    if (n == 0) return GRAVIS_OK;
    if (any_null({x, y, z, m, out})) return fail(c, GRAVIS_INVALID_ARG, "null array");
    GRAVIS_HIP(c, hipSetDevice(c->device));
    return gravis::render_pack(c, n, x, y, z, m, u, ptype, out);
}

gravis_status gravis_render_pack(gravis_ctx* g, size_t n, const double* x, const double* y,
                                const double* z, const double* m, const double* u,
                                const uint8_t* ptype, void* out) {
    if (!g) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    if (n == 0) return GRAVIS_OK;
    if (any_null({x, y, z, m, out})) return fail(c, GRAVIS_INVALID_ARG, "null array");
    GRAVIS_HIP(c, hipSetDevice(c->device));
    double* base;
    GRAVIS_TRY(gravis::buf(c, "render_in", 5 * n, &base));
    uint8_t* dt = nullptr;
    float* dout;
    GRAVIS_TRY(gravis::buf(c, "render_out", 8 * n, &dout));
    GRAVIS_TRY(h2d(c, base, x, n)); GRAVIS_TRY(h2d(c, base + n, y, n));
    GRAVIS_TRY(h2d(c, base + 2 * n, z, n)); GRAVIS_TRY(h2d(c, base + 3 * n, m, n));
    if (u) GRAVIS_TRY(h2d(c, base + 4 * n, u, n));
    if (ptype) {
        GRAVIS_TRY(gravis::buf(c, "render_type", n, &dt));
        GRAVIS_HIP(c, hipMemcpyAsync(dt, ptype, n, hipMemcpyHostToDevice, c->stream));
    }
    GRAVIS_TRY(gravis::render_pack(c, n, base, base + n, base + 2 * n, base + 3 * n,
                                   u ? base + 4 * n : nullptr, dt, dout));
    GRAVIS_HIP(c, hipMemcpyAsync(out, dout, 32 * n, hipMemcpyDeviceToHost, c->stream));
    GRAVIS_HIP(c, hipStreamSynchronize(c->stream));
    return GRAVIS_OK;
}

gravis_status gravis_diagnostics(gravis_ctx* g, size_t n, double softening,
                                 int compute_potential, const double* x, const double* y,
                                 const double* z, const double* vx, const double* vy,
                                 const double* vz, const double* m, const double* u,
                                 const uint8_t* ptype, double time, uint64_t step,
                                 gravis_diag* out) {
    if (!g || !out) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    if (!(softening >= 0.0)) return fail(c, GRAVIS_INVALID_ARG, "softening must be >= 0");
    if (n == 0) {
        double r[19] = {0};
        fill_diag(r, time, step, out);
        return GRAVIS_OK;
    }
    if (any_null({x, y, z, vx, vy, vz, m})) return fail(c, GRAVIS_INVALID_ARG, "null array");
    GRAVIS_HIP(c, hipSetDevice(c->device));
    HostState d;
    GRAVIS_TRY(stage(c, n, &d));
    const double* hin[7] = {x, y, z, vx, vy, vz, m};
    double* din[7] = {d.x, d.y, d.z, d.vx, d.vy, d.vz, d.m};
    for (int k = 0; k < 7; ++k) GRAVIS_TRY(h2d(c, din[k], hin[k], n));
    double* du = nullptr;
    uint8_t* dt = nullptr;
    if (u && ptype) {
        GRAVIS_TRY(gravis::buf(c, "host_u", n, &du));
        GRAVIS_TRY(gravis::buf(c, "host_type", n, &dt));
        GRAVIS_TRY(h2d(c, du, u, n));
        GRAVIS_HIP(c, hipMemcpyAsync(dt, ptype, n, hipMemcpyHostToDevice, c->stream));
    }
    double r[19] = {0};
    GRAVIS_TRY(gravis::diag_device(c, n, softening, compute_potential, d.x, d.y, d.z, d.vx, d.vy,
                                   d.vz, d.m, du, dt, r));
    fill_diag(r, time, step, out);
    return GRAVIS_OK;
}

gravis_status gravis_ensemble_accel_device(gravis_ctx* g, const gravis_solver* s, int nsys,
                                           const uint64_t* offsets, const double* x,
                                           const double* y, const double* z, const double* m,
                                           double* ax, double* ay, double* az) {
    if (!g) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    if (any_null({offsets, x, y, z, m, ax, ay, az}))
        return fail(c, GRAVIS_INVALID_ARG, "null array");
    GRAVIS_HIP(c, hipSetDevice(c->device));
    return gravis::ensemble_accel(c, s, nsys, offsets, x, y, z, m, ax, ay, az);
}

gravis_status gravis_ensemble_accel(gravis_ctx* g, const gravis_solver* s, int nsys,
                                    const uint64_t* offsets, const double* x, const double* y,
                                    const double* z, const double* m, double* ax, double* ay,
                                    double* az) {
    if (!g) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    if (any_null({offsets, x, y, z, m, ax, ay, az}) || nsys <= 0)
        return fail(c, GRAVIS_INVALID_ARG, "null array or empty ensemble");
    const size_t n = offsets[nsys];
    if (n == 0) return GRAVIS_OK;
    GRAVIS_HIP(c, hipSetDevice(c->device));
    HostState d;
    GRAVIS_TRY(stage(c, n, &d));
    GRAVIS_TRY(h2d(c, d.x, x, n)); GRAVIS_TRY(h2d(c, d.y, y, n)); GRAVIS_TRY(h2d(c, d.z, z, n));
    GRAVIS_TRY(h2d(c, d.m, m, n));
    GRAVIS_TRY(gravis::ensemble_accel(c, s, nsys, offsets, d.x, d.y, d.z, d.m, d.ax, d.ay,
                                      d.az));
    GRAVIS_TRY(d2h(c, ax, d.ax, n)); GRAVIS_TRY(d2h(c, ay, d.ay, n)); GRAVIS_TRY(d2h(c, az, d.az, n));
    GRAVIS_HIP(c, hipStreamSynchronize(c->stream));
    return GRAVIS_OK;
}

gravis_status gravis_ensemble_kdk_device(gravis_ctx* g, const gravis_solver* s, int nsys,
                                         const uint64_t* offsets, double* x, double* y,
                                         double* z, double* vx, double* vy, double* vz,
                                         const double* m, double* ax, double* ay, double* az,
                                         double dt, uint64_t nsteps, int use_graph) {
    if (!g) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    if (any_null({offsets, x, y, z, vx, vy, vz, m, ax, ay, az}))
        return fail(c, GRAVIS_INVALID_ARG, "null array");
    GRAVIS_HIP(c, hipSetDevice(c->device));
    return gravis::ensemble_kdk(c, s, nsys, offsets, x, y, z, vx, vy, vz, m, ax, ay, az, dt,
                                nsteps, use_graph);
}

gravis_status gravis_ensemble_kdk(gravis_ctx* g, const gravis_solver* s, int nsys,
                                  const uint64_t* offsets, double* x, double* y, double* z,
                                  double* vx, double* vy, double* vz, const double* m,
                                  double* ax, double* ay, double* az, double dt,
                                  uint64_t nsteps, int use_graph) {
    if (!g) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    if (any_null({offsets, x, y, z, vx, vy, vz, m, ax, ay, az}) || nsys <= 0)
        return fail(c, GRAVIS_INVALID_ARG, "null array or empty ensemble");
    const size_t n = offsets[nsys];
    if (n == 0) return GRAVIS_OK;
    GRAVIS_HIP(c, hipSetDevice(c->device));
    HostState d;
    GRAVIS_TRY(stage(c, n, &d));
    const double* hin[10] = {x, y, z, vx, vy, vz, m, ax, ay, az};
    double* din[10] = {d.x, d.y, d.z, d.vx, d.vy, d.vz, d.m, d.ax, d.ay, d.az};
    for (int k = 0; k < 10; ++k) GRAVIS_TRY(h2d(c, din[k], hin[k], n));
    GRAVIS_TRY(gravis::ensemble_kdk(c, s, nsys, offsets, d.x, d.y, d.z, d.vx, d.vy, d.vz, d.m,
                                    d.ax, d.ay, d.az, dt, nsteps, use_graph));
    double* hout[9] = {x, y, z, vx, vy, vz, ax, ay, az};
    double* dout[9] = {d.x, d.y, d.z, d.vx, d.vy, d.vz, d.ax, d.ay, d.az};
    for (int k = 0; k < 9; ++k) GRAVIS_TRY(d2h(c, hout[k], dout[k], n));
    GRAVIS_HIP(c, hipStreamSynchronize(c->stream));
    return GRAVIS_OK;
}

gravis_status gravis_ensemble_adjoint_device(
    gravis_ctx* g, const gravis_solver* s, int nsys, const uint64_t* offsets, const double* x0,
    const double* y0, const double* z0, const double* vx0, const double* vy0, const double* vz0,
    const double* m, double dt, uint64_t nsteps, uint32_t ckpt, const double* tx,
    const double* ty, const double* tz, const double* gx, const double* gy, const double* gz,
    const double* gvx, const double* gvy, const double* gvz, double* loss, double* final_state,
    double* ox, double* oy, double* oz, double* ovx, double* ovy, double* ovz, double* om) {
    if (!g) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    if (any_null({offsets, x0, y0, z0, vx0, vy0, vz0, m, ox, oy, oz, ovx, ovy, ovz, om}))
        return fail(c, GRAVIS_INVALID_ARG, "null array");
    if (tx ? (!ty || !tz) : (!gx || !gy || !gz))
        return fail(c, GRAVIS_INVALID_ARG, "need all three targets or all three cotangents");
    if (gvx && (!gvy || !gvz)) return fail(c, GRAVIS_INVALID_ARG, "partial velocity cotangent");
    GRAVIS_HIP(c, hipSetDevice(c->device));
    return gravis::ensemble_adjoint(c, s, nsys, offsets, x0, y0, z0, vx0, vy0, vz0, m, dt, nsteps,
                                    ckpt, tx, ty, tz, gx, gy, gz, gvx, gvy, gvz, loss,
                                    final_state, ox, oy, oz, ovx, ovy, ovz, om);
}

gravis_status gravis_ensemble_adjoint(
    gravis_ctx* g, const gravis_solver* s, int nsys, const uint64_t* offsets, const double* x0,
    const double* y0, const double* z0, const double* vx0, const double* vy0, const double* vz0,
    const double* m, double dt, uint64_t nsteps, uint32_t ckpt, const double* tx,
    const double* ty, const double* tz, const double* gx, const double* gy, const double* gz,
    const double* gvx, const double* gvy, const double* gvz, double* loss, double* final_state,
    double* ox, double* oy, double* oz, double* ovx, double* ovy, double* ovz, double* om) {
    if (!g) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    if (!offsets || nsys <= 0) return fail(c, GRAVIS_INVALID_ARG, "empty ensemble");
    if (any_null({x0, y0, z0, vx0, vy0, vz0, m, ox, oy, oz, ovx, ovy, ovz, om}))
        return fail(c, GRAVIS_INVALID_ARG, "null array");
    if (tx ? (!ty || !tz) : (!gx || !gy || !gz))
        return fail(c, GRAVIS_INVALID_ARG, "need all three targets or all three cotangents");
    if (gvx && (!gvy || !gvz)) return fail(c, GRAVIS_INVALID_ARG, "partial velocity cotangent");
    const size_t n = offsets[nsys];
    if (n == 0) return GRAVIS_OK;
    GRAVIS_HIP(c, hipSetDevice(c->device));
    // inputs: 7 state + 3 target/cotangent + 3 velocity cotangent; outputs 7 + 6
    double* base;
    GRAVIS_TRY(gravis::buf(c, "adj_host", 26 * n + (size_t)nsys, &base));
    double* din[7];
    const double* hin[7] = {x0, y0, z0, vx0, vy0, vz0, m};
    for (int k = 0; k < 7; ++k) {
        din[k] = base + k * n;
        GRAVIS_TRY(h2d(c, din[k], hin[k], n));
    }
    double* dt3 = base + 7 * n;
    const double* hsrc[3] = {tx ? tx : gx, tx ? ty : gy, tx ? tz : gz};
    for (int k = 0; k < 3; ++k) GRAVIS_TRY(h2d(c, dt3 + k * n, hsrc[k], n));
    double* dgv = base + 10 * n;
    if (gvx) {
        const double* hv[3] = {gvx, gvy, gvz};
        for (int k = 0; k < 3; ++k) GRAVIS_TRY(h2d(c, dgv + k * n, hv[k], n));
    }
    double* dout = base + 13 * n;       // 7 gradients
    double* dfin = base + 20 * n;       // 6 final state
    double* dloss = base + 26 * n;
    GRAVIS_TRY(gravis::ensemble_adjoint(
        c, s, nsys, offsets, din[0], din[1], din[2], din[3], din[4], din[5], din[6], dt, nsteps,
        ckpt, tx ? dt3 : nullptr, tx ? dt3 + n : nullptr, tx ? dt3 + 2 * n : nullptr,
        tx ? nullptr : dt3, tx ? nullptr : dt3 + n, tx ? nullptr : dt3 + 2 * n,
        gvx ? dgv : nullptr, gvx ? dgv + n : nullptr, gvx ? dgv + 2 * n : nullptr, dloss, dfin,
        dout, dout + n, dout + 2 * n, dout + 3 * n, dout + 4 * n, dout + 5 * n, dout + 6 * n));
    double* hout[7] = {ox, oy, oz, ovx, ovy, ovz, om};
    for (int k = 0; k < 7; ++k) GRAVIS_TRY(d2h(c, hout[k], dout + k * n, n));
    if (final_state) GRAVIS_TRY(d2h(c, final_state, dfin, 6 * n));
    if (loss && tx) GRAVIS_TRY(d2h(c, loss, dloss, (size_t)nsys));
    GRAVIS_HIP(c, hipStreamSynchronize(c->stream));
    return GRAVIS_OK;
}

gravis_status gravis_set_tuning(gravis_ctx* g, const char* key, int64_t value) {
    if (!g || !key) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    if (std::strcmp(key, "bh_frontier_cap") == 0) {
        if (value < 64 || value > (1ll << 26))
            return fail(c, GRAVIS_INVALID_ARG, "bh_frontier_cap must be in [64, 2^26]");
        c->bh_frontier_cap = (uint32_t)value;
        return GRAVIS_OK;
    }
    if (std::strcmp(key, "sph_query_leaves") == 0) {
        if (value < 0 || value > 512)
            return fail(c, GRAVIS_INVALID_ARG, "sph_query_leaves must be in [0, 512]");
        c->sph_range_cap = (uint32_t)value;
        return GRAVIS_OK;
    }
    if (std::strcmp(key, "sph_pool_init") == 0) {
        if (value < 1 || value > 4096)
            return fail(c, GRAVIS_INVALID_ARG, "sph_pool_init must be in [1, 4096]");
        c->sph_pool_init = (uint32_t)value;
        return GRAVIS_OK;
    }
    if (std::strcmp(key, "kdk_coop_max_n") == 0) {
        if (value < 0 || value > gravis::kCoopMaxN)
            return fail(c, GRAVIS_INVALID_ARG, "kdk_coop_max_n must be in [0, %d]",
                        gravis::kCoopMaxN);
        c->kdk_coop_max_n = (uint32_t)value;
        return GRAVIS_OK;
    }
    if (std::strcmp(key, "kdk_fused_max_n") == 0) {
        if (value < 0 || value > gravis::kSmallN)
            return fail(c, GRAVIS_INVALID_ARG, "kdk_fused_max_n must be in [0, %d]",
                        gravis::kSmallN);
        c->kdk_fused_max_n = (uint32_t)value;
        return GRAVIS_OK;
    }
    if (std::strcmp(key, "bf_ipt") == 0) {
        if (value != 1 && value != 2)
            return fail(c, GRAVIS_INVALID_ARG, "bf_ipt must be 1 or 2");
        c->bf_ipt = (int)value;
        return GRAVIS_OK;
    }
    if (std::strcmp(key, "bf_matrix") == 0) {
        if (value < 0 || value > 2)
            return fail(c, GRAVIS_INVALID_ARG, "bf_matrix must be 0, 1 or 2");
        c->bf_matrix = (int)value;
        return GRAVIS_OK;
    }
    if (std::strcmp(key, "bf_near_64ths") == 0) {
        if (value < 0 || value > 64)
            return fail(c, GRAVIS_INVALID_ARG, "bf_near_64ths must be in [0, 64]");
        c->bf_near_fraction = (double)value / 64.0;
        return GRAVIS_OK;
    }
    if (std::strcmp(key, "bf_tile") == 0) {
        if (value != 128 && value != 256)
            return fail(c, GRAVIS_INVALID_ARG, "bf_tile must be 128 or 256");
        c->bf_tile = (int)value;
        return GRAVIS_OK;
    }
    if (std::strcmp(key, "bh_group_size") == 0) {
        if (value < 1 || value > 64)
            return fail(c, GRAVIS_INVALID_ARG, "bh_group_size must be in [1, 64]");
        c->bh_group = (uint32_t)value;
        return GRAVIS_OK;
    }
    if (std::strcmp(key, "bf_waves_per_cu") == 0) {
        if (value < 1 || value > 1024)
            return fail(c, GRAVIS_INVALID_ARG, "bf_waves_per_cu must be in [1, 1024]");
        c->bf_waves_per_cu = (int)value;
        return GRAVIS_OK;
    }
    return fail(c, GRAVIS_INVALID_ARG, "unknown tuning key '%s'", key);
}

gravis_status gravis_get_tuning(gravis_ctx* g, gravis_tuning* out) {
    if (!g || !out) return GRAVIS_INVALID_ARG;
    const Ctx* c = &g->c;
    out->bf_ipt = c->bf_ipt;
    out->bf_tile = c->bf_tile;
    out->bf_waves_per_cu = c->bf_waves_per_cu;
    out->bh_group_size = c->bh_group;
    out->bf_matrix = c->bf_matrix;
    out->best_ms = 0.0;
    return GRAVIS_OK;
}

gravis_status gravis_autotune(gravis_ctx* g, const gravis_solver* s, size_t n, const double* x,
                              const double* y, const double* z, const double* m,
                              gravis_tuning* out) {
    if (!g) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    GRAVIS_TRY(check_solver(c, s));
    if (n == 0) return fail(c, GRAVIS_INVALID_ARG, "autotune needs particles");
    if (any_null({x, y, z, m})) return fail(c, GRAVIS_INVALID_ARG, "null array");
    GRAVIS_HIP(c, hipSetDevice(c->device));
    double* base;
    GRAVIS_TRY(gravis::buf(c, "tune_state", 7 * n, &base));
    double *dx = base, *dy = base + n, *dz = base + 2 * n, *dm = base + 3 * n;
    GRAVIS_TRY(h2d(c, dx, x, n)); GRAVIS_TRY(h2d(c, dy, y, n)); GRAVIS_TRY(h2d(c, dz, z, n));
    GRAVIS_TRY(h2d(c, dm, m, n));
    gravis::FinishArgs fa;
    fa.mode = gravis::kAccumulate;
    fa.ax = base + 4 * n; fa.ay = base + 5 * n; fa.az = base + 6 * n;
    GRAVIS_HIP(c, hipMemsetAsync(fa.ax, 0, 3 * n * sizeof(double), c->stream));

    // one candidate: warm-up call, then the mean of `reps` timed calls
    auto time_it = [&](double* ms, int reps) -> gravis_status {
        GRAVIS_TRY(force_device(c, s, n, dx, dy, dz, dm, fa));
        GRAVIS_HIP(c, hipEventRecord(c->ev0, c->stream));
        for (int r = 0; r < reps; ++r) GRAVIS_TRY(force_device(c, s, n, dx, dy, dz, dm, fa));
        GRAVIS_HIP(c, hipEventRecord(c->ev1, c->stream));
        GRAVIS_HIP(c, hipEventSynchronize(c->ev1));
        float t = 0.f;
        GRAVIS_HIP(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
        *ms = (double)t / reps;
        return GRAVIS_OK;
    };
    const bool profiling = c->profiling;
    c->profiling = false;  // the timed calls must not touch the caller's kernel timers
    // the shapes in force before the sweep, restored if a candidate fails
    const int s_ipt = c->bf_ipt, s_tile = c->bf_tile, s_wpc = c->bf_waves_per_cu,
              s_matrix = c->bf_matrix;
    const uint32_t s_group = c->bh_group;
    double best = -1.0;
    gravis_status st = GRAVIS_OK;
    struct Shape { int ipt, tile, wpc, matrix; double ms; };
    // Barnes-Hut: among the shapes within 2 % of the fastest, the first
    // listed wins
    auto pick = [&](std::vector<Shape>& v) {
        double fastest = v[0].ms;
        for (const Shape& x : v) fastest = std::min(fastest, x.ms);
        for (const Shape& x : v)
            if (x.ms <= 1.02 * fastest) return x;
        return v[0];
    };
    // Brute force: every shape is timed once (two calls); the four fastest
    // are then timed again in three interleaved rounds of three calls each,
    // so a drift of the clock between them cancels, and the fastest total
    // wins. Shapes are listed with the fewest j chunks first (a finer split
    // balances better but writes 3 doubles per particle per chunk), and a
    // later one must be more than 0.5 % faster to be taken.
    auto set_shape = [&](const Shape& b) {
        c->bf_ipt = b.ipt; c->bf_tile = b.tile; c->bf_waves_per_cu = b.wpc;
        c->bf_matrix = b.matrix;
    };
    auto refine = [&](std::vector<Shape>& v, Shape* out) -> gravis_status {
        std::vector<size_t> idx(v.size());
        for (size_t i = 0; i < v.size(); ++i) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(),
                         [&](size_t a, size_t b) { return v[a].ms < v[b].ms; });
        idx.resize(std::min<size_t>(4, idx.size()));
        std::sort(idx.begin(), idx.end());  // back to the order of preference
        std::vector<double> tot(idx.size(), 0.0);
        for (int round = 0; round < 3; ++round)
            for (size_t k = 0; k < idx.size(); ++k) {
                set_shape(v[idx[k]]);
                double ms;
                GRAVIS_TRY(time_it(&ms, 3));
                tot[k] += ms / 3.0;
            }
        size_t b = 0;
        for (size_t k = 1; k < idx.size(); ++k)
            if (tot[k] < tot[b] / 1.005) b = k;
        *out = v[idx[b]];
        out->ms = tot[b];
        return GRAVIS_OK;
    };
    if (s->algorithm == GRAVIS_BRUTE_FORCE) {
        const bool mx = s->precision == GRAVIS_MIXED;
        std::vector<Shape> v;
        // kernel 0, 1: VALU kernel with bf_ipt 2, 1; kernel 2, 3: matrix core
        // with FP32 and split-bf16 operands
        // (768 measured fastest for the split-bf16 kernel at N = 262,144:
        // 15.90 against 16.30 ms at 512 and 16.03 at 1024, same box)
        for (int wpc : {32, 64, 128, 256, 512, 768, 1024})
            for (int kern : {0, 1, 2, 3})
                for (int tile : {128, 256}) {
                    if (!mx && (kern != 1 || tile != 256)) continue;  // FP64: split only
                    if (mx && wpc == 32) continue;
                    if (mx) {
                        c->bf_matrix = kern >= 2 ? kern - 1 : 0;
                        if (kern < 2) c->bf_ipt = 2 - kern;
                        c->bf_tile = tile;
                    }
                    c->bf_waves_per_cu = wpc;
                    double ms;
                    st = time_it(&ms, 2);
                    if (st != GRAVIS_OK) break;
                    v.push_back({c->bf_ipt, c->bf_tile, wpc, c->bf_matrix, ms});
                }
        Shape b{};
        if (st == GRAVIS_OK) st = refine(v, &b);
        if (st == GRAVIS_OK) {
            set_shape(b);
            best = b.ms;
        }
    } else if (s->algorithm == GRAVIS_BARNES_HUT && s->group_size == 0) {
        // 64 first (measured fastest over many C3-class steps), then 32 and
        // 16: another is kept only when it is more than 2 % faster. The walk
        // takes groups from a shared queue and a group whose frontier
        // overflows is walked one target at a time, so one call's time varies
        // with the order the groups are taken in (64-target groups: up to
        // 25 % call to call); four timed calls per size.
        std::vector<Shape> v;
        for (uint32_t gs : {64u, 32u, 16u}) {
            c->bh_group = gs;
            double ms;
            st = time_it(&ms, 4);
            if (st != GRAVIS_OK) break;
            v.push_back({(int)gs, 0, 0, 0, ms});
        }
        if (st == GRAVIS_OK) {
            const Shape b = pick(v);
            c->bh_group = (uint32_t)b.ipt;
            best = b.ms;
        }
    } else {
        st = time_it(&best, 2);
    }
    c->profiling = profiling;
    if (st != GRAVIS_OK) {
        c->bf_ipt = s_ipt; c->bf_tile = s_tile; c->bf_waves_per_cu = s_wpc;
        c->bf_matrix = s_matrix; c->bh_group = s_group;
        return st;
    }
    if (out) {
        GRAVIS_TRY(gravis_get_tuning(g, out));
        out->best_ms = best;
    }
    return GRAVIS_OK;
}

gravis_status gravis_bh_stats(gravis_ctx* g, uint64_t out[5]) {
    if (!g || !out) return GRAVIS_INVALID_ARG;
    for (int k = 0; k < 5; ++k) out[k] = g->c.bh_stats[k];
    return GRAVIS_OK;
}

gravis_status gravis_morton_sort(gravis_ctx* g, size_t n, const double* x, const double* y,
                                 const double* z, uint64_t* keys_sorted, uint32_t* perm,
                                 double* box_out) {
    if (!g) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    if (n == 0) return fail(c, GRAVIS_INVALID_ARG, "n must be > 0 (octree.rs:16)");
    if (any_null({x, y, z, keys_sorted, perm, box_out}))
        return fail(c, GRAVIS_INVALID_ARG, "null array");
    GRAVIS_HIP(c, hipSetDevice(c->device));
    HostState d;
    GRAVIS_TRY(stage(c, n, &d));
    GRAVIS_TRY(h2d(c, d.x, x, n)); GRAVIS_TRY(h2d(c, d.y, y, n)); GRAVIS_TRY(h2d(c, d.z, z, n));
    double* box;
    GRAVIS_TRY(gravis::buf(c, "ms_box", 4, &box));
    GRAVIS_TRY(gravis::root_box(c, n, d.x, d.y, d.z, box));
    uint64_t* k;
    uint32_t* p;
    GRAVIS_TRY(gravis::morton_sort(c, n, d.x, d.y, d.z, box, &k, &p));
    GRAVIS_HIP(c, hipMemcpyAsync(keys_sorted, k, n * 8, hipMemcpyDeviceToHost, c->stream));
    GRAVIS_HIP(c, hipMemcpyAsync(perm, p, n * 4, hipMemcpyDeviceToHost, c->stream));
    GRAVIS_HIP(c, hipMemcpyAsync(box_out, box, 32, hipMemcpyDeviceToHost, c->stream));
    GRAVIS_HIP(c, hipStreamSynchronize(c->stream));
    return GRAVIS_OK;
}

void gravis_sph_default_params(gravis_sph_params* o) {
    if (!o) return;
    o->gamma = 5.0 / 3.0;     // solver.rs:42
    o->eta = 1.2;             // solver.rs:43
    o->h_max = INFINITY;      // solver.rs:47
    o->cfl = 0.3;             // solver.rs:45
    o->force_factor = 0.25;   // solver.rs:46
    o->beta_visc = 2.0;       // forces.rs:34
    o->alpha_min = 0.1;       // forces.rs:35
    o->alpha_max = 2.0;       // forces.rs:36
    o->alpha_cond = 1.0;      // forces.rs:37
    o->sigma_visc = 0.1;      // forces.rs:38
    o->h_tol = 1e-3;          // density.rs:16
    o->kernel = 0;
    o->h_solver = 0;
    o->max_h_iter = 5;        // density.rs:13
    o->alpha_switch = 1;
    o->pair_fp32 = 0;
    o->soft_h = 0;
}

void gravis_sph_variant_params(gravis_sph_params* o) {
    if (!o) return;
    gravis_sph_default_params(o);
    o->kernel = 1;
    o->h_solver = 1;
    o->h_tol = 1e-4;
    o->max_h_iter = 30;
    o->alpha_switch = 0;
    o->alpha_cond = 0.0;
    o->soft_h = 1;
    o->pair_fp32 = 1;
}

gravis_status gravis_sph_compute_device(gravis_ctx* g, const gravis_sph_params* P,
                                        const gravis_sph_state* S, double* div_v,
                                        double* max_mu, double* dalpha_dt, double* omega,
                                        uint32_t* nneigh) {
    if (!g) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    GRAVIS_TRY(check_sph(c, P, S));
    if (S->n == 0) return GRAVIS_OK;
    if (any_null({div_v, max_mu, dalpha_dt}))
        return fail(c, GRAVIS_INVALID_ARG, "null intermediate array");
    GRAVIS_HIP(c, hipSetDevice(c->device));
    return gravis::sph_compute(c, P, S, nullptr, div_v, max_mu, dalpha_dt, omega, nneigh);
}

gravis_status gravis_sph_compute(gravis_ctx* g, const gravis_sph_params* P,
                                 const gravis_sph_state* S, double* div_v, double* max_mu,
                                 double* dalpha_dt, double* omega, uint32_t* nneigh) {
    if (!g) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    GRAVIS_TRY(check_sph(c, P, S));
    const size_t n = S->n;
    if (n == 0) return GRAVIS_OK;
    if (any_null({div_v, max_mu, dalpha_dt}))
        return fail(c, GRAVIS_INVALID_ARG, "null intermediate array");
    GRAVIS_HIP(c, hipSetDevice(c->device));
    gravis_sph_state D;
    GRAVIS_TRY(sph_stage_in(c, S, &D));
    double* iv;
    uint32_t* nn;
    GRAVIS_TRY(gravis::buf(c, "sph_host_interm", 4 * n, &iv));
    GRAVIS_TRY(gravis::buf(c, "sph_host_nneigh", n, &nn));
    GRAVIS_TRY(gravis::sph_compute(c, P, &D, nullptr, iv, iv + n, iv + 2 * n,
                                   omega ? iv + 3 * n : nullptr, nneigh ? nn : nullptr));
    gravis_sph_state H = *S;
    GRAVIS_TRY(sph_stage_out(c, &D, &H, false));
    GRAVIS_TRY(d2h(c, div_v, iv, n));
    GRAVIS_TRY(d2h(c, max_mu, iv + n, n));
    GRAVIS_TRY(d2h(c, dalpha_dt, iv + 2 * n, n));
    if (omega) GRAVIS_TRY(d2h(c, omega, iv + 3 * n, n));
    if (nneigh)
        GRAVIS_HIP(c, hipMemcpyAsync(nneigh, nn, n * 4, hipMemcpyDeviceToHost, c->stream));
    GRAVIS_HIP(c, hipStreamSynchronize(c->stream));
    return GRAVIS_OK;
}

static gravis_status sph_run_checked(Ctx* c, const gravis_sph_params* P, const gravis_solver* s,
                                     const gravis_sph_state* S, double dt) {
    GRAVIS_TRY(check_sph(c, P, S));
    if (s && s->algorithm != GRAVIS_NO_GRAVITY) GRAVIS_TRY(check_solver(c, s));
    if (!(dt > 0.0)) return fail(c, GRAVIS_INVALID_ARG, "dt must be positive, got %g", dt);
    return GRAVIS_OK;
}

gravis_status gravis_sph_run_device(gravis_ctx* g, const gravis_sph_params* P,
                                    const gravis_solver* s, gravis_sph_state* S, uint64_t nsteps,
                                    double* dt_io, double* time_io) {
    if (!g || !dt_io) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    GRAVIS_TRY(sph_run_checked(c, P, s, S, *dt_io));
    if (S->n == 0 || nsteps == 0) return GRAVIS_OK;
    GRAVIS_HIP(c, hipSetDevice(c->device));
    return gravis::sph_run(c, P, s, S, nsteps, dt_io, time_io, nullptr);
}

gravis_status gravis_sph_run(gravis_ctx* g, const gravis_sph_params* P, const gravis_solver* s,
                             gravis_sph_state* S, uint64_t nsteps, double* dt_io,
                             double* time_io) {
    if (!g || !dt_io) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    GRAVIS_TRY(sph_run_checked(c, P, s, S, *dt_io));
    if (S->n == 0 || nsteps == 0) return GRAVIS_OK;
    GRAVIS_HIP(c, hipSetDevice(c->device));
    gravis_sph_state D;
    GRAVIS_TRY(sph_stage_in(c, S, &D));
    GRAVIS_TRY(gravis::sph_run(c, P, s, &D, nsteps, dt_io, time_io, nullptr));
    GRAVIS_TRY(sph_stage_out(c, &D, S, true));
    GRAVIS_HIP(c, hipStreamSynchronize(c->stream));
    return GRAVIS_OK;
}

gravis_status gravis_sph_step(gravis_ctx* g, const gravis_sph_params* P, const gravis_solver* s,
                              gravis_sph_state* S, double dt, double* dt_next) {
    if (!g || !dt_next) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    GRAVIS_TRY(sph_run_checked(c, P, s, S, dt));
    if (S->n == 0) {
        *dt_next = INFINITY;  // no gas particle: dt_min stays +inf (solver.rs:96,130)
        return GRAVIS_OK;
    }
    GRAVIS_HIP(c, hipSetDevice(c->device));
    gravis_sph_state D;
    GRAVIS_TRY(sph_stage_in(c, S, &D));
    double dtc = dt;
    GRAVIS_TRY(gravis::sph_run(c, P, s, &D, 1, &dtc, nullptr, dt_next));
    GRAVIS_TRY(sph_stage_out(c, &D, S, true));
    GRAVIS_HIP(c, hipStreamSynchronize(c->stream));
    return GRAVIS_OK;
}

gravis_status gravis_sph_accel_device(gravis_ctx* g, const gravis_sph_params* P,
                                      const gravis_solver* s, const gravis_sph_state* S) {
    if (!g) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    GRAVIS_TRY(check_sph(c, P, S));
    if (S->n == 0 || !s || s->algorithm == GRAVIS_NO_GRAVITY) return GRAVIS_OK;
    GRAVIS_TRY(check_solver(c, s));
    GRAVIS_HIP(c, hipSetDevice(c->device));
    gravis::Octree t;
    return gravis::sph_gravity(c, P, s, S, &t, nullptr);
}

gravis_status gravis_sph_accel(gravis_ctx* g, const gravis_sph_params* P, const gravis_solver* s,
                               gravis_sph_state* S) {
    if (!g) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    GRAVIS_TRY(check_sph(c, P, S));
    if (S->n == 0 || !s || s->algorithm == GRAVIS_NO_GRAVITY) return GRAVIS_OK;
    GRAVIS_TRY(check_solver(c, s));
    GRAVIS_HIP(c, hipSetDevice(c->device));
    gravis_sph_state D;
    GRAVIS_TRY(sph_stage_in(c, S, &D));
    gravis::Octree t;
    GRAVIS_TRY(gravis::sph_gravity(c, P, s, &D, &t, nullptr));
    GRAVIS_TRY(d2h(c, S->ax, D.ax, S->n));
    GRAVIS_TRY(d2h(c, S->ay, D.ay, S->n));
    GRAVIS_TRY(d2h(c, S->az, D.az, S->n));
    GRAVIS_HIP(c, hipStreamSynchronize(c->stream));
    return GRAVIS_OK;
}

gravis_status gravis_sph_stats(gravis_ctx* g, uint64_t out[4]) {
    if (!g || !out) return GRAVIS_INVALID_ARG;
    for (int k = 0; k < 4; ++k) out[k] = g->c.sph_stats[k];
    return GRAVIS_OK;
}

}  // extern "C"
