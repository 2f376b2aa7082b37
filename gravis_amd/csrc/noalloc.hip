// noalloc.hip -- the brute-force force call for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc). It runs the same kernels as gravis_accel_device,
// with every scratch buffer (gathered records, Morton-sort buffers, j-chunk
// partials, tile boxes) carved out of one static device array instead of the
// context's allocated buffers.
//
// Limits: n <= kNoAllocMaxN particles and at most kNoAllocChunks j chunks,
// which fit the array (at most 248 bytes per particle with the Morton-sort
// buffers of the mixed path, plus small per-call tables).
//
// Each GPU holds one copy of the array until the process ends, shared by
// every context and call on that GPU: calls that overlap on one GPU (two
// host threads, or two streams) race on it. A call leaves its contents
// behind; nothing in these kernels reads a range before writing it in the
// same call.
#include <mutex>

#include "internal.h"

namespace {

constexpr size_t kNoAllocMaxN = 65536;
constexpr int kNoAllocChunks = 8;
constexpr size_t kScratchBytes = 20u << 20;  // 20 MiB

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
bool g_looked[kMaxDevices];
void* g_ptr[kMaxDevices];
hipError_t g_err[kMaxDevices];

}  // namespace

namespace gravis {

// Called by gravis_create with c->device current: the device address of this
// device's copy of `scratch`, looked up once per device. The array's name is
// a host-side stand-in; only this address may reach kernels or copies.
void noalloc_setup(Ctx* c) {
    const int d = c->device;
    if (d < 0 || d >= kMaxDevices) {
        c->scratch = nullptr;
        c->scratch_err = hipErrorInvalidDevice;
        return;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_looked[d]) {
        void* p = nullptr;
        g_err[d] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (g_err[d] != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
        }
        g_ptr[d] = p;
        g_looked[d] = true;
    }
    c->scratch = g_ptr[d];
    c->scratch_err = g_err[d];
}

}  // namespace gravis

using gravis::Ctx;
using gravis::fail;

extern "C" gravis_status gravis_accel_device_noalloc(gravis_ctx* g, const gravis_solver* s,
                                                     size_t n, const double* x,
                                                     const double* y, const double* z,
                                                     const double* m, double* ax, double* ay,
                                                     double* az) {
    if (!g) return GRAVIS_INVALID_ARG;
    Ctx* c = &g->c;
    if (!s) return fail(c, GRAVIS_INVALID_ARG, "solver is NULL");
    if (!(s->softening >= 0.0))
        return fail(c, GRAVIS_INVALID_ARG, "softening must be non-negative, got %g", s->softening);
    if (s->precision != GRAVIS_FP64 && s->precision != GRAVIS_MIXED)
        return fail(c, GRAVIS_INVALID_ARG, "unknown precision %d", s->precision);
    if (s->algorithm == GRAVIS_NO_GRAVITY) return GRAVIS_OK;  // accumulates nothing
    if (s->algorithm != GRAVIS_BRUTE_FORCE)
        return fail(c, GRAVIS_UNSUPPORTED,
                    "gravis_accel_device_noalloc runs brute force only (algorithm %d)",
                    s->algorithm);
    if (n == 0) return GRAVIS_OK;
    if (n > kNoAllocMaxN)
        return fail(c, GRAVIS_INVALID_ARG,
                    "gravis_accel_device_noalloc: n = %zu exceeds its limit of %zu particles", n,
                    kNoAllocMaxN);
    if (!x || !y || !z || !m || !ax || !ay || !az)
        return fail(c, GRAVIS_INVALID_ARG, "null array");
    if (!c->scratch)
        return fail(c, GRAVIS_HIP_ERROR, "no-allocation scratch address: %s",
                    hipGetErrorString(c->scratch_err));
    GRAVIS_HIP(c, hipSetDevice(c->device));
    gravis::FinishArgs fa;
    fa.mode = gravis::kAccumulate;
    fa.ax = ax;
    fa.ay = ay;
    fa.az = az;
    const bool profiling = c->profiling;
    c->profiling = false;  // the kernel timers keep host tables: off for this call
    c->arena = static_cast<char*>(c->scratch);
    c->arena_cap = kScratchBytes;
    c->arena_used = 0;
    c->arena_n = 0;
    c->arena_max_chunks = kNoAllocChunks;
    const gravis_status st = gravis::bf_force(c, s, n, x, y, z, m, fa);
    c->arena = nullptr;
    c->arena_max_chunks = 0;
    c->profiling = profiling;
    return st;
}
