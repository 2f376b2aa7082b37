// internal.h -- shared definitions of the gfx950 engine (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/gravis.h"

namespace gravis {

constexpr int kWave = 64;  // CDNA wavefront

// Device buffer that only grows; freed with the context.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct KernelTimer {
    double ms = 0.0;
    uint64_t launches = 0;
};

// One Barnes-Hut tree arena (device), kept between calls as sim-core keeps
// its Octree in BarnesHut::tree_cache (barnes_hut.rs:27,146-156).
struct TreeArena {
    size_t node_cap = 0;
};

struct Ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    std::map<std::string, DevBuf> bufs;
    bool profiling = false;
    std::map<std::string, KernelTimer> timers;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int num_cus = 256;
    // last Barnes-Hut call: node tests, accepted nodes, leaf particles,
    // pair interactions, tree nodes
    unsigned long long bh_stats[5] = {0, 0, 0, 0, 0};
    // Barnes-Hut walk: breadth-first frontier capacity per wave (nodes)
    uint32_t bh_frontier_cap = 16384;
    // last SPH compute: groups searched depth first, re-queries, query
    // rounds, neighbour-pool entries used
    unsigned long long sph_stats[4] = {0, 0, 0, 0};
    // SPH neighbour search: overlapping leaves one wave may hold in LDS
    // before its group is searched depth first (at most 512)
    uint32_t sph_range_cap = 512;
    // initial neighbour-pool capacity in entries per particle (the pool
    // grows past it, keeping what is written, when a query round runs out)
    uint32_t sph_pool_init = 96;
    // FP64 brute-force KDK of at most this many particles runs fused in one
    // workgroup (0 disables; at most kSmallN)
    uint32_t kdk_fused_max_n = 256;
    // ... and of more than that, up to this many (several steps, stream not
    // being captured) in one cooperative launch (0 disables; at most
    // kCoopMaxN)
    uint32_t kdk_coop_max_n = 2040;
    // Brute-force launch shape (gravis_set_tuning / gravis_autotune):
    // targets per lane of the mixed kernel (wavefront groups of 64 or 128),
    // j records per staged tile (128 or 256), and the wavefronts per CU the
    // j-chunk split aims for.
    int bf_ipt = 2;
    int bf_tile = 128;
    // 1, 2: the mixed brute force runs the matrix-core kernel (64 targets
    // per wavefront; r^2 on v_mfma_f32_16x16x4f32, or on
    // v_mfma_f32_16x16x32_bf16 with FP32 values split in three bf16 pieces);
    // bf_ipt is then unused
    int bf_matrix = 0;
    int bf_waves_per_cu = 256;
    // Barnes-Hut target group when the solver leaves group_size at 0
    uint32_t bh_group = 32;
    // Mixed brute force: a source tile closer to the target box than this
    // fraction q of the box diagonal gets FP64 pair separations, unless the
    // softening exceeds q times the diagonal.
    double bf_near_fraction = 1.0 / 64.0;
    // No-allocation mode (gravis_accel_device_noalloc): while `arena` is
    // set, buffer() hands out 256-byte-aligned ranges of it by name (the
    // same name gives the same range within one call) instead of
    // allocating, and the brute-force j split is capped at
    // arena_max_chunks chunks.
    struct ArenaEntry {
        const char* name;
        size_t off, bytes;
    };
    static constexpr int kArenaEntries = 32;
    char* arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
    int arena_n = 0;
    ArenaEntry arena_tab[kArenaEntries];
    int arena_max_chunks = 0;
    // this device's copy of the library's static scratch array (set by
    // gravis_create; null with scratch_err set if its address was not found)
    void* scratch = nullptr;
    hipError_t scratch_err = hipSuccess;
};

// Looks up (gravis_create) this device's address of the static scratch
// array of the no-allocation entry point.
void noalloc_setup(Ctx* c);

// Status helpers ----------------------------------------------------------
gravis_status fail(Ctx* c, gravis_status s, const char* fmt, ...);

#define GRAVIS_HIP(ctx, call)                                                             \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return ::gravis::fail((ctx), GRAVIS_HIP_ERROR, "%s failed: %s (%s:%d)", #call, \
                                  hipGetErrorString(e_), __FILE__, __LINE__);             \
    } while (0)

#define GRAVIS_TRY(expr)                  \
    do {                                  \
        gravis_status s_ = (expr);        \
        if (s_ != GRAVIS_OK) return s_;   \
    } while (0)

// Returns a device buffer of at least `bytes` under `name` (grow only).
gravis_status buffer(Ctx* c, const char* name, size_t bytes, void** out);

template <typename T>
inline gravis_status buf(Ctx* c, const char* name, size_t count, T** out) {
    void* p = nullptr;
    gravis_status s = buffer(c, name, count * sizeof(T), &p);
    *out = static_cast<T*>(p);
    return s;
}

// Kernel launch check (launch errors are reported by hipGetLastError).
gravis_status check_launch(Ctx* c, const char* what);

// A wave-uniform double moved to scalar registers (frees two VGPRs).
__device__ __forceinline__ double uniform(double v) {
    const uint64_t b = __double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Workgroup -> work-range map for gfx950's 8 XCDs. The dispatcher places
// workgroup b on XCD b % 8, each XCD with its own L2; this bijection gives
// XCD x the contiguous range of work indices [start_x, start_x + count_x),
// so kernels whose neighbouring work items read overlapping data (Morton-
// ordered particles and their neighbour lists) reuse it in one L2 instead
// of pulling the same lines into all eight.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
    const uint32_t per = nb / kXcds, rem = nb % kXcds;
    const uint32_t x = b % kXcds, k = b / kXcds;
    return x * per + (x < rem ? x : rem) + k;
}

// Dynamic work distribution with the same XCD affinity: kXcds counters,
// counter x handing out the work indices of range x of xcd_block's split
// (so Morton-consecutive work stays in one L2). A wavefront drains the
// counter of its workgroup's XCD (home = blockIdx.x % kXcds) first, then
// moves round the others; each index is handed out exactly once, and the
// call returns ~0u once every counter is drained. cx carries the counter
// being drained between calls (start it at home). Every lane of the
// wavefront calls it; the result is wavefront-uniform.
__device__ __forceinline__ uint32_t xcd_take(uint32_t* __restrict__ counters, uint32_t n,
                                             uint32_t home, uint32_t& cx, int lane) {
    const uint32_t per = n / kXcds, rem = n % kXcds;
    for (;;) {
        const uint32_t k = __builtin_amdgcn_readfirstlane(
            atomicAdd(&counters[cx], lane == 0 ? 1u : 0u));
        if (k < per + (cx < rem ? 1u : 0u)) return cx * per + (cx < rem ? cx : rem) + k;
        cx = (cx + 1) % kXcds;
        if (cx == home) return ~0u;
    }
}

// A wavefront-uniform float held in a vector register. On gfx950 a VALU
// operation with an SGPR operand issues at half the FP32 rate of one whose
// operands are all VGPRs or literals (measured: 57 against 103 lane
// operations per cycle per CU), so loop-invariant scalars that feed the pair
// loops are moved to VGPRs once. The empty asm emits no instruction; it only
// hides the value's uniformity from the compiler.
__device__ __forceinline__ float in_vgpr(float v) {
    asm("" : "+v"(v));
    return v;
}
__device__ __forceinline__ double in_vgpr(double v) {
    asm("" : "+v"(v));
    return v;
}

// ---- pair-force helpers shared by the brute-force kernels ----------------
// 1/sqrt(x) for finite x > 0: v_rsq_f64 (relative error about 2^-24 on
// gfx950, measured) refined by one third-order step,
// y' = y (1 + e/2 + 3e^2/8) with e = 1 - x y^2, whose truncation error
// (~e^3) is far below an ulp. Five FP64 operations and the transcendental,
// against about nine for the library rsqrt, whose infinity and zero
// handling a pair never needs (r^2 = 0 is masked by the caller).
__device__ __forceinline__ double rsqrt_refined(double x) {
    const double r = __builtin_amdgcn_rsq(x);
    const double h = x * r;
    const double e = __builtin_fma(-h, r, 1.0);
    const double p = __builtin_fma(e, 0.375, 0.5);
    return __builtin_fma(r * e, p, r);
}

// m r^-3 for finite r^2 > 0 in FP64: v_rsq_f64 and one third-order
// refinement of r^-3 itself, m y^3 (1 + 3e/2 + 15e^2/8) with e = 1 - r^2 y^2.
// Seven operations; measured over 4.2 M values (r^2 in 2^-30..2^30) at
// 2.2 ulp max and 0.44 ulp mean, against 5.1 and 0.96 for the refined r^-1
// cubed (eight operations). r^2 = 0 gives NaN, for the caller to mask.
// The constants are VGPR values made once before the pair loop
// (rinv3_consts): FP64 FMA takes no literal, and a constant addend would be
// copied into the destination of v_fmac_f64 on every pair.
struct Rinv3Consts {
    double one, c15, c1875;
};
__device__ __forceinline__ Rinv3Consts rinv3_consts() {
    return {in_vgpr(1.0), in_vgpr(1.5), in_vgpr(1.875)};
}
__device__ __forceinline__ double m_rinv3(double r2, double m, const Rinv3Consts& k) {
    const double y = __builtin_amdgcn_rsq(r2);
    const double y2 = y * y;
    const double e = __builtin_fma(-r2, y2, k.one);
    const double p = __builtin_fma(e, k.c1875, k.c15);
    const double z = (y2 * y) * m;
    return __builtin_fma(z * e, p, z);
}

// Softening below which the FP32 pair loops mask pairs of zero separation
// (the self pair, coincident points). At or above it the self pair's
// m eps^-3 <= 1e18 m stays finite in FP32 for any mass below 3e20 and adds
// m eps^-3 * 0 = 0, so the unmasked loops need no test.
constexpr float kMaskEps2F = 1e-12f;
// The same for the FP64 pair loops: m eps^-3 <= 1e150 m.
constexpr double kMaskEps2D = 1e-100;

// Pair force factor in FP32: m * r^-3 with r^2 = d.d + eps^2. MASK gives 0
// for d = 0 (whatever eps is).
template <bool MASK>
__device__ __forceinline__ float pair_f32(float dx, float dy, float dz, float m, float eps2) {
    float ri;
    if (MASK) {
        const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        ri = __builtin_amdgcn_rsqf(d2 + eps2);
        ri = d2 > 0.0f ? ri : 0.0f;
    } else {
        const float r2 =
            __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
        ri = __builtin_amdgcn_rsqf(r2);
    }
    const float mr = m * ri;
    const float ri2 = ri * ri;
    return mr * ri2;
}

// Record j of a tile (j < end), else pad. The load reads record
// min(j, end - 1) (end > 0) and the select is on values: a select on the
// address would place pad in scratch and make the load a flat load.
__device__ __forceinline__ double4 record_or(const double4* __restrict__ rec, int j, int end,
                                             const double4 pad) {
    const double4 r = rec[min(j, end - 1)];
    return j < end ? r : pad;
}

__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// Bounding box of the wavefront's IPT*64 targets (FP64); every lane ends
// with the same box, held in scalar registers.
template <int IPT>
__device__ __forceinline__ void wave_box(const double (&x)[IPT], const double (&y)[IPT],
                                         const double (&z)[IPT], double (&lo)[3],
                                         double (&hi)[3]) {
    lo[0] = hi[0] = x[0];
    lo[1] = hi[1] = y[0];
    lo[2] = hi[2] = z[0];
#pragma unroll
    for (int k = 1; k < IPT; ++k) {
        lo[0] = fmin(lo[0], x[k]); hi[0] = fmax(hi[0], x[k]);
        lo[1] = fmin(lo[1], y[k]); hi[1] = fmax(hi[1], y[k]);
        lo[2] = fmin(lo[2], z[k]); hi[2] = fmax(hi[2], z[k]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            lo[d] = fmin(lo[d], __shfl_xor(lo[d], off));
            hi[d] = fmax(hi[d], __shfl_xor(hi[d], off));
        }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        lo[d] = uniform(lo[d]);
        hi[d] = uniform(hi[d]);
    }
}

// True when the gap between a source box tb = (lo xyz, hi xyz) and the
// target box (lo, hi) is below sqrt(lim2).
__device__ __forceinline__ bool tile_is_near(const double* __restrict__ tb, const double (&lo)[3],
                                             const double (&hi)[3], double lim2) {
    double gap2 = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double g = fmax(0.0, fmax(tb[d] - hi[d], lo[d] - tb[3 + d]));
        gap2 = __builtin_fma(g, g, gap2);
    }
    return gap2 < lim2;
}

// Inclusive scan of one u32 per lane over the wavefront, in six DPP adds
// (no LDS round trips): row_shr 1, 2, 4, 8 scan each row of 16 lanes
// (sources outside the row read 0), then row_bcast 15 adds row 0's total to
// row 1 and row 2's to row 3, and row_bcast 31 adds lanes 0..31's total to
// rows 2 and 3. Every lane of the wavefront must be active.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
    x += __builtin_amdgcn_update_dpp(0u, x, 0x111, 0xF, 0xF, true);   // row_shr:1
    x += __builtin_amdgcn_update_dpp(0u, x, 0x112, 0xF, 0xF, true);   // row_shr:2
    x += __builtin_amdgcn_update_dpp(0u, x, 0x114, 0xF, 0xF, true);   // row_shr:4
    x += __builtin_amdgcn_update_dpp(0u, x, 0x118, 0xF, 0xF, true);   // row_shr:8
    x += __builtin_amdgcn_update_dpp(0u, x, 0x142, 0xA, 0xF, false);  // row_bcast:15
    x += __builtin_amdgcn_update_dpp(0u, x, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return x;
}

// Exclusive scan of one u32 per thread over a block of BLOCK threads (a
// multiple of the wavefront); *total receives the block sum. Every thread of
// the block must call it (it synchronises twice).
template <int BLOCK>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[BLOCK / kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    const uint32_t inc = wave_inclusive_scan(v);
    if (lane == kWave - 1) wsum[w] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int k = 0; k < BLOCK / kWave; ++k) {
        const uint32_t x = wsum[k];
        before += k < w ? x : 0u;
        all += x;
    }
    __syncthreads();  // wsum is reused by the next call
    *total = all;
    return before + inc - v;
}

// Optional event bracket around the dominant kernel of a call.
void timer_begin(Ctx* c);
void timer_end(Ctx* c, const char* name);

// ---- spatial sort (sort.hip) --------------------------------------------
// Root cube (octree.rs:14-44) of n positions, written to d_box[4] =
// (cx, cy, cz, half_width) on the device.
gravis_status root_box(Ctx* c, size_t n, const double* x, const double* y, const double* z,
                       double* d_box);
// Morton keys over d_box (3 * levels bits: the top `levels` of the 21
// levels per axis; 21 gives the octree's 63-bit keys) and the stable sort
// permutation. Outputs: d_keys_sorted[n], d_perm[n] (sorted position ->
// original index).
gravis_status morton_sort(Ctx* c, size_t n, const double* x, const double* y, const double* z,
                          const double* d_box, uint64_t** d_keys_sorted, uint32_t** d_perm,
                          int levels = 21);
// Stable LSD radix sort of n 64-bit keys with u32 values over the low
// `bits` bits. The caller fills the buffers of sort_buffers first; the
// sorted keys and values are returned (either set of ping-pong buffers).
gravis_status sort_buffers(Ctx* c, size_t n, uint64_t** keys, uint32_t** vals);
gravis_status radix_sort(Ctx* c, size_t n, int bits, uint64_t** d_keys_sorted,
                         uint32_t** d_vals_sorted);
// Spreads the low 21 bits of v to every third bit (Morton interleave).
__device__ __forceinline__ uint64_t morton_spread3(uint32_t v) {
    uint64_t x = v & 0x1FFFFFu;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}
// x,y,z,m as double4 records in d_perm order (d_perm: sorted -> caller
// index; identity if null). With a permutation the inverse (caller ->
// sorted) is formed first and the records are scattered through it (reads
// in the caller's order, one 32-byte write per particle, instead of four
// line-sized random reads); *inv_out (if given) receives it, or null.
gravis_status gather_sorted(Ctx* c, size_t n, const double* x, const double* y,
                            const double* z, const double* m, const uint32_t* d_perm,
                            double4* d_out, const uint32_t** inv_out = nullptr);

// ---- linear octree (tree.hip) -------------------------------------------
constexpr int kTreeMaxLevel = 21;  // 63-bit keys: 21 levels below the root
constexpr int kTreeBlock = 256;
constexpr int kTreeGrid = 1024;
// meta.w = level | nchild << 8 | kNodeGroupRoot (the node holds more
// particles than a target group).
constexpr uint32_t kNodeGroupRoot = 1u << 16;

// Device view of one built tree. Nodes are numbered breadth first; the
// children of a node are contiguous and in Morton digit order; every node is
// the run [first, first + count) of the key-sorted particles.
struct Octree {
    size_t n = 0;                    // particles
    size_t cap = 0;                  // node capacity
    uint32_t leaf = 16, group = 64;  // leaf capacity, target group size
    const double* box = nullptr;     // root cube (cx, cy, cz, half_width)
    const uint64_t* keys = nullptr;  // sorted Morton keys
    const uint32_t* perm = nullptr;  // sorted position -> caller index
    const uint32_t* inv = nullptr;   // caller index -> sorted position
    double4* src = nullptr;          // [n sorted (x, y, z, m) | cap node (com, mass)]
    const uint4* meta = nullptr;     // {first, count, first child, flags}
    const double4* bmin = nullptr;   // particle bounding box of each node
    const double4* bmax = nullptr;
    const uint32_t* lvl = nullptr;   // level ranges; lvl[2 (kTreeMaxLevel + 1)] = node count
    uint32_t* flags = nullptr;       // [0] walk frontier overflow, [1] node capacity exceeded
    const uint2* groups = nullptr;   // target groups {first, count}
    uint32_t* ngroups = nullptr;     // number of groups (device)
};

// Root cube, Morton sort and level-by-level build of the tree of n
// particles (device SoA), asynchronous on the context stream. The buffers
// are the context's and are reused by the next build.
// split_siblings: target groups cut from each run of consecutive children
// of at most `group` particles (else whole children packed while they fit).
gravis_status build_octree(Ctx* c, size_t n, const double* x, const double* y, const double* z,
                           const double* m, uint32_t leaf, uint32_t group, Octree* t,
                           bool split_siblings = false);

// ---- forces (bf.hip, bh.hip) --------------------------------------------
// How the finishing kernel writes the result in the caller's order.
enum FinishMode : int {
    kAccumulate = 0,  // a += F  (GravitySolver contract)
    kAssignKick = 1,  // a = F; v += a * half_dt  (second half-kick of KDK)
};

struct FinishArgs {
    int mode = kAccumulate;
    double* ax = nullptr;
    double* ay = nullptr;
    double* az = nullptr;
    double* vx = nullptr;
    double* vy = nullptr;
    double* vz = nullptr;
    double half_dt = 0.0;
};

// Sums nchunks partial force arrays part[q][3][n] (sorted order, chunk order
// fixed) and delivers them in the caller's order (perm: sorted -> caller
// index, identity if null; inv: its inverse if already formed) per
// FinishArgs.
gravis_status finish_forces(Ctx* c, const double* part, int nchunks, int n,
                            const uint32_t* perm, const FinishArgs& fa,
                            const uint32_t* inv = nullptr);
// The same for one partial array of (ax, ay, az, 0) records in sorted order
// (one line read per particle), delivered through inv (sorted position of
// each caller index).
gravis_status finish_forces4(Ctx* c, const double4* part4, int n, const uint32_t* inv,
                             const FinishArgs& fa);

// Brute-force force on device SoA arrays, result delivered per FinishArgs.
// Particles up to which a FP64 brute-force KDK runs all its steps in one
// workgroup launch (kdk_small, bf.hip).
constexpr int kSmallN = 256;
gravis_status kdk_small(Ctx* c, size_t n, double* x, double* y, double* z, double* vx,
                        double* vy, double* vz, const double* m, double* ax, double* ay,
                        double* az, double softening, double dt, uint64_t nsteps);
// Particles up to which a FP64 brute-force KDK of several steps runs all of
// them in one cooperative launch with a grid barrier per step (kdk_coop,
// bf.hip; the step's records are held in LDS, 32 B each, with the
// kernel's 4-byte flag within 64 KB). The call synchronises the stream to
// read the barrier's abort word. *nsteps_left counts down the steps done;
// if the runtime refuses the cooperative launch it returns GRAVIS_OK with
// steps left, for the caller to run launch by launch.
constexpr int kCoopMaxN = 2040;
gravis_status kdk_coop(Ctx* c, size_t n, double* x, double* y, double* z, double* vx,
                       double* vy, double* vz, const double* m, double* ax, double* ay,
                       double* az, double softening, double dt, uint64_t* nsteps_left);

gravis_status bf_force(Ctx* c, const gravis_solver* s, size_t n, const double* x,
                       const double* y, const double* z, const double* m, const FinishArgs& fa);

// Barnes-Hut force (linear octree, group walk).
gravis_status bh_force(Ctx* c, const gravis_solver* s, size_t n, const double* x,
                       const double* y, const double* z, const double* m, const FinishArgs& fa);
// The walk alone, on a tree built by build_octree (synchronises the stream
// to check the capacity flags). eps2_sorted (tree order, may be null): a
// squared Plummer softening per target instead of the solver's.
gravis_status bh_walk(Ctx* c, const gravis_solver* s, const Octree& t, const FinishArgs& fa,
                      const double* eps2_sorted = nullptr);

// ---- smoothed particle hydrodynamics (sph.hip), device pointers ---------
// SphSolver::compute on device state S; `tree` (built on S's positions with
// target groups of 64) is reused when given. div_v/max_mu/dalpha: n each.
gravis_status sph_compute(Ctx* c, const gravis_sph_params* P, const gravis_sph_state* S,
                          const Octree* tree, double* div_v, double* max_mu, double* dalpha,
                          double* omega, uint32_t* nneigh);
// Gravity for an SPH state (accumulates into S's a): the solver's softening,
// or eps_i = h_i when P->soft_h (Barnes-Hut only). With Barnes-Hut the tree
// is built into *t and, if shared is not null and the groups have 64
// targets, *shared = t for the neighbour search.
gravis_status sph_gravity(Ctx* c, const gravis_sph_params* P, const gravis_solver* g,
                          const gravis_sph_state* S, Octree* t, const Octree** shared);
// nsteps of step_with_sph with headless's dt rule; *dt_io in/out, time_io
// and dt_next_out may be null.
gravis_status sph_run(Ctx* c, const gravis_sph_params* P, const gravis_solver* g,
                      const gravis_sph_state* S, uint64_t nsteps, double* dt_io,
                      double* time_io, double* dt_next_out);

// ---- integrator / diagnostics -------------------------------------------
gravis_status kick_drift(Ctx* c, size_t n, double* x, double* y, double* z, double* vx,
                         double* vy, double* vz, const double* ax, const double* ay,
                         const double* az, double half_dt, double dt);

// GpuParticle records (32 B: f32 position, mass, colour) of n particles
// into out (render.hip).
gravis_status render_pack(Ctx* c, size_t n, const double* x, const double* y, const double* z,
                          const double* m, const double* u, const uint8_t* type, void* out);

gravis_status diag_device(Ctx* c, size_t n, double eps, int with_potential, const double* x,
                          const double* y, const double* z, const double* vx, const double* vy,
                          const double* vz, const double* m, const double* u,
                          const uint8_t* ptype, double out[19]);

// ---- ensemble (ensemble.hip), device pointers ---------------------------
gravis_status ensemble_kdk(Ctx* c, const gravis_solver* s, int nsys, const uint64_t* offsets,
                           double* x, double* y, double* z, double* vx, double* vy, double* vz,
                           const double* m, double* ax, double* ay, double* az, double dt,
                           uint64_t nsteps, int use_graph);
gravis_status ensemble_accel(Ctx* c, const gravis_solver* s, int nsys, const uint64_t* offsets,
                             const double* x, const double* y, const double* z,
                             const double* m, double* ax, double* ay, double* az);
gravis_status ensemble_adjoint(Ctx* c, const gravis_solver* s, int nsys, const uint64_t* offsets,
                               const double* x0, const double* y0, const double* z0,
                               const double* vx0, const double* vy0, const double* vz0,
                               const double* m, double dt, uint64_t nsteps, uint32_t ckpt,
                               const double* tx, const double* ty, const double* tz,
                               const double* gx, const double* gy, const double* gz,
                               const double* gvx, const double* gvy, const double* gvz,
                               double* loss, double* xT, double* ox, double* oy, double* oz,
                               double* ovx, double* ovy, double* ovz, double* om);

inline unsigned grid_for(size_t n, unsigned block, unsigned cap = 1u << 30) {
    size_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return static_cast<unsigned>(g);
}

}  // namespace gravis

// The opaque context of the C ABI.
struct gravis_ctx {
    gravis::Ctx c;
};
