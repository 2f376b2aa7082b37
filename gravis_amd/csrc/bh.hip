// bh.hip -- Barnes-Hut gravity on gfx950: Morton-ordered linear octree and
// a wavefront-coherent group walk.
//
// Reference: BarnesHut::compute_accelerations (crates/sim-core/src/
// barnes_hut.rs:139-192) over the arena octree of octree.rs:145-239. The
// reference inserts particles one by one (one particle per leaf) and walks
// the tree per particle with a 512-entry stack, accepting a node when
// s^2 < theta^2 d^2 (s = node side, d = distance to the node's COM).
//
// Here the same root cube (octree.rs:14-44) is cut by 63-bit Morton keys, so
// every node is a run of the key-sorted particles and its cell is the
// reference's cell at that depth:
//   build   level by level: each internal node (more than `leaf` particles,
//           level < 21) finds its child runs by binary search on the next
//           3-bit digit; nodes are numbered breadth first, children of a
//           node contiguous in digit order.
//   COM     bottom up in FP64: leaves sum their particles in sorted order,
//           internal nodes sum their children in order (m, m*x; no FMA);
//           com = sum(m x) / sum(m).
//   walk    one wavefront per target group: the first tree node on each
//           root-to-leaf path holding at most `group` (<= 64) particles.
//           The frontier is processed breadth first, 64 nodes at a time
//           (one per lane). A node is accepted for the whole group when
//           s^2 < theta^2 d_min^2 with d_min the distance from its COM to the
//           group's bounding box (so never less accurate than the
//           reference's per-particle test); a rejected leaf contributes its
//           particles directly; a rejected internal node pushes its children.
//           If a group's frontier would exceed the per-wave capacity, the
//           group is walked again one target at a time (point boxes).
//           Accepted nodes and the particles of opened leaves are appended,
//           in frontier order, as source indices to a per-wave LDS list;
//           each flush round stages 64 source records in LDS (one load per
//           lane) that all 64 lanes then read back as broadcasts. Pair arithmetic is
//           barnes_hut.rs:114-135 in FP64 (sqrt and divide correctly
//           rounded, no contraction); a source at r^2 = 0 is skipped.
// The accumulation order is the list order, so results are bit-for-bit the
// CPU oracle's group walk with the same groups.
#include <algorithm>

#include "internal.h"

namespace gravis {
namespace {

constexpr int kMaxLevel = 21;
constexpr uint32_t kLeafFlag = 0x80000000u;
constexpr int kTreeBlock = 256;
constexpr int kTreeGrid = 1024;

// meta = {first, count, first child, flags}; flags = level | nchild << 8 |
// kGroupRoot (the parent holds more particles than a target group).
constexpr uint32_t kGroupRoot = 1u << 16;

__global__ void tree_init_kernel(uint4* __restrict__ meta, uint32_t* __restrict__ lvl,
                                 uint32_t n) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    meta[0] = make_uint4(0u, n, 0u, kGroupRoot);
    lvl[0] = 0u;
    lvl[1] = 1u;
    for (int L = 1; L <= kMaxLevel + 1; ++L) {
        lvl[2 * L] = 1u;
        lvl[2 * L + 1] = 1u;
    }
}

__device__ __forceinline__ uint32_t digit_of(uint64_t key, int shift) {
    return static_cast<uint32_t>((key >> shift) & 7ull);
}

// Number of non-empty children of each node of level L.
__global__ __launch_bounds__(kTreeBlock) void tree_count_kernel(const uint64_t* __restrict__ keys,
                                                               const uint4* __restrict__ meta,
                                                               const uint32_t* __restrict__ lvl,
                                                               int L, uint32_t leaf,
                                                               uint32_t* __restrict__ nch) {
    const uint32_t b = lvl[2 * L], e = lvl[2 * L + 1];
    const int shift = 3 * (kMaxLevel - (L + 1));
    for (uint32_t v = b + blockIdx.x * kTreeBlock + threadIdx.x; v < e;
         v += gridDim.x * kTreeBlock) {
        const uint4 md = meta[v];
        uint32_t c = 0;
        if (md.y > leaf && L < kMaxLevel) {
            uint32_t pos = md.x;
            const uint32_t end = md.x + md.y;
            while (pos < end) {
                const uint32_t d = digit_of(keys[pos], shift);
                uint32_t lo = pos + 1, hi = end;  // first index with a larger digit
                while (lo < hi) {
                    const uint32_t mid = lo + (hi - lo) / 2;
                    if (digit_of(keys[mid], shift) > d) hi = mid;
                    else lo = mid + 1;
                }
                ++c;
                pos = lo;
            }
        }
        nch[v - b] = c;
    }
}

// Exclusive scan of the level's child counts (one block) and the range of
// level L+1.
__global__ __launch_bounds__(1024) void tree_scan_kernel(const uint32_t* __restrict__ nch,
                                                         uint32_t* __restrict__ off,
                                                         uint32_t* __restrict__ lvl, int L) {
    __shared__ uint32_t s[1024];
    const uint32_t b = lvl[2 * L], e = lvl[2 * L + 1];
    const uint32_t m = e - b;
    const uint32_t per = (m + 1023) / 1024;
    const uint32_t lo = threadIdx.x * per;
    const uint32_t hi = min(m, lo + per);
    uint32_t sum = 0;
    for (uint32_t i = lo; i < hi; ++i) sum += nch[i];
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const uint32_t t = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0u;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t run = s[threadIdx.x] - sum;
    for (uint32_t i = lo; i < hi; ++i) {
        off[i] = run;
        run += nch[i];
    }
    if (threadIdx.x == 1023) {
        lvl[2 * (L + 1)] = e;
        lvl[2 * (L + 1) + 1] = e + s[1023];
    }
}

__global__ __launch_bounds__(kTreeBlock) void tree_write_kernel(
    const uint64_t* __restrict__ keys, uint4* __restrict__ meta, const uint32_t* __restrict__ lvl,
    int L, const uint32_t* __restrict__ nch, const uint32_t* __restrict__ off, uint32_t cap,
    uint32_t gsize, uint32_t* __restrict__ flag) {
    const uint32_t b = lvl[2 * L], e = lvl[2 * L + 1];
    const int shift = 3 * (kMaxLevel - (L + 1));
    for (uint32_t v = b + blockIdx.x * kTreeBlock + threadIdx.x; v < e;
         v += gridDim.x * kTreeBlock) {
        const uint32_t nc = nch[v - b];
        if (nc == 0) continue;
        const uint4 md = meta[v];
        const uint32_t c0 = e + off[v - b];
        if (c0 + nc > cap) {
            atomicOr(flag, 2u);
            continue;
        }
        uint32_t pos = md.x, k = 0;
        const uint32_t end = md.x + md.y;
        while (pos < end) {
            const uint32_t d = digit_of(keys[pos], shift);
            uint32_t lo = pos + 1, hi = end;
            while (lo < hi) {
                const uint32_t mid = lo + (hi - lo) / 2;
                if (digit_of(keys[mid], shift) > d) hi = mid;
                else lo = mid + 1;
            }
            meta[c0 + k] = make_uint4(pos, lo - pos, 0u,
                                      (uint32_t)(L + 1) | (md.y > gsize ? kGroupRoot : 0u));
            ++k;
            pos = lo;
        }
        meta[v] = make_uint4(md.x, md.y, c0, (md.w & kGroupRoot) | (nc << 8) | (uint32_t)L);
    }
}

// Raw FP64 sums (sum m x, sum m y, sum m z, sum m) and the particle bounding
// box of the nodes of level L.
__global__ __launch_bounds__(kTreeBlock) void tree_sums_kernel(const double4* __restrict__ src,
                                                              const uint4* __restrict__ meta,
                                                              const uint32_t* __restrict__ lvl,
                                                              int L,
                                                              double4* __restrict__ sums,
                                                              double4* __restrict__ bmin,
                                                              double4* __restrict__ bmax) {
    const uint32_t b = lvl[2 * L], e = lvl[2 * L + 1];
    for (uint32_t v = b + blockIdx.x * kTreeBlock + threadIdx.x; v < e;
         v += gridDim.x * kTreeBlock) {
        const uint4 md = meta[v];
        const uint32_t nc = (md.w >> 8) & 255u;
        double M = 0.0, X = 0.0, Y = 0.0, Z = 0.0;
        double lx = INFINITY, ly = INFINITY, lz = INFINITY;
        double hx = -INFINITY, hy = -INFINITY, hz = -INFINITY;
        if (nc == 0) {
            for (uint32_t k = md.x; k < md.x + md.y; ++k) {
                const double4 p = src[k];
                M = M + p.w;
                X = X + p.w * p.x;
                Y = Y + p.w * p.y;
                Z = Z + p.w * p.z;
                lx = fmin(lx, p.x); hx = fmax(hx, p.x);
                ly = fmin(ly, p.y); hy = fmax(hy, p.y);
                lz = fmin(lz, p.z); hz = fmax(hz, p.z);
            }
        } else {
            for (uint32_t c = md.z; c < md.z + nc; ++c) {
                const double4 q = sums[c];
                M = M + q.w;
                X = X + q.x;
                Y = Y + q.y;
                Z = Z + q.z;
                const double4 a = bmin[c], b = bmax[c];
                lx = fmin(lx, a.x); hx = fmax(hx, b.x);
                ly = fmin(ly, a.y); hy = fmax(hy, b.y);
                lz = fmin(lz, a.z); hz = fmax(hz, b.z);
            }
        }
        sums[v] = make_double4(X, Y, Z, M);
        bmin[v] = make_double4(lx, ly, lz, 0.0);
        bmax[v] = make_double4(hx, hy, hz, 0.0);
    }
}

// Node source records src[n + v] = (com, mass).
__global__ __launch_bounds__(kTreeBlock) void tree_com_kernel(double4* __restrict__ src, int n,
                                                             const uint4* __restrict__ meta,
                                                             const uint32_t* __restrict__ lvl,
                                                             const double4* __restrict__ sums) {
    const uint32_t total = lvl[2 * (kMaxLevel + 1)];
    for (uint32_t v = blockIdx.x * kTreeBlock + threadIdx.x; v < total;
         v += gridDim.x * kTreeBlock) {
        const double4 q = sums[v];
        double4 r;
        if (q.w != 0.0) {
            r = make_double4(q.x / q.w, q.y / q.w, q.z / q.w, q.w);
        } else {
            const double4 p = src[meta[v].x];
            r = make_double4(p.x, p.y, p.z, 0.0);
        }
        src[(size_t)n + v] = r;
    }
}

// Target groups of at most gsize particles, contiguous in sorted order: the
// children of every internal node holding more than gsize particles are
// visited in order and consecutive children holding at most gsize are packed
// greedily while the total stays within gsize and the packed particles'
// bounding box is no wider (largest axis extent) than twice the widest
// packed child's; a child leaf holding more (coincident points at level 21)
// is cut into runs of gsize; a root holding at most gsize (or a root leaf) is
// cut the same way. Appended in any order: each group's result depends only
// on its own particles and the tree.
__device__ __forceinline__ void emit_runs(uint2* groups, uint32_t* ngroups, uint32_t first,
                                          uint32_t cnt, uint32_t gsize) {
    const uint32_t k = (cnt + gsize - 1) / gsize;
    const uint32_t at = atomicAdd(ngroups, k);
    for (uint32_t q = 0; q < k; ++q)
        groups[at + q] = make_uint2(first + q * gsize, min(gsize, cnt - q * gsize));
}

__device__ __forceinline__ double extent(double4 lo, double4 hi) {
    return fmax(fmax(hi.x - lo.x, hi.y - lo.y), hi.z - lo.z);
}

__global__ __launch_bounds__(kTreeBlock) void group_emit_kernel(const uint4* __restrict__ meta,
                                                               const uint32_t* __restrict__ lvl,
                                                               const double4* __restrict__ bmin,
                                                               const double4* __restrict__ bmax,
                                                               uint32_t gsize,
                                                               uint2* __restrict__ groups,
                                                               uint32_t* __restrict__ ngroups) {
    const uint32_t total = lvl[2 * (kMaxLevel + 1)];
    for (uint32_t p = blockIdx.x * kTreeBlock + threadIdx.x; p < total;
         p += gridDim.x * kTreeBlock) {
        const uint4 md = meta[p];
        const uint32_t nc = (md.w >> 8) & 255u;
        if (p == 0 && (md.y <= gsize || nc == 0)) {
            emit_runs(groups, ngroups, md.x, md.y, gsize);
            continue;
        }
        if (md.y <= gsize || nc == 0) continue;
        uint32_t rf = 0, rc = 0;
        double4 rlo = make_double4(0, 0, 0, 0), rhi = rlo;
        double rmax = 0.0;
        for (uint32_t q = 0; q < nc; ++q) {
            const uint32_t cv = md.z + q;
            const uint4 cm = meta[cv];
            const uint32_t cc = (cm.w >> 8) & 255u;
            if (cm.y <= gsize) {
                const double4 clo = bmin[cv], chi = bmax[cv];
                const double cext = extent(clo, chi);
                if (rc > 0) {
                    const double4 ulo = make_double4(fmin(rlo.x, clo.x), fmin(rlo.y, clo.y),
                                                     fmin(rlo.z, clo.z), 0.0);
                    const double4 uhi = make_double4(fmax(rhi.x, chi.x), fmax(rhi.y, chi.y),
                                                     fmax(rhi.z, chi.z), 0.0);
                    if (rc + cm.y <= gsize && extent(ulo, uhi) <= 2.0 * fmax(rmax, cext)) {
                        rc += cm.y;
                        rlo = ulo;
                        rhi = uhi;
                        rmax = fmax(rmax, cext);
                        continue;
                    }
                    emit_runs(groups, ngroups, rf, rc, gsize);
                }
                rf = cm.x;
                rc = cm.y;
                rlo = clo;
                rhi = chi;
                rmax = cext;
            } else {
                if (rc > 0) emit_runs(groups, ngroups, rf, rc, gsize);
                rc = 0;
                if (cc == 0) emit_runs(groups, ngroups, cm.x, cm.y, gsize);
            }
        }
        if (rc > 0) emit_runs(groups, ngroups, rf, rc, gsize);
    }
}

__global__ void walk_reset_kernel(uint32_t* __restrict__ counter,
                                  unsigned long long* __restrict__ stats) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        counter[0] = 0u;
        counter[1] = 0u;
        for (int k = 0; k < 4; ++k) stats[k] = 0ull;
    }
}

constexpr int kWalkWaves = 4;
constexpr uint32_t kListCap = 1024;  // source indices per wave (LDS)

__device__ __forceinline__ void pair_ref(const double4 s, double px, double py, double pz,
                                         double eps2, double& ax, double& ay, double& az) {
    const double dx = s.x - px, dy = s.y - py, dz = s.z - pz;
    const double r2 = dx * dx + dy * dy + dz * dz + eps2;
    const double inv_r3 = 1.0 / (r2 * sqrt(r2));
    const double f = s.w * inv_r3;
    const bool use = r2 != 0.0;
    ax = use ? ax + f * dx : ax;
    ay = use ? ay + f * dy : ay;
    az = use ? az + f * dz : az;
}

// The interaction list holds indices into src: a particle k < n (a source of
// an opened leaf) or n + v (accepted node v), appended in frontier order.
// A flush sweeps it in rounds of 64: every lane loads one record into LDS,
// then all lanes read the 64 records back as broadcasts.
// MIXED: each staged record is rounded to FP32 relative to the group's box
// centre (separations formed in FP64 first), pair forces are FP32 (v_rsq_f32)
// and every round's FP32 sum is Kahan-added; the opening test, tree sums and
// output stay FP64.
template <bool MIXED>
__global__ __launch_bounds__(kWalkWaves* kWave) void bh_walk_kernel(
    const double4* __restrict__ src, const uint4* __restrict__ meta, int n,
    const uint2* __restrict__ groups, const uint32_t* __restrict__ ngroups,
    const double* __restrict__ box, double theta2, double eps2, uint32_t* __restrict__ frontier,
    uint32_t fcap, uint32_t* __restrict__ counter, uint32_t* __restrict__ flag,
    double* __restrict__ part, unsigned long long* __restrict__ stats) {
    __shared__ uint32_t lists[kWalkWaves][kListCap];
    __shared__ double4 stages[kWalkWaves][kWave];
    unsigned long long t_nodes = 0, t_acc = 0, t_leafp = 0, t_pairs = 0;
    const float eps2f = (float)eps2;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    uint32_t* list = lists[w];
    double4* stage = stages[w];
    float4* stage32 = reinterpret_cast<float4*>(stage);
    const uint64_t lt = (1ull << lane) - 1ull;
    const size_t slot = (size_t)blockIdx.x * kWalkWaves + w;
    uint32_t* const fa0 = frontier + slot * 2 * (size_t)fcap;
    uint32_t* const fb0 = fa0 + fcap;
    const double side0 = 2.0 * box[3];  // root side, octree.rs:81-85

    for (;;) {
        uint32_t g = 0;
        if (lane == 0) g = atomicAdd(counter, 1u);
        g = __shfl(g, 0);
        if (g >= ngroups[0]) break;
        const uint2 gr = groups[g];
        const int i0 = (int)gr.x;
        const int cnt = (int)gr.y;
        const bool active = lane < cnt;
        const double4 me = src[i0 + (active ? lane : 0)];
        double glo[3] = {me.x, me.y, me.z}, ghi[3] = {me.x, me.y, me.z};
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                glo[d] = fmin(glo[d], __shfl_xor(glo[d], o));
                ghi[d] = fmax(ghi[d], __shfl_xor(ghi[d], o));
            }
        // MIXED: records and target are rounded to FP32 relative to the
        // group's box centre
        const double gcx = 0.5 * (glo[0] + ghi[0]), gcy = 0.5 * (glo[1] + ghi[1]),
                     gcz = 0.5 * (glo[2] + ghi[2]);
        const float fx = (float)(me.x - gcx), fy = (float)(me.y - gcy), fz = (float)(me.z - gcz);
        double ax = 0.0, ay = 0.0, az = 0.0;
        float sx = 0.f, sy = 0.f, sz = 0.f, kx = 0.f, ky = 0.f, kz = 0.f;
        uint32_t nl = 0;    // list length (wave uniform)
        bool keep = true;   // this lane's target is being walked

        auto flush = [&]() {
            for (uint32_t q0 = 0; q0 < nl; q0 += kWave) {
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                __builtin_amdgcn_wave_barrier();
                const uint32_t q = q0 + lane;
                if (q < nl) {
                    const double4 r = src[list[q]];
                    if (MIXED)
                        stage32[lane] = make_float4((float)(r.x - gcx), (float)(r.y - gcy),
                                                    (float)(r.z - gcz), (float)r.w);
                    else
                        stage[lane] = r;
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                __builtin_amdgcn_wave_barrier();
                const uint32_t m = min(nl - q0, (uint32_t)kWave);
                if (MIXED) {
                    float tx = 0.f, ty = 0.f, tz = 0.f;
                    for (uint32_t k = 0; k < m; ++k) {
                        const float4 r = stage32[k];
                        const float dx = r.x - fx, dy = r.y - fy, dz = r.z - fz;
                        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2f)));
                        float ri = __builtin_amdgcn_rsqf(r2);
                        ri = r2 > 0.f ? ri : 0.f;
                        const float f = (r.w * ri) * (ri * ri);
                        tx = __builtin_fmaf(f, dx, tx);
                        ty = __builtin_fmaf(f, dy, ty);
                        tz = __builtin_fmaf(f, dz, tz);
                    }
                    if (keep) {
                        float y = tx - kx, t = sx + y; kx = (t - sx) - y; sx = t;
                        y = ty - ky; t = sy + y; ky = (t - sy) - y; sy = t;
                        y = tz - kz; t = sz + y; kz = (t - sz) - y; sz = t;
                    }
                } else {
                    double tx = ax, ty = ay, tz = az;
                    for (uint32_t k = 0; k < m; ++k)
                        pair_ref(stage[k], me.x, me.y, me.z, eps2, tx, ty, tz);
                    if (keep) {
                        ax = tx;
                        ay = ty;
                        az = tz;
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            __builtin_amdgcn_wave_barrier();
            nl = 0;
        };
        // appends the uniform index range [first, first + count) in order
        auto append_range = [&](uint32_t first, uint32_t count) {
            while (count > 0) {
                if (nl == kListCap) flush();
                const uint32_t k = min(count, kListCap - nl);
                for (uint32_t r = lane; r < k; r += kWave) list[nl + r] = first + r;
                nl += k;
                first += k;
                count -= k;
            }
        };
        // Breadth-first walk with opening box (lo, hi) for `ntargets`
        // targets; returns true if a next frontier would exceed fcap.
        auto walk = [&](const double (&lo)[3], const double (&hi)[3], int ntargets) -> bool {
            uint32_t* fa = fa0;
            uint32_t* fb = fb0;
            if (lane == 0) fa[0] = 0u;
            uint32_t nf = 1;
            bool overflow = false;
            __threadfence_block();
            while (nf > 0 && !overflow) {
                uint32_t nn = 0;
                for (uint32_t base = 0; base < nf; base += kWave) {
                    const uint32_t idx = base + lane;
                    const bool valid = idx < nf;
                    t_nodes += min(nf - base, (uint32_t)kWave);
                    const uint32_t v = valid ? fa[idx] : 0u;
                    const uint4 md = meta[v];
                    const double4 c = src[(size_t)n + v];
                    const uint32_t level = md.w & 255u;
                    const uint32_t nchild = (md.w >> 8) & 255u;
                    const double s = ldexp(side0, -(int)level);
                    const double ex = fmax(fmax(lo[0] - c.x, c.x - hi[0]), 0.0);
                    const double ey = fmax(fmax(lo[1] - c.y, c.y - hi[1]), 0.0);
                    const double ez = fmax(fmax(lo[2] - c.z, c.z - hi[2]), 0.0);
                    const double d2 = ex * ex + ey * ey + ez * ez;
                    const bool accept = valid && (s * s < theta2 * d2);
                    const bool open = valid && !accept && nchild > 0;
                    const bool direct = valid && !accept && nchild == 0;
                    // sources of this chunk in lane order: 1 per accepted
                    // node, count per opened leaf
                    const uint32_t add = accept ? 1u : (direct ? md.y : 0u);
                    uint32_t sc = add;
#pragma unroll
                    for (int o = 1; o < kWave; o <<= 1) {
                        const uint32_t t = __shfl_up(sc, o);
                        if (lane >= o) sc += t;
                    }
                    const uint32_t total = __shfl(sc, kWave - 1);
                    const uint64_t accb = __ballot(accept);
                    t_acc += (unsigned long long)__popcll(accb);
                    t_leafp += total - (uint32_t)__popcll(accb);
                    t_pairs += (unsigned long long)total * (unsigned long long)ntargets;
                    if (total > 0) {
                        if (nl + total > kListCap) flush();
                        if (total <= kListCap) {
                            const uint32_t pos = nl + sc - add;
                            if (accept) list[pos] = (uint32_t)n + v;
                            else if (direct)
                                for (uint32_t q = 0; q < add; ++q) list[pos + q] = md.x + q;
                            nl += total;
                        } else {
                            // a leaf larger than the list (coincident points):
                            // lane by lane, same order
                            uint64_t todo = __ballot(add > 0);
                            while (todo) {
                                const int l = __ffsll((long long)todo) - 1;
                                todo &= todo - 1;
                                const uint32_t lacc = __shfl((uint32_t)accept, l);
                                const uint32_t lv = __shfl(v, l);
                                const uint32_t lf = __shfl(md.x, l);
                                const uint32_t la = __shfl(add, l);
                                if (lacc) append_range((uint32_t)n + lv, 1u);
                                else append_range(lf, la);
                            }
                        }
                    }
                    // children of opened internal nodes -> next frontier
                    const uint32_t k = open ? nchild : 0u;
                    uint32_t kc = k;
#pragma unroll
                    for (int o = 1; o < kWave; o <<= 1) {
                        const uint32_t t = __shfl_up(kc, o);
                        if (lane >= o) kc += t;
                    }
                    const uint32_t ktot = __shfl(kc, kWave - 1);
                    if (nn + ktot > fcap) {
                        overflow = true;
                        break;
                    }
                    const uint32_t ex0 = nn + kc - k;
                    for (uint32_t q = 0; q < k; ++q) fb[ex0 + q] = md.z + q;
                    nn += ktot;
                }
                __threadfence_block();
                __builtin_amdgcn_wave_barrier();
                uint32_t* t = fa;
                fa = fb;
                fb = t;
                nf = nn;
            }
            flush();
            return overflow;
        };

        if (walk(glo, ghi, cnt)) {
            // the group box opens too much: walk every target alone
            ax = ay = az = 0.0;
            sx = sy = sz = kx = ky = kz = 0.f;
            for (int l = 0; l < cnt; ++l) {
                const double px = __shfl(me.x, l), py = __shfl(me.y, l), pz = __shfl(me.z, l);
                const double p[3] = {px, py, pz};
                keep = lane == l;
                if (walk(p, p, 1) && lane == 0) atomicOr(flag, 1u);
            }
            keep = true;
        }
        if (MIXED) {
            ax = (double)sx - (double)kx;
            ay = (double)sy - (double)ky;
            az = (double)sz - (double)kz;
        }
        if (active) {
            const size_t i = (size_t)i0 + lane;
            part[i] = ax;
            part[(size_t)n + i] = ay;
            part[2 * (size_t)n + i] = az;
        }
    }
    if (lane == 0) {
        atomicAdd(&stats[0], t_nodes);
        atomicAdd(&stats[1], t_acc);
        atomicAdd(&stats[2], t_leafp);
        atomicAdd(&stats[3], t_pairs);
    }
}

}  // namespace

gravis_status bh_force(Ctx* c, const gravis_solver* s, size_t n, const double* x,
                       const double* y, const double* z, const double* m, const FinishArgs& fa) {
    if (n == 0) return GRAVIS_OK;
    if (n > 0x7FFFFFFFull) return fail(c, GRAVIS_INVALID_ARG, "n too large");
    const uint32_t leaf = s->leaf_size ? s->leaf_size : 16u;
    const int gs = s->group_size ? (int)s->group_size : kWave;
    if (gs < 1 || gs > kWave) return fail(c, GRAVIS_INVALID_ARG, "group_size must be 1..64");
    if (leaf < 1 || leaf > 4096) return fail(c, GRAVIS_INVALID_ARG, "leaf_size must be 1..4096");
    const int ni = (int)n;

    double* box;
    GRAVIS_TRY(buf(c, "bh_box", 4, &box));
    GRAVIS_TRY(root_box(c, n, x, y, z, box));
    uint64_t* keys;
    uint32_t* perm;
    GRAVIS_TRY(morton_sort(c, n, x, y, z, box, &keys, &perm));

    // Internal nodes of one level are disjoint runs of more than `leaf`
    // particles and have at most 8 children each.
    const size_t cap = 1 + (size_t)kMaxLevel * 8 * (n / (leaf + 1) + 1);
    if (cap > 0x7FFFFFFFull) return fail(c, GRAVIS_INVALID_ARG, "tree too large");
    double4 *src, *sums, *bmin, *bmax;
    uint4* meta;
    uint32_t *lvl, *nch, *off, *flags;
    GRAVIS_TRY(buf(c, "bh_src", n + cap, &src));
    GRAVIS_TRY(buf(c, "bh_sums", cap, &sums));
    GRAVIS_TRY(buf(c, "bh_bmin", cap, &bmin));
    GRAVIS_TRY(buf(c, "bh_bmax", cap, &bmax));
    GRAVIS_TRY(buf(c, "bh_meta", cap, &meta));
    GRAVIS_TRY(buf(c, "bh_lvl", 2 * (kMaxLevel + 2), &lvl));
    GRAVIS_TRY(buf(c, "bh_nch", cap, &nch));
    GRAVIS_TRY(buf(c, "bh_off", cap, &off));
    GRAVIS_TRY(buf(c, "bh_flags", 4, &flags));
    GRAVIS_TRY(gather_sorted(c, n, x, y, z, m, perm, src));

    hipStream_t st = c->stream;
    GRAVIS_HIP(c, hipMemsetAsync(flags, 0, 4 * sizeof(uint32_t), st));
    tree_init_kernel<<<1, 64, 0, st>>>(meta, lvl, (uint32_t)n);
    const unsigned tg = grid_for(cap, kTreeBlock, kTreeGrid);
    for (int L = 0; L < kMaxLevel; ++L) {
        tree_count_kernel<<<tg, kTreeBlock, 0, st>>>(keys, meta, lvl, L, leaf, nch);
        tree_scan_kernel<<<1, 1024, 0, st>>>(nch, off, lvl, L);
        tree_write_kernel<<<tg, kTreeBlock, 0, st>>>(keys, meta, lvl, L, nch, off, (uint32_t)cap,
                                                    (uint32_t)gs, flags + 1);
    }
    // level 21 holds leaves only (zero children); its scan closes the node range
    tree_count_kernel<<<tg, kTreeBlock, 0, st>>>(keys, meta, lvl, kMaxLevel, leaf, nch);
    tree_scan_kernel<<<1, 1024, 0, st>>>(nch, off, lvl, kMaxLevel);
    GRAVIS_TRY(check_launch(c, "tree build"));
    for (int L = kMaxLevel; L >= 0; --L)
        tree_sums_kernel<<<tg, kTreeBlock, 0, st>>>(src, meta, lvl, L, sums, bmin, bmax);
    tree_com_kernel<<<tg, kTreeBlock, 0, st>>>(src, ni, meta, lvl, sums);
    GRAVIS_TRY(check_launch(c, "tree com"));

    // target groups, then the walk
    const int blocks = (int)std::min<long>((long)(n / 64 / kWalkWaves) + 1,
                                           (long)c->num_cus * 4);
    const uint32_t fcap = (uint32_t)std::min<size_t>(cap, c->bh_frontier_cap);
    uint32_t *front, *counter;
    uint2* groups;
    GRAVIS_TRY(buf(c, "bh_front", (size_t)blocks * kWalkWaves * 2 * fcap, &front));
    GRAVIS_TRY(buf(c, "bh_counter", 4, &counter));
    GRAVIS_TRY(buf(c, "bh_groups", n, &groups));
    double* part;
    GRAVIS_TRY(buf(c, "bh_part", 3 * n, &part));
    const double theta2 = s->theta * s->theta;
    const double eps2 = s->softening * s->softening;
    unsigned long long* stats;
    GRAVIS_TRY(buf(c, "bh_stats", 4, &stats));
    walk_reset_kernel<<<1, 64, 0, st>>>(counter, stats);
    group_emit_kernel<<<tg, kTreeBlock, 0, st>>>(meta, lvl, bmin, bmax, (uint32_t)gs, groups,
                                                 counter + 1);
    timer_begin(c);
    if (s->precision == GRAVIS_MIXED)
        bh_walk_kernel<true><<<blocks, kWalkWaves * kWave, 0, st>>>(
            src, meta, ni, groups, counter + 1, box, theta2, eps2, front, fcap, counter, flags,
            part, stats);
    else
        bh_walk_kernel<false><<<blocks, kWalkWaves * kWave, 0, st>>>(
            src, meta, ni, groups, counter + 1, box, theta2, eps2, front, fcap, counter, flags,
            part, stats);
    timer_end(c, "bh_walk");
    GRAVIS_TRY(check_launch(c, "bh_walk"));
    uint32_t hflags[2] = {0, 0};
    GRAVIS_HIP(c, hipMemcpyAsync(hflags, flags, sizeof hflags, hipMemcpyDeviceToHost, st));
    GRAVIS_HIP(c, hipMemcpyAsync(c->bh_stats, stats, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    uint32_t hlvl[2 * (kMaxLevel + 2)];
    GRAVIS_HIP(c, hipMemcpyAsync(hlvl, lvl, sizeof hlvl, hipMemcpyDeviceToHost, st));
    GRAVIS_HIP(c, hipStreamSynchronize(st));
    c->bh_stats[4] = hlvl[2 * (kMaxLevel + 1)];  // node count
    if (hflags[1])
        return fail(c, GRAVIS_INTERNAL, "octree node capacity exceeded (%zu nodes)", cap);
    if (hflags[0])
        return fail(c, GRAVIS_INTERNAL,
                    "Barnes-Hut walk frontier of a single target exceeded %u nodes (theta %g)",
                    fcap, s->theta);
    return finish_forces(c, part, 1, ni, perm, fa);
}

}  // namespace gravis
