// tree.hip -- Morton-ordered linear octree on gfx950 (shared by the
// Barnes-Hut walk and the SPH neighbour search).
//
// Reference: the arena octree of crates/sim-core/src/octree.rs:145-239,
// which inserts particles one by one (one particle per leaf) into the root
// cube of octree.rs:14-44. Here the same root cube is cut by 63-bit Morton
// keys, so every node is a run of the key-sorted particles and its cell is
// the reference's cell at that depth:
//   build   level by level, one launch per level: each internal node (more
//           than `leaf` particles, level < 21) finds its child runs by
//           binary search on the next 3-bit digit; the nodes of a level
//           follow those of the level above, and the children of a node
//           are contiguous in digit order.
//   sums    bottom up in FP64: leaves sum their particles in sorted order,
//           internal nodes sum their children in order (m, m*x; no FMA);
//           com = sum(m x) / sum(m). The particle bounding box of every node
//           is kept beside it.
//   groups  target groups of at most `group` (<= 64) particles, one per
//           wavefront, cut from the tree (group_emit_kernel).
#include <algorithm>

#include "internal.h"

namespace gravis {
namespace {

constexpr int kMaxLevel = kTreeMaxLevel;
constexpr uint32_t kGroupRoot = kNodeGroupRoot;

__global__ void tree_init_kernel(uint4* __restrict__ meta, uint32_t* __restrict__ lvl,
                                 uint32_t n, uint32_t* __restrict__ flags,
                                 uint32_t* __restrict__ ngroups, uint32_t* __restrict__ cnt) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int k = 0; k < 4; ++k) flags[k] = 0u;
    for (int L = 0; L <= kMaxLevel; ++L) cnt[L] = 0u;
    *ngroups = 0u;
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

// Child runs of one node, 8 lanes per node (lane d = digit d; the node's
// lanes are 8 consecutive, 8-aligned lanes of one wavefront): lane d finds
// by binary search the first particle of the node whose next 3-bit digit is
// >= d, so the 8 searches run side by side instead of one after another.
// Returns this lane's run [start, stop) (empty when the digit is absent) and
// the node's bit mask of non-empty digits. split = false: no children.
struct NodeRun {
    uint32_t start, stop, bits;
};
__device__ __forceinline__ NodeRun node_split(const uint64_t* __restrict__ keys, uint4 md,
                                              int shift, bool split, uint32_t d, int lane) {
    const uint32_t end = md.x + md.y;
    uint32_t lo = md.x, hi = end;
    if (split && d > 0) {
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (digit_of(keys[mid], shift) >= d) hi = mid;
            else lo = mid + 1;
        }
    }
    const uint32_t nxt = __shfl_down(lo, 1);
    const uint32_t stop = d == 7 ? end : nxt;
    const bool ne = split && stop > lo;
    const uint64_t bal = __ballot(ne);
    return {lo, ne ? stop : lo, (uint32_t)(bal >> (lane & ~7)) & 0xFFu};
}

// One level of the build in one launch. Level L's range is the end of
// level L-1 plus level L-1's children (cnt[L-1]); block 0 records it in lvl.
// Each node of the level (8 lanes per node; the grid-stride loop is uniform
// across each block, so the wavefront-wide shuffles, ballots and the block
// scan see every lane) finds its child runs, reserves slots for its
// children after the level (one atomic add on cnt[L] per block and
// iteration) and writes them: the children of a node are contiguous in
// digit order, and the order of the nodes within a level follows the
// reservations (every consumer reads a node's children through its first
// child and count, so none depends on that order).
__global__ __launch_bounds__(kTreeBlock) void tree_level_kernel(
    const uint64_t* __restrict__ keys, uint4* __restrict__ meta, uint32_t* __restrict__ lvl,
    uint32_t* __restrict__ cnt, int L, uint32_t leaf, uint32_t cap, uint32_t gsize,
    uint32_t* __restrict__ flag) {
    __shared__ uint32_t sbase;
    // a level that ran out of node capacity ends the build (the caller
    // reports it); the levels below keep their empty ranges
    if (__builtin_amdgcn_readfirstlane(*flag) & 2u) return;
    uint32_t b = 0, e = 1;
    if (L > 0) {
        b = lvl[2 * L - 1];
        e = min(b + cnt[L - 1], cap);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        lvl[2 * L] = b;
        lvl[2 * L + 1] = e;
    }
    const int shift = 3 * (kMaxLevel - (L + 1));
    const uint32_t total = 8u * (e - b);
    const int lane = threadIdx.x & (kWave - 1);
    for (uint32_t t0 = blockIdx.x * kTreeBlock; t0 < total; t0 += gridDim.x * kTreeBlock) {
        const uint32_t t = t0 + threadIdx.x;
        const bool valid = t < total;
        const uint32_t v = b + (valid ? t / 8u : 0u);
        const uint32_t d = t & 7u;
        const uint4 md = valid ? meta[v] : make_uint4(0u, 0u, 0u, 0u);
        const bool split = valid && md.y > leaf && L < kMaxLevel;
        // every lane of the node reads meta[v] above before lane 0 rewrites it
        const NodeRun r = node_split(keys, md, shift, split, d, lane);
        const uint32_t nc = (uint32_t)__popc(r.bits);
        uint32_t all;
        const uint32_t ex = block_exclusive_scan<kTreeBlock>(d == 0 ? nc : 0u, &all);
        if (threadIdx.x == 0) sbase = all > 0u ? atomicAdd(&cnt[L], all) : 0u;
        __syncthreads();
        const uint32_t c0 = e + sbase + __shfl(ex, lane & ~7);
        __syncthreads();  // sbase is rewritten by the next iteration
        if (nc == 0) continue;
        if (c0 + nc > cap) {
            if (d == 0) atomicOr(flag, 2u);
            continue;
        }
        // children in digit order: this run's rank among the non-empty digits
        if (r.stop > r.start) {
            const uint32_t k = (uint32_t)__popc(r.bits & ((1u << d) - 1u));
            meta[c0 + k] = make_uint4(r.start, r.stop - r.start, 0u,
                                      (uint32_t)(L + 1) | (md.y > gsize ? kGroupRoot : 0u));
        }
        if (d == 0)
            meta[v] = make_uint4(md.x, md.y, c0, (md.w & kGroupRoot) | (nc << 8) | (uint32_t)L);
    }
}

// Node count after the last level: lvl[2 (kMaxLevel + 1)] = lvl[.. + 1] =
// the end of level kMaxLevel (whose nodes have no children).
__global__ void tree_close_kernel(uint32_t* __restrict__ lvl) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        lvl[2 * (kMaxLevel + 1)] = lvl[2 * kMaxLevel + 1];
        lvl[2 * (kMaxLevel + 1) + 1] = lvl[2 * kMaxLevel + 1];
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
// greedily while the total stays within gsize (they lie in their parent's
// cell, so a group stays within one cell) or, with `split`, each maximal
// run of consecutive such children is cut into runs of gsize particles
// (fuller wavefronts; the Barnes-Hut gravity path, where it measured faster,
// while the SPH neighbour search keeps whole children); a child leaf holding more
// (coincident points at level 21) is cut into runs of gsize; a root holding
// at most gsize (or a root leaf) is cut the same way. Appended in any order:
// each group's result depends only on its own particles and the tree.
// (A former rule also stopped a pack when its bounding box grew wider than
// twice its widest child's: groups then held 31 targets per list entry on
// the two-sphere 4M workload against 47 now, and the walk made 45 % more
// node tests; dropping it took that walk from 10.1 to 9.1 ms.)
// Groups of a run of cnt particles from first: ceil(cnt / gsize) runs of
// at most gsize, written at groups[at..] when EMIT; returns their number.
template <bool EMIT>
__device__ __forceinline__ uint32_t emit_runs(uint2* groups, uint32_t at, uint32_t first,
                                              uint32_t cnt, uint32_t gsize) {
    const uint32_t k = (cnt + gsize - 1) / gsize;
    if (EMIT)
        for (uint32_t q = 0; q < k; ++q)
            groups[at + q] = make_uint2(first + q * gsize, min(gsize, cnt - q * gsize));
    return k;
}

// The groups node p cuts (see above): counted, or written from groups[at].
template <bool EMIT>
__device__ __forceinline__ uint32_t node_groups(uint32_t p, const uint4 md,
                                                const uint4* __restrict__ meta, uint32_t gsize,
                                                bool split, uint2* __restrict__ groups,
                                                uint32_t at) {
    const uint32_t nc = (md.w >> 8) & 255u;
    if (p == 0 && (md.y <= gsize || nc == 0))
        return emit_runs<EMIT>(groups, at, md.x, md.y, gsize);
    if (md.y <= gsize || nc == 0) return 0u;
    uint32_t k = 0;
    uint32_t rf = 0, rc = 0;
    for (uint32_t q = 0; q < nc; ++q) {
        const uint4 cm = meta[md.z + q];
        const uint32_t cc = (cm.w >> 8) & 255u;
        if (cm.y <= gsize) {
            if (split) {
                // the segment of consecutive small children grows; it is cut
                // into runs of gsize when it ends
                if (rc == 0) rf = cm.x;
                rc += cm.y;
                continue;
            }
            if (rc > 0) {
                if (rc + cm.y <= gsize) {
                    rc += cm.y;
                    continue;
                }
                k += emit_runs<EMIT>(groups, at + k, rf, rc, gsize);
            }
            rf = cm.x;
            rc = cm.y;
        } else {
            if (rc > 0) k += emit_runs<EMIT>(groups, at + k, rf, rc, gsize);
            rc = 0;
            if (cc == 0) k += emit_runs<EMIT>(groups, at + k, cm.x, cm.y, gsize);
        }
    }
    if (rc > 0) k += emit_runs<EMIT>(groups, at + k, rf, rc, gsize);
    return k;
}

// Every node counts its groups, the wavefront reserves them with one atomic
// add and each node then writes its own at its offset. The loop is uniform
// across each block (every lane takes part in the scan).
__global__ __launch_bounds__(kTreeBlock) void group_emit_kernel(const uint4* __restrict__ meta,
                                                               const uint32_t* __restrict__ lvl,
                                                               uint32_t gsize, bool split,
                                                               uint2* __restrict__ groups,
                                                               uint32_t* __restrict__ ngroups) {
    const uint32_t total = lvl[2 * (kMaxLevel + 1)];
    const int lane = threadIdx.x & (kWave - 1);
    for (uint32_t t0 = blockIdx.x * kTreeBlock; t0 < total; t0 += gridDim.x * kTreeBlock) {
        const uint32_t p = t0 + threadIdx.x;
        const bool valid = p < total;
        const uint4 md = valid ? meta[p] : make_uint4(0u, 0u, 0u, 0u);
        const uint32_t k = valid ? node_groups<false>(p, md, meta, gsize, split, groups, 0u)
                                 : 0u;
        const uint32_t inc = wave_inclusive_scan(k);
        const uint32_t wtot = __builtin_amdgcn_readlane(inc, kWave - 1);
        uint32_t base = 0;
        if (wtot > 0) {
            if (lane == 0) base = atomicAdd(ngroups, wtot);
            base = __builtin_amdgcn_readfirstlane(base);
        }
        if (k > 0) (void)node_groups<true>(p, md, meta, gsize, split, groups, base + inc - k);
    }
}

}  // namespace

gravis_status build_octree(Ctx* c, size_t n, const double* x, const double* y, const double* z,
                           const double* m, uint32_t leaf, uint32_t group, Octree* t,
                           bool split_siblings) {
    if (n == 0) return fail(c, GRAVIS_INVALID_ARG, "octree of zero particles");
    if (n > 0x7FFFFFFFull) return fail(c, GRAVIS_INVALID_ARG, "n too large");
    if (leaf < 1 || leaf > 4096) return fail(c, GRAVIS_INVALID_ARG, "leaf_size must be 1..4096");
    if (group < 1 || group > (uint32_t)kWave)
        return fail(c, GRAVIS_INVALID_ARG, "group_size must be 1..64");
    double* box;
    GRAVIS_TRY(buf(c, "tree_box", 4, &box));
    timer_begin(c);
    GRAVIS_TRY(root_box(c, n, x, y, z, box));
    uint64_t* keys;
    uint32_t* perm;
    GRAVIS_TRY(morton_sort(c, n, x, y, z, box, &keys, &perm));
    timer_end(c, "morton_sort");

    // Internal nodes of one level are disjoint runs of more than `leaf`
    // particles and have at most 8 children each.
    const size_t cap = 1 + (size_t)kMaxLevel * 8 * (n / (leaf + 1) + 1);
    if (cap > 0x7FFFFFFFull) return fail(c, GRAVIS_INVALID_ARG, "tree too large");
    double4 *src, *sums, *bmin, *bmax;
    uint4* meta;
    uint32_t *lvl, *cnt, *flags, *ngroups;
    uint2* groups;
    GRAVIS_TRY(buf(c, "tree_src", n + cap, &src));
    GRAVIS_TRY(buf(c, "tree_sums", cap, &sums));
    GRAVIS_TRY(buf(c, "tree_bmin", cap, &bmin));
    GRAVIS_TRY(buf(c, "tree_bmax", cap, &bmax));
    GRAVIS_TRY(buf(c, "tree_meta", cap, &meta));
    GRAVIS_TRY(buf(c, "tree_lvl", 2 * (kMaxLevel + 2), &lvl));
    GRAVIS_TRY(buf(c, "tree_cnt", kMaxLevel + 1, &cnt));
    GRAVIS_TRY(buf(c, "tree_flags", 4, &flags));
    GRAVIS_TRY(buf(c, "tree_ngroups", 1, &ngroups));
    GRAVIS_TRY(buf(c, "tree_groups", n, &groups));
    timer_begin(c);
    const uint32_t* inv = nullptr;
    GRAVIS_TRY(gather_sorted(c, n, x, y, z, m, perm, src, &inv));

    hipStream_t st = c->stream;
    tree_init_kernel<<<1, 64, 0, st>>>(meta, lvl, (uint32_t)n, flags, ngroups, cnt);
    const unsigned tg = grid_for(cap, kTreeBlock, kTreeGrid);
    // level 21 holds leaves only (no children); the close kernel ends the range
    for (int L = 0; L <= kMaxLevel; ++L)
        tree_level_kernel<<<tg, kTreeBlock, 0, st>>>(keys, meta, lvl, cnt, L, leaf,
                                                     (uint32_t)cap, group, flags + 1);
    tree_close_kernel<<<1, 64, 0, st>>>(lvl);
    GRAVIS_TRY(check_launch(c, "tree build"));
    for (int L = kMaxLevel; L >= 0; --L)
        tree_sums_kernel<<<tg, kTreeBlock, 0, st>>>(src, meta, lvl, L, sums, bmin, bmax);
    tree_com_kernel<<<tg, kTreeBlock, 0, st>>>(src, (int)n, meta, lvl, sums);
    group_emit_kernel<<<tg, kTreeBlock, 0, st>>>(meta, lvl, group, split_siblings, groups,
                                                 ngroups);
    timer_end(c, "tree_build");
    GRAVIS_TRY(check_launch(c, "tree sums and groups"));

    t->n = n;
    t->cap = cap;
    t->leaf = leaf;
    t->group = group;
    t->box = box;
    t->keys = keys;
    t->perm = perm;
    t->inv = inv;
    t->src = src;
    t->meta = meta;
    t->bmin = bmin;
    t->bmax = bmax;
    t->lvl = lvl;
    t->flags = flags;
    t->groups = groups;
    t->ngroups = ngroups;
    return GRAVIS_OK;
}

}  // namespace gravis
