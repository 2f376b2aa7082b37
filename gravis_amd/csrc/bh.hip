// bh.hip -- Barnes-Hut gravity on gfx950: a wavefront-coherent group walk
// over the Morton-ordered linear octree of tree.hip.
//
// Reference: BarnesHut::compute_accelerations (crates/sim-core/src/
// barnes_hut.rs:139-192) over the arena octree of octree.rs:145-239. The
// reference walks the tree per particle with a 512-entry stack, accepting a
// node when s^2 < theta^2 d^2 (s = node side, d = distance to the node's
// COM).
//
// Here one wavefront walks one target group (the first tree node on each
// root-to-leaf path holding at most `group` (<= 64) particles, siblings
// packed). The frontier is processed breadth first, 64 nodes at a time (one
// per lane), from LDS (levels of up to 512 nodes; larger levels spill their
// tail to a per-wave global buffer). A node is accepted for the whole group when
// s^2 < theta^2 d_min^2 with d_min the distance from its COM to the group's
// bounding box (so never less accurate than the reference's per-particle
// test); a rejected leaf contributes its particles directly; a rejected
// internal node pushes its children. If a group's frontier would exceed
// kGroupFrontierCap (or the per-wave capacity, if smaller), the group is
// walked again one target at a time (point boxes). Accepted nodes and the particles of opened leaves are appended, in
// frontier order, as source indices to a per-wave LDS list; each flush round
// stages 64 source records in LDS (one load per lane) that all 64 lanes then
// read back as broadcasts. A group of at most 32 (16) targets is held by 2
// (4) lanes per target: copy r sums the list positions p with p % R == r,
// in order (the mixed walk: a contiguous block of each round of 64), and
// the copies are added as (r0 + r1) [+ (r2 + r3)]. Pair
// arithmetic is barnes_hut.rs:114-135 in FP64 (sqrt and divide correctly
// rounded, no contraction); a source at r^2 = 0 is skipped. The oracle's
// group walk restates the list positions (64-node chunks, 1024-entry list)
// and the copies, so results are bit-for-bit the same with the same groups.
#include <algorithm>
#include <type_traits>

#include "internal.h"

namespace gravis {
namespace {

__global__ void walk_reset_kernel(uint32_t* __restrict__ counter,
                                  unsigned long long* __restrict__ stats) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        for (uint32_t k = 0; k < kXcds; ++k) counter[k] = 0u;
        for (int k = 0; k < 4; ++k) stats[k] = 0ull;
    }
}

constexpr int kWalkWaves = 4;
constexpr uint32_t kListCap = 1024;  // source indices per wave (LDS), FP64 walk
constexpr uint32_t kListCapMixed = 768;  // the same, mixed walk
// Frontier entries per level held in LDS; a level's entries past this spill
// to the wave's global frontier buffer. On the two-sphere 4M workload 91 %
// of the node tests come from levels of at most 512 nodes.
constexpr uint32_t kLdsFront = 512;
static_assert(kLdsFront % 64 == 0, "a 64-node chunk is wholly in LDS or wholly in memory");
// Frontier at which a group's breadth-first walk stops and its targets are
// walked one at a time (measured: 2048-8192 all give the shortest C3 and
// single-sphere walks; 16384 lets a few groups whose box spans a sphere's
// core run breadth first through most of the tree and set the kernel's tail).
constexpr uint32_t kGroupFrontierCap = 4096;

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
// The breadth-first frontier of each level lives in LDS up to kLdsFront
// entries (the rest in the wave's global buffer). MIXED: 96 VGPRs, no spill,
// 32 KB of LDS per block (768-entry lists, FP32 stages): 5 wavefronts per
// SIMD. FP64: about 121 VGPRs and 40 KB: 4 wavefronts per SIMD.
template <bool MIXED>
__global__ __launch_bounds__(kWalkWaves* kWave)
__attribute__((amdgpu_waves_per_eu(MIXED ? 5 : 4, 8))) void bh_walk_kernel(
    const double4* __restrict__ src, const uint4* __restrict__ meta, int n,
    const uint2* __restrict__ groups, const uint32_t* __restrict__ ngroups,
    const double* __restrict__ box, double theta2, double eps2_all,
    const double* __restrict__ eps2_t, uint32_t* __restrict__ frontier,
    uint32_t fcap, uint32_t* __restrict__ counter, uint32_t* __restrict__ flag,
    double4* __restrict__ part, unsigned long long* __restrict__ stats) {
    // LDS per block: 32 KB (mixed: 768-entry lists, FP32 stages) for five
    // blocks per CU, 40 KB (FP64) for four
    constexpr uint32_t kList = MIXED ? kListCapMixed : kListCap;
    __shared__ uint32_t lists[kWalkWaves][kList];
    __shared__ uint32_t fronts[kWalkWaves][2][kLdsFront];
    __shared__ __align__(16) unsigned char stage_mem[kWalkWaves][kWave * (MIXED ? 16 : 32)];
    unsigned long long t_nodes = 0, t_acc = 0, t_leafp = 0, t_pairs = 0;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    uint32_t* list = lists[w];
    double4* stage = reinterpret_cast<double4*>(stage_mem[w]);
    float4* stage32 = reinterpret_cast<float4*>(stage_mem[w]);
    const size_t slot = (size_t)blockIdx.x * kWalkWaves + w;
    // level buffer b: entry p < kLdsFront in fronts[w][b], the rest in
    // the wave's global buffer at the same index
    uint32_t* const fg0 = frontier + slot * 2 * (size_t)fcap;
    uint32_t* const fg1 = fg0 + fcap;
    const double side0 = 2.0 * box[3];  // root side, octree.rs:81-85
    // a group walk gives up at a smaller frontier than a one-target walk:
    // the groups that reach it (outlying targets whose box spans the core)
    // are cheaper walked one target at a time than breadth first
    const uint32_t gcap = min(fcap, kGroupFrontierCap);

    const uint32_t ng = __builtin_amdgcn_readfirstlane(ngroups[0]);
    // groups are Morton ordered: each XCD drains its own contiguous range
    // first (L2 reuse of the nodes and leaves near it), then helps the others
    const uint32_t home = blockIdx.x % kXcds;
    uint32_t cx = home;
    for (;;) {
        const uint32_t g = xcd_take(counter, ng, home, cx, lane);
        if (g >= ng) break;
        const uint2 gr = groups[g];
        const int i0 = (int)gr.x;
        const int cnt = (int)gr.y;
        // A group of at most 32 (16) targets is held twice (four times)
        // across the wave; replica r sums the sources at list positions
        // p % R == r (FP64) or a block of each round (mixed), and the
        // replicas are added in a fixed order at the end.
        const int R = cnt <= 16 ? 4 : (cnt <= 32 ? 2 : 1);
        const int tpr = kWave / R;          // lanes per replica
        const int tl = lane % tpr;          // this lane's target
        const int rep = lane / tpr;         // this lane's replica
        const bool active = tl < cnt;
        const double4 me = src[i0 + (active ? tl : 0)];
        // softening of this lane's target: one value, or one per target
        const double eps2 = eps2_t ? eps2_t[i0 + (active ? tl : 0)] : eps2_all;
        const float eps2f = in_vgpr((float)eps2);  // full-rate FMA operand
        const bool soft = __ballot(!(eps2f >= kMaskEps2F)) == 0ull;  // wave uniform
        double glo[3] = {me.x, me.y, me.z}, ghi[3] = {me.x, me.y, me.z};
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                glo[d] = fmin(glo[d], __shfl_xor(glo[d], o));
                ghi[d] = fmax(ghi[d], __shfl_xor(ghi[d], o));
            }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            glo[d] = uniform(glo[d]);
            ghi[d] = uniform(ghi[d]);
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
                } else if (MIXED) {
                    // past the round's end: zero mass (adds exactly zero)
                    stage32[lane] = make_float4(0.f, 0.f, 0.f, 0.f);
                } else {
                    // past the round's end: a zero-mass copy of the round's
                    // record 0 (adds exactly zero: that pair is evaluated
                    // anyway, so r^2 is non-zero and finite or skipped)
                    double4 r = src[list[q0]];
                    r.w = 0.0;
                    stage[lane] = r;
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                __builtin_amdgcn_wave_barrier();
                const uint32_t m = min(nl - q0, (uint32_t)kWave);
                if (MIXED) {
                    // replica rep takes the records [rep per, (rep + 1) per)
                    // in order, four per iteration, per = ceil(m / R) rounded
                    // up to a multiple of 4: R (1, 2 or 4) divides 64/4, so
                    // R per <= 64, and the records past m have zero mass
                    const uint32_t per = __builtin_amdgcn_readfirstlane(
                        (((m + (uint32_t)R - 1) / (uint32_t)R) + 3u) & ~3u);
                    const float4* mine = stage32 + (uint32_t)rep * per;
                    float tx = 0.f, ty = 0.f, tz = 0.f;
                    auto sweep = [&](auto mask) {
                        constexpr bool MK = decltype(mask)::value;
#pragma unroll 1
                        for (uint32_t i = 0; i < per; i += 4) {
                            float4 r[4];
#pragma unroll
                            for (int j = 0; j < 4; ++j) r[j] = mine[i + (uint32_t)j];
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                const float dx = r[j].x - fx, dy = r[j].y - fy,
                                            dz = r[j].z - fz;
                                const float f = pair_f32<MK>(dx, dy, dz, r[j].w, eps2f);
                                tx = __builtin_fmaf(f, dx, tx);
                                ty = __builtin_fmaf(f, dy, ty);
                                tz = __builtin_fmaf(f, dz, tz);
                            }
                        }
                    };
                    // every softening at least kMaskEps2F: the self pair is finite, no mask
                    if (soft) sweep(std::false_type{});
                    else sweep(std::true_type{});
                    if (keep) {
                        float y = tx - kx, t = sx + y; kx = (t - sx) - y; sx = t;
                        y = ty - ky; t = sy + y; ky = (t - sy) - y; sy = t;
                        y = tz - kz; t = sz + y; kz = (t - sz) - y; sz = t;
                    }
                } else {
                    // replica rep takes records rep, rep + R, ... in order
                    // (the oracle's order), per = ceil(m / R) rounded up to a
                    // multiple of kU of them: R divides 64/kU, so the last
                    // index stays below 64, and the records past m have
                    // zero mass
                    constexpr uint32_t kU = 2;
                    const uint32_t per = __builtin_amdgcn_readfirstlane(
                        (((m + (uint32_t)R - 1) / (uint32_t)R) + kU - 1) & ~(kU - 1));
                    double tx = ax, ty = ay, tz = az;
#pragma unroll 1
                    for (uint32_t i = 0; i < per; i += kU) {
                        double4 r[kU];
#pragma unroll
                        for (uint32_t j = 0; j < kU; ++j)
                            r[j] = stage[(i + j) * (uint32_t)R + (uint32_t)rep];
#pragma unroll
                        for (uint32_t j = 0; j < kU; ++j)
                            pair_ref(r[j], me.x, me.y, me.z, eps2, tx, ty, tz);
                    }
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
                if (nl == kList) flush();
                const uint32_t k = min(count, kList - nl);
                for (uint32_t r = lane; r < k; r += kWave) list[nl + r] = first + r;
                nl += k;
                first += k;
                count -= k;
            }
        };
        // Breadth-first walk with opening box (lo, hi) for `ntargets`
        // targets; returns true if a next frontier would exceed cap.
        auto walk = [&](const double (&lo)[3], const double (&hi)[3], int ntargets,
                        uint32_t cap) -> bool {
            int cur = 0;  // wave-uniform: level buffer being read
            if (lane == 0) fronts[w][0][0] = 0u;
            uint32_t nf = 1;
            bool overflow = false;
            __threadfence_block();
            while (nf > 0 && !overflow) {
                uint32_t nn = 0;
                for (uint32_t base = 0; base < nf; base += kWave) {
                    const uint32_t idx = base + lane;
                    const bool valid = idx < nf;
                    t_nodes += min(nf - base, (uint32_t)kWave);
                    const uint32_t* fl = fronts[w][cur];
                    const uint32_t* fgr = cur ? fg1 : fg0;
                    // base is a multiple of 64: the chunk is in one place
                    const uint32_t v = !valid ? 0u : (base < kLdsFront ? fl[idx] : fgr[idx]);
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
                    const uint32_t sc = wave_inclusive_scan(add);
                    const uint32_t total = __builtin_amdgcn_readlane(sc, kWave - 1);
                    const uint64_t accb = __ballot(accept);
                    t_acc += (unsigned long long)__popcll(accb);
                    t_leafp += total - (uint32_t)__popcll(accb);
                    t_pairs += (unsigned long long)total * (unsigned long long)ntargets;
                    if (total > 0) {
                        if (nl + total > kList) flush();
                        if (total <= kList) {
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
                    const uint32_t kc = wave_inclusive_scan(k);
                    const uint32_t ktot = __builtin_amdgcn_readlane(kc, kWave - 1);
                    if (nn + ktot > cap) {
                        overflow = true;
                        break;
                    }
                    const uint32_t ex0 = nn + kc - k;
                    uint32_t* flw = fronts[w][cur ^ 1];
                    uint32_t* fgw = cur ? fg0 : fg1;
                    for (uint32_t q = 0; q < k; ++q) {
                        const uint32_t pos = ex0 + q;
                        if (pos < kLdsFront) flw[pos] = md.z + q;
                        else fgw[pos] = md.z + q;
                    }
                    nn += ktot;
                }
                __threadfence_block();
                __builtin_amdgcn_wave_barrier();
                cur ^= 1;
                nf = nn;
            }
            flush();
            return overflow;
        };

        if (walk(glo, ghi, cnt, gcap)) {
            // the group box opens too much: walk every target alone
            ax = ay = az = 0.0;
            sx = sy = sz = kx = ky = kz = 0.f;
            for (int l = 0; l < cnt; ++l) {
                const double px = __shfl(me.x, l), py = __shfl(me.y, l), pz = __shfl(me.z, l);
                const double p[3] = {px, py, pz};
                keep = tl == l;
                if (walk(p, p, 1, fcap) && lane == 0) atomicOr(flag, 1u);
            }
            keep = true;
        }
        if (MIXED) {
            ax = (double)sx - (double)kx;
            ay = (double)sy - (double)ky;
            az = (double)sz - (double)kz;
        }
        // replicas in a fixed pairing: (r0 + r1) [+ (r2 + r3)]; each sum is
        // formed identically in both lanes (addition commutes)
        for (int o = tpr; o < kWave; o <<= 1) {
            ax += __shfl_xor(ax, o);
            ay += __shfl_xor(ay, o);
            az += __shfl_xor(az, o);
        }
        if (active && rep == 0) part[(size_t)i0 + tl] = make_double4(ax, ay, az, 0.0);
    }
    if (lane == 0) {
        atomicAdd(&stats[0], t_nodes);
        atomicAdd(&stats[1], t_acc);
        atomicAdd(&stats[2], t_leafp);
        atomicAdd(&stats[3], t_pairs);
    }
}

}  // namespace

gravis_status bh_walk(Ctx* c, const gravis_solver* s, const Octree& t, const FinishArgs& fa,
                      const double* eps2_sorted) {
    const size_t n = t.n;
    const int ni = (int)n;
    hipStream_t st = c->stream;
    // resident blocks per CU: 32 KB of LDS each (mixed), 40 KB (FP64)
    const int per_cu = s->precision == GRAVIS_MIXED ? 5 : 4;
    const int blocks = (int)std::min<long>((long)(n / 64 / kWalkWaves) + 1,
                                           (long)c->num_cus * per_cu);
    const uint32_t fcap = (uint32_t)std::min<size_t>(t.cap, c->bh_frontier_cap);
    uint32_t *front, *counter;
    GRAVIS_TRY(buf(c, "bh_front", (size_t)blocks * kWalkWaves * 2 * fcap, &front));
    GRAVIS_TRY(buf(c, "bh_counter", kXcds, &counter));
    double4* part;
    GRAVIS_TRY(buf(c, "bh_part4", n, &part));
    const double theta2 = s->theta * s->theta;
    const double eps2 = s->softening * s->softening;
    unsigned long long* stats;
    GRAVIS_TRY(buf(c, "bh_stats", 4, &stats));
    walk_reset_kernel<<<1, 64, 0, st>>>(counter, stats);
    timer_begin(c);
    if (s->precision == GRAVIS_MIXED)
        bh_walk_kernel<true><<<blocks, kWalkWaves * kWave, 0, st>>>(
            t.src, t.meta, ni, t.groups, t.ngroups, t.box, theta2, eps2, eps2_sorted, front, fcap,
            counter, t.flags, part, stats);
    else
        bh_walk_kernel<false><<<blocks, kWalkWaves * kWave, 0, st>>>(
            t.src, t.meta, ni, t.groups, t.ngroups, t.box, theta2, eps2, eps2_sorted, front, fcap,
            counter, t.flags, part, stats);
    timer_end(c, "bh_walk");
    GRAVIS_TRY(check_launch(c, "bh_walk"));
    uint32_t hflags[2] = {0, 0};
    GRAVIS_HIP(c, hipMemcpyAsync(hflags, t.flags, sizeof hflags, hipMemcpyDeviceToHost, st));
    GRAVIS_HIP(c, hipMemcpyAsync(c->bh_stats, stats, 4 * sizeof(unsigned long long),
                                 hipMemcpyDeviceToHost, st));
    uint32_t nodes = 0;
    GRAVIS_HIP(c, hipMemcpyAsync(&nodes, t.lvl + 2 * (kTreeMaxLevel + 1), sizeof nodes,
                                 hipMemcpyDeviceToHost, st));
    GRAVIS_HIP(c, hipStreamSynchronize(st));
    c->bh_stats[4] = nodes;
    if (hflags[1])
        return fail(c, GRAVIS_INTERNAL, "octree node capacity exceeded (%zu nodes)", t.cap);
    if (hflags[0])
        return fail(c, GRAVIS_INTERNAL,
                    "Barnes-Hut walk frontier of a single target exceeded %u nodes (theta %g)",
                    fcap, s->theta);
    return finish_forces4(c, part, ni, t.inv, fa);
}

gravis_status bh_force(Ctx* c, const gravis_solver* s, size_t n, const double* x,
                       const double* y, const double* z, const double* m, const FinishArgs& fa) {
    if (n == 0) return GRAVIS_OK;
    const uint32_t leaf = s->leaf_size ? s->leaf_size : 16u;
    const uint32_t gs = s->group_size ? s->group_size : c->bh_group;
    Octree t;
    GRAVIS_TRY(build_octree(c, n, x, y, z, m, leaf, gs, &t, true));
    return bh_walk(c, s, t, fa);
}

}  // namespace gravis
