// sph.hip -- grad-h smoothed particle hydrodynamics on gfx950.
//
// Reference: SphSolver::compute / compute_timestep and step_with_sph
// (crates/sim-core/src/sph/solver.rs:61-212), compute_density
// (sph/density.rs:42-157), compute_forces (sph/forces.rs:69-232), the
// Wendland C2 kernel (sph/kernel.rs:17-108), the neighbour lists of
// sph/neighbors.rs:56-119 and the ball query of octree.rs:245-308.
//
// Layout. Particles are handled in the Morton order of the linear octree
// (tree.hip); per sorted particle s the kernels keep
//   pos[s]  = (x, y, z, m)            (the tree's source records)
//   vh[s]   = (vx, vy, vz, h)
//   aux[s]  = (rho, p, c, alpha)
//   aux2[s] = (omega, u)
// and a neighbour list {start, count} into one index pool, as the
// reference's flat NeighborList (neighbors.rs:13-19); a re-query appends a
// new list to the pool (NeighborList::replace_neighbors, neighbors.rs:36-44).
//
// Neighbour search. One wavefront per tree target group: the group's box
// (over the targets that need a query) grown by the largest query radius is
// tested against the tree's per-node particle boxes breadth first; the
// leaves that overlap are sorted by first particle, and their particles are
// staged in LDS 64 at a time and tested by every lane against its own target
// with the reference's exact test |x_j - x_i|^2 <= R^2 (octree.rs:275-283).
// Lists therefore come out in ascending Morton order, the order of the
// reference's depth-first octant walk. The box test is conservative in
// floating point (every rounding is monotone), so no neighbour is lost. A
// group whose frontier or leaf list exceeds its LDS capacity is searched
// again one target at a time by a depth-first walk with an LDS stack.
//
// Density. One lane per gas particle runs the reference's iteration
// (<= 5 passes, tolerance 1e-3) over its list; when the reference would
// re-query (2 h_new > 2.2 h) the lane records the new radius and stops; the
// host then re-queries those particles and resumes them, until none is left.
//
// Arithmetic is FP64 in the reference's operation order (no contraction;
// division and sqrt correctly rounded); only values that depend on h alone
// (sigma / h^3, -3 / h) are hoisted out of the pair loops.
#include <algorithm>
#include <cmath>

#include "internal.h"

namespace gravis {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kNorm3d = 21.0 / (16.0 * kPi);  // kernel.rs:17
constexpr double kDensityFloor = 1e-10;          // forces.rs:12
constexpr double kEnergyFloor = 1e-10;           // solver.rs:13
constexpr double kDtFloor = 1e-6;                // solver.rs:17
constexpr double kPad = 1.1;                     // density.rs:52
constexpr double kCsNorm = 1.0 / kPi;            // cubic spline M4, support 2h
constexpr uint8_t kGas = 4;                      // particle.rs:9

// per-particle progress of the density phase
constexpr uint8_t kQNone = 0;     // not gas
constexpr uint8_t kQQuery = 1;    // needs a (re-)query at radius[s]
constexpr uint8_t kQDensity = 2;  // list ready, density iteration pending
constexpr uint8_t kQDone = 3;

// control words: [0] pool entries used, [1] pool overflow, [2] groups sent to
// the depth-first search, [3] re-queries requested, [4] target ranges
// searched one target at a time
constexpr int kCtlWords = 8;

constexpr int kQWaves = 4;
constexpr uint32_t kQFront = 512;   // breadth-first frontier (nodes) per wave
constexpr uint32_t kQRanges = 512;  // overlapping leaves per wave
constexpr uint32_t kQStack = 256;   // depth-first stack (>= 1 + 7 * 22)
// candidate rounds (of 64) whose hit masks stay in registers; their
// candidate table (4 B per entry) is written over the leaf ranges, which
// are not needed once the candidates are counted
constexpr uint32_t kQMaskRounds = 16;
static_assert(kQMaskRounds * 64 * 4 <= kQRanges * 8, "candidate table must fit the leaf buffer");

__device__ __forceinline__ void lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, int lane, uint32_t* total) {
    (void)lane;
    const uint32_t s = wave_inclusive_scan(v);
    *total = __builtin_amdgcn_readlane(s, kWave - 1);
    return s - v;
}

// kernel.rs:44-62
__device__ __forceinline__ void grad_w(double dx, double dy, double dz, double r, double h,
                                       double& gx, double& gy, double& gz) {
    gx = gy = gz = 0.0;
    if (r < 1e-30) return;
    const double q = r / h;
    if (q >= 2.0) return;
    const double h3 = h * h * h;
    const double s = 1.0 - 0.5 * q;
    const double s3 = s * s * s;
    const double dwdq = kNorm3d / h3 * (-5.0 * q) * s3;
    const double f = dwdq / (h * r);
    gx = f * dx;
    gy = f * dy;
    gz = f * dz;
}

// kernel.rs:93-107
__device__ __forceinline__ double abs_grad_w(double r, double h) {
    if (r < 1e-30) return 0.0;
    const double q = r / h;
    if (q >= 2.0) return 0.0;
    const double h3 = h * h * h;
    const double s = 1.0 - 0.5 * q;
    const double s3 = s * s * s;
    const double dwdq = kNorm3d / h3 * (-5.0 * q) * s3;
    return -dwdq / h;
}

// Cubic spline M4 shape f(q) and f'(q) (support 2h): 1 - 1.5q^2 + 0.75q^3
// for q < 1, 0.25 (2 - q)^3 for 1 <= q < 2.
template <typename T>
__device__ __forceinline__ void cs_f_df(T q, T& f, T& df) {
    if (q < T(1)) {
        f = T(1) - T(1.5) * q * q + T(0.75) * q * q * q;
        df = T(-3) * q + T(2.25) * q * q;
    } else if (q < T(2)) {
        const T t = T(2) - q;
        f = T(0.25) * t * t * t;
        df = T(-0.75) * t * t;
    } else {
        f = T(0);
        df = T(0);
    }
}

// ---------------------------------------------------------------------------
// Gather of the caller's arrays into sorted records, and the first query
// radius 2 h 1.1 of every gas particle (neighbors.rs:67-70 with
// density.rs:52).
__global__ __launch_bounds__(256) void sph_prep_kernel(
    int n, const uint32_t* __restrict__ inv, const double* __restrict__ vx,
    const double* __restrict__ vy, const double* __restrict__ vz, const double* __restrict__ h,
    const double* __restrict__ rho, const double* __restrict__ pr, const double* __restrict__ cs,
    const double* __restrict__ alpha, const double* __restrict__ u,
    const uint8_t* __restrict__ type, double4* __restrict__ vh, double4* __restrict__ aux,
    double2* __restrict__ aux2, uint8_t* __restrict__ qstate, uint8_t* __restrict__ iter,
    double* __restrict__ hcur, double* __restrict__ radius, uint32_t* __restrict__ ctl) {
    // caller particle p (reads in the caller's order) -> sorted slot s
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0)
        for (int k = 0; k < kCtlWords; ++k) ctl[k] = 0u;
    if (p >= n) return;
    const uint32_t s = inv[p];
    const double hp = h[p];
    vh[s] = make_double4(vx[p], vy[p], vz[p], hp);
    aux[s] = make_double4(rho[p], pr[p], cs[p], alpha[p]);
    aux2[s] = make_double2(1.0, u[p]);
    const bool gas = type[p] == kGas;
    qstate[s] = gas ? kQQuery : kQNone;
    iter[s] = 0;
    hcur[s] = hp;
    radius[s] = 2.0 * hp * kPad;
}

// Stages candidates [c0, c0 + 64) of the group's leaf ranges (rg: first
// particle, exclusive offset; nr ranges, total candidates), keeping only
// those within sqrt(R2g) of the group box [lo, hi] (no target's exact test
// can accept the others: every rounding in the box distance is monotone and
// R2g is the largest R^2), compacted in order: stage[k] = (x, y, z, index)
// for k < the returned count.
__device__ __forceinline__ uint32_t stage_round(uint32_t c0, uint32_t total, uint32_t nr,
                                                const uint2* rg,
                                                const double4* __restrict__ pos,
                                                const double (&lo)[3], const double (&hi)[3],
                                                double R2g, int lane, uint64_t below,
                                                double4* stage) {
    const uint32_t c = c0 + (uint32_t)lane;
    bool keep = false;
    double4 pj = make_double4(0.0, 0.0, 0.0, 0.0);
    uint32_t j = 0;
    if (c < total) {
        uint32_t a = 0, b = nr;  // last leaf whose offset <= c
        while (b - a > 1) {
            const uint32_t mid = (a + b) >> 1;
            if (rg[mid].y <= c) a = mid;
            else b = mid;
        }
        j = rg[a].x + (c - rg[a].y);
        pj = pos[j];
        const double ex = fmax(fmax(lo[0] - pj.x, pj.x - hi[0]), 0.0);
        const double ey = fmax(fmax(lo[1] - pj.y, pj.y - hi[1]), 0.0);
        const double ez = fmax(fmax(lo[2] - pj.z, pj.z - hi[2]), 0.0);
        keep = ex * ex + ey * ey + ez * ez <= R2g;
    }
    const uint64_t kb = __ballot(keep);
    if (keep) stage[__popcll(kb & below)] = make_double4(pj.x, pj.y, pj.z, (double)j);
    lds_sync();
    return (uint32_t)__popcll(kb);
}

// Breadth-first group query of one target group (see the file comment).
// Targets with qstate == kQQuery get a list {start, count} in nbr and move
// on to kQDensity (kQDone if their iteration already ended); a group that
// exceeds the LDS capacities is appended to dfs_groups instead. All control
// flow is wave uniform: g and every count are scalar values.
// Returns false (and writes nothing) when the search exceeds the LDS
// capacities; the caller then splits the range or searches depth first.
__device__ __forceinline__ bool query_range(
    uint2 gr, int lane, const double4* __restrict__ pos, const uint4* __restrict__ meta,
    const double4* __restrict__ bmin, const double4* __restrict__ bmax,
    uint8_t* __restrict__ qstate, const uint8_t* __restrict__ iter,
    const double* __restrict__ radius, uint2* __restrict__ nbr, uint32_t* __restrict__ pool,
    uint32_t pool_cap, uint32_t* __restrict__ ctl, uint32_t rcap, int maxit, uint32_t* fa,
    uint32_t* fb, uint2* rg, double4* stage) {
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t s = gr.x + (uint32_t)lane;
    const bool inside = (uint32_t)lane < gr.y;
    const bool act = inside && qstate[s] == kQQuery;
    if (__ballot(act) == 0ull) return true;
    const double4 me = pos[inside ? s : gr.x];
    const double R = act ? radius[s] : 0.0;
    const double R2 = R * R;
    double lo[3], hi[3];
    lo[0] = act ? me.x : INFINITY; hi[0] = act ? me.x : -INFINITY;
    lo[1] = act ? me.y : INFINITY; hi[1] = act ? me.y : -INFINITY;
    lo[2] = act ? me.z : INFINITY; hi[2] = act ? me.z : -INFINITY;
    double R2g = act ? R2 : 0.0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            lo[d] = fmin(lo[d], __shfl_xor(lo[d], o));
            hi[d] = fmax(hi[d], __shfl_xor(hi[d], o));
        }
        R2g = fmax(R2g, __shfl_xor(R2g, o));
    }

    // breadth-first search for the overlapping leaves
    if (lane == 0) fa[0] = 0u;
    lds_sync();
    uint32_t nf = 1, nr = 0;
    bool ovf = rcap == 0;  // 0: every group searched depth first
    while (nf > 0 && !ovf) {
        uint32_t nn = 0;
        for (uint32_t base = 0; base < nf; base += kWave) {
            const uint32_t idx = base + (uint32_t)lane;
            const bool valid = idx < nf;
            const uint32_t v = valid ? fa[idx] : 0u;
            const uint4 md = meta[v];
            const double4 a = bmin[v], b = bmax[v];
            const double ex = fmax(fmax(a.x - hi[0], lo[0] - b.x), 0.0);
            const double ey = fmax(fmax(a.y - hi[1], lo[1] - b.y), 0.0);
            const double ez = fmax(fmax(a.z - hi[2], lo[2] - b.z), 0.0);
            const double d2 = ex * ex + ey * ey + ez * ez;
            const bool hit = valid && d2 <= R2g;
            const uint32_t nchild = (md.w >> 8) & 255u;
            // a node is kept only if some target's own ball reaches its box
            // (the box test is conservative, as above): a node that only the
            // grown group box reaches holds no neighbour of any target
            uint64_t hb = __ballot(hit);
            {
                uint64_t todo = hb, keep = 0ull;
                while (todo) {
                    const int l = __ffsll((long long)todo) - 1;
                    todo &= todo - 1ull;
                    const double ax = __shfl(a.x, l), ay = __shfl(a.y, l), az = __shfl(a.z, l);
                    const double bx = __shfl(b.x, l), by = __shfl(b.y, l), bz = __shfl(b.z, l);
                    const double fx = fmax(fmax(ax - me.x, me.x - bx), 0.0);
                    const double fy = fmax(fmax(ay - me.y, me.y - by), 0.0);
                    const double fz = fmax(fmax(az - me.z, me.z - bz), 0.0);
                    if (__ballot(act && fx * fx + fy * fy + fz * fz <= R2) != 0ull)
                        keep |= 1ull << l;
                }
                hb = keep;
            }
            const bool kept = ((hb >> lane) & 1ull) != 0ull;
            const bool leafhit = kept && nchild == 0;
            const uint64_t lb = __ballot(leafhit);
            const uint32_t nl = (uint32_t)__popcll(lb);
            if (nr + nl > rcap) {
                ovf = true;
                break;
            }
            if (leafhit) rg[nr + (uint32_t)__popcll(lb & below)] = make_uint2(md.x, md.y);
            nr += nl;
            const uint32_t k = (kept && nchild > 0) ? nchild : 0u;
            uint32_t ktot;
            const uint32_t kx = wave_excl_scan(k, lane, &ktot);
            if (nn + ktot > kQFront) {
                ovf = true;
                break;
            }
            for (uint32_t q = 0; q < k; ++q) fb[nn + kx + q] = md.z + q;
            nn += ktot;
        }
        lds_sync();
        uint32_t* t = fa;
        fa = fb;
        fb = t;
        nf = nn;
    }
    if (ovf) return false;

    // sort the leaves by first particle (bitonic, one wave)
    uint32_t np = 1;
    while (np < nr) np <<= 1;
    for (uint32_t i = nr + (uint32_t)lane; i < np; i += kWave) rg[i] = make_uint2(0xFFFFFFFFu, 0u);
    lds_sync();
    for (uint32_t k = 2; k <= np; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = (uint32_t)lane; i < np; i += kWave) {
                const uint32_t ij = i ^ j;
                if (ij > i) {
                    const uint2 x = rg[i], y = rg[ij];
                    const bool up = (i & k) == 0;
                    if ((x.x > y.x) == up) {
                        rg[i] = y;
                        rg[ij] = x;
                    }
                }
            }
            lds_sync();
        }
    }
    // counts -> exclusive offsets
    uint32_t total = 0;
    for (uint32_t base = 0; base < nr; base += kWave) {
        const uint32_t i = base + (uint32_t)lane;
        const uint32_t cnt = i < nr ? rg[i].y : 0u;
        uint32_t t;
        const uint32_t ex = wave_excl_scan(cnt, lane, &t);
        if (i < nr) rg[i].y = total + ex;
        total += t;
    }
    lds_sync();

    uint32_t mine = 0, wpos = 0;
    bool failed = false;
    const uint32_t nrounds = (total + kWave - 1) / kWave;
    // reserves this lane's list in the pool (wave-wide prefix of the counts)
    auto allocate = [&]() -> bool {
        uint32_t gtot;
        const uint32_t ex = wave_excl_scan(act ? mine : 0u, lane, &gtot);
        uint32_t base = 0;
        if (lane == 0) {
            base = atomicAdd(&ctl[0], gtot);
            if ((uint64_t)base + gtot > (uint64_t)pool_cap) atomicOr(&ctl[1], 1u);
        }
        base = __builtin_amdgcn_readfirstlane(base);
        if ((uint64_t)base + gtot > (uint64_t)pool_cap) return false;
        wpos = base + ex;
        if (act) nbr[s] = make_uint2(wpos, mine);
        return true;
    };
    if (nrounds <= kQMaskRounds) {
        // one pass over the candidates: per round of 64, each lane keeps the
        // mask of its hits and the candidate it staged; the lists are then
        // written from the masks through a table of the candidates
        uint64_t hit[kQMaskRounds];
        uint32_t cj[kQMaskRounds];
#pragma unroll
        for (uint32_t q = 0; q < kQMaskRounds; ++q) {
            hit[q] = 0ull;
            cj[q] = 0u;
            if (q < nrounds) {
                const uint32_t m = stage_round(q * kWave, total, nr, rg, pos, lo, hi, R2g, lane,
                                               below, stage);
                cj[q] = (uint32_t)lane < m ? (uint32_t)stage[lane].w : 0u;
                if (act) {
                    for (uint32_t k = 0; k < m; ++k) {
                        const double4 c4 = stage[k];
                        const double dx = c4.x - me.x, dy = c4.y - me.y, dz = c4.z - me.z;
                        if (dx * dx + dy * dy + dz * dz <= R2) hit[q] |= 1ull << k;
                    }
                }
                mine += (uint32_t)__popcll(hit[q]);
                lds_sync();
            }
        }
        failed = !allocate();
        if (!failed) {
            uint32_t* cand = reinterpret_cast<uint32_t*>(rg);  // kQMaskRounds * 64 entries
#pragma unroll
            for (uint32_t q = 0; q < kQMaskRounds; ++q)
                if (q < nrounds) cand[q * kWave + (uint32_t)lane] = cj[q];
            lds_sync();
            if (act) {
#pragma unroll
                for (uint32_t q = 0; q < kQMaskRounds; ++q) {
                    uint64_t mk = hit[q];
                    while (mk) {
                        const uint32_t b = (uint32_t)__ffsll((long long)mk) - 1u;
                        mk &= mk - 1ull;
                        pool[wpos++] = cand[q * kWave + b];
                    }
                }
            }
            lds_sync();
        }
        if (act && !failed) qstate[s] = iter[s] >= maxit ? kQDone : kQDensity;
        return true;
    }
    // two passes over the candidates: count, then write
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            uint32_t gtot;
            const uint32_t ex = wave_excl_scan(act ? mine : 0u, lane, &gtot);
            uint32_t base = 0;
            if (lane == 0) {
                base = atomicAdd(&ctl[0], gtot);
                if ((uint64_t)base + gtot > (uint64_t)pool_cap) atomicOr(&ctl[1], 1u);
            }
            base = __builtin_amdgcn_readfirstlane(base);
            if ((uint64_t)base + gtot > (uint64_t)pool_cap) {
                failed = true;
                break;
            }
            wpos = base + ex;
            if (act) nbr[s] = make_uint2(wpos, mine);
        }
        for (uint32_t q0 = 0; q0 < total; q0 += kWave) {
            const uint32_t m = stage_round(q0, total, nr, rg, pos, lo, hi, R2g, lane, below, stage);
            if (act) {
                for (uint32_t k = 0; k < m; ++k) {
                    const double4 c = stage[k];
                    const double dx = c.x - me.x, dy = c.y - me.y, dz = c.z - me.z;
                    const double dsq = dx * dx + dy * dy + dz * dz;
                    if (dsq <= R2) {
                        if (pass == 0) ++mine;
                        else pool[wpos++] = (uint32_t)c.w;
                    }
                }
            }
            lds_sync();
        }
    }
    if (act && !failed) qstate[s] = iter[s] >= maxit ? kQDone : kQDensity;
    return true;
}

// 128 VGPRs at most: four wavefronts per SIMD, as many as the 40 KB of LDS
// per block allows (the search is bound by the latency of its node and
// candidate loads; three per SIMD measured about 10 % slower on C4)
__global__ __launch_bounds__(kQWaves* kWave) __attribute__((amdgpu_waves_per_eu(4, 8))) void sph_query_kernel(
    const double4* __restrict__ pos, const uint4* __restrict__ meta,
    const double4* __restrict__ bmin, const double4* __restrict__ bmax,
    const uint2* __restrict__ groups, const uint32_t* __restrict__ ngroups,
    uint32_t* __restrict__ counter, uint8_t* __restrict__ qstate,
    const uint8_t* __restrict__ iter, const double* __restrict__ radius,
    uint2* __restrict__ nbr, uint32_t* __restrict__ pool, uint32_t pool_cap,
    uint32_t* __restrict__ ctl, uint32_t* __restrict__ dfs_groups, uint32_t rcap, int maxit) {
    __shared__ uint32_t fronts[kQWaves][2][kQFront];
    __shared__ uint2 leaves[kQWaves][kQRanges];
    __shared__ double4 stages[kQWaves][kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    const uint32_t ng = __builtin_amdgcn_readfirstlane(ngroups[0]);
    for (;;) {
        // every lane takes part (lane 0 adds 1), so the fetch sits in uniform
        // control flow and g is a scalar
        const uint32_t g =
            __builtin_amdgcn_readfirstlane(atomicAdd(counter, lane == 0 ? 1u : 0u));
        if (g >= ng) break;
        const bool ok = query_range(groups[g], lane, pos, meta, bmin, bmax, qstate, iter,
                                    radius, nbr, pool, pool_cap, ctl, rcap, maxit, fronts[w][0],
                                    fronts[w][1], leaves[w], stages[w]);
        if (!ok && lane == 0) dfs_groups[atomicAdd(&ctl[2], 1u)] = g;
    }
}

// Groups the breadth-first kernel could not hold in LDS are split into
// ranges of kQSub targets, each searched breadth first again with its own
// (smaller) box; a range that still does not fit is searched depth first,
// one target at a time. Children are pushed in reverse digit order, so
// leaves are reached in ascending Morton order.
constexpr uint32_t kQSub = 8;

__global__ __launch_bounds__(kWave) void sph_query_dfs_kernel(
    const double4* __restrict__ pos, const uint4* __restrict__ meta,
    const double4* __restrict__ bmin, const double4* __restrict__ bmax,
    const uint2* __restrict__ groups, const uint32_t* __restrict__ dfs_groups,
    uint8_t* __restrict__ qstate, const uint8_t* __restrict__ iter,
    const double* __restrict__ radius, uint2* __restrict__ nbr, uint32_t* __restrict__ pool,
    uint32_t pool_cap, uint32_t* __restrict__ ctl, uint32_t rcap, int maxit) {
    __shared__ uint32_t stack[kQStack];
    __shared__ uint32_t fronts[2][kQFront];
    __shared__ uint2 leaves[kQRanges];
    __shared__ double4 stage[kWave];
    const int lane = threadIdx.x;
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t ndfs = ctl[2];
    constexpr uint32_t kSubs = kWave / kQSub;
    for (uint32_t item = blockIdx.x; item < ndfs * kSubs; item += gridDim.x) {
        const uint2 grp = groups[dfs_groups[item / kSubs]];
        const uint32_t off = (item % kSubs) * kQSub;
        if (off >= grp.y) continue;
        const uint2 gr = make_uint2(grp.x + off, min(kQSub, grp.y - off));
        if (query_range(gr, lane, pos, meta, bmin, bmax, qstate, iter, radius, nbr, pool,
                        pool_cap, ctl, rcap, maxit, fronts[0], fronts[1], leaves, stage))
            continue;
        if (lane == 0) atomicAdd(&ctl[4], 1u);
        for (uint32_t l = 0; l < gr.y; ++l) {
            const uint32_t s = gr.x + l;
            if (qstate[s] != kQQuery) continue;
            const double4 me = pos[s];
            const double R = radius[s];
            const double R2 = R * R;
            uint32_t cnt = 0, wpos = 0;
            bool failed = false;
            for (int pass = 0; pass < 2 && !failed; ++pass) {
                if (pass == 1) {
                    uint32_t base = 0;
                    if (lane == 0) {
                        base = atomicAdd(&ctl[0], cnt);
                        if ((uint64_t)base + cnt > (uint64_t)pool_cap) atomicOr(&ctl[1], 1u);
                    }
                    base = __builtin_amdgcn_readfirstlane(base);
                    if ((uint64_t)base + cnt > (uint64_t)pool_cap) {
                        failed = true;
                        break;
                    }
                    wpos = base;
                    if (lane == 0) nbr[s] = make_uint2(base, cnt);
                }
                uint32_t sp = 0;
                {
                    const double4 a = bmin[0], b = bmax[0];
                    const double ex = fmax(fmax(a.x - me.x, me.x - b.x), 0.0);
                    const double ey = fmax(fmax(a.y - me.y, me.y - b.y), 0.0);
                    const double ez = fmax(fmax(a.z - me.z, me.z - b.z), 0.0);
                    if (ex * ex + ey * ey + ez * ez <= R2) {
                        if (lane == 0) stack[0] = 0u;
                        sp = 1;
                    }
                }
                lds_sync();
                while (sp > 0) {
                    const uint32_t v = stack[sp - 1];
                    --sp;
                    lds_sync();
                    const uint4 md = meta[v];
                    const uint32_t nchild = (md.w >> 8) & 255u;
                    if (nchild == 0) {
                        for (uint32_t k0 = 0; k0 < md.y; k0 += kWave) {
                            const uint32_t k = k0 + (uint32_t)lane;
                            bool hit = false;
                            if (k < md.y) {
                                const double4 c = pos[md.x + k];
                                const double dx = c.x - me.x, dy = c.y - me.y, dz = c.z - me.z;
                                hit = dx * dx + dy * dy + dz * dz <= R2;
                            }
                            const uint64_t hb = __ballot(hit);
                            if (pass == 1 && hit)
                                pool[wpos + (uint32_t)__popcll(hb & below)] = md.x + k;
                            if (pass == 0) cnt += (uint32_t)__popcll(hb);
                            else wpos += (uint32_t)__popcll(hb);
                        }
                    } else {
                        bool acc = false;
                        const uint32_t cv = md.z + (uint32_t)lane;
                        if ((uint32_t)lane < nchild) {
                            const double4 a = bmin[cv], b = bmax[cv];
                            const double ex = fmax(fmax(a.x - me.x, me.x - b.x), 0.0);
                            const double ey = fmax(fmax(a.y - me.y, me.y - b.y), 0.0);
                            const double ez = fmax(fmax(a.z - me.z, me.z - b.z), 0.0);
                            acc = ex * ex + ey * ey + ez * ez <= R2;
                        }
                        const uint64_t ab = __ballot(acc);
                        const uint32_t na = (uint32_t)__popcll(ab);
                        if (acc) stack[sp + na - 1 - (uint32_t)__popcll(ab & below)] = cv;
                        sp += na;
                        lds_sync();
                    }
                }
            }
            if (!failed && lane == 0) qstate[s] = iter[s] >= maxit ? kQDone : kQDensity;
        }
    }
}

// Gathers of kGather list members: the indices first, then the records, so
// the dependent index -> record loads of a batch overlap. Members past the
// list end read record 0 and are skipped by the caller.
constexpr uint32_t kGather = 8;

__device__ __forceinline__ void gather_batch(const double4* __restrict__ pos,
                                             const uint32_t* __restrict__ pool, uint2 nb,
                                             uint32_t k0, double4 (&qb)[kGather]) {
    uint32_t jb[kGather];
#pragma unroll
    for (uint32_t e = 0; e < kGather; ++e) jb[e] = k0 + e < nb.y ? pool[nb.x + k0 + e] : 0u;
#pragma unroll
    for (uint32_t e = 0; e < kGather; ++e) qb[e] = pos[jb[e]];
}

// Everything the force pass reads of one particle, packed into one 128 B
// line (sorted order): position and mass, velocity and h, (rho, P, c,
// alpha), (Omega, u). Written once after the density pass, so a neighbour
// costs one cache line instead of four.
struct ForceRec {
    double4 p, v, a, b;  // b = (Omega, u, 0, 0)
};

__global__ __launch_bounds__(256) void sph_pack_kernel(int n, const double4* __restrict__ pos,
                                                       const double4* __restrict__ vh,
                                                       const double4* __restrict__ aux,
                                                       const double2* __restrict__ aux2,
                                                       ForceRec* __restrict__ rec) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const double2 b = aux2[s];
    ForceRec& r = rec[s];  // member stores: an aggregate temporary went to scratch
    r.p = pos[s];
    r.v = vh[s];
    r.a = aux[s];
    r.b = make_double4(b.x, b.y, 0.0, 0.0);
}

// The FP32-pair force pass reads a 64-byte record per neighbour: the FP64
// position (separations are formed in FP64) and the FP32 values its pair
// terms use, with the pressure term P / (Omega rho^2) formed once per
// particle (the expression the pair loop used to evaluate per pair, in the
// same FP32 operations). P and u, needed by conduction only, sit beside.
struct __align__(16) ForceRec32 {
    double x, y, z;
    float m, vx, vy, vz, h, rho, pr, c, alpha, spare;
};
static_assert(sizeof(ForceRec32) == 64, "one 64-byte record per particle");

__global__ __launch_bounds__(256) void sph_pack32_kernel(
    int n, const double4* __restrict__ pos, const double4* __restrict__ vh,
    const double4* __restrict__ aux, const double2* __restrict__ aux2,
    ForceRec32* __restrict__ rec, float2* __restrict__ cond) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const double4 p = pos[s], v = vh[s], a = aux[s];
    const double2 b = aux2[s];
    const float rho = fmaxf((float)a.x, (float)kDensityFloor);
    const float pf = (float)a.y, om = (float)b.x;
    ForceRec32 r;
    r.x = p.x; r.y = p.y; r.z = p.z;
    r.m = (float)p.w;
    r.vx = (float)v.x; r.vy = (float)v.y; r.vz = (float)v.z;
    r.h = (float)v.w;
    r.rho = rho;
    r.pr = pf / (om * rho * rho);
    r.c = (float)a.z;
    r.alpha = (float)a.w;
    r.spare = 0.0f;
    rec[s] = r;
    cond[s] = make_float2(pf, (float)b.y);
}

// As gather_batch for the force pass: kFGather members' index and packed
// record (kFGather32 of the 64-byte records).
constexpr uint32_t kFGather = 2;
constexpr uint32_t kFGather32 = 4;

__device__ __forceinline__ void gather_force_batch(
    const ForceRec* __restrict__ rec, const uint32_t* __restrict__ pool, uint2 nb, uint32_t k0,
    uint32_t (&jb)[kFGather], double4 (&pb)[kFGather], double4 (&vb)[kFGather],
    double4 (&ab)[kFGather], double2 (&bb)[kFGather]) {
#pragma unroll
    for (uint32_t e = 0; e < kFGather; ++e) jb[e] = k0 + e < nb.y ? pool[nb.x + k0 + e] : 0u;
#pragma unroll
    for (uint32_t e = 0; e < kFGather; ++e) {
        const ForceRec& r = rec[jb[e]];
        pb[e] = r.p;
        vb[e] = r.v;
        ab[e] = r.a;
        bb[e] = make_double2(r.b.x, r.b.y);
    }
}

// Density sums over a list at smoothing length h: rho = sum m_j W(r, h) and
// S = sum m_j dW/dh(r, h) over list members with r < 2h (density.rs:130-157).
// K = 0: Wendland C2 in sim-core's operation order (only sigma / h^3 and
// -3 / h hoisted); K = 1: cubic spline M4 in the oracle's order. F32: the
// separation is formed in FP64, rounded, and the pair terms are FP32 with
// reciprocals; the sums stay FP64.
template <int K, bool F32>
__device__ __forceinline__ void density_sums(const double4 me, double h,
                                             const double4* __restrict__ pos,
                                             const uint32_t* __restrict__ pool, uint2 nb,
                                             double& rho, double& dsum) {
    double rs = 0.0, ds = 0.0;
    if (!F32) {
        const double h2 = 2.0 * h;
        const double nh3 = (K == 0 ? kNorm3d : kCsNorm) / (h * h * h);
        const double nh4 = kCsNorm / (h * h * h * h);
        const double m3h = -3.0 / h;
        // loads in batches of kGather (indices, then records) so that the
        // gathers of a batch are in flight together; pairs in list order
        for (uint32_t k0 = 0; k0 < nb.y; k0 += kGather) {
            double4 qb[kGather];
            gather_batch(pos, pool, nb, k0, qb);
#pragma unroll
            for (uint32_t e = 0; e < kGather; ++e) {
            if (k0 + e >= nb.y) break;
            const double4 q4 = qb[e];
            const double dx = me.x - q4.x, dy = me.y - q4.y, dz = me.z - q4.z;
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (r < h2) {
                const double q = r / h;
                double wv = 0.0, dwh = 0.0;
                if (K == 0) {
                    if (q < 2.0) {
                        const double sq = 1.0 - 0.5 * q;
                        const double s3 = sq * sq * sq;
                        const double s4 = s3 * sq;
                        wv = nh3 * s4 * (1.0 + 2.0 * q);
                        const double dwdq = nh3 * (-5.0 * q) * s3;
                        dwh = m3h * wv - q / h * dwdq;
                    }
                } else {
                    double f, df;
                    cs_f_df(q, f, df);
                    wv = nh3 * f;
                    dwh = -nh4 * (3.0 * f + q * df);
                }
                rs += q4.w * wv;
                ds += q4.w * dwh;
            }
            }
        }
    } else {
        const float hf = (float)h;
        const float ih = 1.0f / hf;
        const float h2sq = 4.0f * hf * hf;
        const float nh3 = (float)(K == 0 ? kNorm3d : kCsNorm) * ih * ih * ih;
        const float nh4 = nh3 * ih;
        for (uint32_t k0 = 0; k0 < nb.y; k0 += kGather) {
            double4 qb[kGather];
            gather_batch(pos, pool, nb, k0, qb);
#pragma unroll
            for (uint32_t e = 0; e < kGather; ++e) {
            if (k0 + e >= nb.y) break;
            const double4 q4 = qb[e];
            const float dx = (float)(me.x - q4.x), dy = (float)(me.y - q4.y),
                        dz = (float)(me.z - q4.z);
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < h2sq) {
                const float r = __builtin_sqrtf(r2);
                const float q = r * ih;
                float wv, dwh;
                if (K == 0) {
                    const float sq = fmaxf(1.0f - 0.5f * q, 0.0f);
                    const float s3 = sq * sq * sq;
                    wv = nh3 * s3 * sq * (1.0f + 2.0f * q);
                    dwh = -nh4 * (3.0f * s3 * sq * (1.0f + 2.0f * q) - 5.0f * q * q * s3);
                } else {
                    float f, df;
                    cs_f_df(q, f, df);
                    wv = nh3 * f;
                    dwh = -nh4 * (3.0f * f + q * df);
                }
                const float mj = (float)q4.w;
                rs += (double)(mj * wv);
                ds += (double)(mj * dwh);
            }
            }
        }
    }
    rho = rs;
    dsum = ds;
}

struct DensityArgs {
    double gamma, eta, h_max, h_tol;
    int h_solver, max_it;
};

// The density / smoothing-length iteration of density.rs:58-107 (h_solver 0)
// or the Newton iteration of the variant (h_solver 1) for every particle in
// kQDensity, then the equation of state (density.rs:111-119).
template <int K, bool F32>
__global__ __launch_bounds__(256) void sph_density_kernel(
    int n, const double4* __restrict__ pos, const uint32_t* __restrict__ pool,
    const uint2* __restrict__ nbr, uint8_t* __restrict__ qstate, uint8_t* __restrict__ iter,
    double* __restrict__ hcur, double* __restrict__ radius, double4* __restrict__ vh,
    double4* __restrict__ aux, double2* __restrict__ aux2, DensityArgs A,
    uint32_t* __restrict__ ctl) {
    const int s = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    bool paused = false;
    if (s < n && qstate[s] == kQDensity) {
        const double4 me = pos[s];
        const double mi = me.w;
        const uint2 nb = nbr[s];
        double h = hcur[s];
        int it = iter[s];
        double omega = aux2[s].x;
        double rho = 0.0, hn = h;
        while (it < A.max_it) {
            double ds;
            density_sums<K, F32>(me, h, pos, pool, nb, rho, ds);
            if (A.h_solver == 0) {
                hn = rho > 0.0 ? fmin(A.eta * cbrt(mi / rho), A.h_max) : fmin(h, A.h_max);
            } else {
                const double e = A.eta / h;
                const double rt = mi * e * e * e;
                const double gp = ds + 3.0 * rt / h;
                double hx = h - (rho - rt) / gp;
                if (!(gp > 0.0) || !(rho > 0.0)) hx = rho > 0.0 ? A.eta * cbrt(mi / rho) : 2.0 * h;
                if (hx < 0.5 * h) hx = 0.5 * h;
                if (hx > 2.0 * h) hx = 2.0 * h;
                hn = fmin(hx, A.h_max);
            }
            if (rho > 0.0) {
                const double dh = -hn / (3.0 * rho);
                omega = 1.0 - dh * ds;
                if (omega < 0.1) omega = 0.1;
            }
            ++it;
            if (fabs(hn - h) / fmax(h, 1e-30) < A.h_tol) break;
            // sim-core compares with the current h (density.rs:96-97); the
            // variant with the radius the list was built with
            const double sr = A.h_solver == 0 ? 2.0 * h * kPad : radius[s];
            if (2.0 * hn > sr) {
                radius[s] = 2.0 * hn * kPad;
                h = hn;
                paused = true;
                break;
            }
            h = hn;
        }
        const double4 v = vh[s];
        vh[s] = make_double4(v.x, v.y, v.z, hn);
        const double u = aux2[s].y;
        aux2[s] = make_double2(omega, u);
        const double4 a = aux[s];
        const bool ended = !paused || it >= A.max_it;
        const double p = ended ? (A.gamma - 1.0) * rho * u : a.y;
        const double c = ended ? sqrt(A.gamma * (A.gamma - 1.0) * fmax(u, 0.0)) : a.z;
        aux[s] = make_double4(rho, p, c, a.w);
        hcur[s] = h;
        iter[s] = (uint8_t)it;
        qstate[s] = paused ? kQQuery : kQDone;
    }
    const uint64_t pb = __ballot(paused);
    if (pb != 0ull && (threadIdx.x & (kWave - 1)) == 0) atomicAdd(&ctl[3], (uint32_t)__popcll(pb));
}

struct ForceOut {
    double *ax, *ay, *az, *dudt;           // += (caller order)
    double *rho, *h, *p, *c;               // density phase results (gas)
    double *div_v, *max_mu, *dalpha;       // ForceIntermediates (assigned)
    double* omega;                         // may be null
    uint32_t* nneigh;                      // may be null
};

// Kernel gradient factor F with grad W = F d (d = x_i - x_j) in FP64:
// K = 0 sim-core's grad_w (kernel.rs:44-62), K = 1 the cubic spline in the
// oracle's order.
template <int K>
__device__ __forceinline__ double grad_factor(double r, double h) {
    if (K == 0) {
        if (r < 1e-30) return 0.0;
        const double q = r / h;
        if (q >= 2.0) return 0.0;
        const double h3 = h * h * h;
        const double sq = 1.0 - 0.5 * q;
        const double s3 = sq * sq * sq;
        const double dwdq = kNorm3d / h3 * (-5.0 * q) * s3;
        return dwdq / (h * r);
    } else {
        if (r < 1e-30) return 0.0;
        const double q = r / h;
        if (q >= 2.0) return 0.0;
        double f, df;
        cs_f_df(q, f, df);
        return kCsNorm / (h * h * h * h) * df / r;
    }
}

template <int K>
__device__ __forceinline__ double abs_grad(double r, double h) {
    if (K == 0) return abs_grad_w(r, h);
    if (r < 1e-30) return 0.0;
    const double q = r / h;
    if (q >= 2.0) return 0.0;
    double f, df;
    cs_f_df(q, f, df);
    return -kCsNorm / (h * h * h * h) * df;
}

// FP32 gradient factor from r, 1/r and 1/h.
template <int K>
__device__ __forceinline__ float grad_factor_f(float r, float ir, float ih) {
    const float q = r * ih;
    if (!(q < 2.0f) || r < 1e-30f) return 0.0f;
    const float ih2 = ih * ih;
    if (K == 0) {
        const float sq = 1.0f - 0.5f * q;
        return (float)kNorm3d * ih2 * ih2 * (-5.0f * q) * sq * sq * sq * ir;
    } else {
        float f, df;
        cs_f_df(q, f, df);
        return (float)kCsNorm * ih2 * ih2 * df * ir;
    }
}

// compute_forces (forces.rs:80-230), one lane per particle over its list.
// F32: separations formed in FP64 and rounded, pair terms in FP32, sums in
// FP64. alpha_switch 0: dalpha/dt = 0.
struct ForceIn {
    const ForceRec* rec;      // FP64 pairs
    const ForceRec32* rec32;  // FP32 pairs
    const float2* cond;       // FP32 pairs: (P, u)
    const double4 *pos, *vh, *aux;
    const double2* aux2;
};

template <int K, bool F32>
__global__ __launch_bounds__(256) void sph_force_kernel(
    int n, const uint32_t* __restrict__ perm, ForceIn in,
    const uint8_t* __restrict__ qstate, const uint32_t* __restrict__ pool,
    const uint2* __restrict__ nbr, gravis_sph_params P, ForceOut o) {
    const int s = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t pi = perm[s];
    if (qstate[s] == kQNone) {
        o.div_v[pi] = 0.0;
        o.max_mu[pi] = 0.0;
        o.dalpha[pi] = 0.0;
        if (o.omega) o.omega[pi] = 1.0;
        if (o.nneigh) o.nneigh[pi] = 0u;
        return;
    }
    const ForceRec* __restrict__ rec = in.rec;
    const double4 me = F32 ? in.pos[s] : rec[s].p;
    const double4 vi = F32 ? in.vh[s] : rec[s].v;
    const double4 ai = F32 ? in.aux[s] : rec[s].a;
    const double2 bi = F32 ? in.aux2[s] : make_double2(rec[s].b.x, rec[s].b.y);
    const double hi = vi.w;
    const double rho_i = fmax(ai.x, kDensityFloor);
    const double p_i = ai.y, c_i = ai.z, alpha_i = ai.w;
    const double omega_i = bi.x, u_i = bi.y;
    const double pr_i = rho_i > 0.0 ? p_i / (omega_i * rho_i * rho_i) : 0.0;
    double axs = 0.0, ays = 0.0, azs = 0.0, du = 0.0, dv = 0.0, mmu = 0.0;
    const uint2 nb = nbr[s];
    if (!F32) {
        for (uint32_t k0 = 0; k0 < nb.y; k0 += kFGather) {
            uint32_t jb[kFGather];
            double4 pb[kFGather], vb[kFGather], ab[kFGather];
            double2 bb[kFGather];
            gather_force_batch(rec, pool, nb, k0, jb, pb, vb, ab, bb);
#pragma unroll
            for (uint32_t e = 0; e < kFGather; ++e) {
            if (k0 + e >= nb.y) break;
            const uint32_t j = jb[e];
            if (j == (uint32_t)s) continue;
            const double4 pj = pb[e];
            const double dx = me.x - pj.x, dy = me.y - pj.y, dz = me.z - pj.z;
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            const double4 vj = vb[e];
            const double hj = vj.w;
            const double havg = 0.5 * (hi + hj);
            if (r >= 2.0 * hi && r >= 2.0 * hj) continue;
            const double mj = pj.w;
            const double4 aj = ab[e];
            const double2 bj = bb[e];
            const double rho_j = fmax(aj.x, kDensityFloor);
            const double p_j = aj.y, c_j = aj.z, alpha_j = aj.w, omega_j = bj.x;
            const double dvx = vi.x - vj.x, dvy = vi.y - vj.y, dvz = vi.z - vj.z;
            const double vr = dvx * dx + dvy * dy + dvz * dz;
            const double fi = grad_factor<K>(r, hi), fj = grad_factor<K>(r, hj),
                         fa = grad_factor<K>(r, havg);
            const double gi0 = fi * dx, gi1 = fi * dy, gi2 = fi * dz;
            const double gj0 = fj * dx, gj1 = fj * dy, gj2 = fj * dz;
            const double ga0 = fa * dx, ga1 = fa * dy, ga2 = fa * dz;
            const double pr_j = rho_j > 0.0 ? p_j / (omega_j * rho_j * rho_j) : 0.0;
            double pij = 0.0;
            if (vr < 0.0) {
                const double aavg = 0.5 * (alpha_i + alpha_j);
                const double cavg = 0.5 * (c_i + c_j);
                const double ravg = 0.5 * (rho_i + rho_j);
                const double eta2 = 0.01 * havg * havg;
                const double mu = havg * vr / (r * r + eta2);
                mmu = fmax(mmu, fabs(mu));
                pij = ravg > 0.0 ? (-aavg * cavg * mu + P.beta_visc * mu * mu) / ravg : 0.0;
            }
            axs -= mj * (pr_i * gi0 + pr_j * gj0 + pij * ga0);
            ays -= mj * (pr_i * gi1 + pr_j * gj1 + pij * ga1);
            azs -= mj * (pr_i * gi2 + pr_j * gj2 + pij * ga2);
            const double vgi = dvx * gi0 + dvy * gi1 + dvz * gi2;
            const double vga = dvx * ga0 + dvy * ga1 + dvz * ga2;
            du += mj * (pr_i * vgi + 0.5 * pij * vga);
            dv -= mj * vgi;
            if (P.alpha_cond > 0.0) {
                const double u_j = bj.y;
                const double rsum = rho_i + rho_j;
                if (rsum > 0.0) {
                    const double ravg = 0.5 * rsum;
                    const double dp = fabs(p_i - p_j);
                    const double vsig = sqrt(dp / ravg);
                    const double agw = abs_grad<K>(r, havg);
                    du += mj * P.alpha_cond * vsig * (u_i - u_j) / rsum * agw;
                }
            }
            }
        }
    } else {
        const float hif = (float)hi, ihi = 1.0f / hif;
        const float pri = (float)pr_i, rhoi = (float)rho_i, ci = (float)c_i;
        const float ali = (float)alpha_i, ui = (float)u_i, pif = (float)p_i;
        const float beta = (float)P.beta_visc, acond = (float)P.alpha_cond;
        const float vxi = (float)vi.x, vyi = (float)vi.y, vzi = (float)vi.z;
        float mmuf = 0.0f;
        const ForceRec32* __restrict__ r32 = in.rec32;
        for (uint32_t k0 = 0; k0 < nb.y; k0 += kFGather32) {
            // the batch's indices, then its records, then the pairs in order
            uint32_t jb[kFGather32];
            ForceRec32 rb[kFGather32];
#pragma unroll
            for (uint32_t e = 0; e < kFGather32; ++e)
                jb[e] = k0 + e < nb.y ? pool[nb.x + k0 + e] : 0u;
#pragma unroll
            for (uint32_t e = 0; e < kFGather32; ++e) rb[e] = r32[jb[e]];
#pragma unroll
            for (uint32_t e = 0; e < kFGather32; ++e) {
            if (k0 + e >= nb.y) break;
            const uint32_t j = jb[e];
            if (j == (uint32_t)s) continue;
            const ForceRec32& q = rb[e];
            const float dx = (float)(me.x - q.x), dy = (float)(me.y - q.y),
                        dz = (float)(me.z - q.z);
            const float r2 = dx * dx + dy * dy + dz * dz;
            // r and 1/r from one v_rsq_f32 (0 for coincident points)
            const float rr = __builtin_amdgcn_rsqf(r2);
            const float ir = r2 > 0.0f ? rr : 0.0f;
            const float r = r2 * ir;
            const float hj = q.h;
            if (r >= 2.0f * hif && r >= 2.0f * hj) continue;
            const float havg = 0.5f * (hif + hj);
            const float mj = q.m;
            const float rhoj = q.rho;
            const float dvx = vxi - q.vx, dvy = vyi - q.vy, dvz = vzi - q.vz;
            const float vr = dvx * dx + dvy * dy + dvz * dz;
            const float iha = __builtin_amdgcn_rcpf(havg);
            const float fi = grad_factor_f<K>(r, ir, ihi);
            const float fj = grad_factor_f<K>(r, ir, __builtin_amdgcn_rcpf(hj));
            const float fa = grad_factor_f<K>(r, ir, iha);
            const float prj = q.pr;
            float pij = 0.0f;
            if (vr < 0.0f) {
                const float mu = havg * vr * __builtin_amdgcn_rcpf(r2 + 0.01f * havg * havg);
                mmuf = fmaxf(mmuf, fabsf(mu));
                pij = (-0.5f * (ali + q.alpha) * 0.5f * (ci + q.c) * mu + beta * mu * mu) *
                      __builtin_amdgcn_rcpf(0.5f * (rhoi + rhoj));
            }
            const float fsum = pri * fi + prj * fj + pij * fa;
            axs -= (double)(mj * fsum * dx);
            ays -= (double)(mj * fsum * dy);
            azs -= (double)(mj * fsum * dz);
            const float vgi = fi * vr;  // v_ij . grad W(h_i)
            du += (double)(mj * (pri * vgi + 0.5f * pij * fa * vr));
            dv -= (double)(mj * vgi);
            if (acond > 0.0f) {
                const float2 cj = in.cond[j];
                const float rsum = rhoi + rhoj;
                const float vsig = __builtin_sqrtf(fabsf(pif - cj.x) / (0.5f * rsum));
                const float agw = -fa * r;  // |grad W(h_avg)|
                du += (double)(mj * acond * vsig * (ui - cj.y) / rsum * agw);
            }
            }
        }
        mmu = (double)mmuf;
    }
    o.ax[pi] += axs;
    o.ay[pi] += ays;
    o.az[pi] += azs;
    o.dudt[pi] += du;
    const double divv = rho_i > 0.0 ? dv / rho_i : 0.0;
    o.div_v[pi] = divv;
    o.max_mu[pi] = mmu;
    const double tau = c_i > 0.0 ? hi / (2.0 * P.sigma_visc * c_i) : 1e10;
    const double src = fmax(-divv, 0.0) * (P.alpha_max - alpha_i);
    o.dalpha[pi] = P.alpha_switch ? -(alpha_i - P.alpha_min) / tau + src : 0.0;
    o.rho[pi] = ai.x;
    o.h[pi] = hi;
    o.p[pi] = p_i;
    o.c[pi] = c_i;
    if (o.omega) o.omega[pi] = omega_i;
    if (o.nneigh) o.nneigh[pi] = nb.y;
}

// ---- step_with_sph (solver.rs:145-212) ----------------------------------

// First half kicks of v and u (gas, floored), drift, and the clears of
// a and du/dt (solver.rs:151-177). dt from device memory.
__global__ __launch_bounds__(256) void sph_kick_drift_kernel(
    int n, const double* __restrict__ dtp, double* __restrict__ x, double* __restrict__ y,
    double* __restrict__ z, double* __restrict__ vx, double* __restrict__ vy,
    double* __restrict__ vz, double* __restrict__ ax, double* __restrict__ ay,
    double* __restrict__ az, double* __restrict__ u, double* __restrict__ dudt,
    const uint8_t* __restrict__ type) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double dt = *dtp;
    const double hdt = 0.5 * dt;
    const double vxi = vx[i] + ax[i] * hdt;
    const double vyi = vy[i] + ay[i] * hdt;
    const double vzi = vz[i] + az[i] * hdt;
    vx[i] = vxi;
    vy[i] = vyi;
    vz[i] = vzi;
    if (type[i] == kGas) {
        double ui = u[i] + dudt[i] * hdt;
        if (ui < kEnergyFloor) ui = kEnergyFloor;
        u[i] = ui;
    }
    x[i] = x[i] + vxi * dt;
    y[i] = y[i] + vyi * dt;
    z[i] = z[i] + vzi * dt;
    ax[i] = 0.0;
    ay[i] = 0.0;
    az[i] = 0.0;
    dudt[i] = 0.0;
}

// Second half kicks, the viscosity switch update (solver.rs:191-208) and
// the per-particle timestep of compute_timestep (solver.rs:90-131), reduced
// to the minimum (non-negative doubles order as their bit patterns; NaN
// candidates are ignored as f64::min ignores them).
__global__ __launch_bounds__(256) void sph_kick_kernel(
    int n, const double* __restrict__ dtp, double* __restrict__ vx, double* __restrict__ vy,
    double* __restrict__ vz, const double* __restrict__ ax, const double* __restrict__ ay,
    const double* __restrict__ az, double* __restrict__ u, const double* __restrict__ dudt,
    double* __restrict__ alpha, const double* __restrict__ h, const double* __restrict__ cs,
    const double* __restrict__ div_v, const double* __restrict__ max_mu,
    const double* __restrict__ dalpha, const uint8_t* __restrict__ type, gravis_sph_params P,
    unsigned long long* __restrict__ dtmin) {
    __shared__ unsigned long long red[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double cand = INFINITY;
    if (i < n) {
        const double dt = *dtp;
        const double hdt = 0.5 * dt;
        const double axi = ax[i], ayi = ay[i], azi = az[i];
        vx[i] = vx[i] + axi * hdt;
        vy[i] = vy[i] + ayi * hdt;
        vz[i] = vz[i] + azi * hdt;
        if (type[i] == kGas) {
            double ui = u[i] + dudt[i] * hdt;
            if (ui < kEnergyFloor) ui = kEnergyFloor;
            u[i] = ui;
            double al = alpha[i];
            if (P.alpha_switch) {
                al = al + dalpha[i] * dt;
                al = al < P.alpha_min ? P.alpha_min : (al > P.alpha_max ? P.alpha_max : al);
                alpha[i] = al;
            }
            const double hi = h[i];
            const double am = sqrt(axi * axi + ayi * ayi + azi * azi);
            if (am > 0.0) cand = fmin(cand, P.force_factor * sqrt(hi / am));
            const double c = cs[i];
            const double vsig = c + hi * fabs(div_v[i]) + 1.2 * (al * c + P.beta_visc * max_mu[i]);
            if (vsig > 0.0) cand = fmin(cand, P.cfl * hi / vsig);
        }
    }
    unsigned long long bits = (unsigned long long)__double_as_longlong(cand);
    if (!(cand >= 0.0)) bits = 0x7FF0000000000000ull;  // NaN: ignored
    red[threadIdx.x] = bits;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            const unsigned long long a = red[threadIdx.x], b = red[threadIdx.x + o];
            red[threadIdx.x] = a < b ? a : b;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMin(dtmin, red[0]);
}

// Squared softening h^2 of every particle in the tree's order (eps = h).
// e2[sorted slot] = h^2 (caller particle p -> slot inv[p])
__global__ __launch_bounds__(256) void sph_eps2_kernel(int n, const uint32_t* __restrict__ inv,
                                                       const double* __restrict__ h,
                                                       double* __restrict__ e2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) {
        const double hp = h[p];
        e2[inv[p]] = hp * hp;
    }
}

__global__ void sph_query_reset_kernel(uint32_t* __restrict__ counter, uint32_t* __restrict__ ctl) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    counter[0] = 0u;
    ctl[1] = 0u;
    ctl[2] = 0u;
    ctl[3] = 0u;
    ctl[4] = 0u;
}

__global__ void sph_pool_rewind_kernel(uint32_t* __restrict__ ctl, uint32_t used) {
    if (threadIdx.x == 0 && blockIdx.x == 0) ctl[0] = used;
}

// dt_next = max(dt_min, 1e-6) (solver.rs:130); the headless loop's
// dt = min(dt_next, 1.5 dt) and sim_time += dt (main.rs:178-187).
__global__ void sph_dt_kernel(unsigned long long* __restrict__ dtmin, double* __restrict__ dtp,
                              double* __restrict__ dt_next, double* __restrict__ time) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double dm = __longlong_as_double(*dtmin);
    const double nx = dm < kDtFloor ? kDtFloor : dm;
    const double cur = *dtp;
    *dt_next = nx;
    *time = *time + cur;
    *dtp = fmin(nx, cur * 1.5);
    *dtmin = 0x7FF0000000000000ull;
}

__global__ void sph_dt_init_kernel(double* __restrict__ dtp, double dt, double* __restrict__ time,
                                   double t0, unsigned long long* __restrict__ dtmin) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    *dtp = dt;
    *time = t0;
    *dtmin = 0x7FF0000000000000ull;
}

// Device buffer that keeps its first keep_bytes when it grows.
gravis_status grow_keep(Ctx* c, const char* name, size_t bytes, size_t keep_bytes, void** out) {
    DevBuf& b = c->bufs[name];
    if (b.bytes >= bytes) {
        *out = b.p;
        return GRAVIS_OK;
    }
    void* np = nullptr;
    const size_t want = bytes + bytes / 4;
    GRAVIS_HIP(c, hipStreamSynchronize(c->stream));
    hipError_t e = hipMalloc(&np, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, GRAVIS_OUT_OF_MEMORY, "hipMalloc(%zu) for %s: %s", want, name,
                    hipGetErrorString(e));
    }
    if (b.p && keep_bytes > 0)
        GRAVIS_HIP(c, hipMemcpyAsync(np, b.p, std::min(keep_bytes, b.bytes),
                                     hipMemcpyDeviceToDevice, c->stream));
    GRAVIS_HIP(c, hipStreamSynchronize(c->stream));
    if (b.p) (void)hipFree(b.p);
    b.p = np;
    b.bytes = want;
    *out = np;
    return GRAVIS_OK;
}

}  // namespace

gravis_status sph_compute(Ctx* c, const gravis_sph_params* P, const gravis_sph_state* S,
                          const Octree* tree, double* div_v, double* max_mu, double* dalpha,
                          double* omega, uint32_t* nneigh) {
    const size_t n = S->n;
    if (n == 0) return GRAVIS_OK;
    if (n > 0x7FFFFFFFull) return fail(c, GRAVIS_INVALID_ARG, "n too large");
    const int ni = (int)n;
    if (!P || !(P->gamma > 0.0) || !(P->eta > 0.0) || !(P->h_max > 0.0) || !(P->h_tol > 0.0))
        return fail(c, GRAVIS_INVALID_ARG, "SPH parameters: gamma, eta, h_max, h_tol must be > 0");
    if (P->kernel < 0 || P->kernel > 1 || P->h_solver < 0 || P->h_solver > 1 ||
        P->max_h_iter < 1 || P->max_h_iter > 200)
        return fail(c, GRAVIS_INVALID_ARG,
                    "SPH parameters: kernel and h_solver must be 0 or 1, max_h_iter 1..200");
    for (int k = 0; k < 4; ++k) c->sph_stats[k] = 0;
    Octree own;
    if (!tree) {
        GRAVIS_TRY(build_octree(c, n, S->x, S->y, S->z, S->m, 16u, (uint32_t)kWave, &own));
        tree = &own;
    }
    const Octree& t = *tree;
    hipStream_t st = c->stream;
    double4 *vh, *aux;
    double2* aux2;
    uint8_t *qstate, *iter;
    double *hcur, *radius;
    uint2* nbr;
    uint32_t *ctl, *dfs_groups, *counter;
    GRAVIS_TRY(buf(c, "sph_vh", n, &vh));
    GRAVIS_TRY(buf(c, "sph_aux", n, &aux));
    GRAVIS_TRY(buf(c, "sph_aux2", n, &aux2));
    GRAVIS_TRY(buf(c, "sph_qstate", n, &qstate));
    GRAVIS_TRY(buf(c, "sph_iter", n, &iter));
    GRAVIS_TRY(buf(c, "sph_hcur", n, &hcur));
    GRAVIS_TRY(buf(c, "sph_radius", n, &radius));
    GRAVIS_TRY(buf(c, "sph_nbr", n, &nbr));
    GRAVIS_TRY(buf(c, "sph_ctl", kCtlWords, &ctl));
    GRAVIS_TRY(buf(c, "sph_dfs", n, &dfs_groups));
    GRAVIS_TRY(buf(c, "sph_counter", 2, &counter));
    // pool: grows (keeping what is written) when a round runs out
    size_t pool_cap = std::max<size_t>(c->bufs["sph_pool"].bytes / sizeof(uint32_t),
                                       n * c->sph_pool_init);
    pool_cap = std::min<size_t>(pool_cap, 0xFFFFFFFFull);
    void* pv = nullptr;
    GRAVIS_TRY(grow_keep(c, "sph_pool", pool_cap * sizeof(uint32_t), 0, &pv));
    pool_cap = std::min<size_t>(c->bufs["sph_pool"].bytes / sizeof(uint32_t), 0xFFFFFFFFull);
    uint32_t* pool = static_cast<uint32_t*>(pv);

    const unsigned g256 = grid_for(n, 256);
    sph_prep_kernel<<<g256, 256, 0, st>>>(ni, t.inv, S->vx, S->vy, S->vz, S->h, S->rho, S->p,
                                          S->c, S->alpha, S->u, S->type, vh, aux, aux2, qstate,
                                          iter, hcur, radius, ctl);
    GRAVIS_TRY(check_launch(c, "sph_prep"));
    const int qblocks = (int)std::min<long>((long)(n / 64 / kQWaves) + 1, (long)c->num_cus * 8);
    // one wavefront per target range of the overflowed groups (8 per group),
    // 16 resident per CU
    const int dblocks = c->num_cus * 16;
    for (int round = 0;; ++round) {
        if (round > 64) return fail(c, GRAVIS_INTERNAL, "SPH neighbour rounds did not settle");
        sph_query_reset_kernel<<<1, 64, 0, st>>>(counter, ctl);
        timer_begin(c);
        sph_query_kernel<<<qblocks, kQWaves * kWave, 0, st>>>(
            t.src, t.meta, t.bmin, t.bmax, t.groups, t.ngroups, counter, qstate, iter, radius,
            nbr, pool, (uint32_t)pool_cap, ctl, dfs_groups,
            std::min<uint32_t>(c->sph_range_cap, kQRanges), P->max_h_iter);
        sph_query_dfs_kernel<<<dblocks, kWave, 0, st>>>(
            t.src, t.meta, t.bmin, t.bmax, t.groups, dfs_groups, qstate, iter, radius, nbr, pool,
            (uint32_t)pool_cap, ctl, std::min<uint32_t>(c->sph_range_cap, kQRanges),
            P->max_h_iter);
        timer_end(c, "sph_query");
        GRAVIS_TRY(check_launch(c, "sph_query"));
        timer_begin(c);
        const DensityArgs da{P->gamma, P->eta, P->h_max, P->h_tol, P->h_solver, P->max_h_iter};
#define GRAVIS_DENSITY(K, F) \
    sph_density_kernel<K, F><<<g256, 256, 0, st>>>(ni, t.src, pool, nbr, qstate, iter, hcur, \
                                                   radius, vh, aux, aux2, da, ctl)
        if (P->kernel == 0 && !P->pair_fp32) GRAVIS_DENSITY(0, false);
        else if (P->kernel == 0) GRAVIS_DENSITY(0, true);
        else if (!P->pair_fp32) GRAVIS_DENSITY(1, false);
        else GRAVIS_DENSITY(1, true);
#undef GRAVIS_DENSITY
        timer_end(c, "sph_density");
        GRAVIS_TRY(check_launch(c, "sph_density"));
        uint32_t h_ctl[kCtlWords];
        GRAVIS_HIP(c, hipMemcpyAsync(h_ctl, ctl, sizeof h_ctl, hipMemcpyDeviceToHost, st));
        GRAVIS_HIP(c, hipStreamSynchronize(st));
        c->sph_stats[0] += h_ctl[2];
        c->sph_stats[1] += h_ctl[3];
        if (h_ctl[1]) {
            // some lists did not fit: keep the written prefix, grow, query again
            const size_t used = std::min<size_t>(h_ctl[0], pool_cap);
            const size_t want = std::max<size_t>(2 * pool_cap, (size_t)h_ctl[0] + n * 8);
            if (want > 0xFFFFFFFFull)
                return fail(c, GRAVIS_OUT_OF_MEMORY, "SPH neighbour pool exceeds 2^32 entries");
            GRAVIS_TRY(grow_keep(c, "sph_pool", want * sizeof(uint32_t), used * sizeof(uint32_t),
                                 &pv));
            pool = static_cast<uint32_t*>(pv);
            pool_cap = std::min<size_t>(c->bufs["sph_pool"].bytes / sizeof(uint32_t),
                                        0xFFFFFFFFull);
            sph_pool_rewind_kernel<<<1, 64, 0, st>>>(ctl, (uint32_t)used);
            continue;
        }
        c->sph_stats[2] += 1;
        if (h_ctl[3] == 0) {
            c->sph_stats[3] = h_ctl[0];
            break;
        }
    }
    ForceOut o{S->ax, S->ay, S->az, S->dudt, S->rho, S->h, S->p, S->c,
               div_v, max_mu, dalpha, omega, nneigh};
    ForceIn in{nullptr, nullptr, nullptr, t.src, vh, aux, aux2};
    timer_begin(c);
    if (P->pair_fp32) {
        ForceRec32* r32;
        float2* cond;
        GRAVIS_TRY(buf(c, "sph_frec32", n, &r32));
        GRAVIS_TRY(buf(c, "sph_cond", n, &cond));
        sph_pack32_kernel<<<g256, 256, 0, st>>>(ni, t.src, vh, aux, aux2, r32, cond);
        in.rec32 = r32;
        in.cond = cond;
    } else {
        ForceRec* frec;
        GRAVIS_TRY(buf(c, "sph_frec", n, &frec));
        sph_pack_kernel<<<g256, 256, 0, st>>>(ni, t.src, vh, aux, aux2, frec);
        in.rec = frec;
    }
#define GRAVIS_FORCE(K, F)                                                               \
    sph_force_kernel<K, F><<<g256, 256, 0, st>>>(ni, t.perm, in, qstate, pool, nbr, *P, o)
    if (P->kernel == 0 && !P->pair_fp32) GRAVIS_FORCE(0, false);
    else if (P->kernel == 0) GRAVIS_FORCE(0, true);
    else if (!P->pair_fp32) GRAVIS_FORCE(1, false);
    else GRAVIS_FORCE(1, true);
#undef GRAVIS_FORCE
    timer_end(c, "sph_force");
    return check_launch(c, "sph_force");
}

gravis_status sph_gravity(Ctx* c, const gravis_sph_params* P, const gravis_solver* g,
                          const gravis_sph_state* S, Octree* t, const Octree** shared) {
    const size_t n = S->n;
    if (n == 0 || !g || g->algorithm == GRAVIS_NO_GRAVITY) return GRAVIS_OK;
    FinishArgs fa;
    fa.mode = kAccumulate;
    fa.ax = S->ax;
    fa.ay = S->ay;
    fa.az = S->az;
    if (g->algorithm == GRAVIS_BARNES_HUT) {
        const uint32_t leaf = g->leaf_size ? g->leaf_size : 16u;
        const uint32_t gs = g->group_size ? g->group_size : (uint32_t)kWave;
        GRAVIS_TRY(build_octree(c, n, S->x, S->y, S->z, S->m, leaf, gs, t));
        double* e2 = nullptr;
        if (P->soft_h) {
            GRAVIS_TRY(buf(c, "sph_eps2", n, &e2));
            sph_eps2_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>((int)n, t->inv, S->h, e2);
            GRAVIS_TRY(check_launch(c, "sph_eps2"));
        }
        GRAVIS_TRY(bh_walk(c, g, *t, fa, e2));
        if (shared && gs == (uint32_t)kWave) *shared = t;  // the same tree serves the search
        return GRAVIS_OK;
    }
    if (P->soft_h)
        return fail(c, GRAVIS_UNSUPPORTED,
                    "gravity softening eps = h is implemented for Barnes-Hut only");
    return bf_force(c, g, n, S->x, S->y, S->z, S->m, fa);
}

gravis_status sph_run(Ctx* c, const gravis_sph_params* P, const gravis_solver* g,
                      const gravis_sph_state* S, uint64_t nsteps, double* dt_io,
                      double* time_io, double* dt_next_out) {
    const size_t n = S->n;
    if (n == 0 || nsteps == 0) return GRAVIS_OK;
    const int ni = (int)n;
    hipStream_t st = c->stream;
    double *dts, *iv;
    unsigned long long* dtmin;
    GRAVIS_TRY(buf(c, "sph_dts", 4, &dts));  // dt, time, dt_next
    GRAVIS_TRY(buf(c, "sph_dtmin", 1, &dtmin));
    GRAVIS_TRY(buf(c, "sph_interm", 3 * n, &iv));
    double *div_v = iv, *max_mu = iv + n, *dalpha = iv + 2 * n;
    sph_dt_init_kernel<<<1, 64, 0, st>>>(dts, *dt_io, dts + 1, time_io ? *time_io : 0.0, dtmin);
    const unsigned g256 = grid_for(n, 256);
    const bool grav = g && g->algorithm != GRAVIS_NO_GRAVITY;
    for (uint64_t k = 0; k < nsteps; ++k) {
        sph_kick_drift_kernel<<<g256, 256, 0, st>>>(ni, dts, S->x, S->y, S->z, S->vx, S->vy,
                                                    S->vz, S->ax, S->ay, S->az, S->u, S->dudt,
                                                    S->type);
        GRAVIS_TRY(check_launch(c, "sph_kick_drift"));
        const Octree* shared = nullptr;
        Octree t;
        if (grav) GRAVIS_TRY(sph_gravity(c, P, g, S, &t, &shared));
        GRAVIS_TRY(sph_compute(c, P, S, shared, div_v, max_mu, dalpha, nullptr, nullptr));
        sph_kick_kernel<<<g256, 256, 0, st>>>(ni, dts, S->vx, S->vy, S->vz, S->ax, S->ay, S->az,
                                              S->u, S->dudt, S->alpha, S->h, S->c, div_v,
                                              max_mu, dalpha, S->type, *P, dtmin);
        sph_dt_kernel<<<1, 64, 0, st>>>(dtmin, dts, dts + 2, dts + 1);
        GRAVIS_TRY(check_launch(c, "sph_kick"));
    }
    double h[3];
    GRAVIS_HIP(c, hipMemcpyAsync(h, dts, sizeof h, hipMemcpyDeviceToHost, st));
    GRAVIS_HIP(c, hipStreamSynchronize(st));
    *dt_io = h[0];
    if (time_io) *time_io = h[1];
    if (dt_next_out) *dt_next_out = h[2];
    return GRAVIS_OK;
}

}  // namespace gravis
