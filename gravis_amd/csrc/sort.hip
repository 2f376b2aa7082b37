// sort.hip -- root cube, 63-bit Morton keys and a stable LSD radix sort.
//
// The root cube restates BoundingBox::containing (crates/sim-core/src/
// octree.rs:14-44) bit for bit: min/max are exact and the cube arithmetic
// runs in one thread in the same operation order with contraction off.
// The key quantises each axis to 21 bits over that cube with the x/y/z bit
// order of BoundingBox::octant (octree.rs:67-79), so key prefixes are the
// octree cells level by level.
//
// The radix sort is stable (LSD, 9-bit digits, 7 passes for the 63-bit
// keys): equal keys keep their input order, so the permutation is bit-exact
// with any stable CPU sort of the same keys.
#include "internal.h"

namespace gravis {
namespace {

constexpr int kBoxBlock = 256;
constexpr int kBoxMaxBlocks = 1024;

__global__ __launch_bounds__(kBoxBlock) void box_partial_kernel(const double* __restrict__ x,
                                                                const double* __restrict__ y,
                                                                const double* __restrict__ z,
                                                                size_t n,
                                                                double* __restrict__ part) {
    double mn[3] = {INFINITY, INFINITY, INFINITY};
    double mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = blockIdx.x * (size_t)kBoxBlock + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kBoxBlock) {
        double p[3] = {x[i], y[i], z[i]};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            mn[d] = fmin(mn[d], p[d]);
            mx[d] = fmax(mx[d], p[d]);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            mn[d] = fmin(mn[d], __shfl_xor(mn[d], off));
            mx[d] = fmax(mx[d], __shfl_xor(mx[d], off));
        }
    }
    __shared__ double s[kBoxBlock / kWave][6];
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        for (int d = 0; d < 3; ++d) {
            s[w][d] = mn[d];
            s[w][3 + d] = mx[d];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBoxBlock / kWave; ++k)
            for (int d = 0; d < 3; ++d) {
                s[0][d] = fmin(s[0][d], s[k][d]);
                s[0][3 + d] = fmax(s[0][3 + d], s[k][3 + d]);
            }
        for (int d = 0; d < 6; ++d) part[blockIdx.x * 6 + d] = s[0][d];
    }
}

// One block of 256 threads folds the partial boxes, then thread 0 applies
// the cube arithmetic.
__global__ __launch_bounds__(kBoxBlock) void box_final_kernel(const double* __restrict__ part,
                                                              int nparts,
                                                              double* __restrict__ box) {
    double mn[3] = {INFINITY, INFINITY, INFINITY};
    double mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < nparts; b += kBoxBlock)
        for (int d = 0; d < 3; ++d) {
            mn[d] = fmin(mn[d], part[b * 6 + d]);
            mx[d] = fmax(mx[d], part[b * 6 + 3 + d]);
        }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            mn[d] = fmin(mn[d], __shfl_xor(mn[d], off));
            mx[d] = fmax(mx[d], __shfl_xor(mx[d], off));
        }
    __shared__ double s[kBoxBlock / kWave][6];
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0)
        for (int d = 0; d < 3; ++d) {
            s[w][d] = mn[d];
            s[w][3 + d] = mx[d];
        }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int k = 1; k < kBoxBlock / kWave; ++k)
        for (int d = 0; d < 3; ++d) {
            mn[d] = fmin(mn[d], s[k][d]);
            mx[d] = fmax(mx[d], s[k][3 + d]);
        }
    // octree.rs:29-41, same operation order (no contraction in this build).
    box[0] = 0.5 * (mn[0] + mx[0]);
    box[1] = 0.5 * (mn[1] + mx[1]);
    box[2] = 0.5 * (mn[2] + mx[2]);
    double ext = fmax(fmax(mx[0] - mn[0], mx[1] - mn[1]), mx[2] - mn[2]);
    box[3] = 0.5 * ext * 1.001 + 1e-10;
}

__device__ __forceinline__ uint32_t quantise(double v, double lo, double inv) {
    double q = (v - lo) * inv;
    if (!(q >= 0.0)) return 0u;
    if (q >= 2097151.0) return 2097151u;
    return static_cast<uint32_t>(q);
}

// 63-bit keys (21 levels per axis) shifted right by `shift` (a multiple of
// 3: the key of the coarser grid of 21 - shift / 3 levels).
__global__ __launch_bounds__(256) void morton_kernel(const double* __restrict__ x,
                                                     const double* __restrict__ y,
                                                     const double* __restrict__ z, size_t n,
                                                     const double* __restrict__ box,
                                                     uint64_t* __restrict__ keys,
                                                     uint32_t* __restrict__ vals, int shift) {
    const double hw = box[3];
    const double inv = 2097152.0 / (2.0 * hw);
    const double lx = box[0] - hw, ly = box[1] - hw, lz = box[2] - hw;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        uint32_t qx = quantise(x[i], lx, inv), qy = quantise(y[i], ly, inv),
                 qz = quantise(z[i], lz, inv);
        keys[i] = (morton_spread3(qx) | (morton_spread3(qy) << 1) | (morton_spread3(qz) << 2)) >>
                  shift;
        vals[i] = static_cast<uint32_t>(i);
    }
}

// ---- radix sort -----------------------------------------------------------
// LSD passes of kRsBits-bit digits. Per pass: radix_hist_kernel counts the
// digits of every 4096-key tile; radix_row_scan_kernel and
// radix_digit_base_kernel turn the counts into each (digit, tile)'s first
// output position; radix_scatter_kernel ranks every tile's keys by digit in
// LDS (stable: element order within a digit) and writes each digit's run of
// the tile to consecutive output positions, so the stores are contiguous
// runs rather than one scattered key per lane.
constexpr int kRsBits = 9;
constexpr int kRsDigits = 1 << kRsBits;
constexpr uint32_t kRsMask = kRsDigits - 1;
constexpr int kRsBlock = 256;               // 4 wavefronts
constexpr int kRsWaves = kRsBlock / kWave;  // 4
constexpr int kRsRounds = 16;               // 64 keys per wave per round
constexpr int kRsWaveTile = kWave * kRsRounds;      // 1024
constexpr int kRsTile = kRsWaveTile * kRsWaves;     // 4096 keys per block
constexpr int kRsPer = kRsDigits / kRsBlock;        // digits per thread
static_assert(kRsDigits % kRsBlock == 0, "whole digits per thread");

__global__ __launch_bounds__(kRsBlock) void radix_hist_kernel(const uint64_t* __restrict__ keys,
                                                              size_t n, int shift,
                                                              uint32_t* __restrict__ counts,
                                                              int nblocks) {
    __shared__ uint32_t h[kRsDigits];
    for (int d = threadIdx.x; d < kRsDigits; d += kRsBlock) h[d] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kRsTile;
    for (int e = threadIdx.x; e < kRsTile; e += kRsBlock) {
        size_t i = base + e;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & kRsMask], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < kRsDigits; d += kRsBlock)
        counts[(size_t)d * nblocks + blockIdx.x] = h[d];
}

// Exclusive scan of the per-tile counts in (digit, tile) order, in two
// launches: one block per digit scans that digit's row of nblocks counts in
// place and stores the row total; one block then turns the totals into
// exclusive digit bases. The scatter adds base[digit] to the row offsets.
constexpr int kScanBlock = 256;
constexpr int kScanPer = 4;  // consecutive counts per thread per chunk

__global__ __launch_bounds__(kScanBlock) void radix_row_scan_kernel(uint32_t* __restrict__ counts,
                                                                    int nblocks,
                                                                    uint32_t* __restrict__ totals) {
    uint32_t* row = counts + (size_t)blockIdx.x * nblocks;
    uint32_t carry = 0;
    for (int base = 0; base < nblocks; base += kScanBlock * kScanPer) {
        const int i0 = base + threadIdx.x * kScanPer;
        uint32_t v[kScanPer];
        uint32_t sum = 0;
#pragma unroll
        for (int e = 0; e < kScanPer; ++e) {
            v[e] = i0 + e < nblocks ? row[i0 + e] : 0u;
            sum += v[e];
        }
        uint32_t chunk_total;
        uint32_t run = carry + block_exclusive_scan<kScanBlock>(sum, &chunk_total);
#pragma unroll
        for (int e = 0; e < kScanPer; ++e)
            if (i0 + e < nblocks) {
                row[i0 + e] = run;
                run += v[e];
            }
        carry += chunk_total;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// kRsPer consecutive digits per thread
__global__ __launch_bounds__(kRsBlock) void radix_digit_base_kernel(
    uint32_t* __restrict__ totals) {
    uint32_t v[kRsPer], sum = 0;
#pragma unroll
    for (int e = 0; e < kRsPer; ++e) {
        v[e] = totals[threadIdx.x * kRsPer + e];
        sum += v[e];
    }
    uint32_t all;
    uint32_t run = block_exclusive_scan<kRsBlock>(sum, &all);
#pragma unroll
    for (int e = 0; e < kRsPer; ++e) {
        totals[threadIdx.x * kRsPer + e] = run;
        run += v[e];
    }
}

__device__ __forceinline__ uint64_t digit_match(uint32_t d, uint64_t active) {
    uint64_t mask = active;
#pragma unroll
    for (int b = 0; b < kRsBits; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bal = __ballot(bit);
        mask &= bit ? bal : ~bal;
    }
    return mask;
}

__global__ __launch_bounds__(kRsBlock) void radix_scatter_kernel(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin, size_t n, int shift,
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ dbase, int nblocks,
    uint64_t* __restrict__ kout, uint32_t* __restrict__ vout) {
    __shared__ uint32_t cnt[kRsWaves][kRsDigits];  // per-wave counts, then LDS positions
    __shared__ uint32_t lbase[kRsDigits];          // first LDS position of each digit
    __shared__ uint32_t gbase[kRsDigits];          // first output position of each digit
    __shared__ uint64_t lkey[kRsTile];             // the tile's keys, ranked by digit
    __shared__ uint32_t lval[kRsTile];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (int e = threadIdx.x; e < kRsWaves * kRsDigits; e += kRsBlock) (&cnt[0][0])[e] = 0u;
    __syncthreads();

    // consecutive tiles on one XCD: their runs of a digit are adjacent in
    // the output and share cache lines, which then fill in one L2
    const uint32_t tile = xcd_block(blockIdx.x, gridDim.x);
    const size_t tbase = (size_t)tile * kRsTile;
    const size_t wbase = tbase + (size_t)w * kRsWaveTile;
    // Phase A: per-wave digit counts (each digit has one leader lane per
    // round). Rounds are not unrolled: a round's ballots are used at once
    // (unrolled, the masks of every round stayed live and the kernel ran
    // at one wavefront per SIMD).
    // Each round's key is loaded one round ahead (a tile of few blocks has
    // no other wavefronts to cover the load latency).
    uint64_t knext = wbase + lane < n ? kin[wbase + lane] : 0ull;
#pragma unroll 1
    for (int r = 0; r < kRsRounds; ++r) {
        const size_t i = wbase + (size_t)r * kWave + lane;
        const uint64_t kcur = knext;
        const size_t inext = i + kWave;
        knext = r + 1 < kRsRounds && inext < n ? kin[inext] : 0ull;
        const uint64_t active = __ballot(i < n);
        const uint32_t dig = i < n ? static_cast<uint32_t>((kcur >> shift) & kRsMask) : 0u;
        const uint64_t mask = digit_match(dig, active);
        if (i < n && (mask & lt) == 0ull) cnt[w][dig] += (uint32_t)__popcll(mask);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // Phase B: digit runs of the tile in LDS (exclusive scan of the digit
    // totals), each wave's first LDS position per digit, and each digit's
    // first output position (the tile's row offset plus the digit base).
    {
        uint32_t tot[kRsPer], sum = 0;
#pragma unroll
        for (int e = 0; e < kRsPer; ++e) {
            const int d = threadIdx.x * kRsPer + e;
            tot[e] = 0;
#pragma unroll
            for (int k = 0; k < kRsWaves; ++k) tot[e] += cnt[k][d];
            sum += tot[e];
        }
        uint32_t all;
        uint32_t run = block_exclusive_scan<kRsBlock>(sum, &all);
#pragma unroll
        for (int e = 0; e < kRsPer; ++e) {
            const int d = threadIdx.x * kRsPer + e;
            lbase[d] = run;
            gbase[d] = dbase[d] + offsets[(size_t)d * nblocks + tile];
            uint32_t p = run;
#pragma unroll
            for (int k = 0; k < kRsWaves; ++k) {
                const uint32_t c = cnt[k][d];
                cnt[k][d] = p;
                p += c;
            }
            run += tot[e];
        }
    }
    __syncthreads();
    // Phase C: stable ranks in element order, into LDS (keys read again:
    // the tile is in L2).
    knext = wbase + lane < n ? kin[wbase + lane] : 0ull;
    uint32_t vnext = wbase + lane < n ? vin[wbase + lane] : 0u;
#pragma unroll 1
    for (int r = 0; r < kRsRounds; ++r) {
        const size_t i = wbase + (size_t)r * kWave + lane;
        const uint64_t key = knext;
        const uint32_t val = vnext;
        const size_t inext = i + kWave;
        const bool more = r + 1 < kRsRounds && inext < n;
        knext = more ? kin[inext] : 0ull;
        vnext = more ? vin[inext] : 0u;
        const uint64_t active = __ballot(i < n);
        const uint32_t dig = static_cast<uint32_t>((key >> shift) & kRsMask);
        const uint64_t mask = digit_match(dig, active);
        if (i < n) {
            const uint32_t lp = cnt[w][dig] + (uint32_t)__popcll(mask & lt);
            lkey[lp] = key;
            lval[lp] = val;
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
        __builtin_amdgcn_wave_barrier();
        if (i < n && (mask & lt) == 0ull) cnt[w][dig] += (uint32_t)__popcll(mask);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // Phase D: the tile in digit order, each digit's run to consecutive
    // output positions.
    const int count = (int)min((size_t)kRsTile, n - tbase);
    for (int e = threadIdx.x; e < count; e += kRsBlock) {
        const uint64_t k = lkey[e];
        const uint32_t d = static_cast<uint32_t>((k >> shift) & kRsMask);
        const uint32_t gp = gbase[d] + ((uint32_t)e - lbase[d]);
        kout[gp] = k;
        vout[gp] = lval[e];
    }
}

// inv[perm[k]] = k
__global__ __launch_bounds__(256) void invert_perm_kernel(const uint32_t* __restrict__ perm,
                                                          size_t n, uint32_t* __restrict__ inv) {
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
         k += (size_t)gridDim.x * blockDim.x)
        inv[perm[k]] = (uint32_t)k;
}

// out[inv[i]] = (x_i, y_i, z_i, m_i)
__global__ __launch_bounds__(256) void scatter_sorted_kernel(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ m, const uint32_t* __restrict__ inv, size_t n,
    double4* __restrict__ out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        out[inv[i]] = make_double4(x[i], y[i], z[i], m[i]);
}

__global__ __launch_bounds__(256) void gather_kernel(const double* __restrict__ x,
                                                     const double* __restrict__ y,
                                                     const double* __restrict__ z,
                                                     const double* __restrict__ m,
                                                     const uint32_t* __restrict__ perm, size_t n,
                                                     double4* __restrict__ out) {
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t p = perm ? perm[k] : k;
        out[k] = make_double4(x[p], y[p], z[p], m[p]);
    }
}

}  // namespace

gravis_status root_box(Ctx* c, size_t n, const double* x, const double* y, const double* z,
                       double* d_box) {
    double* part = nullptr;
    GRAVIS_TRY(buf(c, "box_part", kBoxMaxBlocks * 6, &part));
    const unsigned g = grid_for(n, kBoxBlock, kBoxMaxBlocks);
    box_partial_kernel<<<g, kBoxBlock, 0, c->stream>>>(x, y, z, n, part);
    box_final_kernel<<<1, kBoxBlock, 0, c->stream>>>(part, (int)g, d_box);
    return check_launch(c, "root_box");
}

gravis_status sort_buffers(Ctx* c, size_t n, uint64_t** keys, uint32_t** vals) {
    if (n > 0xFFFFFFFFull) return fail(c, GRAVIS_INVALID_ARG, "n too large for 32-bit indices");
    GRAVIS_TRY(buf(c, "sort_k0", n, keys));
    GRAVIS_TRY(buf(c, "sort_v0", n, vals));
    return GRAVIS_OK;
}

gravis_status radix_sort(Ctx* c, size_t n, int bits, uint64_t** d_keys_sorted,
                         uint32_t** d_vals_sorted) {
    if (n > 0xFFFFFFFFull) return fail(c, GRAVIS_INVALID_ARG, "n too large for 32-bit indices");
    if (bits < 1 || bits > 64) return fail(c, GRAVIS_INTERNAL, "radix sort bits must be 1..64");
    const int passes = (bits + kRsBits - 1) / kRsBits;
    uint64_t *k0, *k1;
    uint32_t *v0, *v1, *counts, *dbase;
    const int nblocks = (int)((n + kRsTile - 1) / kRsTile);
    const size_t m = (size_t)kRsDigits * nblocks;
    GRAVIS_TRY(buf(c, "sort_k0", n, &k0));
    GRAVIS_TRY(buf(c, "sort_k1", n, &k1));
    GRAVIS_TRY(buf(c, "sort_v0", n, &v0));
    GRAVIS_TRY(buf(c, "sort_v1", n, &v1));
    GRAVIS_TRY(buf(c, "sort_counts", m, &counts));
    GRAVIS_TRY(buf(c, "sort_dbase", kRsDigits, &dbase));
    for (int pass = 0; pass < passes; ++pass) {
        const int shift = kRsBits * pass;
        radix_hist_kernel<<<nblocks, kRsBlock, 0, c->stream>>>(k0, n, shift, counts, nblocks);
        radix_row_scan_kernel<<<kRsDigits, kScanBlock, 0, c->stream>>>(counts, nblocks, dbase);
        radix_digit_base_kernel<<<1, kRsBlock, 0, c->stream>>>(dbase);
        radix_scatter_kernel<<<nblocks, kRsBlock, 0, c->stream>>>(k0, v0, n, shift, counts,
                                                                  dbase, nblocks, k1, v1);
        std::swap(k0, k1);
        std::swap(v0, v1);
    }
    *d_keys_sorted = k0;
    *d_vals_sorted = v0;
    return check_launch(c, "radix_sort");
}

gravis_status morton_sort(Ctx* c, size_t n, const double* x, const double* y, const double* z,
                          const double* d_box, uint64_t** d_keys_sorted, uint32_t** d_perm,
                          int levels) {
    if (levels < 1 || levels > 21) return fail(c, GRAVIS_INTERNAL, "morton_sort: %d levels", levels);
    uint64_t* k0;
    uint32_t* v0;
    GRAVIS_TRY(sort_buffers(c, n, &k0, &v0));
    morton_kernel<<<grid_for(n, 256, 8192), 256, 0, c->stream>>>(x, y, z, n, d_box, k0, v0,
                                                                 3 * (21 - levels));
    return radix_sort(c, n, 3 * levels, d_keys_sorted, d_perm);
}

gravis_status gather_sorted(Ctx* c, size_t n, const double* x, const double* y,
                            const double* z, const double* m, const uint32_t* d_perm,
                            double4* d_out, const uint32_t** inv_out) {
    if (inv_out) *inv_out = nullptr;
    if (!d_perm || n == 0) {
        gather_kernel<<<grid_for(n, 256, 8192), 256, 0, c->stream>>>(x, y, z, m, d_perm, n,
                                                                     d_out);
        return check_launch(c, "gather_sorted");
    }
    uint32_t* inv;
    GRAVIS_TRY(buf(c, "sort_inv", n, &inv));
    invert_perm_kernel<<<grid_for(n, 256, 8192), 256, 0, c->stream>>>(d_perm, n, inv);
    scatter_sorted_kernel<<<grid_for(n, 256, 8192), 256, 0, c->stream>>>(x, y, z, m, inv, n,
                                                                         d_out);
    if (inv_out) *inv_out = inv;
    return check_launch(c, "gather_sorted");
}

}  // namespace gravis
