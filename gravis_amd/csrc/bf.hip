// bf.hip -- brute-force gravity on gfx950 (FP64 and mixed precision).
//
// Reference: BruteForce::compute_accelerations (crates/sim-core/src/
// gravity.rs:47-80): a_i = sum_{j != i} G m_j d_ij / (|d_ij|^2 + eps^2)^(3/2),
// d_ij = x_j - x_i, accumulated into the caller's acceleration arrays.
//
// The reference visits pairs i<j once (Newton's third law, sequential). Here
// every target i gathers over all j (no write conflicts, no atomics): a
// workgroup owns a run of the Morton-sorted particle list as targets and a
// contiguous j range; j records are staged through LDS in tiles and read
// back as wavefront-uniform broadcasts. Pairwise gravity is mostly VALU
// work: only |x_j - x_i|^2 is a contraction (of length 4), which the
// matrix-core variant below computes with v_mfma_f32_16x16x4f32. An FP32
// MFMA occupies the SIMD for its 32 cycles (no overlap with the VALU was
// measured), so the variant trades six VALU operations per pair for an
// eighth of an MFMA and gains a few per cent; the autotuner picks between
// them. The j range is split in S chunks so the grid holds many wavefronts
// per CU; the chunk partials are summed in a fixed order by the finishing
// kernel (deterministic).
//
// Mixed precision: separations are formed in FP64 relative to the centre of
// the wavefront's target box, rounded to FP32, and the pair force is FP32
// (v_rsq_f32); sources close to the group get FP64 separations per pair.
// Each lane sums a tile of pair forces in FP32 and adds the tile sum into a
// Kahan-compensated FP32 accumulator; the result leaves as FP64
// sum - compensation.
#include <type_traits>

#include "internal.h"

namespace gravis {
namespace {

// MASK gives 0 for d = 0 (whatever eps is); the unmasked form needs
// eps^2 >= kMaskEps2D, where the self pair's m eps^-3 is finite.
template <bool MASK>
__device__ __forceinline__ double pair_f64(double dx, double dy, double dz, double m,
                                           double eps2, const Rinv3Consts& k3) {
    if (MASK) {
        const double d2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz));
        const double f = m_rinv3(d2 + eps2, m, k3);
        return d2 > 0.0 ? f : 0.0;
    }
    const double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, __builtin_fma(dz, dz, eps2)));
    return m_rinv3(r2, m, k3);
}

// Bounding box of each TJ-wide tile of the sorted records: tbox[kTileBox t
// ..] = (lo x, lo y, lo z, hi x, hi y, hi z).
constexpr int kTileBox = 6;

template <int TJ>
__global__ __launch_bounds__(256) void tile_box_kernel(const double4* __restrict__ pm, int n,
                                                       double* __restrict__ tbox) {
    const int base = blockIdx.x * TJ;
    const int last = min(n, base + TJ) - 1;
    double lo[3], hi[3];
    {
        const double4 p = pm[base];
        lo[0] = hi[0] = p.x; lo[1] = hi[1] = p.y; lo[2] = hi[2] = p.z;
    }
    for (int j = base + threadIdx.x; j <= last; j += blockDim.x) {
        const double4 p = pm[j];
        lo[0] = fmin(lo[0], p.x); hi[0] = fmax(hi[0], p.x);
        lo[1] = fmin(lo[1], p.y); hi[1] = fmax(hi[1], p.y);
        lo[2] = fmin(lo[2], p.z); hi[2] = fmax(hi[2], p.z);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            lo[d] = fmin(lo[d], __shfl_xor(lo[d], off));
            hi[d] = fmax(hi[d], __shfl_xor(hi[d], off));
        }
    __shared__ double sh[4][6];
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0)
        for (int d = 0; d < 3; ++d) { sh[w][d] = lo[d]; sh[w][3 + d] = hi[d]; }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int d = threadIdx.x;
        double v = sh[0][d];
        for (int k = 1; k < (int)(blockDim.x / kWave); ++k)
            v = d < 3 ? fmin(v, sh[k][d]) : fmax(v, sh[k][d]);
        tbox[(size_t)blockIdx.x * kTileBox + d] = v;
    }
}

// Far/near split of the mixed kernel. Source coordinates are rounded to
// FP32 relative to the centre c of the target group's box, so a separation
// d = x_j - x_i carries an absolute error of about u (|x_i - c| + |x_j - c|)
// <= u (diag + |d|), u = 2^-24. The pair force is m d / (|d|^2 + eps^2)^(3/2),
// so the error matters only through u diag / max(|d|, eps). A tile is "near"
// when both its gap to the target box and eps are below q * diag: there the
// separations are formed in FP64 per pair. Every FP32 pair then has
// max(|d|, eps) >= q * diag and a relative force error of at most about
// 3 u (1 + 1 / q). q is the context's bf_near_fraction (default 1/64).
// Mixed precision: one wavefront per workgroup owns IPT*64 Morton-consecutive
// targets and one j chunk. Each TJ-record tile of the chunk is staged through
// LDS by the wave itself: far tiles as FP32 records relative to the wave's
// own target-box centre (so the rounding error scales with the size of 64 or
// 128 targets, not of a larger block), near tiles as FP64 records. The far
// loop is 12 VALU operations and one v_rsq_f32 per pair. On gfx950 FP32
// VALU operations with VGPR or literal operands issue at about 103 lane
// operations per cycle per CU and do not overlap the transcendental; in
// this loop the rsq costs about 3.9 FMA slots (measured: the loop with the
// rsq replaced by a multiply takes 14.95 against 18.3 ms at N = 262,144).
// Tile sums are FP32, added into a Kahan-compensated FP32 accumulator.
template <int IPT, int TJ, bool MASK>
__global__ __launch_bounds__(kWave) void bf_mixed_kernel(const double4* __restrict__ pm, int n,
                                                         int jchunk, float eps2, double near2,
                                                         const double* __restrict__ tbox,
                                                         double* __restrict__ part) {
    static_assert(TJ % kWave == 0, "tile must be a multiple of the wavefront");
    constexpr int LPT = TJ / kWave;  // j records each lane stages per tile
    __shared__ double4 tile[TJ];     // near: FP64 records; far: float4 [0, TJ)

    const int lane = threadIdx.x;
    const int ibase = blockIdx.x * kWave * IPT;
    int iidx[IPT];
    double lo[3], hi[3];
    float xi[IPT], yi[IPT], zi[IPT];
    double cx, cy, cz, lim2;
    {
        double xd[IPT], yd[IPT], zd[IPT];
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            const int i = ibase + k * kWave + lane;
            iidx[k] = i < n ? i : ibase;  // out-of-range targets mirror a valid one
            const double4 p = pm[iidx[k]];
            xd[k] = p.x; yd[k] = p.y; zd[k] = p.z;
        }
        wave_box<IPT>(xd, yd, zd, lo, hi);
        cx = 0.5 * (lo[0] + hi[0]);
        cy = 0.5 * (lo[1] + hi[1]);
        cz = 0.5 * (lo[2] + hi[2]);
        double di2 = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) di2 = __builtin_fma(hi[d] - lo[d], hi[d] - lo[d], di2);
        lim2 = near2 * di2;
        // softening above q * diag bounds every pair's error: no near tiles
        if ((double)eps2 >= lim2) lim2 = 0.0;
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            xi[k] = (float)(xd[k] - cx);
            yi[k] = (float)(yd[k] - cy);
            zi[k] = (float)(zd[k] - cz);
        }
    }
    float sx[IPT], sy[IPT], sz[IPT], kx[IPT], ky[IPT], kz[IPT];
#pragma unroll
    for (int k = 0; k < IPT; ++k) { sx[k] = sy[k] = sz[k] = 0.f; kx[k] = ky[k] = kz[k] = 0.f; }
    const float e2 = in_vgpr(eps2);  // full-rate operand of the r^2 FMA

    const int j0 = blockIdx.y * jchunk;
    const int j1 = min(n, j0 + jchunk);
    const int ntiles = (j1 - j0 + TJ - 1) / TJ;
    const int tile0 = j0 / TJ;

    double4 stage[LPT];
    auto load = [&](int t) {
#pragma unroll
        for (int e = 0; e < LPT; ++e) {
            const int j = j0 + t * TJ + e * kWave + lane;
            // padding: zero mass at the box centre (no contribution)
            stage[e] = j < j1 ? pm[j] : make_double4(cx, cy, cz, 0.0);
        }
    };
    auto store = [&](bool near) {
        if (near) {
#pragma unroll
            for (int e = 0; e < LPT; ++e) tile[e * kWave + lane] = stage[e];
        } else {
            float4* t32 = reinterpret_cast<float4*>(&tile[0]);
#pragma unroll
            for (int e = 0; e < LPT; ++e)
                t32[e * kWave + lane] =
                    make_float4((float)(stage[e].x - cx), (float)(stage[e].y - cy),
                                (float)(stage[e].z - cz), (float)stage[e].w);
        }
    };

    bool near_cur = false;
    if (ntiles > 0) {
        near_cur = tile_is_near(tbox + (size_t)tile0 * kTileBox, lo, hi, lim2);
        load(0);
    }
    for (int t = 0; t < ntiles; ++t) {
        store(near_cur);
        __syncthreads();
        bool near_next = false;
        if (t + 1 < ntiles) {
            near_next = tile_is_near(tbox + (size_t)(tile0 + t + 1) * kTileBox, lo, hi, lim2);
            load(t + 1);
        }
        float ax[IPT], ay[IPT], az[IPT];
#pragma unroll
        for (int k = 0; k < IPT; ++k) ax[k] = ay[k] = az[k] = 0.f;
        if (!near_cur) {
            const float4* t32 = reinterpret_cast<const float4*>(&tile[0]);
            // unroll depth chosen for register pressure: 63 VGPRs (8 waves
            // per SIMD) at IPT = 1, 84 (6 waves) at IPT = 2
            constexpr int kUnroll = IPT == 1 ? 4 : 8;
#pragma unroll kUnroll
            for (int j = 0; j < TJ; ++j) {
                const float4 s = t32[j];
#pragma unroll
                for (int k = 0; k < IPT; ++k) {
                    const float dx = s.x - xi[k], dy = s.y - yi[k], dz = s.z - zi[k];
                    const float f = pair_f32<MASK>(dx, dy, dz, s.w, e2);
                    ax[k] = __builtin_fmaf(f, dx, ax[k]);
                    ay[k] = __builtin_fmaf(f, dy, ay[k]);
                    az[k] = __builtin_fmaf(f, dz, az[k]);
                }
            }
        } else {
            double px[IPT], py[IPT], pz[IPT];
#pragma unroll
            for (int k = 0; k < IPT; ++k) {
                const double4 p = pm[iidx[k]];
                px[k] = p.x; py[k] = p.y; pz[k] = p.z;
            }
#pragma unroll 1
            for (int j = 0; j < TJ; ++j) {
                const double4 s = tile[j];
                const float m = (float)s.w;
#pragma unroll
                for (int k = 0; k < IPT; ++k) {
                    const float dx = (float)(s.x - px[k]);
                    const float dy = (float)(s.y - py[k]);
                    const float dz = (float)(s.z - pz[k]);
                    const float d2 = MASK ? __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)) : 0.f;
                    const float r2 = MASK ? d2 + e2
                                          : __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, e2)));
                    float ri = __builtin_amdgcn_rsqf(r2);
                    const float h = (r2 * ri) * ri;
                    ri = ri * __builtin_fmaf(-0.5f, h, 1.5f);
                    if (MASK) ri = d2 > 0.0f ? ri : 0.0f;
                    const float f = (m * ri) * (ri * ri);
                    ax[k] = __builtin_fmaf(f, dx, ax[k]);
                    ay[k] = __builtin_fmaf(f, dy, ay[k]);
                    az[k] = __builtin_fmaf(f, dz, az[k]);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            kahan_add(sx[k], kx[k], ax[k]);
            kahan_add(sy[k], ky[k], ay[k]);
            kahan_add(sz[k], kz[k], az[k]);
        }
        near_cur = near_next;
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const int i = ibase + k * kWave + lane;
        if (i < n) {
            const size_t base = (size_t)blockIdx.y * 3 * n;
            part[base + i] = (double)sx[k] - (double)kx[k];
            part[base + n + i] = (double)sy[k] - (double)ky[k];
            part[base + 2 * (size_t)n + i] = (double)sz[k] - (double)kz[k];
        }
    }
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// The r^2 product of one 16 x 16 block: FP32 operands on
// v_mfma_f32_16x16x4f32 (K = 4), or bf16 operands on
// v_mfma_f32_16x16x32_bf16 (K = 32, each FP32 value split in three exact
// bf16 pieces; half the matrix-core cycles).
__device__ __forceinline__ f32x4 mx_mfma(float a, float b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mx_mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// x = hi + mid + lo exactly for FP32 x (8 significant bits each; the
// residuals are exact in FP32).
__device__ __forceinline__ void split_bf16x3(float x, __bf16& hi, __bf16& mid, __bf16& lo) {
    hi = (__bf16)x;
    const float r1 = x - (float)hi;
    mid = (__bf16)r1;
    lo = (__bf16)(r1 - (float)mid);
}

// K slots of the bf16 form: for component c, slots 6c..6c+5 hold the
// products (b, a') pieces (hi,hi) (hi,mid) (mid,hi) (hi,lo) (mid,mid)
// (lo,hi) -- every product above 2^-24 of |b||a'| -- slots 18..20 hold
// 1 x (|a|^2 + eps^2) in three pieces, slots 21..31 are zero.
__device__ constexpr int kSplitB[6] = {0, 0, 1, 0, 1, 2};
__device__ constexpr int kSplitA[6] = {0, 1, 0, 2, 1, 0};

// One source block of the matrix-core far loop: the pair forces of block
// sbc (products dc; its source terms are read from LDS here rather than a
// block ahead, so the split-bf16 kernel fits 128 VGPRs, four wavefronts per
// SIMD: 0.5-1.5 % faster than three with the early read) are summed into acc
// while the products of block sbn are issued (MORE), one product per
// 16-target block between the VALU work of the previous one. T is the
// operand type (float or bf16x8).
template <bool MORE, bool MASK, typename T>
__device__ __forceinline__ void mx_block(const T* __restrict__ aop,
                                         const float4* __restrict__ qop,
                                         const float* __restrict__ bbq, int sbc, int sbn,
                                         int lane, int grp, const T (&bop)[4],
                                         const f32x4 (&dc)[4], f32x4 (&dn)[4],
                                         float (&acc)[4][4]) {
    T avn = {};
    f32x4 cbn = {0.f, 0.f, 0.f, 0.f};
    float4 qc[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) qc[v] = qop[sbc * 16 + 4 * grp + v];
    if (MORE) {
        avn = aop[sbn * 64 + lane];
        const float4 c4 = *reinterpret_cast<const float4*>(bbq + sbn * 16 + 4 * grp);
        cbn = f32x4{c4.x, c4.y, c4.z, c4.w};
    }
#pragma unroll
    for (int tb = 0; tb < 4; ++tb) {
        if (MORE) dn[tb] = mx_mfma(avn, bop[tb], cbn);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const float r2 = dc[tb][v];
            float ri = __builtin_amdgcn_rsqf(r2);
            if (MASK) ri = r2 > 0.0f ? ri : 0.0f;
            const float f = (ri * ri) * ri;
            acc[tb][0] = __builtin_fmaf(f, qc[v].x, acc[tb][0]);
            acc[tb][1] = __builtin_fmaf(f, qc[v].y, acc[tb][1]);
            acc[tb][2] = __builtin_fmaf(f, qc[v].z, acc[tb][2]);
            acc[tb][3] = __builtin_fmaf(f, qc[v].w, acc[tb][3]);
        }
    }
    if (MORE) {
#pragma unroll
        for (int tb = 0; tb < 4; ++tb) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, MASK ? 32 : 28, 0);
        }
    }
}

// Matrix-core variant of the mixed kernel. With a = x_i - c and b = x_j - c
// (FP32, c the wavefront's target-box centre),
//   |x_j - x_i|^2 + eps^2 = (b, 1) . (-2 a, |a|^2 + eps^2) + |b|^2,
// one v_mfma_f32_16x16x4f32 per 16 sources x 16 targets: A rows are sources
// (b_x, b_y, b_z, 1), B columns are targets (-2a_x, -2a_y, -2a_z, |a|^2 +
// eps^2), C is |b|^2 of the row's source. The acceleration is then
//   sum_j f_ij m_j (b_j - a_i) = sum_j f_ij (m_j b_j) - a_i sum_j f_ij m_j,
// f = r^-3, so each pair costs one v_rsq_f32, two multiplies and four FMAs
// on the VALU (six operations instead of twelve). Lane l of the product holds
// rows 4 (l / 16) + v, v = 0..3, of target column l % 16: each lane sums for
// one target per 16-target block (64 targets, four blocks) over four sources
// per block; the four lane groups are added at the end of every tile and the
// lane of target l keeps its Kahan-compensated sums. The products of the next
// source block are issued between the VALU work of the current one, so the
// matrix core runs beside the VALU.
// Error: the expansion rounds r^2 + eps^2 with an absolute error of about
// u (|a| + |b|)^2 <= u (diag + |d|)^2, i.e. a relative error of
// u (1 + 1 / q')^2 when max(|d|, eps) >= q' diag. The caller passes q' with
// (1 + 1 / q')^2 = 2 (1 + 1 / q), which keeps the per-pair bound of the VALU
// kernel for the context's q; nearer tiles take the FP64-separation path.
template <int TJ, bool MASK, bool BF, int WPE = (TJ == 128 ? 4 : 3)>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(WPE, 8))) void bf_matrix_kernel(const double4* __restrict__ pm, int n,
                                                          int jchunk, float eps2, double near2,
                                                          const double* __restrict__ tbox,
                                                          double* __restrict__ part) {
    static_assert(TJ % kWave == 0, "tile must be a multiple of the wavefront");
    constexpr int LPT = TJ / kWave;  // records each lane stages per tile
    constexpr int SB = TJ / 16;      // 16-source blocks per tile (even)
    using T = typename std::conditional<BF, bf16x8, float>::type;
    // near: TJ FP64 records; far: A operands [SB][4][16] of type T, then
    // (m b, m) and |b|^2 per source (36 or 84 TJ bytes)
    constexpr int kA = BF ? 4 * TJ : TJ;  // float4 units of the A operands
    __shared__ float4 lds[kA + TJ + TJ / 4];
    double4* tile = reinterpret_cast<double4*>(&lds[0]);
    T* aop = reinterpret_cast<T*>(&lds[0]);
    float4* qop = &lds[kA];
    float* bbq = reinterpret_cast<float*>(&lds[kA + TJ]);

    const int lane = threadIdx.x;
    const int col = lane & 15, grp = lane >> 4;
    const int ibase = blockIdx.x * kWave;
    const int iself = ibase + lane < n ? ibase + lane : ibase;
    double lo[3], hi[3];
    double cx, cy, cz, lim2;
    float ax, ay, az;  // this lane's target relative to c (set-up only)
    {
        const double4 p = pm[iself];
        const double xd[1] = {p.x}, yd[1] = {p.y}, zd[1] = {p.z};
        wave_box<1>(xd, yd, zd, lo, hi);
        cx = 0.5 * (lo[0] + hi[0]);
        cy = 0.5 * (lo[1] + hi[1]);
        cz = 0.5 * (lo[2] + hi[2]);
        double di2 = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) di2 = __builtin_fma(hi[d] - lo[d], hi[d] - lo[d], di2);
        lim2 = near2 * di2;
        if ((double)eps2 >= lim2) lim2 = 0.0;
        ax = (float)(p.x - cx);
        ay = (float)(p.y - cy);
        az = (float)(p.z - cz);
    }
    // B operand of each 16-target block
    T bop[4];
#pragma unroll
    for (int tb = 0; tb < 4; ++tb) {
        const float tx = __shfl(ax, 16 * tb + col);
        const float ty = __shfl(ay, 16 * tb + col);
        const float tz = __shfl(az, 16 * tb + col);
        const float aa = __builtin_fmaf(tx, tx, __builtin_fmaf(ty, ty, __builtin_fmaf(tz, tz, eps2)));
        if constexpr (BF) {
            // slots 8 grp .. 8 grp + 7 of this lane's target column
            __bf16 pc[4][3];
            split_bf16x3(-2.0f * tx, pc[0][0], pc[0][1], pc[0][2]);
            split_bf16x3(-2.0f * ty, pc[1][0], pc[1][1], pc[1][2]);
            split_bf16x3(-2.0f * tz, pc[2][0], pc[2][1], pc[2][2]);
            split_bf16x3(aa, pc[3][0], pc[3][1], pc[3][2]);
            // every slot index is a compile-time constant (a runtime index
            // into pc would place it in scratch memory); grp selects
            bf16x8 v;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                __bf16 c4[4];
#pragma unroll
                for (int gs = 0; gs < 4; ++gs) {
                    const int kk = 8 * gs + j;
                    c4[gs] = kk < 18 ? pc[kk / 6][kSplitA[kk % 6]]
                                     : (kk < 21 ? pc[3][kk - 18] : (__bf16)0.0f);
                }
                v[j] = grp == 0 ? c4[0] : grp == 1 ? c4[1] : grp == 2 ? c4[2] : c4[3];
            }
            bop[tb] = v;
        } else {
            bop[tb] = grp == 0 ? -2.0f * tx : grp == 1 ? -2.0f * ty : grp == 2 ? -2.0f * tz : aa;
        }
    }
    float s[4], k[4];  // target ibase + lane: x, y, z, sum f m
#pragma unroll
    for (int q = 0; q < 4; ++q) s[q] = k[q] = 0.f;

    const int j0 = blockIdx.y * jchunk;
    const int j1 = min(n, j0 + jchunk);
    const int ntiles = (j1 - j0 + TJ - 1) / TJ;
    const int tile0 = j0 / TJ;

    double4 stage[LPT];
    auto load = [&](int t) {
#pragma unroll
        for (int e = 0; e < LPT; ++e) {
            const int j = j0 + t * TJ + e * kWave + lane;
            stage[e] = j < j1 ? pm[j] : make_double4(cx, cy, cz, 0.0);
        }
    };
    auto store = [&](bool near) {
        if (near) {
#pragma unroll
            for (int e = 0; e < LPT; ++e) tile[e * kWave + lane] = stage[e];
        } else {
#pragma unroll
            for (int e = 0; e < LPT; ++e) {
                const int sj = e * kWave + lane;
                const float bx = (float)(stage[e].x - cx);
                const float by = (float)(stage[e].y - cy);
                const float bz = (float)(stage[e].z - cz);
                const float m = (float)stage[e].w;
                if constexpr (BF) {
                    // row sj & 15 of block sj >> 4, slot groups 0..3
                    __bf16 pc[3][3];
                    split_bf16x3(bx, pc[0][0], pc[0][1], pc[0][2]);
                    split_bf16x3(by, pc[1][0], pc[1][1], pc[1][2]);
                    split_bf16x3(bz, pc[2][0], pc[2][1], pc[2][2]);
                    T* a = aop + (sj >> 4) * 64 + (sj & 15);
#pragma unroll
                    for (int kg = 0; kg < 4; ++kg) {
                        bf16x8 v;
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            const int kk = 8 * kg + j;
                            v[j] = kk < 18 ? pc[kk / 6][kSplitB[kk % 6]]
                                           : (kk < 21 ? (__bf16)1.0f : (__bf16)0.0f);
                        }
                        a[16 * kg] = v;
                    }
                } else {
                    float* a = aop + (sj >> 4) * 64 + (sj & 15);
                    a[0] = bx;
                    a[16] = by;
                    a[32] = bz;
                    a[48] = 1.0f;
                }
                qop[sj] = make_float4(m * bx, m * by, m * bz, m);
                bbq[sj] = __builtin_fmaf(bx, bx, __builtin_fmaf(by, by, bz * bz));
            }
        }
    };

    // records are loaded at the top of each tile (no register prefetch: the
    // registers go to occupancy, and other wavefronts cover the latency)
    for (int t = 0; t < ntiles; ++t) {
        const bool near_cur = tile_is_near(tbox + (size_t)(tile0 + t) * kTileBox, lo, hi, lim2);
        load(t);
        store(near_cur);
        __syncthreads();
        float own[4];
        if (!near_cur) {
            float acc[4][4];
#pragma unroll
            for (int tb = 0; tb < 4; ++tb)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[tb][q] = 0.f;
            f32x4 dA[4], dB[4];
            {
                const T av = aop[lane];
                const float4 c4 = *reinterpret_cast<const float4*>(bbq + 4 * grp);
                const f32x4 cb = {c4.x, c4.y, c4.z, c4.w};
#pragma unroll
                for (int tb = 0; tb < 4; ++tb)
                    dA[tb] = mx_mfma(av, bop[tb], cb);
            }
            // blocks in pairs (A current, B next; then B current, A next)
            int sb = 0;
#pragma unroll 1
            for (; sb + 2 < SB; sb += 2) {
                mx_block<true, MASK>(aop, qop, bbq, sb, sb + 1, lane, grp, bop, dA, dB, acc);
                mx_block<true, MASK>(aop, qop, bbq, sb + 1, sb + 2, lane, grp, bop, dB, dA, acc);
            }
            mx_block<true, MASK>(aop, qop, bbq, SB - 2, SB - 1, lane, grp, bop, dA, dB, acc);
            mx_block<false, MASK>(aop, qop, bbq, SB - 1, 0, lane, grp, bop, dB, dA, acc);
            // add the four lane groups; lane l keeps block grp (target l)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
#pragma unroll
                for (int tb = 0; tb < 4; ++tb) {
                    float v = acc[tb][q];
                    v += __shfl_xor(v, 16);
                    v += __shfl_xor(v, 32);
                    acc[tb][q] = v;
                }
                own[q] = grp == 0 ? acc[0][q] : grp == 1 ? acc[1][q] : grp == 2 ? acc[2][q] : acc[3][q];
            }
        } else {
            // FP64 separations, one target per lane (target ibase + lane)
            const double4 p = pm[iself];
            float nx = 0.f, ny = 0.f, nz = 0.f;
#pragma unroll 1
            for (int j = 0; j < TJ; ++j) {
                const double4 sj = tile[j];
                const float m = (float)sj.w;
                const float dx = (float)(sj.x - p.x);
                const float dy = (float)(sj.y - p.y);
                const float dz = (float)(sj.z - p.z);
                const float d2 = MASK ? __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)) : 0.f;
                const float r2 = MASK ? d2 + eps2
                                      : __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
                float ri = __builtin_amdgcn_rsqf(r2);
                const float h = (r2 * ri) * ri;
                ri = ri * __builtin_fmaf(-0.5f, h, 1.5f);
                if (MASK) ri = d2 > 0.0f ? ri : 0.0f;
                const float f = (m * ri) * (ri * ri);
                nx = __builtin_fmaf(f, dx, nx);
                ny = __builtin_fmaf(f, dy, ny);
                nz = __builtin_fmaf(f, dz, nz);
            }
            own[0] = nx;
            own[1] = ny;
            own[2] = nz;
            own[3] = 0.f;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) kahan_add(s[q], k[q], own[q]);
        __syncthreads();
    }
    const int i = ibase + lane;
    if (i < n) {
        double o[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) o[q] = (double)s[q] - (double)k[q];
        // the target offset a (FP32, as in the products), formed again here
        // so it holds no registers through the loop
        const double4 p = pm[i];
        const size_t base = (size_t)blockIdx.y * 3 * n;
        part[base + i] = o[0] - (double)(float)(p.x - cx) * o[3];
        part[base + n + i] = o[1] - (double)(float)(p.y - cy) * o[3];
        part[base + 2 * (size_t)n + i] = o[2] - (double)(float)(p.z - cz) * o[3];
    }
}

template <int BLOCK, int IPT, int TJ, bool MASK>
__global__ __launch_bounds__(BLOCK) void bf_fp64_kernel(const double4* __restrict__ pm, int n,
                                                        int jchunk, double eps2,
                                                        double* __restrict__ part) {
    static_assert(TJ % BLOCK == 0, "tile must be a multiple of the block");
    constexpr int LPT = TJ / BLOCK;
    __shared__ double4 tile[2][TJ];

    const int ibase = blockIdx.x * BLOCK * IPT;
    double xi[IPT], yi[IPT], zi[IPT], sx[IPT], sy[IPT], sz[IPT];
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        int i = ibase + k * BLOCK + threadIdx.x;
        i = i < n ? i : ibase;
        const double4 p = pm[i];
        xi[k] = p.x; yi[k] = p.y; zi[k] = p.z;
        sx[k] = sy[k] = sz[k] = 0.0;
    }
    const int j0 = blockIdx.y * jchunk;
    const int j1 = min(n, j0 + jchunk);
    const int ntiles = (j1 - j0 + TJ - 1) / TJ;
    // Padding records: zero mass placed far away (no contribution, r^2 > 0).
    const double4 pad = make_double4(1e150, 1e150, 1e150, 0.0);
    const double e2 = in_vgpr(eps2);  // VGPR operand of the r^2 FMA
    const Rinv3Consts k3 = rinv3_consts();

    double4 stage[LPT];
    auto load = [&](int t) {
#pragma unroll
        for (int e = 0; e < LPT; ++e) {
            const int j = j0 + t * TJ + e * BLOCK + threadIdx.x;
            stage[e] = j < j1 ? pm[j] : pad;
        }
    };
    auto store = [&](int b) {
#pragma unroll
        for (int e = 0; e < LPT; ++e) tile[b][e * BLOCK + threadIdx.x] = stage[e];
    };
    if (ntiles > 0) {
        load(0);
        store(0);
    }
    __syncthreads();
    for (int t = 0; t < ntiles; ++t) {
        const int cur = t & 1;
        if (t + 1 < ntiles) load(t + 1);
        double ax[IPT], ay[IPT], az[IPT];
#pragma unroll
        for (int k = 0; k < IPT; ++k) ax[k] = ay[k] = az[k] = 0.0;
#pragma unroll 4
        for (int j = 0; j < TJ; ++j) {
            const double4 s = tile[cur][j];
#pragma unroll
            for (int k = 0; k < IPT; ++k) {
                const double dx = s.x - xi[k], dy = s.y - yi[k], dz = s.z - zi[k];
                const double f = pair_f64<MASK>(dx, dy, dz, s.w, e2, k3);
                ax[k] = __builtin_fma(f, dx, ax[k]);
                ay[k] = __builtin_fma(f, dy, ay[k]);
                az[k] = __builtin_fma(f, dz, az[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < IPT; ++k) { sx[k] += ax[k]; sy[k] += ay[k]; sz[k] += az[k]; }
        if (t + 1 < ntiles) store(cur ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const int i = ibase + k * BLOCK + threadIdx.x;
        if (i < n) {
            const size_t base = (size_t)blockIdx.y * 3 * n;
            part[base + i] = sx[k];
            part[base + n + i] = sy[k];
            part[base + 2 * (size_t)n + i] = sz[k];
        }
    }
}

// Leapfrog KDK of a small system (n <= kSmallN) with FP64 brute force, all
// nsteps in one launch of one workgroup: thread i keeps particle i's state
// in registers, the positions and masses of the step go through LDS. Every
// operation is the multi-launch path's (kick_drift_kernel, bf_fp64_kernel
// with its single 256-record tile and zero-mass padding, finish_kernel's
// 0 + partial and kick), in the same order, so the results are
// bit-identical to it; only the launches between the steps are gone.
template <bool MASK>
__global__ __launch_bounds__(kSmallN) void kdk_small_kernel(
    int n, double* __restrict__ x, double* __restrict__ y, double* __restrict__ z,
    double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
    const double* __restrict__ m, double* __restrict__ ax, double* __restrict__ ay,
    double* __restrict__ az, double eps2, double h, double dt, uint32_t nsteps) {
    __shared__ double4 tile[kSmallN];
    const int i = threadIdx.x;
    const bool live = i < n;
    double px = 0, py = 0, pz = 0, pvx = 0, pvy = 0, pvz = 0, pm = 0, a0 = 0, a1 = 0, a2 = 0;
    if (live) {
        px = x[i]; py = y[i]; pz = z[i];
        pvx = vx[i]; pvy = vy[i]; pvz = vz[i];
        pm = m[i];
        a0 = ax[i]; a1 = ay[i]; a2 = az[i];
    }
    const double e2 = in_vgpr(eps2);
    const Rinv3Consts k3 = rinv3_consts();
    for (uint32_t step = 0; step < nsteps; ++step) {
        // v += a h; x += v dt (integrator.rs:53-64)
        pvx = pvx + a0 * h;
        pvy = pvy + a1 * h;
        pvz = pvz + a2 * h;
        px = px + pvx * dt;
        py = py + pvy * dt;
        pz = pz + pvz * dt;
        if (live) tile[i] = make_double4(px, py, pz, pm);
        __syncthreads();
        double sx = 0.0, sy = 0.0, sz = 0.0;
        // the records past n (zero-mass padding in bf_fp64_kernel's tile)
        // add m r^-3 d = 0 to a non-zero or +0 sum: skipping them changes
        // no bit
#pragma unroll 4
        for (int j = 0; j < n; ++j) {
            const double4 s = tile[j];
            const double dx = s.x - px, dy = s.y - py, dz = s.z - pz;
            const double f = pair_f64<MASK>(dx, dy, dz, s.w, e2, k3);
            sx = __builtin_fma(f, dx, sx);
            sy = __builtin_fma(f, dy, sy);
            sz = __builtin_fma(f, dz, sz);
        }
        // the partial of one chunk (0 + tile sum), then v += a h
        a0 = 0.0 + (0.0 + sx);
        a1 = 0.0 + (0.0 + sy);
        a2 = 0.0 + (0.0 + sz);
        pvx = pvx + a0 * h;
        pvy = pvy + a1 * h;
        pvz = pvz + a2 * h;
        __syncthreads();  // the tile is rewritten by the next step
    }
    if (live) {
        x[i] = px; y[i] = py; z[i] = pz;
        vx[i] = pvx; vy[i] = pvy; vz[i] = pvz;
        ax[i] = a0; ay[i] = a1; az[i] = a2;
    }
}

constexpr size_t kSortMinN = 2048;

template <int IPT, int TJ>
void launch_mixed(dim3 g, hipStream_t st, bool mask, const double4* pm, int n, int jchunk,
                  float eps2, double near2, const double* tbox, double* part) {
    if (mask)
        bf_mixed_kernel<IPT, TJ, true><<<g, kWave, 0, st>>>(pm, n, jchunk, eps2, near2, tbox,
                                                            part);
    else
        bf_mixed_kernel<IPT, TJ, false><<<g, kWave, 0, st>>>(pm, n, jchunk, eps2, near2, tbox,
                                                             part);
}

template <int TJ, bool BF>
void launch_matrix(dim3 g, hipStream_t st, bool mask, const double4* pm, int n, int jchunk,
                   float eps2, double near2, const double* tbox, double* part) {
    if (mask)
        bf_matrix_kernel<TJ, true, BF><<<g, kWave, 0, st>>>(pm, n, jchunk, eps2, near2, tbox,
                                                            part);
    else
        bf_matrix_kernel<TJ, false, BF><<<g, kWave, 0, st>>>(pm, n, jchunk, eps2, near2, tbox,
                                                             part);
}

template <int BLOCK, int IPT, int TJ>
void launch_fp64(dim3 g, hipStream_t st, bool mask, const double4* pm, int n, int jchunk,
                 double eps2, double* part) {
    if (mask)
        bf_fp64_kernel<BLOCK, IPT, TJ, true><<<g, BLOCK, 0, st>>>(pm, n, jchunk, eps2, part);
    else
        bf_fp64_kernel<BLOCK, IPT, TJ, false><<<g, BLOCK, 0, st>>>(pm, n, jchunk, eps2, part);
}

}  // namespace

namespace {
// Sum the S chunk partials in chunk order (G = 1) and deliver in the
// caller's particle order.
__global__ __launch_bounds__(256) void finish_kernel(const double* __restrict__ part,
                                                        int nchunks, int n,
                                                        const uint32_t* __restrict__ perm,
                                                        FinishArgs fa) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double a[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double s = 0.0;
        for (int q = 0; q < nchunks; ++q) s += part[((size_t)q * 3 + c) * n + k];
        a[c] = s;
    }
    const uint32_t o = perm ? perm[k] : (uint32_t)k;
    if (fa.mode == kAccumulate) {
        fa.ax[o] += a[0];
        fa.ay[o] += a[1];
        fa.az[o] += a[2];
    } else {
        fa.ax[o] = a[0];
        fa.ay[o] = a[1];
        fa.az[o] = a[2];
        fa.vx[o] = fa.vx[o] + a[0] * fa.half_dt;
        fa.vy[o] = fa.vy[o] + a[1] * fa.half_dt;
        fa.vz[o] = fa.vz[o] + a[2] * fa.half_dt;
    }
}

// inv[perm[k]] = k
__global__ __launch_bounds__(256) void finish_invert_kernel(const uint32_t* __restrict__ perm,
                                                            int n, uint32_t* __restrict__ inv) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) inv[perm[k]] = (uint32_t)k;
}

// One partial array, delivered by caller index o: the sorted position
// inv[o] is read at random and the caller's arrays are written (and, for
// the kick, read) in order, where finish_kernel scatters six arrays.
__global__ __launch_bounds__(256) void finish_inv_kernel(const double* __restrict__ part, int n,
                                                         const uint32_t* __restrict__ inv,
                                                         FinishArgs fa) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n) return;
    const size_t k = inv[o];
    // 0 + p, as finish_kernel's sum over one chunk forms it (-0 becomes +0)
    const double a0 = 0.0 + part[k], a1 = 0.0 + part[(size_t)n + k],
                 a2 = 0.0 + part[2 * (size_t)n + k];
    if (fa.mode == kAccumulate) {
        fa.ax[o] += a0;
        fa.ay[o] += a1;
        fa.az[o] += a2;
    } else {
        fa.ax[o] = a0;
        fa.ay[o] = a1;
        fa.az[o] = a2;
        fa.vx[o] = fa.vx[o] + a0 * fa.half_dt;
        fa.vy[o] = fa.vy[o] + a1 * fa.half_dt;
        fa.vz[o] = fa.vz[o] + a2 * fa.half_dt;
    }
}

}  // namespace

gravis_status finish_forces(Ctx* c, const double* part, int nchunks, int n,
                            const uint32_t* perm, const FinishArgs& fa, const uint32_t* inv) {
    if (perm && nchunks == 1 && n > 0) {
        // one partial array: deliver in the caller's order through the
        // inverse permutation (same values, same arithmetic per particle)
        if (!inv) {
            uint32_t* w;
            GRAVIS_TRY(buf(c, "finish_inv", (size_t)n, &w));
            finish_invert_kernel<<<grid_for((size_t)n, 256), 256, 0, c->stream>>>(perm, n, w);
            GRAVIS_TRY(check_launch(c, "invert_perm"));
            inv = w;
        }
        finish_inv_kernel<<<grid_for((size_t)n, 256), 256, 0, c->stream>>>(part, n, inv, fa);
        return check_launch(c, "finish_forces");
    }
    finish_kernel<<<grid_for((size_t)n, 256), 256, 0, c->stream>>>(part, nchunks, n, perm, fa);
    return check_launch(c, "finish_forces");
}

gravis_status bf_force(Ctx* c, const gravis_solver* s, size_t n, const double* x,
                       const double* y, const double* z, const double* m, const FinishArgs& fa) {
    if (n == 0) return GRAVIS_OK;
    if (n > 0x7FFFFFFFull) return fail(c, GRAVIS_INVALID_ARG, "n too large");
    const int ni = (int)n;
    const bool mixed = s->precision == GRAVIS_MIXED;
    // Mixed precision forms separations relative to each workgroup's target
    // box, so targets are Morton sorted to make those boxes compact. FP64
    // needs no ordering and runs on the caller's order.
    uint32_t* perm = nullptr;
    if (mixed && n >= kSortMinN) {
        double* box;
        GRAVIS_TRY(buf(c, "bf_box", 4, &box));
        GRAVIS_TRY(root_box(c, n, x, y, z, box));
        uint64_t* keys;
        GRAVIS_TRY(morton_sort(c, n, x, y, z, box, &keys, &perm));
    }
    double4* pm;
    GRAVIS_TRY(buf(c, "bf_pm", n, &pm));
    const uint32_t* inv = nullptr;
    GRAVIS_TRY(gather_sorted(c, n, x, y, z, m, perm, pm, &inv));

    // FP64: 256-thread workgroups of 256 targets. Mixed: one wavefront per
    // workgroup with bf_ipt * 64 targets (64 for the matrix-core kernel) and
    // bf_tile-record tiles.
    const bool matrix = mixed && c->bf_matrix;
    const int ipt = mixed && !matrix ? c->bf_ipt : 1;
    const int group = mixed ? kWave * ipt : 256;
    const int tj = mixed ? c->bf_tile : 256;
    const int gx = (ni + group - 1) / group;
    // Split the j range so the grid holds about bf_waves_per_cu wavefronts
    // per CU, keeping at least one tile per chunk.
    const int waves_per_group = mixed ? 1 : 4;
    const int target = (c->bf_waves_per_cu * c->num_cus + waves_per_group - 1) / waves_per_group;
    int S = (target + gx - 1) / gx;
    const int max_s = (ni + tj - 1) / tj;
    if (S > max_s) S = max_s;
    if (c->arena_max_chunks > 0 && S > c->arena_max_chunks) S = c->arena_max_chunks;
    if (S < 1) S = 1;
    int jchunk = (ni + S - 1) / S;
    jchunk = (jchunk + tj - 1) / tj * tj;
    S = (ni + jchunk - 1) / jchunk;

    double* part;
    GRAVIS_TRY(buf(c, "bf_part", (size_t)S * 3 * n, &part));
    const dim3 grid(gx, S);
    const double eps2 = s->softening * s->softening;
    double* tbox = nullptr;
    if (mixed) {
        const int ntile = (ni + tj - 1) / tj;
        GRAVIS_TRY(buf(c, "bf_tbox", (size_t)ntile * kTileBox, &tbox));
        if (tj == 256)
            tile_box_kernel<256><<<ntile, 256, 0, c->stream>>>(pm, ni, tbox);
        else
            tile_box_kernel<128><<<ntile, 256, 0, c->stream>>>(pm, ni, tbox);
        GRAVIS_TRY(check_launch(c, "tile_box"));
    }
    timer_begin(c);
    if (mixed) {
        const float eps2f = (float)eps2;
        const bool mask = !(eps2f >= kMaskEps2F);
        const double q = c->bf_near_fraction;
        if (matrix) {
            // q' with (1 + 1/q')^2 = 2 (1 + 1/q): the same per-pair bound
            const double qm = q > 0.0 ? 1.0 / (std::sqrt(2.0 * (1.0 + 1.0 / q)) - 1.0) : 0.0;
            if (tj == 128 && c->bf_matrix == 2)
                launch_matrix<128, true>(grid, c->stream, mask, pm, ni, jchunk, eps2f, qm * qm,
                                         tbox, part);
            else if (tj == 128)
                launch_matrix<128, false>(grid, c->stream, mask, pm, ni, jchunk, eps2f, qm * qm,
                                          tbox, part);
            else if (c->bf_matrix == 2)
                launch_matrix<256, true>(grid, c->stream, mask, pm, ni, jchunk, eps2f, qm * qm,
                                         tbox, part);
            else
                launch_matrix<256, false>(grid, c->stream, mask, pm, ni, jchunk, eps2f, qm * qm,
                                          tbox, part);
            timer_end(c, "bf_matrix");
            GRAVIS_TRY(check_launch(c, "bf kernel"));
            return finish_forces(c, part, S, ni, perm, fa, inv);
        }
        const double f2 = q * q;
        if (ipt == 1 && tj == 128)
            launch_mixed<1, 128>(grid, c->stream, mask, pm, ni, jchunk, eps2f, f2, tbox, part);
        else if (ipt == 1)
            launch_mixed<1, 256>(grid, c->stream, mask, pm, ni, jchunk, eps2f, f2, tbox, part);
        else if (tj == 128)
            launch_mixed<2, 128>(grid, c->stream, mask, pm, ni, jchunk, eps2f, f2, tbox, part);
        else
            launch_mixed<2, 256>(grid, c->stream, mask, pm, ni, jchunk, eps2f, f2, tbox, part);
        timer_end(c, "bf_mixed");
    } else {
        const bool mask = !(eps2 >= kMaskEps2D);
        launch_fp64<256, 1, 256>(grid, c->stream, mask, pm, ni, jchunk, eps2, part);
        timer_end(c, "bf_fp64");
    }
    GRAVIS_TRY(check_launch(c, "bf kernel"));
    return finish_forces(c, part, S, ni, perm, fa, inv);
}


gravis_status kdk_small(Ctx* c, size_t n, double* x, double* y, double* z, double* vx,
                        double* vy, double* vz, const double* m, double* ax, double* ay,
                        double* az, double softening, double dt, uint64_t nsteps) {
    if (n == 0 || n > (size_t)kSmallN) return fail(c, GRAVIS_INTERNAL, "kdk_small: n = %zu", n);
    const double eps2 = softening * softening;
    const double h = 0.5 * dt;
    const bool mask = !(eps2 >= kMaskEps2D);
    const unsigned threads = (unsigned)((n + kWave - 1) / kWave * kWave);
    // launches of at most 2^20 steps each
    while (nsteps > 0) {
        const uint32_t k = (uint32_t)std::min<uint64_t>(nsteps, 1u << 20);
        if (mask)
            kdk_small_kernel<true><<<1, threads, 0, c->stream>>>((int)n, x, y, z, vx, vy, vz, m,
                                                                 ax, ay, az, eps2, h, dt, k);
        else
            kdk_small_kernel<false><<<1, threads, 0, c->stream>>>((int)n, x, y, z, vx, vy, vz, m,
                                                                  ax, ay, az, eps2, h, dt, k);
        GRAVIS_TRY(check_launch(c, "kdk_small"));
        nsteps -= k;
    }
    return GRAVIS_OK;
}

}  // namespace gravis
