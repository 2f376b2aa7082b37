// matrix.h -- the matrix-core tile sweep of the mixed-precision brute force,
// shared by the brute-force kernel (bf.hip) and the ensemble (ensemble.hip).
//
// With a = x_i - c and b = x_j - c (FP32, c the wavefront's target-box
// centre),
//   |x_j - x_i|^2 + eps^2 = (b, 1) . (-2 a, |a|^2 + eps^2) + |b|^2,
// one matrix-core product per 16 sources x 16 targets: A rows are sources
// (b_x, b_y, b_z, 1), B columns are targets (-2a_x, -2a_y, -2a_z, |a|^2 +
// eps^2), C is |b|^2 of the row's source. The acceleration is then
//   sum_j f_ij m_j (b_j - a_i) = sum_j f_ij (m_j b_j) - a_i sum_j f_ij m_j,
// f = r^-3, so each pair costs one v_rsq_f32, two multiplies and four FMAs
// on the VALU. Lane l of the product holds rows 4 (l / 16) + v, v = 0..3, of
// target column l % 16: each lane sums for one target per 16-target block
// (64 targets, four blocks) over four sources per block; the four lane
// groups are added at the end of every tile and the lane of target l keeps
// its Kahan-compensated sums. The products of the next source block are
// issued between the VALU work of the current one, so the matrix core runs
// beside the VALU.
// Operands are FP32 on v_mfma_f32_16x16x4f32 or, BF, FP32 values split
// exactly in three bf16 pieces on v_mfma_f32_16x16x32_bf16.
// Error: the expansion rounds r^2 + eps^2 with an absolute error of about
// u (|a| + |b|)^2 <= u (diag + |d|)^2, i.e. a relative error of
// u (1 + 1 / q')^2 when max(|d|, eps) >= q' diag. The caller sorts every
// tile into one of three tiers: far (gap to the target box or eps at least
// q' diag) on the matrix cores; mid (at least q diag, (1 + 1/q')^2 =
// 2 (1 + 1/q)) in the vector-ALU form d = b - a on the FP32 offsets, whose
// per-pair error bound 3 u (1 + 1/q) is the same; near tiles with FP64
// separations, one target per lane.
#pragma once

#include <type_traits>

#include "internal.h"

namespace gravis {

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

// A operands per 16-source block in LDS: the FP32 form keeps all 64 (four
// K slots of 16 rows); the bf16 form keeps the slot groups 0..2 (48), since
// the B operand of slot group 3 (slots 24..31) is zero for every target.
// The lanes that supply group 3 (48..63) read 48 entries on, i.e. the next
// block's group 0 (finite split values) or, after the tile's last block, 16
// zero entries kept past the A operands; their products with the zero B
// slots add exactly zero.
template <typename T>
constexpr int mx_a_stride() {
    return std::is_same<T, float>::value ? 64 : 48;
}

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
        avn = aop[sbn * mx_a_stride<T>() + lane];
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


// LDS of one wavefront's tile in float4 units: near, TJ FP64 records; far,
// the A operands [TJ / 16][4 or 3][16] of type T (FP32: 4 floats, one
// float4, per source; bf16: 3 bf16x8 per source, then 16 zero entries),
// then (m b, m) and |b|^2 per source. 8.75 KB at TJ = 128 in the bf16 form:
// 18 wavefronts' worth per CU, above the 16 that 128 VGPRs allow (with all
// four slot groups, 10.5 KB held a CU to 15).
template <int TJ, bool BF>
constexpr int mx_a_f4() {
    return BF ? 3 * TJ + 16 : TJ;
}
template <int TJ, bool BF>
constexpr int mx_lds_f4() {
    return mx_a_f4<TJ, BF>() + TJ + TJ / 4;
}

// Tiers of a source tile.
constexpr int kTierFar = 0, kTierMid = 1, kTierNear = 2;

// The tile sweep of one wavefront: the lane's target is record iself of pm
// and (ax, ay, az) its FP32 offset from the target-box centre c; sources are
// records [j0, j1) of pm, TJ to a tile; tier(t) gives tile t's tier (wave
// uniform). Every lane of the wavefront takes part. o = (sum f m b_x,
// sum f m b_y, sum f m b_z, sum f m) over the far tiles, with the other
// tiles' accelerations added to o[0..2]: the target's acceleration is
// o[0..2] - a o[3].
template <int TJ, bool MASK, bool BF, class Tier>
__device__ __forceinline__ void mx_sweep(float4* lds, const double4* __restrict__ pm, int iself,
                                         int lane, double cx, double cy, double cz, float ax,
                                         float ay, float az, float eps2, int j0, int j1,
                                         Tier&& tier, double (&o)[4]) {
    static_assert(TJ % kWave == 0, "tile must be a multiple of the wavefront");
    constexpr int LPT = TJ / kWave;  // records each lane stages per tile
    constexpr int SB = TJ / 16;      // 16-source blocks per tile (even)
    using T = typename std::conditional<BF, bf16x8, float>::type;
    constexpr int kA = mx_a_f4<TJ, BF>();  // float4 units of the A operands
    constexpr int kAS = mx_a_stride<T>();  // A operands per 16-source block
    double4* tile = reinterpret_cast<double4*>(&lds[0]);
    T* aop = reinterpret_cast<T*>(&lds[0]);
    float4* qop = &lds[kA];
    float* bbq = reinterpret_cast<float*>(&lds[kA + TJ]);
    const int col = lane & 15, grp = lane >> 4;
    // the zero entries read by slot group 3 after a tile's last block
    // (outside every tier's staging; visible after the first tile's barrier)
    if (BF && lane < 16) aop[SB * kAS + lane] = T{};
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
    float s[4], k[4];  // the lane's target: x, y, z, sum f m
#pragma unroll
    for (int q = 0; q < 4; ++q) s[q] = k[q] = 0.f;

    const int ntiles = (j1 - j0 + TJ - 1) / TJ;

    double4 stage[LPT];
    auto load = [&](int t) {
#pragma unroll
        for (int e = 0; e < LPT; ++e) {
            const int j = j0 + t * TJ + e * kWave + lane;
            // conditional load (measured 0.9 % faster here than an
            // always-issued clamped load followed by a value select)
            stage[e] = j < j1 ? pm[j] : make_double4(cx, cy, cz, 0.0);
        }
    };
    auto store = [&](int tr) {
        if (tr == kTierNear) {
#pragma unroll
            for (int e = 0; e < LPT; ++e) tile[e * kWave + lane] = stage[e];
        } else if (tr == kTierMid) {
            // (b, m) per source
#pragma unroll
            for (int e = 0; e < LPT; ++e)
                lds[e * kWave + lane] =
                    make_float4((float)(stage[e].x - cx), (float)(stage[e].y - cy),
                                (float)(stage[e].z - cz), (float)stage[e].w);
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
                    T* a = aop + (sj >> 4) * kAS + (sj & 15);
#pragma unroll
                    for (int kg = 0; kg < 3; ++kg) {
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
                    float* a = aop + (sj >> 4) * kAS + (sj & 15);
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
        const int tr = tier(t);
        load(t);
        store(tr);
        __syncthreads();
        float own[4];
        if (tr == kTierFar) {
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
        } else if (tr == kTierMid) {
            // d = b - a on the FP32 offsets, one target per lane (a formed
            // again from the record: no register holds it through the sweep)
            const double4 p = pm[iself];
            const float qx = (float)(p.x - cx), qy = (float)(p.y - cy), qz = (float)(p.z - cz);
            const float e2 = in_vgpr(eps2);  // full-rate operand of the r^2 FMA
            float nx = 0.f, ny = 0.f, nz = 0.f;
#pragma unroll 4
            for (int j = 0; j < TJ; ++j) {
                const float4 sj = lds[j];
                const float dx = sj.x - qx, dy = sj.y - qy, dz = sj.z - qz;
                const float f = pair_f32<MASK>(dx, dy, dz, sj.w, e2);
                nx = __builtin_fmaf(f, dx, nx);
                ny = __builtin_fmaf(f, dy, ny);
                nz = __builtin_fmaf(f, dz, nz);
            }
            own[0] = nx;
            own[1] = ny;
            own[2] = nz;
            own[3] = 0.f;
        } else {
            // FP64 separations, one target per lane
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
#pragma unroll
    for (int q = 0; q < 4; ++q) o[q] = (double)s[q] - (double)k[q];
}

}  // namespace gravis
