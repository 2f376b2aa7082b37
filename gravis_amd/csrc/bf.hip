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
#include "matrix.h"

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
            stage[e] = record_or(pm, j, j1, make_double4(cx, cy, cz, 0.0));
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

// Matrix-core variant of the mixed kernel (the sweep, its tiers and error
// bound: matrix.h). One wavefront per 64 targets and j chunk. near2 = q'^2
// with (1 + 1 / q')^2 = 2 (1 + 1 / q), mid2 = q^2 for the context's q: the
// per-pair bound of the VALU kernel.
template <int TJ, bool MASK, bool BF, int WPE = (TJ == 128 ? 4 : 3)>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(WPE, 8))) void bf_matrix_kernel(const double4* __restrict__ pm, int n,
                                                          int jchunk, float eps2, double near2,
                                                          double mid2,
                                                          const double* __restrict__ tbox,
                                                          double* __restrict__ part) {
    __shared__ float4 lds[mx_lds_f4<TJ, BF>()];
    const int lane = threadIdx.x;
    const int ibase = blockIdx.x * kWave;
    const int iself = ibase + lane < n ? ibase + lane : ibase;
    double lo[3], hi[3];
    double cx, cy, cz, lim2, limv2;
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
        limv2 = mid2 * di2;
        if ((double)eps2 >= limv2) limv2 = 0.0;
        ax = (float)(p.x - cx);
        ay = (float)(p.y - cy);
        az = (float)(p.z - cz);
    }
    const int j0 = blockIdx.y * jchunk;
    const int j1 = min(n, j0 + jchunk);
    const int tile0 = j0 / TJ;
    double o[4];
    mx_sweep<TJ, MASK, BF>(
        lds, pm, iself, lane, cx, cy, cz, ax, ay, az, eps2, j0, j1,
        [&](int t) __attribute__((always_inline)) {
            const double* tb = tbox + (size_t)(tile0 + t) * kTileBox;
            if (!tile_is_near(tb, lo, hi, lim2)) return kTierFar;
            return tile_is_near(tb, lo, hi, limv2) ? kTierNear : kTierMid;
        },
        o);
    const int i = ibase + lane;
    if (i < n) {
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
            // (a select on the address: it keeps pad in scratch and the
            // kernel at 88 VGPRs, five wavefronts per SIMD, which measured
            // 1.3 % faster than the value select at 65 VGPRs and seven)
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

// Leapfrog KDK of a system of kSmallN < n <= kCoopMaxN particles with FP64
// brute force, all steps in one cooperative launch: every workgroup is
// resident at once, so the steps are separated by a grid barrier instead of
// kernel boundaries. Wavefront w of workgroup b owns target
// i = b * kCoopWaves + w for the whole launch; its lane 0 holds the
// target's position, velocity and acceleration in registers and writes
// them to the caller's arrays once, at the end. Per step every workgroup
// copies the step's n position records into LDS; each lane sums the
// sources j = lane, lane + 64, ... in order, and the 64 lane sums are added
// in a fixed butterfly, so a target's force is deterministic (it differs
// from the multi-launch path's chunk order in rounding only). Lane 0 then
// applies the step's closing kick and the next step's kick and drift, in
// the operations of kick_drift_kernel and finish_kernel, and stores the new
// record in the other half of pbuf: positions alternate between the
// halves, so one barrier per step suffices (a half is rewritten only after
// the barrier that follows every read of it).
//
// The records are exchanged with device-scope loads and stores, which are
// coherent across the XCDs' L2 caches without cache maintenance (agent-scope
// fences write back and invalidate a whole L2 per barrier: 31 against 9.6 us
// per step at n = 1000). Order: each wavefront waits for its record stores
// to complete before the workgroup barrier that precedes its workgroup's
// arrival (a device-scope add), and the record loads of the next step are
// issued after the workgroup barrier that follows the arrival count being
// seen, so they return every record stored before the count was reached.
// The wait is bounded: past kCoopSpins polls, or once another workgroup has
// given up, the barrier sets the abort word and the kernel ends; the host
// reports the failure.
constexpr int kCoopWaves = 16;
constexpr int kCoopThreads = kCoopWaves * kWave;
constexpr uint32_t kCoopSpins = 1u << 20;  // polls of about 1 us: about 1 s

__device__ __forceinline__ void store_record(double4* p, const double4 v) {
    double* d = reinterpret_cast<double*>(p);
    __hip_atomic_store(d + 0, v.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(d + 1, v.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(d + 2, v.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(d + 3, v.w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double4 load_record(const double4* p) {
    const double* d = reinterpret_cast<const double*>(p);
    return make_double4(__hip_atomic_load(d + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                        __hip_atomic_load(d + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                        __hip_atomic_load(d + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                        __hip_atomic_load(d + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
// this wavefront's stores complete (vector memory count 0); the signal
// fences keep the compiler from moving memory operations across the wait
__device__ __forceinline__ void wait_stores() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Grid barrier on a monotonic arrival count: the k-th barrier (k = 1, 2,
// ...) is passed once bar[0] >= k * gridDim.x. bar[1] is the abort word.
// Returns false if the kernel is to end.
__device__ __forceinline__ bool coop_barrier(uint32_t* __restrict__ bar, uint32_t target,
                                             int* ok) {
    __syncthreads();
    if (threadIdx.x == 0) {
        int good = 1;
        // the last arrival need not poll
        if (atomicAdd(&bar[0], 1u) + 1u < target) {
            for (uint32_t spin = 0;; ++spin) {
                if (__hip_atomic_load(&bar[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >=
                    target)
                    break;
                if ((spin & 63u) == 63u &&
                    __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
                    good = 0;
                    break;
                }
                if (spin >= kCoopSpins) {
                    atomicOr(&bar[1], 1u);
                    good = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        *ok = good;
    }
    __syncthreads();
    return *ok != 0;
}

template <bool MASK>
__global__ __launch_bounds__(kCoopThreads) void kdk_coop_kernel(
    int n, double* __restrict__ x, double* __restrict__ y, double* __restrict__ z,
    double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
    const double* __restrict__ m, double* __restrict__ ax, double* __restrict__ ay,
    double* __restrict__ az, double eps2, double h, double dt, uint32_t nsteps,
    double4* __restrict__ pbuf, uint32_t* __restrict__ bar) {
    extern __shared__ double4 src[];  // the step's n position records
    __shared__ int ok;
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int i = blockIdx.x * kCoopWaves + w;  // wave-uniform
    const bool own = i < n;
    double px = 0.0, py = 0.0, pz = 0.0, pvx = 0.0, pvy = 0.0, pvz = 0.0, pmass = 0.0;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    if (own && lane == 0) {
        px = x[i]; py = y[i]; pz = z[i];
        pvx = vx[i]; pvy = vy[i]; pvz = vz[i];
        a0 = ax[i]; a1 = ay[i]; a2 = az[i];
        pmass = m[i];
        // the first step's kick and drift (integrator.rs:53-64)
        pvx = pvx + a0 * h;
        pvy = pvy + a1 * h;
        pvz = pvz + a2 * h;
        px = px + pvx * dt;
        py = py + pvy * dt;
        pz = pz + pvz * dt;
        store_record(pbuf + i, make_double4(px, py, pz, pmass));
    }
    wait_stores();
    const double e2 = in_vgpr(eps2);
    const Rinv3Consts k3 = rinv3_consts();
    uint32_t target = 0;
    for (uint32_t step = 0; step < nsteps; ++step) {
        target += gridDim.x;
        if (!coop_barrier(bar, target, &ok)) return;
        const double4* cur = pbuf + (size_t)(step & 1u) * n;
        double4* nxt = pbuf + (size_t)((step + 1u) & 1u) * n;
        for (int j = threadIdx.x; j < n; j += kCoopThreads) src[j] = load_record(cur + j);
        __syncthreads();
        if (own) {
            const double4 me = src[i];
            double sx = 0.0, sy = 0.0, sz = 0.0;
#pragma unroll 4
            for (int j = lane; j < n; j += kWave) {
                const double4 s = src[j];
                const double dx = s.x - me.x, dy = s.y - me.y, dz = s.z - me.z;
                const double f = pair_f64<MASK>(dx, dy, dz, s.w, e2, k3);
                sx = __builtin_fma(f, dx, sx);
                sy = __builtin_fma(f, dy, sy);
                sz = __builtin_fma(f, dz, sz);
            }
#pragma unroll
            for (int o = kWave / 2; o > 0; o >>= 1) {
                sx += __shfl_xor(sx, o);
                sy += __shfl_xor(sy, o);
                sz += __shfl_xor(sz, o);
            }
            if (lane == 0) {
                // v += a h closes this step (finish_kernel's 0 + sum)
                a0 = 0.0 + sx;
                a1 = 0.0 + sy;
                a2 = 0.0 + sz;
                pvx = pvx + a0 * h;
                pvy = pvy + a1 * h;
                pvz = pvz + a2 * h;
                if (step + 1u < nsteps) {
                    // the next step's kick and drift
                    pvx = pvx + a0 * h;
                    pvy = pvy + a1 * h;
                    pvz = pvz + a2 * h;
                    px = px + pvx * dt;
                    py = py + pvy * dt;
                    pz = pz + pvz * dt;
                    store_record(nxt + i, make_double4(px, py, pz, pmass));
                }
            }
        }
        wait_stores();
    }
    if (own && lane == 0) {
        x[i] = px; y[i] = py; z[i] = pz;
        vx[i] = pvx; vy[i] = pvy; vz[i] = pvz;
        ax[i] = a0; ay[i] = a1; az[i] = a2;
    }
}

__global__ void coop_reset_kernel(uint32_t* __restrict__ bar) {
    if (threadIdx.x < 2) bar[threadIdx.x] = 0u;
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
                   float eps2, double near2, double mid2, const double* tbox, double* part) {
    if (mask)
        bf_matrix_kernel<TJ, true, BF><<<g, kWave, 0, st>>>(pm, n, jchunk, eps2, near2, mid2,
                                                            tbox, part);
    else
        bf_matrix_kernel<TJ, false, BF><<<g, kWave, 0, st>>>(pm, n, jchunk, eps2, near2, mid2,
                                                             tbox, part);
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

// As finish_inv_kernel for (ax, ay, az, 0) records.
__global__ __launch_bounds__(256) void finish_inv4_kernel(const double4* __restrict__ part4,
                                                          int n, const uint32_t* __restrict__ inv,
                                                          FinishArgs fa) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n) return;
    const double4 p = part4[inv[o]];
    // 0 + p, as finish_kernel's sum over one chunk forms it (-0 becomes +0)
    const double a0 = 0.0 + p.x, a1 = 0.0 + p.y, a2 = 0.0 + p.z;
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

gravis_status finish_forces4(Ctx* c, const double4* part4, int n, const uint32_t* inv,
                             const FinishArgs& fa) {
    if (n == 0) return GRAVIS_OK;
    if (!inv) return fail(c, GRAVIS_INTERNAL, "finish_forces4 needs the inverse permutation");
    finish_inv4_kernel<<<grid_for((size_t)n, 256), 256, 0, c->stream>>>(part4, n, inv, fa);
    return check_launch(c, "finish_forces");
}

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
        // Levels per axis: 9 finer than the mean spacing over the root cube
        // (2^(log2 n / 3) cells per axis), at most the octree's 21. A
        // concentrated core (a Plummer sphere's, about 62 times denser per
        // axis than the cube's mean for the benchmark's N = 262,144) stays
        // resolved, and 64-target groups stay as compact as with 21 levels;
        // N = 262,144 takes 15 levels, 45-bit keys, 5 radix passes instead
        // of 7.
        int lv = 9;
        for (size_t m = 1; m < n; m <<= 3) ++lv;
        GRAVIS_TRY(morton_sort(c, n, x, y, z, box, &keys, &perm, std::min(lv, 21)));
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
                                         q * q, tbox, part);
            else if (tj == 128)
                launch_matrix<128, false>(grid, c->stream, mask, pm, ni, jchunk, eps2f, qm * qm,
                                          q * q, tbox, part);
            else if (c->bf_matrix == 2)
                launch_matrix<256, true>(grid, c->stream, mask, pm, ni, jchunk, eps2f, qm * qm,
                                         q * q, tbox, part);
            else
                launch_matrix<256, false>(grid, c->stream, mask, pm, ni, jchunk, eps2f, qm * qm,
                                          q * q, tbox, part);
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

gravis_status kdk_coop(Ctx* c, size_t n, double* x, double* y, double* z, double* vx,
                       double* vy, double* vz, const double* m, double* ax, double* ay,
                       double* az, double softening, double dt, uint64_t* nsteps_left) {
    if (n == 0 || n > (size_t)kCoopMaxN) return fail(c, GRAVIS_INTERNAL, "kdk_coop: n = %zu", n);
    if (*nsteps_left == 0) return GRAVIS_OK;
    const int ni = (int)n;
    const double eps2 = softening * softening;
    const double h = 0.5 * dt;
    const bool mask = !(eps2 >= kMaskEps2D);
    const size_t lds = n * sizeof(double4);
    double4* pbuf;
    GRAVIS_TRY(buf(c, "coop_pos", 2 * n, &pbuf));
    uint32_t* bar;
    GRAVIS_TRY(buf(c, "coop_bar", 2, &bar));
    const void* fn = mask ? reinterpret_cast<const void*>(&kdk_coop_kernel<true>)
                          : reinterpret_cast<const void*>(&kdk_coop_kernel<false>);
    // one workgroup per kCoopWaves targets, every one resident at once
    // a grid that cannot be resident at once (or an occupancy query that
    // fails) leaves the steps to the caller
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kCoopThreads, lds) !=
        hipSuccess) {
        (void)hipGetLastError();
        return GRAVIS_OK;
    }
    const int grid = (ni + kCoopWaves - 1) / kCoopWaves;
    if (grid > per_cu * c->num_cus) return GRAVIS_OK;
    // launches of at most 2^20 steps each (the arrival count stays below 2^32)
    while (*nsteps_left > 0) {
        uint32_t k = (uint32_t)std::min<uint64_t>(*nsteps_left, 1u << 20);
        coop_reset_kernel<<<1, kWave, 0, c->stream>>>(bar);
        GRAVIS_TRY(check_launch(c, "coop_reset"));
        int na = ni;
        double *px = x, *py = y, *pz = z, *pvx = vx, *pvy = vy, *pvz = vz;
        const double* pm = m;
        double *pax = ax, *pay = ay, *paz = az;
        double e2 = eps2, hh = h, dd = dt;
        double4* pb = pbuf;
        uint32_t* pbar = bar;
        void* args[] = {&na, &px, &py, &pz, &pvx, &pvy, &pvz, &pm, &pax,
                        &pay, &paz, &e2, &hh, &dd, &k, &pb, &pbar};
        // the runtime may refuse a cooperative launch (for one, while another
        // process holds the device's cooperative-launch resources): the
        // state is then untouched and the caller steps the rest launch by
        // launch
        if (hipLaunchCooperativeKernel(fn, dim3(grid), dim3(kCoopThreads), args, (unsigned)lds,
                                       c->stream) != hipSuccess) {
            (void)hipGetLastError();
            return GRAVIS_OK;
        }
        uint32_t abort_word = 0;
        GRAVIS_HIP(c, hipMemcpyAsync(&abort_word, bar + 1, sizeof(uint32_t),
                                     hipMemcpyDeviceToHost, c->stream));
        GRAVIS_HIP(c, hipStreamSynchronize(c->stream));
        if (abort_word != 0)
            return fail(c, GRAVIS_INTERNAL,
                        "kdk_coop: a grid barrier waited past its limit (%d workgroups)", grid);
        *nsteps_left -= k;
    }
    return GRAVIS_OK;
}

}  // namespace gravis
