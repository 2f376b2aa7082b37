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
// back as wavefront-uniform broadcasts. Pairwise gravity is not a dense
// contraction, so this is VALU work (the matrix cores were measured: with
// four useful output columns v_mfma_f32_4x4x1 issues at a quarter of the
// FP32 matrix rate and saves about 5 %). The j range is split in S chunks so
// the grid holds many wavefronts per CU; the chunk partials are summed in a
// fixed order by the finishing kernel (deterministic).
//
// Mixed precision: separations are formed in FP64 relative to the centre of
// the wavefront's target box, rounded to FP32, and the pair force is FP32
// (v_rsq_f32); sources close to the group get FP64 separations per pair.
// Each lane sums a tile of pair forces in FP32 and adds the tile sum into a
// Kahan-compensated FP32 accumulator; the result leaves as FP64
// sum - compensation.
#include "internal.h"

namespace gravis {
namespace {

template <bool MASK>
__device__ __forceinline__ double pair_f64(double dx, double dy, double dz, double m,
                                           double eps2) {
    const double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, __builtin_fma(dz, dz, eps2)));
    double ri = rsqrt(r2);
    if (MASK) ri = r2 > 0.0 ? ri : 0.0;
    const double mr = m * ri;
    const double ri2 = ri * ri;
    return mr * ri2;
}

// Bounding box of each TJ-wide tile of the sorted records: tbox[6 t ..] =
// (lo x, lo y, lo z, hi x, hi y, hi z).
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
        tbox[(size_t)blockIdx.x * 6 + d] = v;
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
// loop is 12 VALU operations and one v_rsq_f32 per pair; the transcendental
// issues beside the FMAs, so the loop is bound by the 12 FMA-class slots.
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
        near_cur = tile_is_near(tbox + (size_t)tile0 * 6, lo, hi, lim2);
        load(0);
    }
    for (int t = 0; t < ntiles; ++t) {
        store(near_cur);
        __syncthreads();
        bool near_next = false;
        if (t + 1 < ntiles) {
            near_next = tile_is_near(tbox + (size_t)(tile0 + t + 1) * 6, lo, hi, lim2);
            load(t + 1);
        }
        float ax[IPT], ay[IPT], az[IPT];
#pragma unroll
        for (int k = 0; k < IPT; ++k) ax[k] = ay[k] = az[k] = 0.f;
        if (!near_cur) {
            const float4* t32 = reinterpret_cast<const float4*>(&tile[0]);
            // unroll depth chosen for register pressure: 63 VGPRs (8 waves
            // per SIMD) at IPT = 1, 83 (6 waves) at IPT = 2
            constexpr int kUnroll = IPT == 1 ? 4 : 8;
#pragma unroll kUnroll
            for (int j = 0; j < TJ; ++j) {
                const float4 s = t32[j];
#pragma unroll
                for (int k = 0; k < IPT; ++k) {
                    const float dx = s.x - xi[k], dy = s.y - yi[k], dz = s.z - zi[k];
                    const float f = pair_f32<MASK>(dx, dy, dz, s.w, eps2);
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
                    const float r2 =
                        __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
                    float ri = __builtin_amdgcn_rsqf(r2);
                    const float h = (r2 * ri) * ri;
                    ri = ri * __builtin_fmaf(-0.5f, h, 1.5f);
                    if (MASK) ri = r2 > 0.0f ? ri : 0.0f;
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
                const double f = pair_f64<MASK>(dx, dy, dz, s.w, eps2);
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

}  // namespace

gravis_status finish_forces(Ctx* c, const double* part, int nchunks, int n,
                            const uint32_t* perm, const FinishArgs& fa) {
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
    GRAVIS_TRY(gather_sorted(c, n, x, y, z, m, perm, pm));

    // FP64: 256-thread workgroups of 256 targets. Mixed: one wavefront per
    // workgroup with bf_ipt * 64 targets and bf_tile-record tiles.
    const int ipt = mixed ? c->bf_ipt : 1;
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
        GRAVIS_TRY(buf(c, "bf_tbox", (size_t)ntile * 6, &tbox));
        if (tj == 256)
            tile_box_kernel<256><<<ntile, 256, 0, c->stream>>>(pm, ni, tbox);
        else
            tile_box_kernel<128><<<ntile, 256, 0, c->stream>>>(pm, ni, tbox);
        GRAVIS_TRY(check_launch(c, "tile_box"));
    }
    timer_begin(c);
    if (mixed) {
        const float eps2f = (float)eps2;
        const bool mask = !(eps2f >= 1.17549435e-38f);
        const double f2 = c->bf_near_fraction * c->bf_near_fraction;
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
        const bool mask = !(eps2 > 0.0);
        launch_fp64<256, 1, 256>(grid, c->stream, mask, pm, ni, jchunk, eps2, part);
        timer_end(c, "bf_fp64");
    }
    GRAVIS_TRY(check_launch(c, "bf kernel"));
    return finish_forces(c, part, S, ni, perm, fa);
}

}  // namespace gravis
