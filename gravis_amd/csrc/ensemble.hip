// ensemble.hip -- batched ensemble of independent N-body systems and the
// discrete adjoint of its KDK trajectory (no counterpart in the reference).
//
// Forward: every system runs sim-core's LeapfrogKDK (crates/sim-core/src/
// integrator.rs:41-77) with softened brute force (gravity.rs:47-80) in the
// gather form. All systems advance together. Mixed precision: ONE launch per
// step, whose workgroups cover every (system, 64-target item), takes the
// forces, closes the step and opens the next (double-buffered positions);
// two steps are captured once as a hipGraph and replayed. FP64: a
// kick-drift launch over the concatenated arrays and a force launch per
// step, captured once and replayed nsteps times.
//
// Reverse: for cotangents (lx_T, lv_T) the adjoint of
//   vh = v + h a(x);  x' = x + dt vh;  v' = vh + h a(x')       (h = dt/2)
// is, per step from the last,
//   lvh = lv' + dt lx'  ;  lx += J(x)^T (w lvh)  ;  lv = lvh
// with w = 2h at interior time levels (the K1 of step t and the K2 of step
// t-1 share a(x_t)), h at t = 0, and a first J^T(x_T)(h lv_T) term. With
// d = x_j - x_k, s = |d|^2 + eps^2, w = s^{-3/2}, v = m_k u_j - m_j u_k:
//   (J_x^T u)_k = sum_j w v - 3 s^{-5/2} d (d . v)
//   (J_m^T u)_k = - sum_j w (u_j . d)
// computed by ens_vjp_kernel (FP64). The trajectory is recomputed from
// checkpoints stored every `ckpt` steps, one segment at a time.
#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.h"

namespace gravis {
namespace {

constexpr int kEnsBlock = 256;

struct EnsBlock {
    int i0, i1, j0, j1;  // targets [i0, i1), sources [j0, j1) of one system
};

template <bool MIXED>
__global__ __launch_bounds__(kEnsBlock) void ens_force_kernel(
    const EnsBlock* __restrict__ blocks, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, const double* __restrict__ m,
    double eps2, double h, int kick, double* __restrict__ ax, double* __restrict__ ay,
    double* __restrict__ az, double* __restrict__ vx, double* __restrict__ vy,
    double* __restrict__ vz) {
    const EnsBlock b = blocks[blockIdx.x];
    const int i = b.i0 + threadIdx.x;
    const bool active = i < b.i1;
    const int ii = active ? i : b.i0;
    const double px = x[ii], py = y[ii], pz = z[ii];
    if (MIXED) {
        // Systems are not spatially sorted, so every pair separation is formed
        // in FP64 and rounded once to FP32; r^2, v_rsq_f32, m r^-3 and the
        // accumulation are FP32 (tile sums Kahan-added).
        __shared__ double4 tile[kEnsBlock];
        const float e2 = (float)eps2;
        float sx = 0.f, sy = 0.f, sz = 0.f, kx = 0.f, ky = 0.f, kz = 0.f;
        for (int jb = b.j0; jb < b.j1; jb += kEnsBlock) {
            const int j = jb + threadIdx.x;
            __syncthreads();
            tile[threadIdx.x] = j < b.j1 ? make_double4(x[j], y[j], z[j], m[j])
                                         : make_double4(px, py, pz, 0.0);
            __syncthreads();
            const int cnt = min(kEnsBlock, b.j1 - jb);
            float tx = 0.f, ty = 0.f, tz = 0.f;
            for (int k = 0; k < cnt; ++k) {
                const double4 s = tile[k];
                const float dx = (float)(s.x - px), dy = (float)(s.y - py), dz = (float)(s.z - pz);
                // zero separation (the self pair, coincident points) adds 0,
                // also under a softening too small for m eps^-3 in FP32
                const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                float ri = __builtin_amdgcn_rsqf(d2 + e2);
                ri = d2 > 0.f ? ri : 0.f;
                const float f = ((float)s.w * ri) * (ri * ri);
                tx = __builtin_fmaf(f, dx, tx);
                ty = __builtin_fmaf(f, dy, ty);
                tz = __builtin_fmaf(f, dz, tz);
            }
            float yk = tx - kx, t = sx + yk; kx = (t - sx) - yk; sx = t;
            yk = ty - ky; t = sy + yk; ky = (t - sy) - yk; sy = t;
            yk = tz - kz; t = sz + yk; kz = (t - sz) - yk; sz = t;
        }
        if (active) {
            const double a0 = (double)sx - (double)kx, a1 = (double)sy - (double)ky,
                         a2 = (double)sz - (double)kz;
            ax[i] = a0; ay[i] = a1; az[i] = a2;
            if (kick) {
                vx[i] = vx[i] + a0 * h;
                vy[i] = vy[i] + a1 * h;
                vz[i] = vz[i] + a2 * h;
            }
        }
    } else {
        __shared__ double4 tile[kEnsBlock];
        double sx = 0.0, sy = 0.0, sz = 0.0;
        const Rinv3Consts k3 = rinv3_consts();
        for (int jb = b.j0; jb < b.j1; jb += kEnsBlock) {
            const int j = jb + threadIdx.x;
            __syncthreads();
            tile[threadIdx.x] = j < b.j1 ? make_double4(x[j], y[j], z[j], m[j])
                                         : make_double4(0, 0, 0, 0);
            __syncthreads();
            const int cnt = min(kEnsBlock, b.j1 - jb);
            for (int k = 0; k < cnt; ++k) {
                const double4 s = tile[k];
                const double dx = s.x - px, dy = s.y - py, dz = s.z - pz;
                const double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, __builtin_fma(dz, dz, eps2)));
                const double f = (jb + k == i) ? 0.0 : m_rinv3(r2, s.w, k3);
                sx = __builtin_fma(f, dx, sx);
                sy = __builtin_fma(f, dy, sy);
                sz = __builtin_fma(f, dz, sz);
            }
        }
        if (active) {
            ax[i] = sx; ay[i] = sy; az[i] = sz;
            if (kick) {
                vx[i] = vx[i] + sx * h;
                vy[i] = vy[i] + sy * h;
                vz[i] = vz[i] + sz * h;
            }
        }
    }
}

// ---- mixed forward in per-system Morton order --------------------------------
// gravis_ensemble_kdk with mixed precision sorts every system's particles
// by Morton key inside the system's own cube (the system index is the key's
// high part, so systems keep their ranges), runs the steps in that order and
// restores the caller's order at the end. In sorted order a wavefront's 64
// targets are spatially compact, so ens_wave_kernel rounds sources to FP32
// relative to the wavefront's own box centre (as bf_mixed_kernel does): 12
// VALU operations and one v_rsq_f32 per pair instead of forming every
// separation in FP64. Near tiles (gap to the target box and eps both below
// q times its diagonal, q = the context's bf_near_fraction) keep FP64
// separations.
constexpr int kEnsTile = 128;       // sources per staged tile
constexpr int kEnsMortonBits = 14;  // per axis inside a system's cube

// Cube of each system: centre and half width of its bounding box.
__global__ __launch_bounds__(256) void ens_cube_kernel(const EnsBlock* __restrict__ sys,
                                                       const double* __restrict__ x,
                                                       const double* __restrict__ y,
                                                       const double* __restrict__ z,
                                                       double4* __restrict__ cube) {
    const EnsBlock b = sys[blockIdx.x];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = b.j0 + threadIdx.x; i < b.j1; i += blockDim.x) {
        lo[0] = fmin(lo[0], x[i]); hi[0] = fmax(hi[0], x[i]);
        lo[1] = fmin(lo[1], y[i]); hi[1] = fmax(hi[1], y[i]);
        lo[2] = fmin(lo[2], z[i]); hi[2] = fmax(hi[2], z[i]);
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
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x / kWave); ++k)
            for (int d = 0; d < 3; ++d) {
                sh[0][d] = fmin(sh[0][d], sh[k][d]);
                sh[0][3 + d] = fmax(sh[0][3 + d], sh[k][3 + d]);
            }
        double ext = 0.0;
        for (int d = 0; d < 3; ++d) ext = fmax(ext, sh[0][3 + d] - sh[0][d]);
        const double hw = b.j1 > b.j0 ? 0.5 * ext + 1e-300 : 1.0;
        cube[blockIdx.x] = make_double4(0.5 * (sh[0][0] + sh[0][3]), 0.5 * (sh[0][1] + sh[0][4]),
                                        0.5 * (sh[0][2] + sh[0][5]), hw);
    }
}

// key = system index << 42 | 42-bit Morton key inside the system's cube.
__global__ __launch_bounds__(256) void ens_key_kernel(const EnsBlock* __restrict__ sys,
                                                      const double* __restrict__ x,
                                                      const double* __restrict__ y,
                                                      const double* __restrict__ z,
                                                      const double4* __restrict__ cube,
                                                      uint64_t* __restrict__ keys,
                                                      uint32_t* __restrict__ vals) {
    const EnsBlock b = sys[blockIdx.x];
    const double4 cb = cube[blockIdx.x];
    const double cells = (double)(1u << kEnsMortonBits);
    const double inv = cells / (2.0 * cb.w);
    auto q = [&](double v, double c) -> uint32_t {
        const double t = (v - (c - cb.w)) * inv;
        if (!(t >= 0.0)) return 0u;
        if (t >= cells - 1.0) return (1u << kEnsMortonBits) - 1u;
        return (uint32_t)t;
    };
    for (int i = b.j0 + threadIdx.x; i < b.j1; i += blockDim.x) {
        const uint64_t mk = morton_spread3(q(x[i], cb.x)) | (morton_spread3(q(y[i], cb.y)) << 1) |
                            (morton_spread3(q(z[i], cb.z)) << 2);
        keys[i] = ((uint64_t)blockIdx.x << (3 * kEnsMortonBits)) | mk;
        vals[i] = (uint32_t)i;
    }
}

// dst[k][i] = src[k][perm[i]] for the state arrays (gather into sorted
// order), or dst[k][perm[i]] = src[k][i] (back to the caller's order).
struct EnsArrays {
    const double* src[10];
    double* dst[10];
    int count;
};

__global__ __launch_bounds__(256) void ens_permute_kernel(EnsArrays a, const uint32_t* __restrict__ perm,
                                                          size_t n, int scatter) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t p = perm[i];
        for (int k = 0; k < a.count; ++k) {
            if (scatter) a.dst[k][p] = a.src[k][i];
            else a.dst[k][i] = a.src[k][p];
        }
    }
}

// ---- mixed forward, one launch per step ------------------------------------
// The state stays in per-system Morton order. ens_pack_kernel (once per call)
// runs the first half of step 1 -- v += h a, x += dt v, kick_drift_kernel's
// arithmetic -- and writes the positions as (x, y, z, m) records P with every
// item's box B. ens_wave_kernel (one launch per step) takes the forces on P
// and applies the step's closing half kick v += h a; with DRIFT it goes on
// with the next step's v += h a, x += dt v, written as records P' with boxes
// B' (positions are double-buffered: every wavefront reads all of its
// system's records while the others move theirs); without, it writes the
// final positions to x, y, z. The rounding of every update is that of the
// separate kernels.

// Box of one item from the lanes' positions (inactive lanes: the item's
// first target, held by lane 0); lane 0 writes bout[6 item ..].
__device__ __forceinline__ void ens_item_box(bool active, double px, double py, double pz,
                                             double* __restrict__ bout, unsigned item) {
    const double qx = active ? px : __shfl(px, 0), qy = active ? py : __shfl(py, 0),
                 qz = active ? pz : __shfl(pz, 0);
    const double xd[1] = {qx}, yd[1] = {qy}, zd[1] = {qz};
    double lo[3], hi[3];
    wave_box<1>(xd, yd, zd, lo, hi);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            bout[(size_t)6 * item + d] = lo[d];
            bout[(size_t)6 * item + 3 + d] = hi[d];
        }
    }
}

__global__ __launch_bounds__(kWave) void ens_pack_kernel(
    const EnsBlock* __restrict__ items, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, const double* __restrict__ m,
    double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
    const double* __restrict__ ax, const double* __restrict__ ay,
    const double* __restrict__ az, double h, double dt, double4* __restrict__ pout,
    double* __restrict__ bout) {
    const EnsBlock b = items[blockIdx.x];
    const int i = b.i0 + (int)threadIdx.x;
    const bool active = i < b.i1;
    double nx = 0.0, ny = 0.0, nz = 0.0;
    if (active) {
        const double nvx = vx[i] + ax[i] * h;
        const double nvy = vy[i] + ay[i] * h;
        const double nvz = vz[i] + az[i] * h;
        vx[i] = nvx;
        vy[i] = nvy;
        vz[i] = nvz;
        nx = x[i] + nvx * dt;
        ny = y[i] + nvy * dt;
        nz = z[i] + nvz * dt;
        pout[i] = make_double4(nx, ny, nz, m[i]);
    }
    ens_item_box(active, nx, ny, nz, bout, blockIdx.x);
}

// One wavefront per item: targets [i0, i1) (at most 64) of one system,
// sources [j0, j1) = that system, records pin with item boxes box; sources
// are rounded to FP32 relative to the item box's centre, 12 VALU operations
// and one v_rsq_f32 per pair; near tiles (gap to the target box and eps both
// below q times its diagonal, near2 = q^2) keep FP64 separations.
template <bool MASK, bool DRIFT>
__global__ __launch_bounds__(kWave) void ens_wave_kernel(
    const EnsBlock* __restrict__ items, const double* __restrict__ box,
    const double4* __restrict__ pin, float eps2, double near2, double h, double dt,
    double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az,
    double4* __restrict__ pout, double* __restrict__ bout, double* __restrict__ x,
    double* __restrict__ y, double* __restrict__ z) {
    static_assert(kEnsTile % kWave == 0, "a tile is whole items");
    constexpr int LPT = kEnsTile / kWave;
    __shared__ double4 tile[kEnsTile];  // near: FP64 records; far: float4 [0, kEnsTile)
    const EnsBlock b = items[blockIdx.x];
    const int lane = threadIdx.x;
    const int i = b.i0 + lane;
    const bool active = i < b.i1;
    const int ii = active ? i : b.i0;
    // the system's first item (items of a system are consecutive)
    const int first = (int)blockIdx.x - (b.i0 - b.j0) / kWave;
    double lo[3], hi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        lo[d] = box[(size_t)6 * blockIdx.x + d];
        hi[d] = box[(size_t)6 * blockIdx.x + 3 + d];
    }
    const double4 me = pin[ii];
    const double px = me.x, py = me.y, pz = me.z;
    const double cx = 0.5 * (lo[0] + hi[0]), cy = 0.5 * (lo[1] + hi[1]), cz = 0.5 * (lo[2] + hi[2]);
    double di2 = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) di2 = __builtin_fma(hi[d] - lo[d], hi[d] - lo[d], di2);
    double lim2 = near2 * di2;
    if ((double)eps2 >= lim2) lim2 = 0.0;  // softening bounds every pair's error
    const float xi = (float)(px - cx), yi = (float)(py - cy), zi = (float)(pz - cz);
    float sx = 0.f, sy = 0.f, sz = 0.f, kx = 0.f, ky = 0.f, kz = 0.f;
    // records of the next tile are loaded while the current one is computed
    double4 nxt[LPT];
    auto load = [&](int jb) {
#pragma unroll
        for (int e = 0; e < LPT; ++e) {
            const int j = jb + e * kWave + lane;
            // padding: zero mass at the box centre (no contribution)
            nxt[e] = record_or(pin, j, b.j1, make_double4(cx, cy, cz, 0.0));
        }
    };
    if (b.j0 < b.j1) load(b.j0);
    for (int jb = b.j0; jb < b.j1; jb += kEnsTile) {
        double4 rec[LPT];
#pragma unroll
        for (int e = 0; e < LPT; ++e) rec[e] = nxt[e];
        bool near = false;
        if (lim2 > 0.0) {
            // the tile's box: union of its items' boxes (records < j1 only)
            const int t0 = first + (jb - b.j0) / kWave;
            double tb[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) tb[q] = box[(size_t)6 * t0 + q];
#pragma unroll
            for (int e = 1; e < LPT; ++e) {
                if (jb + e * kWave < b.j1) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        tb[d] = fmin(tb[d], box[(size_t)6 * (t0 + e) + d]);
                        tb[3 + d] = fmax(tb[3 + d], box[(size_t)6 * (t0 + e) + 3 + d]);
                    }
                }
            }
            near = tile_is_near(tb, lo, hi, lim2);
        }
        near = __builtin_amdgcn_readfirstlane((int)near) != 0;
        if (near) {
#pragma unroll
            for (int e = 0; e < LPT; ++e) tile[e * kWave + lane] = rec[e];
        } else {
            float4* t32 = reinterpret_cast<float4*>(&tile[0]);
#pragma unroll
            for (int e = 0; e < LPT; ++e)
                t32[e * kWave + lane] = make_float4((float)(rec[e].x - cx), (float)(rec[e].y - cy),
                                                    (float)(rec[e].z - cz), (float)rec[e].w);
        }
        __syncthreads();
        if (jb + kEnsTile < b.j1) load(jb + kEnsTile);
        const int cnt = min(kEnsTile, b.j1 - jb);
        float tx = 0.f, ty = 0.f, tz = 0.f;
        if (!near) {
            const float4* t32 = reinterpret_cast<const float4*>(&tile[0]);
            const float e2 = in_vgpr(eps2);  // full-rate operand of the r^2 FMA
#pragma unroll 4
            for (int k = 0; k < cnt; ++k) {
                const float4 s4 = t32[k];
                const float dx = s4.x - xi, dy = s4.y - yi, dz = s4.z - zi;
                const float f = pair_f32<MASK>(dx, dy, dz, s4.w, e2);
                tx = __builtin_fmaf(f, dx, tx);
                ty = __builtin_fmaf(f, dy, ty);
                tz = __builtin_fmaf(f, dz, tz);
            }
        } else {
            for (int k = 0; k < cnt; ++k) {
                const double4 s4 = tile[k];
                const float dx = (float)(s4.x - px), dy = (float)(s4.y - py),
                            dz = (float)(s4.z - pz);
                const float d2 = MASK ? __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)) : 0.f;
                const float r2 = MASK ? d2 + eps2
                                      : __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
                float ri = __builtin_amdgcn_rsqf(r2);
                const float hh = (r2 * ri) * ri;
                ri = ri * __builtin_fmaf(-0.5f, hh, 1.5f);
                if (MASK) ri = d2 > 0.0f ? ri : 0.0f;
                const float f = ((float)s4.w * ri) * (ri * ri);
                tx = __builtin_fmaf(f, dx, tx);
                ty = __builtin_fmaf(f, dy, ty);
                tz = __builtin_fmaf(f, dz, tz);
            }
        }
        kahan_add(sx, kx, tx);
        kahan_add(sy, ky, ty);
        kahan_add(sz, kz, tz);
        __syncthreads();
    }
    double nx = 0.0, ny = 0.0, nz = 0.0;
    if (active) {
        const double a0 = (double)sx - (double)kx, a1 = (double)sy - (double)ky,
                     a2 = (double)sz - (double)kz;
        ax[i] = a0;
        ay[i] = a1;
        az[i] = a2;
        double nvx = vx[i] + a0 * h, nvy = vy[i] + a1 * h, nvz = vz[i] + a2 * h;
        if (DRIFT) {
            nvx = nvx + a0 * h;
            nvy = nvy + a1 * h;
            nvz = nvz + a2 * h;
            nx = px + nvx * dt;
            ny = py + nvy * dt;
            nz = pz + nvz * dt;
            pout[i] = make_double4(nx, ny, nz, me.w);
        } else {
            x[i] = px;
            y[i] = py;
            z[i] = pz;
        }
        vx[i] = nvx;
        vy[i] = nvy;
        vz[i] = nvz;
    }
    if (DRIFT) ens_item_box(active, nx, ny, nz, bout, blockIdx.x);
}

// lx += J_x^T u, lm += J_m^T u (FP64).
// SELF: eps2 = 0, where the self pair must be skipped (r = 0). With eps2 > 0
// it contributes exactly zero -- v = m_k u_k - m_k u_k is a difference of
// identical products and every other term carries d = 0 -- so no per-pair
// branch is taken.
template <bool SELF>
__global__ __launch_bounds__(kEnsBlock) void ens_vjp_kernel(
    const EnsBlock* __restrict__ blocks, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, const double* __restrict__ m,
    const double* __restrict__ ux, const double* __restrict__ uy, const double* __restrict__ uz,
    double eps2, double* __restrict__ lx, double* __restrict__ ly, double* __restrict__ lz,
    double* __restrict__ lm) {
    __shared__ double4 tp[kEnsBlock];
    __shared__ double4 tu[kEnsBlock];
    const EnsBlock b = blocks[blockIdx.x];
    const int k = b.i0 + threadIdx.x;
    const bool active = k < b.i1;
    const int kk = active ? k : b.i0;
    const double px = x[kk], py = y[kk], pz = z[kk], mk = m[kk];
    const double u0 = ux[kk], u1 = uy[kk], u2 = uz[kk];
    double g0 = 0.0, g1 = 0.0, g2 = 0.0, gm = 0.0;
    for (int jb = b.j0; jb < b.j1; jb += kEnsBlock) {
        const int j = jb + threadIdx.x;
        __syncthreads();
        if (j < b.j1) {
            tp[threadIdx.x] = make_double4(x[j], y[j], z[j], m[j]);
            tu[threadIdx.x] = make_double4(ux[j], uy[j], uz[j], 0.0);
        } else {
            tp[threadIdx.x] = make_double4(0, 0, 0, 0);
            tu[threadIdx.x] = make_double4(0, 0, 0, 0);
        }
        __syncthreads();
        const int cnt = min(kEnsBlock, b.j1 - jb);
        for (int q = 0; q < cnt; ++q) {
            if (SELF && jb + q == k) continue;
            const double4 s = tp[q];
            const double4 uj = tu[q];
            const double dx = s.x - px, dy = s.y - py, dz = s.z - pz;
            const double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, __builtin_fma(dz, dz, eps2)));
            const double ri = rsqrt_refined(r2);
            const double ri2 = ri * ri;
            const double w = ri2 * ri;
            const double qq = 3.0 * w * ri2;
            const double v0 = mk * uj.x - s.w * u0;
            const double v1 = mk * uj.y - s.w * u1;
            const double v2 = mk * uj.z - s.w * u2;
            const double dv = __builtin_fma(dx, v0, __builtin_fma(dy, v1, dz * v2));
            const double c = qq * dv;
            g0 = __builtin_fma(w, v0, __builtin_fma(-c, dx, g0));
            g1 = __builtin_fma(w, v1, __builtin_fma(-c, dy, g1));
            g2 = __builtin_fma(w, v2, __builtin_fma(-c, dz, g2));
            gm = __builtin_fma(-w, __builtin_fma(uj.x, dx, __builtin_fma(uj.y, dy, uj.z * dz)), gm);
        }
    }
    if (active) {
        lx[k] += g0;
        ly[k] += g1;
        lz[k] += g2;
        lm[k] += gm;
    }
}

// lv = lv + dt lx (skipped when dt == 0); u = wgt * lv
__global__ __launch_bounds__(256) void adj_level_kernel(size_t n, const double* __restrict__ lx,
                                                        const double* __restrict__ ly,
                                                        const double* __restrict__ lz,
                                                        double* __restrict__ lv, double dt,
                                                        double wgt, double* __restrict__ u) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const double l[3] = {lx[i], ly[i], lz[i]};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double v = lv[c * n + i];
            if (dt != 0.0) v = v + dt * l[c];
            lv[c * n + i] = v;
            u[c * n + i] = wgt * v;
        }
    }
}

// cotangent of L = sum |x_T - target|^2 and per-system loss partials
__global__ __launch_bounds__(256) void adj_loss_kernel(const EnsBlock* __restrict__ sys, int nsys,
                                                       const double* __restrict__ x,
                                                       const double* __restrict__ y,
                                                       const double* __restrict__ z,
                                                       const double* __restrict__ tx,
                                                       const double* __restrict__ ty,
                                                       const double* __restrict__ tz,
                                                       double* __restrict__ gx,
                                                       double* __restrict__ gy,
                                                       double* __restrict__ gz,
                                                       double* __restrict__ loss) {
    const int s = blockIdx.x;
    if (s >= nsys) return;
    const EnsBlock b = sys[s];
    double acc = 0.0;
    for (int i = b.j0 + threadIdx.x; i < b.j1; i += blockDim.x) {
        const double d0 = x[i] - tx[i], d1 = y[i] - ty[i], d2 = z[i] - tz[i];
        gx[i] = 2.0 * d0;
        gy[i] = 2.0 * d1;
        gz[i] = 2.0 * d2;
        acc += d0 * d0 + d1 * d1 + d2 * d2;
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[s] = red[0];
}

struct EnsPlan {
    std::vector<EnsBlock> blocks;  // force/vjp work items, largest systems first
    std::vector<EnsBlock> waves;   // 64-target items of ens_wave_kernel, same order
    std::vector<EnsBlock> systems; // one per system (j0, j1)
    size_t n = 0;
};

gravis_status make_plan(Ctx* c, int nsys, const uint64_t* offsets, EnsPlan* p) {
    if (nsys <= 0 || !offsets) return fail(c, GRAVIS_INVALID_ARG, "empty ensemble");
    if (offsets[0] != 0) return fail(c, GRAVIS_INVALID_ARG, "offsets[0] must be 0");
    std::vector<int> order(nsys);
    for (int s = 0; s < nsys; ++s) {
        if (offsets[s + 1] < offsets[s])
            return fail(c, GRAVIS_INVALID_ARG, "offsets must be non-decreasing");
        order[s] = s;
        p->systems.push_back({(int)offsets[s], (int)offsets[s + 1], (int)offsets[s],
                              (int)offsets[s + 1]});
    }
    if (offsets[nsys] > 0x7FFFFFFFull) return fail(c, GRAVIS_INVALID_ARG, "ensemble too large");
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return offsets[a + 1] - offsets[a] > offsets[b + 1] - offsets[b];
    });
    for (int s : order) {
        const int j0 = (int)offsets[s], j1 = (int)offsets[s + 1];
        for (int i0 = j0; i0 < j1; i0 += kEnsBlock)
            p->blocks.push_back({i0, std::min(i0 + kEnsBlock, j1), j0, j1});
        for (int i0 = j0; i0 < j1; i0 += kWave)
            p->waves.push_back({i0, std::min(i0 + kWave, j1), j0, j1});
    }
    p->n = offsets[nsys];
    return GRAVIS_OK;
}

gravis_status upload(Ctx* c, const char* name, const std::vector<EnsBlock>& v, EnsBlock** out) {
    GRAVIS_TRY(buf(c, name, std::max<size_t>(v.size(), 1), out));
    if (!v.empty())
        GRAVIS_HIP(c, hipMemcpyAsync(*out, v.data(), v.size() * sizeof(EnsBlock),
                                     hipMemcpyHostToDevice, c->stream));
    return GRAVIS_OK;
}

struct EnsArgs {
    double *x, *y, *z, *vx, *vy, *vz, *ax, *ay, *az;
    const double* m;
};

gravis_status ens_force(Ctx* c, bool mixed, const EnsBlock* d_blocks, int nblocks,
                        const EnsArgs& a, double eps2, double h, int kick) {
    if (nblocks == 0) return GRAVIS_OK;
    if (mixed)
        ens_force_kernel<true><<<nblocks, kEnsBlock, 0, c->stream>>>(
            d_blocks, a.x, a.y, a.z, a.m, eps2, h, kick, a.ax, a.ay, a.az, a.vx, a.vy, a.vz);
    else
        ens_force_kernel<false><<<nblocks, kEnsBlock, 0, c->stream>>>(
            d_blocks, a.x, a.y, a.z, a.m, eps2, h, kick, a.ax, a.ay, a.az, a.vx, a.vy, a.vz);
    return check_launch(c, "ens_force");
}

gravis_status ens_step(Ctx* c, bool mixed, const EnsBlock* d_blocks, int nblocks, size_t n,
                       const EnsArgs& a, double eps2, double dt) {
    const double h = 0.5 * dt;
    GRAVIS_TRY(kick_drift(c, n, a.x, a.y, a.z, a.vx, a.vy, a.vz, a.ax, a.ay, a.az, h, dt));
    return ens_force(c, mixed, d_blocks, nblocks, a, eps2, h, 1);
}

// Permutation that sorts every system's particles by Morton key within the
// system (systems keep their ranges): perm[sorted position] = caller index.
gravis_status ens_sort(Ctx* c, const EnsPlan& plan, const EnsBlock* d_sys, const double* x,
                       const double* y, const double* z, uint32_t** perm) {
    const int nsys = (int)plan.systems.size();
    double4* cube;
    GRAVIS_TRY(buf(c, "ens_cube", (size_t)nsys, &cube));
    ens_cube_kernel<<<nsys, 256, 0, c->stream>>>(d_sys, x, y, z, cube);
    uint64_t* keys;
    uint32_t* vals;
    GRAVIS_TRY(sort_buffers(c, plan.n, &keys, &vals));
    ens_key_kernel<<<nsys, 256, 0, c->stream>>>(d_sys, x, y, z, cube, keys, vals);
    GRAVIS_TRY(check_launch(c, "ens_sort keys"));
    int bits = 3 * kEnsMortonBits;
    while ((1ll << (bits - 3 * kEnsMortonBits)) < nsys) ++bits;
    uint64_t* sk;
    return radix_sort(c, plan.n, bits, &sk, perm);
}

gravis_status check_solver_ens(Ctx* c, const gravis_solver* s) {
    if (!s) return fail(c, GRAVIS_INVALID_ARG, "solver is NULL");
    if (s->algorithm != GRAVIS_BRUTE_FORCE)
        return fail(c, GRAVIS_UNSUPPORTED, "the ensemble runs brute-force gravity only");
    if (s->precision != GRAVIS_FP64 && s->precision != GRAVIS_MIXED)
        return fail(c, GRAVIS_INVALID_ARG, "unknown precision %d", s->precision);
    if (!(s->softening >= 0.0)) return fail(c, GRAVIS_INVALID_ARG, "softening must be >= 0");
    return GRAVIS_OK;
}

}  // namespace

gravis_status ensemble_kdk(Ctx* c, const gravis_solver* s, int nsys, const uint64_t* offsets,
                           double* x, double* y, double* z, double* vx, double* vy, double* vz,
                           const double* m, double* ax, double* ay, double* az, double dt,
                           uint64_t nsteps, int use_graph) {
    GRAVIS_TRY(check_solver_ens(c, s));
    if (!(dt > 0.0)) return fail(c, GRAVIS_INVALID_ARG, "dt must be positive");
    EnsPlan plan;
    GRAVIS_TRY(make_plan(c, nsys, offsets, &plan));
    if (plan.n == 0 || nsteps == 0) return GRAVIS_OK;
    EnsBlock* d_blocks;
    GRAVIS_TRY(upload(c, "ens_blocks", plan.blocks, &d_blocks));
    const bool mixed = s->precision == GRAVIS_MIXED;
    const double eps2 = s->softening * s->softening;
    const int nb = (int)plan.blocks.size();
    const size_t n = plan.n;
    EnsArgs a{x, y, z, vx, vy, vz, ax, ay, az, m};
    // Mixed: run in per-system Morton order on a sorted copy of the state.
    EnsBlock* d_waves = nullptr;
    uint32_t* perm = nullptr;
    double* sorted = nullptr;
    // (the sort key holds the system index in 21 bits above the Morton key)
    const bool sorted_run = mixed && plan.systems.size() <= (1u << 21);
    if (sorted_run) {
        EnsBlock* d_sys;
        GRAVIS_TRY(upload(c, "ens_sys", plan.systems, &d_sys));
        GRAVIS_TRY(upload(c, "ens_waves", plan.waves, &d_waves));
        GRAVIS_TRY(ens_sort(c, plan, d_sys, x, y, z, &perm));
        GRAVIS_TRY(buf(c, "ens_sorted", 10 * n, &sorted));
        EnsArrays pa;
        const double* src[10] = {x, y, z, vx, vy, vz, ax, ay, az, m};
        for (int k = 0; k < 10; ++k) {
            pa.src[k] = src[k];
            pa.dst[k] = sorted + (size_t)k * n;
        }
        pa.count = 10;
        ens_permute_kernel<<<grid_for(n, 256, 4096), 256, 0, c->stream>>>(pa, perm, n, 0);
        GRAVIS_TRY(check_launch(c, "ens_permute"));
        a = EnsArgs{sorted, sorted + n, sorted + 2 * n, sorted + 3 * n, sorted + 4 * n,
                    sorted + 5 * n, sorted + 6 * n, sorted + 7 * n, sorted + 8 * n,
                    sorted + 9 * n};
    }
    const int nw = (int)plan.waves.size();
    double* wbox = nullptr;
    double4* rec = nullptr;
    if (sorted_run) {
        GRAVIS_TRY(buf(c, "ens_wbox", (size_t)12 * std::max(nw, 1), &wbox));
        GRAVIS_TRY(buf(c, "ens_rec", 2 * n, &rec));
    }
    const double h = 0.5 * dt;
    const float e2 = (float)eps2;
    const bool mask = !(e2 >= kMaskEps2F);
    const double q2 = c->bf_near_fraction * c->bf_near_fraction;
    // one step on records/boxes buffer `from` (0 or 1)
    auto wave_step = [&](int from, bool drift) -> gravis_status {
        const double4* pin = rec + (size_t)from * n;
        double4* pout = rec + (size_t)(1 - from) * n;
        const double* bin = wbox + (size_t)6 * nw * from;
        double* bout = wbox + (size_t)6 * nw * (1 - from);
#define GRAVIS_ENS_WAVE(M, D)                                                                  \
    ens_wave_kernel<M, D><<<nw, kWave, 0, c->stream>>>(d_waves, bin, pin, e2, q2, h, dt, a.vx, \
                                                       a.vy, a.vz, a.ax, a.ay, a.az, pout,   \
                                                       bout, a.x, a.y, a.z)
        if (mask && drift) GRAVIS_ENS_WAVE(true, true);
        else if (mask) GRAVIS_ENS_WAVE(true, false);
        else if (drift) GRAVIS_ENS_WAVE(false, true);
        else GRAVIS_ENS_WAVE(false, false);
#undef GRAVIS_ENS_WAVE
        return check_launch(c, "ens_wave");
    };
    auto step = [&]() -> gravis_status { return ens_step(c, mixed, d_blocks, nb, n, a, eps2, dt); };
    // back to the caller's order (positions, velocities, accelerations)
    auto restore = [&]() -> gravis_status {
        if (!sorted_run) return GRAVIS_OK;
        EnsArrays pa;
        double* dst[9] = {x, y, z, vx, vy, vz, ax, ay, az};
        for (int k = 0; k < 9; ++k) {
            pa.src[k] = sorted + (size_t)k * n;
            pa.dst[k] = dst[k];
        }
        pa.count = 9;
        ens_permute_kernel<<<grid_for(n, 256, 4096), 256, 0, c->stream>>>(pa, perm, n, 1);
        return check_launch(c, "ens_permute");
    };
    if (sorted_run) {
        // first half step, then nsteps - 1 launches that close a step and
        // open the next (buffers 0 -> 1 -> 0 ...), then the closing launch
        ens_pack_kernel<<<nw, kWave, 0, c->stream>>>(d_waves, a.x, a.y, a.z, a.m, a.vx, a.vy,
                                                     a.vz, a.ax, a.ay, a.az, h, dt, rec, wbox);
        GRAVIS_TRY(check_launch(c, "ens_pack"));
        uint64_t left = nsteps - 1;
        int cur = 0;
        if (use_graph && left >= 2) {
            // two steps (0 -> 1, 1 -> 0) captured once and replayed
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            GRAVIS_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            gravis_status st = wave_step(0, true);
            if (st == GRAVIS_OK) st = wave_step(1, true);
            hipError_t e = hipStreamEndCapture(c->stream, &graph);
            if (st != GRAVIS_OK) {
                if (graph) (void)hipGraphDestroy(graph);
                return st;
            }
            if (e != hipSuccess)
                return fail(c, GRAVIS_HIP_ERROR, "capture: %s", hipGetErrorString(e));
            e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(graph);
                return fail(c, GRAVIS_HIP_ERROR, "graph instantiate: %s", hipGetErrorString(e));
            }
            for (; left >= 2 && e == hipSuccess; left -= 2) e = hipGraphLaunch(exec, c->stream);
            (void)hipGraphExecDestroy(exec);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess)
                return fail(c, GRAVIS_HIP_ERROR, "graph launch: %s", hipGetErrorString(e));
        }
        for (; left > 0; --left) {
            GRAVIS_TRY(wave_step(cur, true));
            cur = 1 - cur;
        }
        GRAVIS_TRY(wave_step(cur, false));
        return restore();
    }
    if (!use_graph) {
        for (uint64_t t = 0; t < nsteps; ++t) GRAVIS_TRY(step());
        return restore();
    }
    // Capture one step (kick-drift + force/kick) and replay it.
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    GRAVIS_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    gravis_status st = step();
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (st != GRAVIS_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    if (e != hipSuccess) return fail(c, GRAVIS_HIP_ERROR, "capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return fail(c, GRAVIS_HIP_ERROR, "graph instantiate: %s", hipGetErrorString(e));
    }
    for (uint64_t t = 0; t < nsteps && e == hipSuccess; ++t) e = hipGraphLaunch(exec, c->stream);
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(c, GRAVIS_HIP_ERROR, "graph launch: %s", hipGetErrorString(e));
    return restore();
}

gravis_status ensemble_accel(Ctx* c, const gravis_solver* s, int nsys, const uint64_t* offsets,
                             const double* x, const double* y, const double* z,
                             const double* m, double* ax, double* ay, double* az) {
    GRAVIS_TRY(check_solver_ens(c, s));
    EnsPlan plan;
    GRAVIS_TRY(make_plan(c, nsys, offsets, &plan));
    if (plan.n == 0) return GRAVIS_OK;
    EnsBlock* d_blocks;
    GRAVIS_TRY(upload(c, "ens_blocks", plan.blocks, &d_blocks));
    const EnsArgs a{const_cast<double*>(x), const_cast<double*>(y), const_cast<double*>(z),
                    nullptr, nullptr, nullptr, ax, ay, az, m};
    return ens_force(c, s->precision == GRAVIS_MIXED, d_blocks, (int)plan.blocks.size(), a,
                     s->softening * s->softening, 0.0, 0);
}

gravis_status ensemble_adjoint(Ctx* c, const gravis_solver* s, int nsys, const uint64_t* offsets,
                               const double* x0, const double* y0, const double* z0,
                               const double* vx0, const double* vy0, const double* vz0,
                               const double* m, double dt, uint64_t nsteps, uint32_t ckpt,
                               const double* tx, const double* ty, const double* tz,
                               const double* gx, const double* gy, const double* gz,
                               const double* gvx, const double* gvy, const double* gvz,
                               double* loss, double* xT, double* ox, double* oy, double* oz,
                               double* ovx, double* ovy, double* ovz, double* om) {
    GRAVIS_TRY(check_solver_ens(c, s));
    if (!(dt > 0.0)) return fail(c, GRAVIS_INVALID_ARG, "dt must be positive");
    if (ckpt == 0) ckpt = 32;
    if (!tx && !gx) return fail(c, GRAVIS_INVALID_ARG, "need targets or position cotangents");
    EnsPlan plan;
    GRAVIS_TRY(make_plan(c, nsys, offsets, &plan));
    const size_t n = plan.n;
    if (n == 0) return GRAVIS_OK;
    EnsBlock *d_blocks, *d_sys;
    GRAVIS_TRY(upload(c, "ens_blocks", plan.blocks, &d_blocks));
    GRAVIS_TRY(upload(c, "ens_sys", plan.systems, &d_sys));
    const int nb = (int)plan.blocks.size();
    const bool mixed = s->precision == GRAVIS_MIXED;
    const double eps2 = s->softening * s->softening, h = 0.5 * dt;
    const size_t T = nsteps;
    const size_t nseg = (T + ckpt - 1) / ckpt;
    const size_t nck = nseg > 0 ? nseg : 1;

    double *st, *ck, *traj, *lam, *u, *gbuf;
    GRAVIS_TRY(buf(c, "adj_state", 9 * n, &st));             // x y z vx vy vz ax ay az
    GRAVIS_TRY(buf(c, "adj_ckpt", nck * 6 * n, &ck));        // (x, v) at segment starts
    GRAVIS_TRY(buf(c, "adj_traj", (size_t)(ckpt + 1) * 3 * n, &traj));
    GRAVIS_TRY(buf(c, "adj_lambda", 7 * n, &lam));           // lx ly lz lvx lvy lvz lm
    GRAVIS_TRY(buf(c, "adj_u", 3 * n, &u));
    GRAVIS_TRY(buf(c, "adj_g", 3 * n, &gbuf));
    EnsArgs a{st, st + n, st + 2 * n, st + 3 * n, st + 4 * n, st + 5 * n, st + 6 * n,
              st + 7 * n, st + 8 * n, m};
    hipStream_t S = c->stream;
    auto dcopy = [&](double* d, const double* src, size_t cnt) -> gravis_status {
        GRAVIS_HIP(c, hipMemcpyAsync(d, src, cnt * sizeof(double), hipMemcpyDeviceToDevice, S));
        return GRAVIS_OK;
    };
    // forward with checkpoints
    GRAVIS_TRY(dcopy(a.x, x0, n)); GRAVIS_TRY(dcopy(a.y, y0, n)); GRAVIS_TRY(dcopy(a.z, z0, n));
    GRAVIS_TRY(dcopy(a.vx, vx0, n)); GRAVIS_TRY(dcopy(a.vy, vy0, n)); GRAVIS_TRY(dcopy(a.vz, vz0, n));
    GRAVIS_TRY(ens_force(c, mixed, d_blocks, nb, a, eps2, h, 0));
    for (size_t t = 0; t < T; ++t) {
        if (t % ckpt == 0) GRAVIS_TRY(dcopy(ck + (t / ckpt) * 6 * n, st, 6 * n));
        GRAVIS_TRY(ens_step(c, mixed, d_blocks, nb, n, a, eps2, dt));
    }
    if (T == 0) GRAVIS_TRY(dcopy(ck, st, 6 * n));
    if (xT) GRAVIS_TRY(dcopy(xT, st, 6 * n));  // final x y z vx vy vz
    // cotangents of the final state
    double *lx = lam, *ly = lam + n, *lz = lam + 2 * n, *lv = lam + 3 * n, *lm = lam + 6 * n;
    if (tx) {
        double* dloss;
        GRAVIS_TRY(buf(c, "adj_loss", (size_t)nsys, &dloss));
        adj_loss_kernel<<<nsys, 256, 0, S>>>(d_sys, nsys, a.x, a.y, a.z, tx, ty, tz, lx, ly, lz,
                                             dloss);
        GRAVIS_TRY(check_launch(c, "adj_loss"));
        if (loss) GRAVIS_TRY(dcopy(loss, dloss, (size_t)nsys));
    } else {
        GRAVIS_TRY(dcopy(lx, gx, n)); GRAVIS_TRY(dcopy(ly, gy, n)); GRAVIS_TRY(dcopy(lz, gz, n));
    }
    if (gvx) {
        GRAVIS_TRY(dcopy(lv, gvx, n)); GRAVIS_TRY(dcopy(lv + n, gvy, n)); GRAVIS_TRY(dcopy(lv + 2 * n, gvz, n));
    } else {
        GRAVIS_HIP(c, hipMemsetAsync(lv, 0, 3 * n * sizeof(double), S));
    }
    GRAVIS_HIP(c, hipMemsetAsync(lm, 0, n * sizeof(double), S));
    const unsigned eg = grid_for(n, 256, 4096);
    // top: K2 of the last step at x_T
    if (T > 0) {
        adj_level_kernel<<<eg, 256, 0, S>>>(n, lx, ly, lz, lv, 0.0, h, u);
        if (eps2 > 0.0)
            ens_vjp_kernel<false><<<nb, kEnsBlock, 0, S>>>(d_blocks, a.x, a.y, a.z, m, u, u + n,
                                                           u + 2 * n, eps2, lx, ly, lz, lm);
        else
            ens_vjp_kernel<true><<<nb, kEnsBlock, 0, S>>>(d_blocks, a.x, a.y, a.z, m, u, u + n,
                                                          u + 2 * n, eps2, lx, ly, lz, lm);
        GRAVIS_TRY(check_launch(c, "adj top"));
    }
    for (size_t sg = nseg; sg-- > 0;) {
        const size_t t0 = sg * ckpt, t1 = std::min(T, t0 + ckpt);
        // recompute the segment, storing positions at every level
        GRAVIS_TRY(dcopy(st, ck + sg * 6 * n, 6 * n));
        GRAVIS_TRY(ens_force(c, mixed, d_blocks, nb, a, eps2, h, 0));
        GRAVIS_TRY(dcopy(traj, st, 3 * n));
        for (size_t t = t0; t < t1; ++t) {
            GRAVIS_TRY(ens_step(c, mixed, d_blocks, nb, n, a, eps2, dt));
            if (t + 1 < t1) GRAVIS_TRY(dcopy(traj + (t + 1 - t0) * 3 * n, st, 3 * n));
        }
        for (size_t t = t1; t-- > t0;) {
            const double* xt = traj + (t - t0) * 3 * n;
            adj_level_kernel<<<eg, 256, 0, S>>>(n, lx, ly, lz, lv, dt, t > 0 ? 2.0 * h : h, u);
            if (eps2 > 0.0)
                ens_vjp_kernel<false><<<nb, kEnsBlock, 0, S>>>(d_blocks, xt, xt + n, xt + 2 * n,
                                                               m, u, u + n, u + 2 * n, eps2, lx,
                                                               ly, lz, lm);
            else
                ens_vjp_kernel<true><<<nb, kEnsBlock, 0, S>>>(d_blocks, xt, xt + n, xt + 2 * n,
                                                              m, u, u + n, u + 2 * n, eps2, lx,
                                                              ly, lz, lm);
        }
        GRAVIS_TRY(check_launch(c, "adj segment"));
    }
    GRAVIS_TRY(dcopy(ox, lx, n)); GRAVIS_TRY(dcopy(oy, ly, n)); GRAVIS_TRY(dcopy(oz, lz, n));
    GRAVIS_TRY(dcopy(ovx, lv, n)); GRAVIS_TRY(dcopy(ovy, lv + n, n)); GRAVIS_TRY(dcopy(ovz, lv + 2 * n, n));
    GRAVIS_TRY(dcopy(om, lm, n));
    return GRAVIS_OK;
}

}  // namespace gravis
