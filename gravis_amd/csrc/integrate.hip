// integrate.hip -- leapfrog KDK pieces and FP64 diagnostics.
//
// LeapfrogKDK::step (crates/sim-core/src/integrator.rs:41-77) is
//   v += a*dt/2;  x += v*dt;  a = F(x);  v += a*dt/2
// as four sequential loops. On the device the first two loops are one
// kernel (kick_drift_kernel) and the last half-kick is fused into the force
// finishing kernel (FinishArgs::kAssignKick). Each update keeps the
// reference's rounding: v + (a*h), x + (v*dt), no contraction.
//
// Diagnostics restate diagnostics.rs:50-148 in FP64. The O(N) sums use a
// fixed block partition and a fixed reduction tree; the O(N^2) softened
// potential is summed per target over all j != i and halved.
#include "internal.h"

namespace gravis {
namespace {

__global__ __launch_bounds__(256) void kick_drift_kernel(double* __restrict__ x,
                                                         double* __restrict__ y,
                                                         double* __restrict__ z,
                                                         double* __restrict__ vx,
                                                         double* __restrict__ vy,
                                                         double* __restrict__ vz,
                                                         const double* __restrict__ ax,
                                                         const double* __restrict__ ay,
                                                         const double* __restrict__ az, size_t n,
                                                         double h, double dt) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const double nvx = vx[i] + ax[i] * h;
        const double nvy = vy[i] + ay[i] * h;
        const double nvz = vz[i] + az[i] * h;
        vx[i] = nvx;
        vy[i] = nvy;
        vz[i] = nvz;
        x[i] = x[i] + nvx * dt;
        y[i] = y[i] + nvy * dt;
        z[i] = z[i] + nvz * dt;
    }
}

constexpr int kDiagBlock = 256;
constexpr int kDiagMaxBlocks = 512;
constexpr int kDiagFields = 16;  // K, px,py,pz, Lx,Ly,Lz, mx,my,mz, mvx,mvy,mvz, M, Eth, U

__global__ __launch_bounds__(kDiagBlock) void diag_sums_kernel(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ vx, const double* __restrict__ vy, const double* __restrict__ vz,
    const double* __restrict__ m, const double* __restrict__ u, const uint8_t* __restrict__ pt,
    size_t n, double* __restrict__ part) {
    double a[15];
#pragma unroll
    for (int k = 0; k < 15; ++k) a[k] = 0.0;
    for (size_t i = blockIdx.x * (size_t)kDiagBlock + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kDiagBlock) {
        const double mi = m[i], X = x[i], Y = y[i], Z = z[i], VX = vx[i], VY = vy[i], VZ = vz[i];
        const double v2 = VX * VX + VY * VY + VZ * VZ;
        a[0] += 0.5 * mi * v2;
        a[1] += mi * VX;
        a[2] += mi * VY;
        a[3] += mi * VZ;
        a[4] += mi * (Y * VZ - Z * VY);
        a[5] += mi * (Z * VX - X * VZ);
        a[6] += mi * (X * VY - Y * VX);
        a[7] += mi * X;
        a[8] += mi * Y;
        a[9] += mi * Z;
        a[10] += mi * VX;
        a[11] += mi * VY;
        a[12] += mi * VZ;
        a[13] += mi;
        if (u && pt && pt[i] == 4) a[14] += mi * u[i];
    }
    __shared__ double s[kDiagBlock][15];
#pragma unroll
    for (int k = 0; k < 15; ++k) s[threadIdx.x][k] = a[k];
    __syncthreads();
    for (int half = kDiagBlock / 2; half > 0; half >>= 1) {
        if (threadIdx.x < (unsigned)half)
            for (int k = 0; k < 15; ++k) s[threadIdx.x][k] += s[threadIdx.x + half][k];
        __syncthreads();
    }
    if (threadIdx.x < 15) part[blockIdx.x * kDiagFields + threadIdx.x] = s[0][threadIdx.x];
}

// phi_i = sum_{j != i} m_j / sqrt(|d|^2 + eps^2), FP64, tiled over j.
constexpr int kPotBlock = 256;
__global__ __launch_bounds__(kPotBlock) void diag_potential_kernel(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ m, size_t n, double eps2, double* __restrict__ phi_m) {
    __shared__ double4 tile[kPotBlock];
    const size_t i = blockIdx.x * (size_t)kPotBlock + threadIdx.x;
    const bool ok = i < n;
    const double xi = ok ? x[i] : 0.0, yi = ok ? y[i] : 0.0, zi = ok ? z[i] : 0.0;
    double phi = 0.0;
    for (size_t jb = 0; jb < n; jb += kPotBlock) {
        const size_t j = jb + threadIdx.x;
        tile[threadIdx.x] = j < n ? make_double4(x[j], y[j], z[j], m[j]) : make_double4(0, 0, 0, 0);
        __syncthreads();
        const int cnt = (int)(n - jb < (size_t)kPotBlock ? n - jb : kPotBlock);
        double part = 0.0;
        for (int k = 0; k < cnt; ++k) {
            const double4 s = tile[k];
            const double dx = s.x - xi, dy = s.y - yi, dz = s.z - zi;
            const double r2 = dx * dx + dy * dy + dz * dz + eps2;
            const double t = s.w / sqrt(r2);
            part += (jb + k == i) ? 0.0 : t;
        }
        phi += part;
        __syncthreads();
    }
    if (ok) phi_m[i] = m[i] * phi;
}

__global__ __launch_bounds__(kDiagBlock) void diag_potential_sum_kernel(
    const double* __restrict__ phi_m, size_t n, double* __restrict__ part) {
    double a = 0.0;
    for (size_t i = blockIdx.x * (size_t)kDiagBlock + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kDiagBlock)
        a += phi_m[i];
    __shared__ double s[kDiagBlock];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int half = kDiagBlock / 2; half > 0; half >>= 1) {
        if (threadIdx.x < (unsigned)half) s[threadIdx.x] += s[threadIdx.x + half];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x * kDiagFields + 15] = s[0];
}

__global__ void diag_final_kernel(const double* __restrict__ part, int nb, int with_pot,
                                  double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a[kDiagFields];
    for (int k = 0; k < kDiagFields; ++k) a[k] = 0.0;
    for (int b = 0; b < nb; ++b)
        for (int k = 0; k < kDiagFields; ++k) a[k] += part[b * kDiagFields + k];
    const double K = a[0];
    const double U = with_pot ? -0.5 * a[15] : 0.0;
    const double Eth = a[14];
    const double mt = a[13];
    out[0] = K;
    out[1] = U;
    out[2] = Eth;
    out[3] = K + U + Eth;
    out[4] = a[1]; out[5] = a[2]; out[6] = a[3];
    out[7] = a[4]; out[8] = a[5]; out[9] = a[6];
    for (int d = 0; d < 3; ++d) {
        out[10 + d] = mt > 0.0 ? a[7 + d] / mt : 0.0;
        out[13 + d] = mt > 0.0 ? a[10 + d] / mt : 0.0;
    }
    out[16] = fabs(U) > 0.0 ? 2.0 * K / fabs(U) : 0.0;
}

}  // namespace

gravis_status kick_drift(Ctx* c, size_t n, double* x, double* y, double* z, double* vx,
                         double* vy, double* vz, const double* ax, const double* ay,
                         const double* az, double half_dt, double dt) {
    kick_drift_kernel<<<grid_for(n, 256, 8192), 256, 0, c->stream>>>(x, y, z, vx, vy, vz, ax, ay,
                                                                     az, n, half_dt, dt);
    return check_launch(c, "kick_drift");
}

gravis_status diag_device(Ctx* c, size_t n, double eps, int with_potential, const double* x,
                          const double* y, const double* z, const double* vx, const double* vy,
                          const double* vz, const double* m, const double* u,
                          const uint8_t* ptype, double out[19]) {
    double *part, *res, *phi;
    GRAVIS_TRY(buf(c, "diag_part", (size_t)kDiagMaxBlocks * kDiagFields, &part));
    GRAVIS_TRY(buf(c, "diag_res", 32, &res));
    const unsigned nb = grid_for(n, kDiagBlock, kDiagMaxBlocks);
    diag_sums_kernel<<<nb, kDiagBlock, 0, c->stream>>>(x, y, z, vx, vy, vz, m, u, ptype, n, part);
    if (with_potential) {
        GRAVIS_TRY(buf(c, "diag_phi", n, &phi));
        diag_potential_kernel<<<grid_for(n, kPotBlock), kPotBlock, 0, c->stream>>>(
            x, y, z, m, n, eps * eps, phi);
        diag_potential_sum_kernel<<<nb, kDiagBlock, 0, c->stream>>>(phi, n, part);
    }
    diag_final_kernel<<<1, 64, 0, c->stream>>>(part, (int)nb, with_potential, res);
    GRAVIS_TRY(check_launch(c, "diagnostics"));
    GRAVIS_HIP(c, hipMemcpyAsync(out, res, 17 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    GRAVIS_HIP(c, hipStreamSynchronize(c->stream));
    return GRAVIS_OK;
}

}  // namespace gravis
