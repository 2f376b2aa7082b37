// render.hip -- the render-side view of the device state: one 32-byte
// GpuParticle record per particle (crates/render-core/src/gpu_types.rs:5-9:
// position f32 x3, mass f32, colour f32 x4), packed from the FP64 state on
// the GPU so a render thread can take it without a host round trip.
//
// Reference: the sim thread converts positions, masses and internal energies
// with `as f32` (crates/native-app/src/sim_thread.rs:250-258) and the
// renderer zips them with particle_color (crates/render-core/src/color.rs:
// 9-52) into GpuParticle (crates/render-core/src/particles.rs:272-282).
// Here the f64 -> f32 conversions round to nearest like `as f32`; the gas
// colour map uses the device's log10f / powf.
//
// HBM bound: 41 bytes read and 32 bytes written per particle.
#include "internal.h"

namespace gravis {
namespace {

// particle_type_to_color (color.rs:9-19)
__device__ __forceinline__ float4 type_color(uint8_t t) {
    switch (t) {
        case 1: return make_float4(0.6f, 0.75f, 1.0f, 1.0f);
        case 2: return make_float4(1.0f, 0.8f, 0.4f, 1.0f);
        case 3: return make_float4(0.3f, 0.2f, 0.5f, 0.15f);
        case 4: return make_float4(0.3f, 0.6f, 1.0f, 0.8f);
        default: return make_float4(0.7f, 0.85f, 1.0f, 1.0f);
    }
}

// gas_temperature_color (color.rs:38-52): log10(u) over [-2, 2] -> t,
// cool blue -> warm yellow -> hot white, hotter more opaque
__device__ __forceinline__ float4 gas_color(float u) {
    const float log_u = log10f(fmaxf(u, 1e-6f));
    const float t = fminf(fmaxf((log_u + 2.0f) / 4.0f, 0.0f), 1.0f);
    const float r = powf(t, 0.5f);
    const float g = powf(t, 1.2f);
    const float b = powf(1.0f - t, 0.7f) * 0.9f + t * 0.5f;
    const float a = 0.7f + 0.3f * t;
    return make_float4(r, g, b, a);
}

__global__ __launch_bounds__(256) void render_pack_kernel(
    size_t n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ m, const double* __restrict__ u,
    const uint8_t* __restrict__ type, float4* __restrict__ out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const uint8_t t = type ? type[i] : (uint8_t)0;
        // particle_color (color.rs:22-28): gas by temperature, else by type
        const float4 col = t == 4 ? gas_color(u ? (float)u[i] : 0.0f) : type_color(t);
        out[2 * i] = make_float4((float)x[i], (float)y[i], (float)z[i], (float)m[i]);
        out[2 * i + 1] = col;
    }
}

}  // namespace

gravis_status render_pack(Ctx* c, size_t n, const double* x, const double* y, const double* z,
                          const double* m, const double* u, const uint8_t* type, void* out) {
    if (n == 0) return GRAVIS_OK;
    render_pack_kernel<<<grid_for(n, 256, 16384), 256, 0, c->stream>>>(
        n, x, y, z, m, u, type, static_cast<float4*>(out));
    return check_launch(c, "render_pack");
}

}  // namespace gravis
