"""CPU restatement of the renderer's particle packing (test infrastructure;
only tests/ import it).

Follows crates/render-core/src/color.rs:9-52 (particle_type_to_color,
particle_color, gas_temperature_color) in float32 and the `as f32`
conversions of crates/native-app/src/sim_thread.rs:250-258, zipped into
GpuParticle records (crates/render-core/src/gpu_types.rs:5-9,
particles.rs:272-282).
"""
from __future__ import annotations

import numpy as np

# color.rs:9-19
PALETTE = {
    0: (0.7, 0.85, 1.0, 1.0),
    1: (0.6, 0.75, 1.0, 1.0),
    2: (1.0, 0.8, 0.4, 1.0),
    3: (0.3, 0.2, 0.5, 0.15),
    4: (0.3, 0.6, 1.0, 0.8),
}
FALLBACK = (0.7, 0.85, 1.0, 1.0)

GPU_PARTICLE = np.dtype([("position", np.float32, 3), ("mass", np.float32),
                         ("color", np.float32, 4)])


def type_color(t: int) -> np.ndarray:
    return np.array(PALETTE.get(int(t), FALLBACK), dtype=np.float32)


def gas_temperature_color(u) -> np.ndarray:
    """color.rs:38-52 in float32 for an array of internal energies (f32)."""
    u = np.asarray(u, dtype=np.float32)
    log_u = np.log10(np.maximum(u, np.float32(1e-6)))
    t = np.clip((log_u + np.float32(2.0)) / np.float32(4.0), np.float32(0.0), np.float32(1.0))
    r = np.power(t, np.float32(0.5))
    g = np.power(t, np.float32(1.2))
    b = np.power(np.float32(1.0) - t, np.float32(0.7)) * np.float32(0.9) + t * np.float32(0.5)
    a = np.float32(0.7) + np.float32(0.3) * t
    return np.stack([r, g, b, a], -1).astype(np.float32)


def gpu_particles(x, y, z, m, u=None, ptype=None) -> np.ndarray:
    """GpuParticle records (structured array of GPU_PARTICLE)."""
    n = len(x)
    out = np.zeros(n, dtype=GPU_PARTICLE)
    out["position"] = np.stack([np.asarray(x), np.asarray(y), np.asarray(z)], 1).astype(np.float32)
    out["mass"] = np.asarray(m).astype(np.float32)
    t = np.zeros(n, dtype=np.uint8) if ptype is None else np.asarray(ptype)
    uf = np.zeros(n, dtype=np.float32) if u is None else np.asarray(u).astype(np.float32)
    col = np.array([type_color(k) for k in t], dtype=np.float32).reshape(n, 4)
    gas = t == 4
    if gas.any():
        col[gas] = gas_temperature_color(uf[gas])
    out["color"] = col
    return out
