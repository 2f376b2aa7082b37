"""Pins of the renderer-packing restatement (oracle/render.py) by closed
forms of crates/render-core/src/color.rs:38-52 and the palette of
color.rs:9-19, and the GpuParticle layout of gpu_types.rs:5-9 (32 bytes:
position, mass, colour)."""
import numpy as np

from oracle import render as R


def test_layout_is_32_bytes_position_mass_color():
    assert R.GPU_PARTICLE.itemsize == 32
    assert R.GPU_PARTICLE.fields["position"][1] == 0
    assert R.GPU_PARTICLE.fields["mass"][1] == 12
    assert R.GPU_PARTICLE.fields["color"][1] == 16


def test_gas_color_closed_forms():
    c = R.gas_temperature_color([0.01, 100.0, 1.0, 1e-9, 1e9])
    assert np.allclose(c[0], [0.0, 0.0, 0.9, 0.7], atol=1e-7)      # log u = -2: t = 0
    assert np.allclose(c[1], [1.0, 1.0, 0.5, 1.0], atol=1e-7)      # log u = 2: t = 1
    h = 0.5
    assert np.allclose(c[2], [h ** 0.5, h ** 1.2, h ** 0.7 * 0.9 + h * 0.5, 0.7 + 0.3 * h],
                       rtol=1e-6)                                    # log u = 0: t = 1/2
    assert np.allclose(c[3], c[0]) and np.allclose(c[4], c[1])     # clamped both ends


def test_palette_and_dispatch():
    p = R.gpu_particles([0.0] * 6, [0.0] * 6, [0.0] * 6, [1.0] * 6,
                        u=[0, 0, 0, 0, 100.0, 0], ptype=[0, 1, 2, 3, 4, 9])
    assert np.allclose(p["color"][1], [0.6, 0.75, 1.0, 1.0])
    assert np.allclose(p["color"][3], [0.3, 0.2, 0.5, 0.15])
    assert np.allclose(p["color"][4], [1.0, 1.0, 0.5, 1.0])        # gas: by temperature
    assert np.allclose(p["color"][5], [0.7, 0.85, 1.0, 1.0])       # unknown: fallback
