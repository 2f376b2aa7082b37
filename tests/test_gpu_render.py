"""The render view packed on the GPU (gravis_render_pack /
gravis_render_pack_device): GpuParticle records (crates/render-core/src/
gpu_types.rs:5-9) against the CPU restatement in oracle/render.py.

Bars: positions and masses bit-identical to numpy's round-to-nearest f64 ->
f32 (Rust's `as f32`); palette colours bit-identical; gas colours (log10f /
powf on the device) within 2e-6 absolute of the float32 restatement; the
device form writes the same bytes as the host form; a galaxy-collision
state (types 1-3) and an Evrard state (gas) both covered.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import render as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g():
    import gravis_amd
    return gravis_amd


def check(g, p):
    got = g.render_pack(p)
    ref = R.gpu_particles(p.x, p.y, p.z, p.mass, p.internal_energy, p.particle_type)
    assert got.dtype.itemsize == 32
    assert np.array_equal(got["position"].view(np.uint32), ref["position"].view(np.uint32))
    assert np.array_equal(got["mass"].view(np.uint32), ref["mass"].view(np.uint32))
    gas = p.particle_type == 4
    assert np.array_equal(got["color"][~gas], ref["color"][~gas])
    if gas.any():
        assert np.abs(got["color"][gas] - ref["color"][gas]).max() <= 2e-6
    return got


def test_render_pack_galaxy_types(g):
    p = g.scenarios.galaxy_collision(n_per_galaxy=5000)
    assert set(np.unique(p.particle_type)) == {1, 2, 3}
    check(g, p)


def test_render_pack_gas_spread_of_temperatures(g):
    p = g.scenarios.evrard(4000)
    p.internal_energy[:] = 10.0 ** np.random.default_rng(3).uniform(-3, 3, p.count)
    p.particle_type[::7] = 0  # mixed population
    check(g, p)


def test_render_pack_device_matches_host(g):
    import torch
    p = g.scenarios.evrard(3000)
    p.internal_energy[:] = np.linspace(1e-3, 50.0, p.count)
    host = check(g, p)
    ctx = g.default_context()
    T = [torch.from_numpy(getattr(p, f).copy()).cuda() for f in ("x", "y", "z", "mass",
                                                                  "internal_energy")]
    ty = torch.from_numpy(p.particle_type.copy()).cuda()
    out = torch.zeros(8 * p.count, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ctx.check(ctx._L.gravis_render_pack_device(ctx.handle, p.count,
                                               *[C.c_void_p(t.data_ptr()) for t in T],
                                               C.c_void_p(ty.data_ptr()),
                                               C.c_void_p(out.data_ptr())))
    ctx.synchronize()
    dev = out.cpu().numpy().view(R.GPU_PARTICLE)
    assert dev.tobytes() == host.tobytes()
