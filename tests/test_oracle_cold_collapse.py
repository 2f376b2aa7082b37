"""Pins of the oracle's cold-collapse generator (ColdCollapse::generate,
crates/sim-core/src/scenarios/cold_collapse.rs:56-102), which has no
reference fixture: the product's generator is compared with it bit for bit
(tests/test_snapshot_cli.py), so it must be right on its own.

  * the samples before the centre-of-mass shift lie inside the sphere, and
    (r/R)^3 is uniform on [0, 1] (Kolmogorov-Smirnov; a radius drawn
    uniformly instead, the usual slip, is rejected by the same test);
  * directions are isotropic (mean unit vector, octant counts);
  * velocities are zero and masses equal M/n; gas particles carry u = 0.05,
    h = 1.5 (V/n)^(1/3), alpha 1 and type 4, collisionless ones type 0;
  * the gas form draws exactly the Evrard sampler's points
    (fill_uniform_gas_sphere, scenarios/mod.rs:33-51, pinned by the
    reference's Evrard tests) for the same seed and radius;
  * the shifted output equals the samples minus their mean position.
"""
import numpy as np
import pytest
from scipy import stats

import oracle as O

N = 20000


@pytest.mark.parametrize("gas", [False, True])
@pytest.mark.parametrize("radius", [1.0, 2.5])
def test_cold_collapse_samples(gas, radius):
    s = O.cold_collapse(N, seed=42, radius=radius, total_mass=3.0, gas=gas, com_shift=False)
    r = np.sqrt(s.x ** 2 + s.y ** 2 + s.z ** 2)
    assert r.max() <= radius
    u = (r / radius) ** 3
    assert stats.kstest(u, "uniform").pvalue > 1e-3
    assert stats.kstest(r / radius, "uniform").pvalue < 1e-12  # power against r ~ U[0, R]
    d = np.stack([s.x, s.y, s.z], 1) / r[:, None]
    assert np.linalg.norm(d.mean(0)) < 4.0 / np.sqrt(N)
    octant = (s.x >= 0) * 1 + (s.y >= 0) * 2 + (s.z >= 0) * 4
    counts = np.bincount(octant, minlength=8)
    assert stats.chisquare(counts).pvalue > 1e-3
    for f in ("vx", "vy", "vz", "ax", "ay", "az", "rho", "p", "dudt", "c"):
        assert np.all(getattr(s, f) == 0.0), f
    assert np.all(s.m == 3.0 / N)
    if gas:
        h = 1.5 * np.cbrt(4.0 / 3.0 * np.pi * radius ** 3 / N)
        assert np.all(s.type == 4) and np.all(s.u == 0.05) and np.all(s.alpha == 1.0)
        assert np.allclose(s.h, h, rtol=1e-15, atol=0)
    else:
        assert np.all(s.type == 0) and np.all(s.u == 0.0) and np.all(s.h == 0.0)


def test_cold_collapse_gas_draws_evrard_points():
    c = O.cold_collapse(5000, seed=7, radius=1.5, gas=True, com_shift=False)
    e = O.evrard(5000, seed=7, radius=1.5)
    for f in ("x", "y", "z"):
        assert np.array_equal(getattr(c, f), getattr(e, f)), f


@pytest.mark.parametrize("gas", [False, True])
def test_cold_collapse_com_shift(gas):
    raw = O.cold_collapse(N, seed=11, gas=gas, com_shift=False)
    out = O.cold_collapse(N, seed=11, gas=gas)
    for f in ("x", "y", "z"):
        a, b = getattr(raw, f), getattr(out, f)
        assert np.abs((a - a.mean()) - b).max() <= 1e-15, f
        assert abs(b.mean()) <= 1e-15
    assert np.all(out.vx == 0.0)
