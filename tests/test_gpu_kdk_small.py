"""Small systems (n <= 256) with FP64 brute force run every KDK step in one
single-workgroup launch (kdk_small_kernel). It repeats the multi-launch
path's operations in the same order, so the two must agree bit for bit;
with the launches gone, the reference's long Kepler tests
(crates/sim-core/tests/kepler.rs) run at their full length on the GPU:

  * kepler.rs:14-58 -- e = 0.5, dt = T/10,000, 1000 orbits (10^7 steps):
    |dE/E| < 1e-4;
  * kepler.rs:61-98 -- e = 0.7, 100 orbits: |dL/L| < 1e-6;
  * kepler.rs:100-150 -- e = 0.3, 10 orbits: the period from the last two
    upward y crossings of particle 1 within 1e-4 of 2 pi;
  * the m1_kepler_diag.csv energies (1e5 steps) are replayed by the
    headless tests through the same path.
"""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def g():
    import gravis_amd
    return gravis_amd


@pytest.mark.parametrize("n,eps", [(2, 5e-4), (100, 0.05), (256, 0.05), (255, 0.0),
                                   (64, 1e-60)])
def test_fused_equals_multi_launch(g, n, eps):
    s = O.plummer(n, 9) if n > 2 else O.two_body(0.5)
    runs = []
    for fused in (256, 0):
        ctx = g.Context(0)
        ctx.set_tuning("kdk_fused_max_n", fused)
        p = g.Particles.from_arrays(s.x, s.y, s.z, s.vx, s.vy, s.vz, s.m)
        bf = g.BruteForce(eps, ctx=ctx)
        bf.compute_accelerations(p)
        g.LeapfrogKDK().run(p, bf, 1e-3, 300)
        ctx.close()
        runs.append(p)
    for f in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"):
        assert np.array_equal(getattr(runs[0], f), getattr(runs[1], f)), f
    assert np.isfinite(runs[0].x).all()


def kepler(g, e, orbits):
    p = g.scenarios.two_body(e)
    T = g.scenarios.two_body_period()
    dt = T / 10_000.0
    eps = (1.0 - e) * 0.001
    bf = g.BruteForce(eps)
    bf.compute_accelerations(p)
    d0 = g.diagnostics.compute(p, eps)
    steps = int(round(T / dt)) * orbits
    g.LeapfrogKDK().run(p, bf, dt, steps)
    d1 = g.diagnostics.compute(p, eps, time=steps * dt, step=steps)
    return d0, d1, steps


def test_kepler_energy_1000_orbits(g):
    d0, d1, steps = kepler(g, 0.5, 1000)
    de = abs((d1.total_energy - d0.total_energy) / d0.total_energy)
    say(f"Kepler e=0.5, {steps} steps: dE/E {de:.3e}")
    assert de < 1e-4


def test_kepler_angular_momentum_100_orbits(g):
    d0, d1, steps = kepler(g, 0.7, 100)
    l0, l1 = np.linalg.norm(d0.angular_momentum), np.linalg.norm(d1.angular_momentum)
    dl = abs((l1 - l0) / l0)
    say(f"Kepler e=0.7, {steps} steps: dL/L {dl:.3e}")
    assert dl < 1e-6


def test_kepler_third_law_period(g):
    p = g.scenarios.two_body(0.3)
    T = g.scenarios.two_body_period()
    dt = T / 10_000.0
    eps = 0.7 * 0.001
    bf = g.BruteForce(eps)
    bf.compute_accelerations(p)
    kdk = g.LeapfrogKDK()
    steps = int(round(10.0 * T / dt))
    last_y, t, crossings = p.y[1], 0.0, []
    for _ in range(steps):
        kdk.step(p, bf, dt)
        t += dt
        y = p.y[1]
        if last_y <= 0.0 < y:
            frac = abs(last_y) / abs(y - last_y)
            crossings.append(t - dt * (1.0 - frac))
        last_y = y
    assert len(crossings) >= 2
    err = abs((crossings[-1] - crossings[-2]) - T) / T
    say(f"Kepler e=0.3: {len(crossings)} crossings, period error {err:.3e}")
    assert err < 1e-4
