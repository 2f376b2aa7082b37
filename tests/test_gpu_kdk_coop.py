"""FP64 brute-force KDK of 256 < n <= 2040 particles over several steps runs
in one cooperative launch with a grid barrier per step (kdk_coop_kernel).
Its forces are summed in another order than the step-by-step path's
(64 strided lane sums and a butterfly instead of 256-record chunks), so the
two agree to rounding, not bit for bit:

  * against the step-by-step path (kdk_coop_max_n = 0), 20 steps of
    LeapfrogKDK::step (integrator.rs:41-77) from the same state: positions
    and velocities within 1e-12 of their scale, accelerations within 1e-11;
  * against the oracle (gravity.rs:47-80, i<j, f64): the accelerations
    returned after the steps are those of the returned positions within
    1e-12 (condition-normalised), and the positions after 5 steps agree
    with the oracle's sim-core loop within 1e-12 of their scale;
  * two runs give identical bits (the barrier fixes every sum's order).
The C1 workload (Plummer 1000, 10,000 steps) runs through this path in
test_gpu_bf.py::test_c1_plummer_1k_10k_steps.
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


FIELDS = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az")


def run(g, s, eps, steps, coop_max=2040):
    ctx = g.Context(0)
    ctx.set_tuning("kdk_coop_max_n", coop_max)
    p = g.Particles.from_arrays(s.x, s.y, s.z, s.vx, s.vy, s.vz, s.m)
    bf = g.BruteForce(eps, ctx=ctx)
    bf.compute_accelerations(p)
    g.LeapfrogKDK().run(p, bf, 1e-3, steps)
    ctx.close()
    return p


def scaled_diff(a, b, fields):
    ref = np.concatenate([getattr(a, f) for f in fields])
    got = np.concatenate([getattr(b, f) for f in fields])
    return float(np.abs(got - ref).max() / np.abs(ref).max())


@pytest.mark.parametrize("n,eps", [(257, 0.05), (1000, 0.05), (2040, 0.05), (1000, 0.0)])
def test_coop_matches_step_by_step(g, n, eps):
    say(f"kdk_coop n={n} eps={eps}: 20 steps, cooperative against step by step")
    s = O.plummer(n, 7)
    a = run(g, s, eps, 20, coop_max=0)
    b = run(g, s, eps, 20)
    dx = scaled_diff(a, b, ("x", "y", "z"))
    dv = scaled_diff(a, b, ("vx", "vy", "vz"))
    da = scaled_diff(a, b, ("ax", "ay", "az"))
    say(f"  position {dx:.2e} velocity {dv:.2e} acceleration {da:.2e}")
    assert dx <= 1e-12 and dv <= 1e-12
    assert da <= 1e-11
    assert all(np.isfinite(getattr(b, f)).all() for f in FIELDS)


def test_coop_against_oracle(g):
    n, eps, steps = 1000, 0.05, 5
    say(f"kdk_coop n={n}: {steps} steps against the oracle loop")
    s = O.plummer(n, 42)
    p = run(g, s, eps, steps)
    # the returned accelerations are the FP64 forces at the returned positions
    fin = s.copy()
    fin.x, fin.y, fin.z = p.x.copy(), p.y.copy(), p.z.copy()
    ref, absum = O.bf_targets(fin, eps, np.arange(n))
    got = np.stack([p.ax, p.ay, p.az], 1)
    cond = float((np.sqrt(((got - ref) ** 2).sum(1)) / absum).max())
    ref_loop = s.copy()
    O.bf_accel(ref_loop, eps)
    O.kdk_steps(ref_loop, 1e-3, steps, eps)
    pos = np.concatenate([p.x, p.y, p.z])
    pref = np.concatenate([ref_loop.x, ref_loop.y, ref_loop.z])
    dpos = float(np.abs(pos - pref).max() / np.abs(pref).max())
    say(f"  force cond {cond:.2e}, position vs oracle loop {dpos:.2e}")
    assert cond <= 1e-12
    assert dpos <= 1e-12


def test_coop_deterministic(g):
    s = O.plummer(1500, 3)
    say("kdk_coop n=1500: two runs of 30 steps")
    a = run(g, s, 0.02, 30)
    b = run(g, s, 0.02, 30)
    for f in FIELDS:
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
