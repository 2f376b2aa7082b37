"""The reference's invariant tests (crates/sim-core/tests/), run on the HIP
backend with the reference's scenarios, step counts and bounds:

  * determinism.rs:28-68 -- 50 steps of a 200-particle Plummer sphere give
    bit-identical positions run to run (brute force and Barnes-Hut, FP64 and
    mixed, on one context and on a fresh one);
  * conservation.rs -- Plummer N=200: |p| < 1e-12 and COM drift < 1e-12
    after 1000 steps; two-body e = 0.8: dL/L < 1e-8 over 50 orbits; two-body
    e = 0.5: |p| < 1e-14 over 10 orbits;
  * plummer.rs:14-101 -- N=500: mean virial ratio within 0.15 of 1 over 20
    dynamical times, every sample in (0.5, 1.8); N=300: dE/E < 1e-3 over 50
    dynamical times;
  * barnes_hut.rs:141-194 -- Plummer N=500 with theta = 0.5 over 10
    dynamical times: virial deviation < 0.5 and dE/E < 0.01;
  * scaling.rs:14-76 -- Barnes-Hut is O(N log N): the log-log slope of the
    walk's pair count over N = 2,000..20,000 (the reference's sizes) lies
    in the reference's [0.8, 1.8]; those of the device-resident force call
    and of its walk kernel over N = 1,048,576..8,388,608 lie in
    [0.5, 1.8] (the GPU's efficiency still grows with N at these sizes).
"""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def g():
    import gravis_amd
    return gravis_amd


def run_sim(g, gravity, steps, dt, n=200, seed=42):
    p = g.scenarios.plummer(n, seed)
    gravity.compute_accelerations(p)
    g.LeapfrogKDK().run(p, gravity, dt, steps)
    return p


@pytest.mark.parametrize("algo,precision", [("bf", "fp64"), ("bf", "mixed"), ("bh", "fp64"),
                                            ("bh", "mixed")])
def test_deterministic_50_steps(g, algo, precision):
    def solver(ctx):
        if algo == "bf":
            return g.BruteForce(0.05, precision, ctx=ctx)
        return g.BarnesHut(0.05, 0.5, precision=precision, ctx=ctx)
    ctx = g.default_context()
    a = run_sim(g, solver(ctx), 50, 0.001)
    b = run_sim(g, solver(ctx), 50, 0.001)
    other = g.Context(0)
    c = run_sim(g, solver(other), 50, 0.001)
    other.close()
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
        assert np.array_equal(getattr(a, f), getattr(c, f)), f


def test_momentum_and_com_plummer_200(g):
    for seed, what in ((77, "momentum"), (99, "com")):
        n = 200
        dt = g.scenarios.plummer_suggested_dt()
        eps = g.scenarios.plummer_suggested_softening(n)
        p = run_sim(g, g.BruteForce(eps), 1000, dt, n, seed)
        d = g.diagnostics.compute(p, eps, time=1000 * dt, step=1000)
        pm = float(np.linalg.norm(d.momentum))
        com = float(np.linalg.norm(d.center_of_mass))
        say(f"Plummer N=200 seed {seed}: |p| {pm:.2e}, |COM| {com:.2e}")
        assert (pm if what == "momentum" else com) < 1e-12


def two_body_run(g, e, orbits):
    p = g.scenarios.two_body(e)
    T = g.scenarios.two_body_period()
    dt = T / 10_000.0
    eps = 1.0 * (1.0 - e) * 0.001
    bf = g.BruteForce(eps)
    bf.compute_accelerations(p)
    d0 = g.diagnostics.compute(p, eps)
    steps = int(round(orbits * T / dt))
    g.LeapfrogKDK().run(p, bf, dt, steps)
    d1 = g.diagnostics.compute(p, eps, time=steps * dt, step=steps)
    return d0, d1, steps


def test_two_body_angular_momentum_e08(g):
    d0, d1, steps = two_body_run(g, 0.8, 50)
    l0, l1 = np.linalg.norm(d0.angular_momentum), np.linalg.norm(d1.angular_momentum)
    dl = abs(l1 - l0) / l0
    say(f"two-body e=0.8, {steps} steps: dL/L {dl:.2e}")
    assert dl < 1e-8


def test_two_body_momentum(g):
    _, d1, steps = two_body_run(g, 0.5, 10)
    pm = float(np.linalg.norm(d1.momentum))
    say(f"two-body e=0.5, {steps} steps: |p| {pm:.2e}")
    assert pm < 1e-14


def test_plummer_virial_20_tdyn(g):
    n = 500
    dt = g.scenarios.plummer_suggested_dt()
    eps = g.scenarios.plummer_suggested_softening(n)
    p = g.scenarios.plummer(n, 42)
    bf = g.BruteForce(eps)
    bf.compute_accelerations(p)
    every = int(round(1.0 / dt))
    total = int(round(20.0 / dt))
    v = []
    for k in range(total // every):
        g.LeapfrogKDK().run(p, bf, dt, every)
        v.append(g.diagnostics.compute(p, eps, time=(k + 1) * every * dt).virial_ratio)
    v = np.array(v)
    say(f"Plummer N=500 virial over 20 t_dyn: mean {v.mean():.4f}, min {v.min():.3f}, "
        f"max {v.max():.3f}")
    assert abs(v.mean() - 1.0) < 0.15 and (v > 0.5).all() and (v < 1.8).all()


def test_plummer_energy_50_tdyn(g):
    n = 300
    dt = g.scenarios.plummer_suggested_dt()
    eps = g.scenarios.plummer_suggested_softening(n)
    p = g.scenarios.plummer(n, 123)
    bf = g.BruteForce(eps)
    bf.compute_accelerations(p)
    d0 = g.diagnostics.compute(p, eps)
    steps = int(round(50.0 / dt))
    g.LeapfrogKDK().run(p, bf, dt, steps)
    d1 = g.diagnostics.compute(p, eps, time=steps * dt, step=steps)
    de = abs(d1.total_energy - d0.total_energy) / abs(d0.total_energy)
    say(f"Plummer N=300, {steps} steps: dE/E {de:.2e}")
    assert de < 1e-3


def test_plummer_barnes_hut_virial(g):
    n = 500
    dt = g.scenarios.plummer_suggested_dt()
    eps = g.scenarios.plummer_suggested_softening(n)
    p = g.scenarios.plummer(n, 42)
    bh = g.BarnesHut(eps, 0.5)
    bh.compute_accelerations(p)
    d0 = g.diagnostics.compute(p, eps)
    steps = int(round(10.0 / dt))
    dev = 0.0
    for k in range(steps // 100):
        g.LeapfrogKDK().run(p, bh, dt, 100)
        d = g.diagnostics.compute(p, eps, time=(k + 1) * 100 * dt)
        dev = max(dev, abs(d.virial_ratio - 1.0))
    de = abs(d.total_energy - d0.total_energy) / abs(d0.total_energy)
    say(f"Plummer N=500 Barnes-Hut: max virial deviation {dev:.4f}, dE/E {de:.2e}")
    assert dev < 0.5 and de < 0.01


def slope(ns, ys):
    x, y = np.log(np.asarray(ns, float)), np.log(np.asarray(ys, float))
    return float(((x - x.mean()) * (y - y.mean())).sum() / ((x - x.mean()) ** 2).sum())


def test_barnes_hut_scaling_is_nlogn(g):
    ctx = g.default_context()
    ns = [1000, 2000, 5000, 10000, 20000]
    pairs = []
    for n in ns:
        p = g.scenarios.plummer(n, 42)
        g.BarnesHut(0.05, 0.5).compute_accelerations(p)
        pairs.append(ctx.bh_stats()["pairs"])
    s_work = slope(ns[1:], pairs[1:])
    # device-resident force calls (no PCIe copies), the mean of 5 after a
    # warm-up; the walk kernel's own time beside the whole call (sort, tree
    # build and walk)
    import torch
    big = [1048576, 2097152, 4194304, 8388608]
    calls, walks = [], []
    stream = torch.cuda.Stream()
    ctx.set_stream(stream.cuda_stream)
    try:
        with torch.cuda.stream(stream):
            for n in big:
                p = g.scenarios.plummer(n, 42)
                T = [torch.from_numpy(a).cuda() for a in (p.x, p.y, p.z, p.mass)]
                A = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(3)]
                bh = g.BarnesHut(0.05, 0.5, precision="mixed")
                ptrs = [t.data_ptr() for t in T + A]
                bh.compute_accelerations_device(n, *ptrs)
                ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
                ev[0].record()
                for _ in range(5):
                    bh.compute_accelerations_device(n, *ptrs)
                ev[1].record()
                torch.cuda.synchronize()
                calls.append(ev[0].elapsed_time(ev[1]) / 5)
                ctx.set_profiling(True)
                ctx.kernel_time("bh_walk")
                for _ in range(3):
                    bh.compute_accelerations_device(n, *ptrs)
                ctx.set_profiling(False)
                torch.cuda.synchronize()
                ms, k = ctx.kernel_time("bh_walk")
                walks.append(ms / k)
    finally:
        ctx.set_stream(None)
    s_call, s_walk = slope(big, calls), slope(big, walks)
    say(f"BH pair counts {pairs}: slope {s_work:.3f}; N = {big}: force call "
        f"{[round(t, 2) for t in calls]} ms, slope {s_call:.3f}; walk kernel "
        f"{[round(t, 2) for t in walks]} ms, slope {s_walk:.3f}")
    # the reference's bound (scaling.rs:57-62) on the work; the device times
    # keep its upper bound (an N^2 walk would show a slope near 2) but may
    # fall below its lower one: the GPU fills better as N grows (more target
    # groups per CU, a shorter tail), so time grows slower than the work
    # over these sizes (measured: 4M -> 8M took 1.69 times as long)
    assert 0.8 <= s_work <= 1.8
    assert 0.5 <= s_call <= 1.8 and 0.5 <= s_walk <= 1.8
