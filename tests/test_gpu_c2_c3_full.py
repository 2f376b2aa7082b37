"""The brute-force and Barnes-Hut workloads at their configured sizes.

Mixed brute force: Plummer N = 262,144 (seed 42), eps 0.01, dt 1e-4, 100 KDK
steps with the launch shape the autotuner picks.
  * at step 0 and after the 100 steps, mixed forces on 4,096 random targets
    against the oracle's f64 brute force: RMS |da|/|a| <= 1e-6, max <= 1e-4,
    condition-normalised |da| / sum_j |a_ij| <= 195 u (the per-pair bound);
  * at every one of the 100 steps, the mixed forces of that step against
    the FP64 GPU forces on the same positions: RMS <= 1e-6, max <= 1e-4;
  * energy changes by less than 1e-6 relative over the 100 steps, and the
    mixed trajectory stays within 1e-9 (RMS position) of the FP64 one.

Barnes-Hut: two Plummer spheres of 2,097,152 particles (seeds 42/43) 10a
apart on a parabolic approach, eps 0.01, theta 0.5, leaf 16, dt 5e-4.
  * forces of the mixed walk (64-target groups, the benchmark's) and of the
    FP64 walk on 16,384 random targets against the oracle's f64 brute force:
    within the reference's theta = 0.5 error (m3_accuracy.csv: RMS
    4.942351e-3, max 3.588391e-2);
  * 50 KDK steps with the mixed walk: energy changes by less than 1e-4
    relative, and the total momentum stays below 1e-6 of sum m |v| (the walk
    is not pairwise symmetric, barnes_hut.rs:114-135, so momentum is not
    exact), the centre of mass moves by less than 1e-6.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

U = 2.0 ** -24


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_errors(ref, got):
    m2 = (ref ** 2).sum(1)
    e2 = ((got - ref) ** 2).sum(1)
    ok = m2 >= 1e-30
    rel2 = e2[ok] / m2[ok]
    return float(np.sqrt(rel2.sum() / len(ref))), float(np.sqrt(rel2.max()))


def oracle_state(x, y, z, m):
    s = O.State(len(x))
    s.x, s.y, s.z, s.m = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, m))
    return s


def test_c2_mixed_100_steps():
    import torch
    import gravis_amd as g
    n, eps, dt, steps = 262144, 0.01, 1e-4, 100
    ctx = g.Context(0)
    # the library's kernels and the torch checks below share one stream
    stream = torch.cuda.Stream()
    ctx.set_stream(stream.cuda_stream)
    with torch.cuda.stream(stream):
        c2_steps(g, ctx, torch, n, eps, dt, steps)
    ctx.close()


def c2_steps(g, ctx, torch, n, eps, dt, steps):
    p = g.scenarios.plummer(n, 42)
    mixed = g.BruteForce(eps, "mixed", ctx=ctx)
    fp64 = g.BruteForce(eps, "fp64", ctx=ctx)
    tuning = ctx.autotune(mixed, p)
    say(f"C2: autotuned shape {tuning}")
    dev = torch.device("cuda:0")
    keys = ("x", "y", "z", "vx", "vy", "vz", "mass", "ax", "ay", "az")
    Tm = {k: torch.from_numpy(getattr(p, k).copy()).to(dev) for k in keys}
    T6 = {k: v.clone() for k, v in Tm.items()}
    L = ctx._L
    ptr = lambda t: C.c_void_p(t.data_ptr())

    def accel(solver, S, out):
        d = solver.desc()
        ctx.check(L.gravis_accel_device(ctx.handle, C.byref(d), n,
                                        *[ptr(S[k]) for k in ("x", "y", "z", "mass")],
                                        *[ptr(o) for o in out]))

    def kdk(solver, S, k):
        d = solver.desc()
        ctx.check(L.gravis_kdk_device(ctx.handle, C.byref(d), n, *[ptr(S[q]) for q in keys],
                                      dt, k))

    tg = np.sort(np.random.default_rng(42).choice(n, 4096, replace=False))

    def vs_oracle(S, label):
        s = oracle_state(*[S[k].cpu().numpy() for k in ("x", "y", "z", "mass")])
        say(f"C2 {label}: oracle brute force on 4096 targets")
        ref, absum = O.bf_targets(s, eps, tg, nthreads=16)
        got = np.stack([S[k].cpu().numpy()[tg] for k in ("ax", "ay", "az")], 1)
        rms, mx = rel_errors(ref, got)
        cond = float((np.sqrt(((got - ref) ** 2).sum(1)) / absum).max())
        say(f"C2 {label}: rms {rms:.3e} max {mx:.3e} cond {cond:.3e}")
        assert rms <= 1e-6 and mx <= 1e-4 and cond <= 195 * U

    accel(mixed, Tm, [Tm["ax"], Tm["ay"], Tm["az"]])
    accel(fp64, T6, [T6["ax"], T6["ay"], T6["az"]])
    torch.cuda.synchronize()
    vs_oracle(Tm, "step 0")
    diag0 = g.diagnostics.compute(p, eps, ctx=ctx)
    a6 = [torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(3)]
    worst_rms = worst_max = 0.0
    for k in range(steps):
        kdk(mixed, Tm, 1)
        kdk(fp64, T6, 1)
        for a in a6:
            a.zero_()
        accel(fp64, Tm, a6)
        A6 = torch.stack(a6, 1)
        AM = torch.stack([Tm["ax"], Tm["ay"], Tm["az"]], 1)
        rel = torch.sqrt(((AM - A6) ** 2).sum(1) / (A6 ** 2).sum(1))
        rms, mx = float(torch.sqrt((rel ** 2).mean())), float(rel.max())
        worst_rms, worst_max = max(worst_rms, rms), max(worst_max, mx)
        assert rms <= 1e-6 and mx <= 1e-4, (k, rms, mx)
    say(f"C2: per-step mixed vs FP64 over {steps} steps: worst RMS {worst_rms:.3e}, "
        f"worst max {worst_max:.3e}")
    torch.cuda.synchronize()
    vs_oracle(Tm, f"after {steps} steps")
    q = g.Particles.from_arrays(*[Tm[k].cpu().numpy() for k in ("x", "y", "z", "vx", "vy",
                                                                   "vz", "mass")])
    diag1 = g.diagnostics.compute(q, eps, ctx=ctx)
    de = abs(diag1.total_energy - diag0.total_energy) / abs(diag0.total_energy)
    dx = float(torch.sqrt(sum((Tm[k] - T6[k]) ** 2 for k in ("x", "y", "z")).mean()))
    say(f"C2: |dE/E| over {steps} mixed steps {de:.2e}; RMS |x_mixed - x_fp64| {dx:.2e}")
    assert de <= 1e-6 and dx <= 1e-9


@pytest.fixture(scope="module")
def C3():
    import gravis_amd as g
    ne = 2097152
    p = g.scenarios.plummer_pair(ne, 42, 43, 10.0, 0.0)
    tg = np.sort(np.random.default_rng(4242).choice(p.count, 16384, replace=False))
    s = oracle_state(p.x, p.y, p.z, p.mass)
    say("C3: oracle brute force on 16384 targets of 4,194,304")
    ref, _ = O.bf_targets(s, 0.01, tg, nthreads=16)
    return dict(g=g, p=p, tg=tg, ref=ref)


@pytest.mark.parametrize("precision,group", [("mixed", 64), ("fp64", 32)])
def test_c3_forces_16384_targets(C3, precision, group):
    g, p = C3["g"], C3["p"].copy()
    p.clear_accelerations()
    g.BarnesHut(0.01, 0.5, 16, group, precision=precision).compute_accelerations(p)
    rms, mx = rel_errors(C3["ref"], p.acc()[C3["tg"]])
    say(f"C3 {precision} walk (group {group}): rms {rms:.3e} max {mx:.3e}")
    assert rms <= 4.942351e-3 and mx <= 3.588391e-2


def test_c3_50_steps(C3):
    g, p = C3["g"], C3["p"].copy()
    ctx = g.default_context()
    bh = g.BarnesHut(0.01, 0.5, 16, 64, precision="mixed")
    p.clear_accelerations()
    bh.compute_accelerations(p)
    d0 = g.diagnostics.compute(p, 0.01, ctx=ctx)
    say(f"C3: E0 {d0.total_energy:.10e}; 50 KDK steps (mixed walk)")
    g.LeapfrogKDK().run(p, bh, 5e-4, 50)
    d1 = g.diagnostics.compute(p, 0.01, ctx=ctx, time=50 * 5e-4, step=50)
    de = abs(d1.total_energy - d0.total_energy) / abs(d0.total_energy)
    pmag = float(np.sqrt(sum(c ** 2 for c in d1.momentum)))
    pscale = float((p.mass * np.sqrt(p.vx ** 2 + p.vy ** 2 + p.vz ** 2)).sum())
    com = float(np.sqrt(sum(c ** 2 for c in d1.center_of_mass)))
    say(f"C3: E1 {d1.total_energy:.10e}, |dE/E| {de:.2e}, |p| / sum m|v| {pmag / pscale:.2e}, "
        f"|COM| {com:.2e}")
    assert np.isfinite(p.x).all()
    assert de <= 1e-4 and pmag <= 1e-6 * pscale and com <= 1e-6
