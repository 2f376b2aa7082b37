"""GPU parity of brute-force gravity, KDK and diagnostics through the C ABI
against the CPU oracle (oracle/gravis_oracle.c, pinned by the reference's
fixtures in tests/test_oracle_pins.py).

Bars (written here, per test):
  * FP64 force: per particle |a_gpu - a_ref| / sum_j |a_ij| <= 1e-12 (the
    condition-normalised error: a gather changes the summation order of the
    reference's i<j loop, gravity.rs:55-78), and RMS of |da|/|a| <= 1e-12.
  * Mixed force vs FP64: RMS |da|/|a| <= 1e-6, max <= 1e-4.
  * KDK: same rounding as integrator.rs:41-77 given the same forces.
"""
import csv
import os
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


def to_particles(g, s):
    return g.Particles.from_arrays(s.x, s.y, s.z, s.vx, s.vy, s.vz, s.m)


def rel_errors(ref, got):
    m2 = (ref ** 2).sum(1)
    e2 = ((got - ref) ** 2).sum(1)
    ok = m2 >= 1e-30
    rel2 = e2[ok] / m2[ok]
    return float(np.sqrt(rel2.sum() / len(ref))), float(np.sqrt(rel2.max()))


def cond_errors(s, eps, got, targets=None):
    if targets is None:
        targets = np.arange(s.n)
    ref, absum = O.bf_targets(s, eps, targets)
    d = np.sqrt(((got[targets] - ref) ** 2).sum(1))
    return float((d / absum).max()), ref


@pytest.mark.parametrize("n,eps", [(1000, 0.05), (257, 0.05), (4099, 0.01)])
def test_bf_fp64_matches_oracle(g, n, eps):
    say(f"bf_fp64 n={n} eps={eps}")
    s = O.plummer(n, 42)
    ref = s.copy()
    O.bf_accel(ref, eps)
    p = to_particles(g, s)
    g.BruteForce(eps).compute_accelerations(p)
    rms, mx = rel_errors(ref.acc(), p.acc())
    cmax, _ = cond_errors(s, eps, p.acc())
    say(f"  rms={rms:.3e} max={mx:.3e} cond_max={cmax:.3e}")
    assert cmax <= 1e-12
    assert rms <= 1e-12


def test_bf_accumulates(g):
    say("bf accumulate contract")
    s = O.plummer(500, 3)
    p = to_particles(g, s)
    p.ax[:] = 1.0
    p.ay[:] = -2.0
    p.az[:] = 0.5
    g.BruteForce(0.05).compute_accelerations(p)
    ref = s.copy()
    O.bf_accel(ref, 0.05)
    assert np.abs(p.ax - 1.0 - ref.ax).max() < 1e-12
    assert np.abs(p.ay + 2.0 - ref.ay).max() < 1e-12
    assert np.abs(p.az - 0.5 - ref.az).max() < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_bf_closed_forms(g, precision):
    say(f"bf closed forms {precision}")
    tol = 1e-14 if precision == "fp64" else 1e-6
    p = g.Particles.from_arrays([0.0, 1.0], [0, 0], [0, 0], [0, 0], [0, 0], [0, 0], [1, 1])
    g.BruteForce(0.0, precision).compute_accelerations(p)
    assert abs(p.ax[0] + p.ax[1]) < tol and p.ax[0] > 0 and p.ax[1] < 0
    p = g.Particles.from_arrays([0.0, 2.0], [0, 0], [0, 0], [0, 0], [0, 0], [0, 0], [1, 1])
    g.BruteForce(0.0, precision).compute_accelerations(p)
    assert abs(p.ax[0] - 0.25) < tol * 0.25
    p = g.Particles.from_arrays([0.0, 0.0], [0, 0], [0, 0], [0, 0], [0, 0], [0, 0], [1, 1])
    g.BruteForce(0.1, precision).compute_accelerations(p)
    assert np.isfinite(p.acc()).all() and np.abs(p.acc()).max() == 0.0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_bf_degenerate_sizes(g, precision):
    say(f"bf sizes 0/1 {precision}")
    p = g.Particles.from_arrays([], [], [], [], [], [], [])
    g.BruteForce(0.05, precision).compute_accelerations(p)
    p = g.Particles.from_arrays([0.3], [0.1], [0.2], [0], [0], [0], [1.0])
    g.BruteForce(0.05, precision).compute_accelerations(p)
    assert (p.acc() == 0.0).all()


@pytest.mark.parametrize("n", [4096, 65537])
def test_bf_mixed_error_budget(g, n):
    say(f"bf mixed vs fp64 n={n}")
    s = O.plummer(n, 42)
    p64 = to_particles(g, s)
    g.BruteForce(0.01).compute_accelerations(p64)
    pm = to_particles(g, s)
    g.BruteForce(0.01, "mixed").compute_accelerations(pm)
    rms, mx = rel_errors(p64.acc(), pm.acc())
    say(f"  mixed rms={rms:.3e} max={mx:.3e}")
    assert rms <= 1e-6
    assert mx <= 1e-4


@pytest.mark.parametrize("kernel", ["valu", "matrix", "matrix_bf16"])
def test_c2_mixed_full_size_vs_oracle(g, kernel):
    """The brute-force benchmark workload at full size (Plummer N=262,144,
    seed 42, eps 0.01): mixed forces of both kernels on 512 random targets
    against the oracle's f64 brute force. RMS |da|/|a| <= 1e-6, max <= 1e-4,
    and the condition-normalised error |da| / sum_j |a_ij| within the
    per-pair bound 3u(1 + 1/q) = 195 u ~ 1.2e-5 at q = 1/64: the pair errors
    sum to at most that bound times sum_j |a_ij|. The FP32 tile sums could add
    up to ~127 u more in the worst case; compensated accumulation keeps the
    measured errors orders of magnitude below either."""
    say(f"C2 full size, mixed {kernel} kernel vs oracle, 512 targets")
    n = 262144
    s = O.plummer(n, 42)
    ctx = g.Context(0)
    ctx.set_tuning("bf_matrix", {"matrix": 1, "matrix_bf16": 2}.get(kernel, 0))
    pm = to_particles(g, s)
    g.BruteForce(0.01, "mixed", ctx=ctx).compute_accelerations(pm)
    ctx.close()
    t = np.sort(np.random.default_rng(42).choice(n, 512, replace=False))
    cmax, ref = cond_errors(s, 0.01, pm.acc(), t)
    rms, mx = rel_errors(ref, pm.acc()[t])
    say(f"  rms={rms:.3e} max={mx:.3e} cond_max={cmax:.3e}")
    assert rms <= 1e-6 and mx <= 1e-4
    assert cmax <= 195 * 2.0 ** -24


@pytest.fixture(scope="module")
def bf_ref():
    """FP64 GPU forces of the inputs the kernel-variant test uses (cached)."""
    return {}


@pytest.mark.parametrize("n,eps", [(4099, 0.01), (65537, 0.01), (20011, 0.0), (1500, 0.05),
                                   (64, 0.0), (100, 0.0)])
@pytest.mark.parametrize("ipt,tile", [(1, 128), (2, 128), (1, 256), (2, 256), ("m", 128),
                                      ("m", 256), ("b", 128), ("b", 256)])
def test_bf_mixed_launch_shapes(g, bf_ref, ipt, tile, n, eps):
    """Every mixed launch shape (64 or 128 targets per wavefront, 128 or 256
    records per tile, j split for 8 or 32 wavefronts per CU; ipt "m" / "b"
    is the matrix-core kernel with FP32 / split-bf16 operands) against the
    FP64 path: RMS |da|/|a| <= 1e-6,
    max <= 1e-4. eps = 0 takes the masked self-pair path (positions
    distinct) and the FP64 near tiles; n <= 1500 runs unsorted (below the
    Morton-sort threshold); n = 64, 100 are one wavefront and one ragged
    wavefront of targets."""
    say(f"bf mixed ipt={ipt} tile={tile} n={n} eps={eps}")
    s = O.plummer(n, 7)
    key = (n, eps)
    if key not in bf_ref:
        p64 = to_particles(g, s)
        g.BruteForce(eps).compute_accelerations(p64)
        bf_ref[key] = p64.acc()
    ctx = g.Context(0)
    ctx.set_tuning("bf_matrix", {"m": 1, "b": 2}.get(ipt, 0))
    if ipt not in ("m", "b"):
        ctx.set_tuning("bf_ipt", ipt)
    ctx.set_tuning("bf_tile", tile)
    for wpc in (8, 32):
        ctx.set_tuning("bf_waves_per_cu", wpc)
        pm = to_particles(g, s)
        g.BruteForce(eps, "mixed", ctx=ctx).compute_accelerations(pm)
        rms, mx = rel_errors(bf_ref[key], pm.acc())
        say(f"  waves/CU {wpc}: rms={rms:.3e} max={mx:.3e}")
        assert rms <= 1e-6
        assert mx <= 1e-4
    ctx.close()


@pytest.mark.parametrize("kernel", ["valu", "matrix", "matrix_bf16"])
def test_bf_mixed_unequal_masses(g, kernel):
    """Source masses spread over a factor of three (Plummer positions, masses
    times U(0.5, 1.5)): both mixed kernels against the FP64 path at the
    launch-shape bars (the matrix-core kernel carries m_j in its per-source
    terms (m b, m), the VALU kernel per pair)."""
    say(f"bf mixed unequal masses, {kernel} kernel, n=20011")
    s = O.plummer(20011, 11)
    s.m[:] = s.m * np.random.default_rng(4).uniform(0.5, 1.5, s.n)
    p64 = to_particles(g, s)
    g.BruteForce(0.01).compute_accelerations(p64)
    ctx = g.Context(0)
    ctx.set_tuning("bf_matrix", {"matrix": 1, "matrix_bf16": 2}.get(kernel, 0))
    pm = to_particles(g, s)
    g.BruteForce(0.01, "mixed", ctx=ctx).compute_accelerations(pm)
    rms, mx = rel_errors(p64.acc(), pm.acc())
    say(f"  rms={rms:.3e} max={mx:.3e}")
    assert rms <= 1e-6 and mx <= 1e-4
    ctx.close()


def test_bf_tuning_keys_validate(g):
    ctx = g.Context(0)
    for key, bad in (("bf_ipt", 3), ("bf_waves_per_cu", 0), ("bf_kernel", 1), ("bf_tile", 64),
                     ("bh_group_size", 65), ("bf_near_64ths", 65), ("bf_matrix", 3)):
        with pytest.raises(g.GravisError):
            ctx.set_tuning(key, bad)
    ctx.close()


def test_bf_invalid_args(g):
    say("bf invalid args")
    p = g.Particles.from_arrays([0.0, 1.0], [0, 0], [0, 0], [0, 0], [0, 0], [0, 0], [1, 1])
    with pytest.raises(g.GravisError) as e:
        g.BruteForce(-1.0).compute_accelerations(p)
    assert e.value.status == 1
    with pytest.raises(g.GravisError):
        g.LeapfrogKDK().run(p, g.BruteForce(0.1), 0.0, 1)


def test_kdk_one_step_matches_oracle(g):
    say("kdk one step n=1000")
    s = O.plummer(1000, 42)
    eps, dt = 0.05, 1e-3
    O.bf_accel(s, eps)  # a(t0)
    ref = s.copy()
    O.kdk_steps(ref, dt, 1, eps)
    p = to_particles(g, s)
    p.ax[:], p.ay[:], p.az[:] = s.ax, s.ay, s.az
    g.LeapfrogKDK().run(p, g.BruteForce(eps), dt, 1)
    for f, rf in (("x", ref.x), ("y", ref.y), ("z", ref.z)):
        assert np.abs(getattr(p, f) - rf).max() <= 1e-15 * np.abs(rf).max() * 10
    rms, _ = rel_errors(ref.acc(), p.acc())
    assert rms <= 1e-12


def test_kdk_kepler_fixture(g, golden_dir):
    """Two-body e=0.5 over 10 orbits (m1_kepler_diag.csv): the GPU trajectory
    reproduces the reference's K, U, E at the fixture's printed precision."""
    say("kdk kepler fixture 1e5 steps")
    rows = list(csv.reader(open(os.path.join(golden_dir, "m1_kepler_diag.csv"))))[1:]
    ref = np.array([[float(v) for v in r] for r in rows])
    s = O.two_body(0.5)
    eps = 0.5 * 0.001
    dt = O.two_body_period() / 10000.0
    p = to_particles(g, s)
    gr = g.BruteForce(eps)
    gr.compute_accelerations(p)
    kdk = g.LeapfrogKDK()
    for r in ref:
        step = int(r[0])
        if step > 0:
            kdk.run(p, gr, dt, 10000)
        d = g.diagnostics.compute(p, eps)
        assert abs(d.kinetic_energy - r[2]) <= 1e-9 * abs(r[2]), (step, d.kinetic_energy, r[2])
        assert abs(d.potential_energy - r[3]) <= 1e-9 * abs(r[3])
        assert abs(d.total_energy - r[4]) <= 1e-9 * abs(r[4])


def test_plummer_1k_closed_loop_conservation(g, golden_dir):
    """headless plummer N=1000, 1000 steps (m1_plummer_1k_diag.csv): the GPU
    trajectory conserves energy like the reference's (|dE/E| at every row
    within 2e-6 of the fixture's E) and keeps the virial ratio within 0.01
    of the fixture row (trajectories diverge chaotically; E does not)."""
    say("plummer 1k closed loop 1000 steps")
    rows = list(csv.reader(open(os.path.join(golden_dir, "m1_plummer_1k_diag.csv"))))[1:]
    ref = np.array([[float(v) for v in r] for r in rows])
    s = O.plummer(1000, 42)
    eps = O.plummer_suggested_softening(1000)
    p = to_particles(g, s)
    gr = g.BruteForce(eps)
    gr.compute_accelerations(p)
    kdk = g.LeapfrogKDK()
    for r in ref:
        if int(r[0]) > 0:
            kdk.run(p, gr, 0.01, 100)
        d = g.diagnostics.compute(p, eps)
        assert abs(d.total_energy - r[4]) <= 2e-6 * abs(r[4]), (r[0], d.total_energy, r[4])
        assert abs(d.virial_ratio - r[8]) <= 0.01
        assert np.linalg.norm(d.momentum) < 1e-12


def test_diagnostics_match_oracle(g):
    say("diagnostics n=2000")
    s = O.plummer(2000, 9)
    eps = 0.05
    ref = O.diagnostics(s, eps, True, 1.5, 7)
    p = to_particles(g, s)
    d = g.diagnostics.compute(p, eps, 1.5, 7)
    assert abs(d.kinetic_energy - ref["K"]) <= 1e-13 * abs(ref["K"])
    assert abs(d.potential_energy - ref["U"]) <= 1e-12 * abs(ref["U"])
    assert abs(d.virial_ratio - ref["virial"]) <= 1e-12
    L = np.array(d.angular_momentum)
    Lr = np.array([ref["Lx"], ref["Ly"], ref["Lz"]])
    assert np.abs(L - Lr).max() <= 1e-13 * np.abs(Lr).max()
    assert np.abs(np.array(d.momentum)).max() < 1e-15
    assert d.time == 1.5 and d.step == 7
    f = g.diagnostics.compute_fast(p)
    assert f.potential_energy == 0.0 and f.virial_ratio == 0.0
    assert abs(f.kinetic_energy - ref["K"]) <= 1e-13 * abs(ref["K"])


@pytest.mark.parametrize("n", [1, 5, 4096, 4097, 100003])
def test_morton_sort_bit_exact(g, n):
    say(f"morton sort n={n}")
    s = O.plummer(max(n, 2), 17)
    x, y, z = s.x[:n].copy(), s.y[:n].copy(), s.z[:n].copy()
    keys, perm, box = g.morton_sort(x, y, z)
    rbox = O.root_box(x, y, z)
    assert (box == rbox).all()
    rk, rp = O.sort_keys(O.morton63(x, y, z, rbox))
    assert (keys == rk).all()
    assert (perm == rp).all()


def test_morton_sort_ties_stable(g):
    say("morton sort ties")
    n = 20000
    rng = np.random.default_rng(5)
    x = rng.integers(0, 4, n).astype(np.float64)
    y = rng.integers(0, 4, n).astype(np.float64)
    z = np.zeros(n)
    keys, perm, box = g.morton_sort(x, y, z)
    rk, rp = O.sort_keys(O.morton63(x, y, z, O.root_box(x, y, z)))
    assert (perm == rp).all() and (keys == rk).all()


def test_conservation_momentum_com(g):
    say("conservation n=200 1000 steps")
    for seed, check in ((77, "p"), (99, "com")):
        s = O.plummer(200, seed)
        eps = O.plummer_suggested_softening(200)
        p = to_particles(g, s)
        gr = g.BruteForce(eps)
        gr.compute_accelerations(p)
        g.LeapfrogKDK().run(p, gr, O.plummer_suggested_dt(), 1000)
        d = g.diagnostics.compute(p, eps)
        if check == "p":
            assert np.linalg.norm(d.momentum) < 1e-12
        else:
            assert np.linalg.norm(d.center_of_mass) < 1e-12


def test_autotune_brute_force_and_barnes_hut(g):
    """The autotuner picks shapes from its candidate sets, keeps them in the
    context, and the tuned forces still meet the mixed error bars (brute
    force) and the reference's theta = 0.5 error (Barnes-Hut, m3_accuracy:
    RMS 4.942351e-3)."""
    say("autotune bf mixed / bf fp64 / bh, n=20000")
    s = O.plummer(20000, 5)
    p64 = to_particles(g, s)
    g.BruteForce(0.01).compute_accelerations(p64)
    ref = p64.acc()
    ctx = g.Context(0)
    t = ctx.autotune(g.BruteForce(0.01, "mixed", ctx=ctx), to_particles(g, s))
    say(f"  mixed: {t}")
    assert t["bf_ipt"] in (1, 2) and t["bf_tile"] in (128, 256) and t["bf_matrix"] in (0, 1, 2)
    assert t["bf_waves_per_cu"] in (64, 128, 256, 512, 768, 1024) and t["best_ms"] > 0
    assert ctx.tuning() == {k: v for k, v in t.items() if k != "best_ms"}
    pm = to_particles(g, s)
    g.BruteForce(0.01, "mixed", ctx=ctx).compute_accelerations(pm)
    rms, mx = rel_errors(ref, pm.acc())
    assert rms <= 1e-6 and mx <= 1e-4
    t = ctx.autotune(g.BruteForce(0.01, ctx=ctx), to_particles(g, s))
    say(f"  fp64: {t}")
    assert t["bf_waves_per_cu"] in (32, 64, 128, 256, 512, 768, 1024)
    t = ctx.autotune(g.BarnesHut(0.01, 0.5, ctx=ctx), to_particles(g, s))
    say(f"  bh: {t}")
    assert t["bh_group_size"] in (16, 32, 64)
    pb = to_particles(g, s)
    g.BarnesHut(0.01, 0.5, ctx=ctx).compute_accelerations(pb)
    rms, _ = rel_errors(ref, pb.acc())
    assert rms <= 4.942351e-3
    ctx.close()


def test_c1_plummer_1k_10k_steps(g):
    """Plummer N=1000 (seed 42), eps 0.05, dt 1e-3, 10,000 KDK steps with
    FP64 brute force, against the oracle's sim-core loop (gravity.rs i<j,
    integrator.rs): forces at t0 within 1e-12 (condition-normalised), and
    after 10,000 steps the total energy within 1e-7 relative of the
    oracle's and the virial ratio within 0.01 (the trajectories separate
    chaotically; the energy does not), both drifting less than 1e-4."""
    say("C1: plummer 1000, 10000 steps")
    eps, dt, T = 0.05, 1e-3, 10000
    s = O.plummer(1000, 42)
    ref = s.copy()
    O.bf_accel(ref, eps)
    p = to_particles(g, s)
    gr = g.BruteForce(eps)
    gr.compute_accelerations(p)
    worst, _ = cond_errors(s, eps, p.acc())
    assert worst <= 1e-12
    e0 = O.diagnostics(ref, eps)["E"]
    O.kdk_steps(ref, dt, T, eps)
    g.LeapfrogKDK().run(p, gr, dt, T)
    dr = O.diagnostics(ref, eps)
    dg = g.diagnostics.compute(p, eps)
    say(f"  dE/E gpu {(dg.total_energy - e0) / abs(e0):.3e} oracle {(dr['E'] - e0) / abs(e0):.3e}")
    assert abs(dg.total_energy - dr["E"]) <= 1e-7 * abs(dr["E"])
    assert abs(dg.virial_ratio - dr["virial"]) <= 0.01
    assert abs((dg.total_energy - e0) / e0) < 1e-4
