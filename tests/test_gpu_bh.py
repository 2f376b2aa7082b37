"""GPU parity of Barnes-Hut (linear octree, group walk) through the C ABI.

Bars:
  * Morton keys, sort permutation, tree and walk are deterministic and the
    GPU forces equal the CPU oracle's group walk (oracle/gravis_oracle.c,
    orc_bh_group_accel) BIT FOR BIT for the same theta, leaf and group.
  * Accuracy vs brute force at theta 0.3/0.5/0.7/1.0 on the reference's
    N=1000 seed-42 Plummer set is within the reference's own Barnes-Hut
    error (m3_accuracy.csv), the accuracy bar for BH forces.
  * theta = 0 reduces to brute force (barnes_hut.rs:219-265), |a| = 0.25 for
    two unit masses at distance 2 (barnes_hut.rs:268-283).
  * KDK with BH conserves energy to 1% and keeps the virial ratio within
    0.5 of 1 (tests/barnes_hut.rs:103-194).
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


@pytest.mark.parametrize("n,theta,eps", [(1000, 0.5, 0.05), (1000, 0.3, 0.05), (1000, 1.0, 0.05),
                                         (257, 0.5, 0.05), (20000, 0.5, 0.01), (65, 0.7, 0.0)])
def test_bh_bit_exact_with_oracle_group_walk(g, n, theta, eps):
    say(f"bh bit-exact n={n} theta={theta} eps={eps}")
    s = O.plummer(n, 42)
    ref = s.copy()
    O.bh_group_accel(ref, eps, theta, 16, 64)
    p = to_particles(g, s)
    g.BarnesHut(eps, theta).compute_accelerations(p)
    diff = np.abs(p.acc() - ref.acc()).max()
    say(f"  max |da| = {diff:.3e}")
    assert np.array_equal(p.acc(), ref.acc())


def test_bh_within_reference_error(g, golden_dir):
    rows = list(csv.reader(open(os.path.join(golden_dir, "m3_accuracy.csv"))))[1:]
    s = O.plummer(1000, 42)
    bf = s.copy()
    O.bf_accel(bf, 0.05)
    for r in rows:
        theta, rms_ref, max_ref = float(r[0]), float(r[2]), float(r[3])
        say(f"bh accuracy theta={theta}")
        p = to_particles(g, s)
        g.BarnesHut(0.05, theta).compute_accelerations(p)
        rms, mx = rel_errors(bf.acc(), p.acc())
        say(f"  rms={rms:.3e} (ref {rms_ref:.3e}) max={mx:.3e} (ref {max_ref:.3e})")
        assert rms <= rms_ref and mx <= max_ref


def test_bh_theta_zero_and_pair(g):
    say("bh closed forms")
    pts = [(0.0, 0.0, 0.0, 1.0), (1.0, 0.0, 0.0, 2.0), (0.0, 1.0, 0.0, 1.5),
           (-1.0, -1.0, 0.5, 0.5), (0.5, 0.5, -0.5, 3.0)]
    s = O.State(5)
    for i, (x, y, z, m) in enumerate(pts):
        s.x[i], s.y[i], s.z[i], s.m[i] = x, y, z, m
    bf = s.copy()
    O.bf_accel(bf, 0.05)
    p = to_particles(g, s)
    g.BarnesHut(0.05, 0.0).compute_accelerations(p)
    assert np.abs(p.acc() - bf.acc()).max() < 1e-12
    q = g.Particles.from_arrays([0.0, 2.0], [0, 0], [0, 0], [0, 0], [0, 0], [0, 0], [1, 1])
    g.BarnesHut(0.0, 0.5).compute_accelerations(q)
    assert abs(q.ax[0] - 0.25) < 1e-14 and abs(q.ax[0] + q.ax[1]) < 1e-14


def test_bh_degenerate_inputs(g):
    say("bh n=1, coincident clusters")
    p = g.Particles.from_arrays([0.5], [0.1], [0.2], [0], [0], [0], [1.0])
    g.BarnesHut(0.05, 0.5).compute_accelerations(p)
    assert (p.acc() == 0.0).all()
    # 40 particles in two coincident clumps: level-21 leaves of 20 particles
    s = O.State(40)
    s.x[:20] = 0.25
    s.x[20:] = -0.25
    s.m[:] = 1.0 / 40
    ref = s.copy()
    O.bh_group_accel(ref, 0.05, 0.5)
    p = to_particles(g, s)
    g.BarnesHut(0.05, 0.5).compute_accelerations(p)
    assert np.array_equal(p.acc(), ref.acc())
    assert np.isfinite(p.acc()).all()


def test_bh_leaf_and_group_sizes(g):
    say("bh leaf 1/8/32, group 32")
    s = O.plummer(3000, 8)
    for leaf, grp in ((1, 64), (8, 64), (32, 64), (16, 32)):
        ref = s.copy()
        O.bh_group_accel(ref, 0.05, 0.5, leaf, grp)
        p = to_particles(g, s)
        g.BarnesHut(0.05, 0.5, leaf_size=leaf, group_size=grp).compute_accelerations(p)
        assert np.array_equal(p.acc(), ref.acc()), (leaf, grp)


def test_bh_accumulates_and_is_deterministic(g):
    say("bh accumulate + determinism")
    s = O.plummer(5000, 4)
    a = to_particles(g, s)
    b = to_particles(g, s)
    b.ax[:] = 1.0
    bh = g.BarnesHut(0.05, 0.5)
    bh.compute_accelerations(a)
    bh.compute_accelerations(b)
    assert np.array_equal(b.ax - 1.0, a.ax) or np.abs(b.ax - 1.0 - a.ax).max() < 1e-15
    c = to_particles(g, s)
    bh.compute_accelerations(c)
    assert np.array_equal(a.acc(), c.acc())


def test_bh_kdk_plummer_conservation(g):
    """tests/barnes_hut.rs:103-194: Plummer N=500 seed 42, suggested eps and
    dt, 10 t_dyn with BH theta 0.5."""
    say("bh kdk plummer 1000 steps")
    s = O.plummer(500, 42)
    eps = O.plummer_suggested_softening(500)
    dt = O.plummer_suggested_dt()
    p = to_particles(g, s)
    bh = g.BarnesHut(eps, 0.5)
    bh.compute_accelerations(p)
    d0 = g.diagnostics.compute(p, eps)
    kdk = g.LeapfrogKDK()
    worst = 0.0
    for _ in range(10):
        kdk.run(p, bh, dt, 100)
        d = g.diagnostics.compute(p, eps)
        worst = max(worst, abs(d.virial_ratio - 1.0))
    de = abs((d.total_energy - d0.total_energy) / d0.total_energy)
    say(f"  dE/E={de:.3e} max|virial-1|={worst:.3f}")
    assert de < 0.01
    assert worst < 0.5


def test_bh_large_n_against_brute_force_sample(g):
    """N=200,000 Plummer, theta 0.5: forces on 512 random targets checked
    against the oracle's f64 brute force (RMS within the reference's
    theta=0.5 error) and the Morton permutation checked bit-exact."""
    say("bh n=200000 sample vs brute force")
    n = 200000
    s = O.plummer(n, 42)
    p = to_particles(g, s)
    g.BarnesHut(0.01, 0.5).compute_accelerations(p)
    rng = np.random.default_rng(42)
    t = np.sort(rng.choice(n, 512, replace=False))
    ref, _ = O.bf_targets(s, 0.01, t, nthreads=8)
    rms, mx = rel_errors(ref, p.acc()[t])
    say(f"  rms={rms:.3e} max={mx:.3e}")
    assert rms <= 4.942351e-3 and mx <= 3.588391e-2


def test_bh_two_sphere_4m_parity(g):
    """The Barnes-Hut benchmark workload at full size: two Plummer spheres of
    2,097,152 particles (seeds 42/43) 10a apart. Morton keys and the sort
    permutation are bit-exact with the oracle; forces of the FP64 and the
    mixed walk on 256 random targets are within the reference's theta=0.5
    error of f64 brute force."""
    say("bh two-sphere N=4194304: keys/permutation and force sample")
    ne = 2097152
    p = g.scenarios.plummer_pair(ne, 42, 43, 10.0, 0.0)
    keys, perm, box = g.morton_sort(p.x, p.y, p.z)
    rbox = O.root_box(p.x, p.y, p.z)
    assert (box == rbox).all()
    rk, rp = O.sort_keys(O.morton63(p.x, p.y, p.z, rbox))
    assert (keys == rk).all() and (perm == rp).all()
    g.BarnesHut(0.01, 0.5).compute_accelerations(p)
    s = O.State(p.count)
    s.x, s.y, s.z, s.m = p.x, p.y, p.z, p.mass
    rng = np.random.default_rng(42)
    t = np.sort(rng.choice(p.count, 256, replace=False))
    ref, _ = O.bf_targets(s, 0.01, t, nthreads=16)
    rms, mx = rel_errors(ref, p.acc()[t])
    say(f"  fp64 walk: rms={rms:.3e} max={mx:.3e} nodes={g.default_context().bh_stats()['nodes']}")
    assert rms <= 4.942351e-3 and mx <= 3.588391e-2
    # the mixed walk (the benchmark's) on the same sample
    pm = g.scenarios.plummer_pair(ne, 42, 43, 10.0, 0.0)
    g.BarnesHut(0.01, 0.5, precision="mixed").compute_accelerations(pm)
    rms_m, mx_m = rel_errors(ref, pm.acc()[t])
    say(f"  mixed walk: rms={rms_m:.3e} max={mx_m:.3e}")
    assert rms_m <= 4.942351e-3 and mx_m <= 3.588391e-2


@pytest.mark.parametrize("n,theta", [(20000, 0.5), (3000, 0.3)])
def test_bh_mixed_matches_fp64_walk(g, n, theta):
    """precision='mixed' walks the same interaction lists with FP32 pair
    forces on FP64 group-relative separations: RMS |da|/|a| <= 1e-6 and max
    <= 1e-4 against the FP64 walk, and still within the reference's BH
    error against brute force."""
    say(f"bh mixed vs fp64 n={n} theta={theta}")
    s = O.plummer(n, 42)
    p64 = to_particles(g, s)
    g.BarnesHut(0.01, theta).compute_accelerations(p64)
    pm = to_particles(g, s)
    g.BarnesHut(0.01, theta, precision="mixed").compute_accelerations(pm)
    rms, mx = rel_errors(p64.acc(), pm.acc())
    say(f"  mixed vs fp64 walk rms={rms:.3e} max={mx:.3e}")
    assert rms <= 1e-6 and mx <= 1e-4


@pytest.mark.parametrize("group", [16, 32, 64])
def test_bh_mixed_replicas_and_fallback(g, group):
    """The mixed walk holds groups of <= 32 (<= 16) targets two (four) times
    across the wavefront and splits each source round between the copies;
    with groups capped at 16/32/64 targets and a frontier capacity that
    forces per-target fallback walks, it still matches the FP64 walk of the
    same context (RMS |da|/|a| <= 1e-6, max <= 1e-4)."""
    say(f"bh mixed replicas group={group} + fallback (cap 350)")
    s = O.plummer(5000, 12)
    ctx = g.Context(0)
    ctx.set_tuning("bh_frontier_cap", 350)
    p64 = to_particles(g, s)
    g.BarnesHut(0.05, 0.5, 16, group, ctx=ctx).compute_accelerations(p64)
    pm = to_particles(g, s)
    g.BarnesHut(0.05, 0.5, 16, group, precision="mixed", ctx=ctx).compute_accelerations(pm)
    rms, mx = rel_errors(p64.acc(), pm.acc())
    say(f"  rms={rms:.3e} max={mx:.3e}")
    assert rms <= 1e-6 and mx <= 1e-4
    ctx.close()


def test_bh_frontier_fallback_bit_exact(g):
    """With a small frontier capacity (forced from the host: 350 nodes, above
    the largest single-target frontier of this input, 327, below the largest
    group frontier, 402) some groups fall back to per-target walks; the GPU
    still matches the oracle, which applies the same rule, bit for bit."""
    say("bh frontier fallback (cap 350)")
    s = O.plummer(5000, 12)
    ref = s.copy()
    O.bh_group_accel(ref, 0.05, 0.5, 16, 64, frontier_cap=350)
    ctx = g.Context(0)
    ctx.set_tuning("bh_frontier_cap", 350)
    p = to_particles(g, s)
    g.BarnesHut(0.05, 0.5, ctx=ctx).compute_accelerations(p)
    assert np.array_equal(p.acc(), ref.acc())
    st = ctx.bh_stats()
    normal = s.copy()
    O.bh_group_accel(normal, 0.05, 0.5)
    assert not np.array_equal(normal.acc(), ref.acc())  # the fallback did run
    ctx.close()


def test_repeated_builds_in_one_context(g):
    """One context rebuilding the octree for alternating inputs (reusing its
    cached buffers, with other work between calls) matches a fresh
    context's result bit for bit every time."""
    say("repeated tree builds in one context")
    a = O.plummer(20000, 21)
    b = O.plummer(20000, 22)
    ctx = g.Context(0)

    def forces(s, c):
        p = to_particles(g, s)
        g.BarnesHut(0.01, 0.5, ctx=c).compute_accelerations(p)
        return p.acc()

    ref_a = forces(a, g.Context(0))
    ref_b = forces(b, g.Context(0))
    runs = [(a, ref_a), (a, ref_a), (b, ref_b), (a, ref_a), (b, ref_b), (b, ref_b)]
    for k, (s, ref) in enumerate(runs):
        got = forces(s, ctx)
        assert np.array_equal(got, ref), k
        g.BruteForce(0.01, ctx=ctx).compute_accelerations(to_particles(g, O.plummer(1000, k)))
    ctx.close()
