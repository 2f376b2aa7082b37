"""Pins the CPU oracle (oracle/gravis_oracle.c) before anything is checked
against it.

What pins what:
  * ChaCha20 block function   -> RFC 8439 A.1 known-answer vectors.
  * seed_from_u64 + rand 0.9 float conversions + Plummer sampler + BF + KDK
    + diagnostics             -> the reference's m1_plummer_1k_diag.csv, every
    row of a 1000-step run (artifacts/benchmarks/m1_plummer_1k_diag.csv:1-12),
    which needs the RNG stream, the IC sampler, the force and the integrator
    to be right together.
  * two-body ICs + BF + KDK   -> m1_kepler_diag.csv (no RNG at all).
  * reference octree + BH walk -> m3_accuracy.csv (RMS/max error vs BF for
    theta 0.3/0.5/0.7/1.0, m3_artifacts.rs:88-146).
  * closed forms from the reference's unit tests (gravity.rs:86-134,
    barnes_hut.rs:200-283, octree.rs:338-454).
"""
import csv
import json
import os

import numpy as np
import pytest

import oracle as O


def _csv(path):
    rows = list(csv.reader(open(path)))
    return rows[0], np.array([[float(v) for v in r] for r in rows[1:]])


def _sig(a, b, digits):
    """Equal to `digits` significant digits (the fixture's printed precision)."""
    a, b = float(a), float(b)
    if b == 0.0:
        return abs(a) < 10.0 ** (-digits)
    return abs(a - b) <= 0.5 * 10.0 ** (1 - digits) * abs(b) * 1.0000001


def test_chacha20_rfc8439(golden_dir):
    kat = json.load(open(os.path.join(golden_dir, "chacha20_rfc8439_a1.json")))
    key = np.frombuffer(bytes.fromhex(kat["key_hex"]), dtype="<u4").astype(np.uint32)
    for v in kat["vectors"]:
        ks = O.chacha20_block(key, v["counter"], 0)
        assert ks.astype("<u4").tobytes().hex() == v["keystream_hex"]


def test_chacha20_planted_mistake_is_caught(golden_dir):
    """A one-bit change in the key must change the keystream: the KAT pins
    the key path, not just the constants."""
    kat = json.load(open(os.path.join(golden_dir, "chacha20_rfc8439_a1.json")))
    key = np.zeros(8, dtype=np.uint32)
    key[3] = 1
    ks = O.chacha20_block(key, 0, 0)
    assert ks.astype("<u4").tobytes().hex() != kat["vectors"][0]["keystream_hex"]


def test_rng_range_bounds_and_determinism():
    s1 = O.rng_stream(42, 1000)
    s2 = O.rng_stream(42, 1000)
    assert (s1 == s2).all()
    assert not (O.rng_stream(43, 16) == s1[:16]).all()
    # seed expansion: 8 distinct PCG32 words
    k = O.rng_seed_key(42)
    assert len(set(k.tolist())) == 8


def test_plummer_initial_row_matches_fixture(golden_dir):
    hdr, ref = _csv(os.path.join(golden_dir, "m1_plummer_1k_diag.csv"))
    s = O.plummer(1000, 42)
    eps = O.plummer_suggested_softening(1000)
    d = O.diagnostics(s, eps)
    r0 = ref[0]
    assert _sig(d["K"], r0[2], 11)
    assert _sig(d["U"], r0[3], 11)
    assert _sig(d["E"], r0[4], 11)
    L = np.sqrt(d["Lx"] ** 2 + d["Ly"] ** 2 + d["Lz"] ** 2)
    assert _sig(L, r0[6], 11)
    assert abs(d["virial"] - r0[8]) <= 5e-7
    # momentum / COM drift are rounding-level (1e-17): order of magnitude only
    p = np.sqrt(d["px"] ** 2 + d["py"] ** 2 + d["pz"] ** 2)
    assert p < 1e-16 and r0[5] < 1e-16


def test_plummer_1k_trajectory_matches_fixture(golden_dir):
    """headless --scenario plummer -n 1000 --steps 1000 --diag-interval 100
    (book/src/ch01_gravity.md:225-228): every row, 11 significant digits."""
    hdr, ref = _csv(os.path.join(golden_dir, "m1_plummer_1k_diag.csv"))
    s = O.plummer(1000, 42)
    eps = O.plummer_suggested_softening(1000)
    dt = O.plummer_suggested_dt()
    assert dt == 0.01
    rows = O.run(s, dt, 1000, eps, 100)
    assert rows.shape[0] == ref.shape[0]
    for r, g in zip(rows, ref):
        assert int(r[18]) == int(g[0])
        assert _sig(r[0], g[2], 11), (r[18], r[0], g[2])
        assert _sig(r[1], g[3], 11), (r[18], r[1], g[3])
        assert _sig(r[3], g[4], 11), (r[18], r[3], g[4])
        assert _sig(np.linalg.norm(r[7:10]), g[6], 11)
        assert abs(r[16] - g[8]) <= 5e-7


def test_kepler_fixture(golden_dir):
    hdr, ref = _csv(os.path.join(golden_dir, "m1_kepler_diag.csv"))
    s = O.two_body(0.5)
    dt = O.two_body_period(1.0) / 10000.0
    eps = 1.0 * (1.0 - 0.5) * 0.001
    rows = O.run(s, dt, 100000, eps, 10000)
    assert rows.shape[0] == ref.shape[0]
    for r, g in zip(rows, ref):
        assert _sig(r[0], g[2], 11) and _sig(r[1], g[3], 11) and _sig(r[3], g[4], 11)
        assert _sig(np.linalg.norm(r[7:10]), g[6], 11)


def _rms_max(ref_acc, acc):
    """m3_artifacts.rs:121-139: skip |a|^2 < 1e-30, RMS divides by N."""
    m2 = (ref_acc ** 2).sum(1)
    e2 = ((acc - ref_acc) ** 2).sum(1)
    ok = m2 >= 1e-30
    rel2 = e2[ok] / m2[ok]
    return float(np.sqrt(rel2.sum() / len(ref_acc))), float(np.sqrt(rel2.max()))


def test_reference_bh_accuracy_matches_fixture(golden_dir):
    hdr, ref = _csv(os.path.join(golden_dir, "m3_accuracy.csv"))
    s = O.plummer(1000, 42)
    bf = s.copy()
    O.bf_accel(bf, 0.05)
    for theta, n, rms_ref, max_ref in ref:
        bh = s.copy()
        O.bh_accel(bh, 0.05, theta)
        rms, mx = _rms_max(bf.acc(), bh.acc())
        assert _sig(rms, rms_ref, 7), (theta, rms, rms_ref)
        assert _sig(mx, max_ref, 7), (theta, mx, max_ref)


# --- closed forms from the reference unit tests ---------------------------

def _pair(x1, eps):
    s = O.State(2)
    s.x[:] = [0.0, x1]
    s.m[:] = [1.0, 1.0]
    O.bf_accel(s, eps)
    return s


def test_bf_closed_forms():
    s = _pair(1.0, 0.0)
    assert abs(s.ax[0] + s.ax[1]) < 1e-14 and s.ax[0] > 0 and s.ax[1] < 0
    s = _pair(2.0, 0.0)
    assert abs(s.ax[0] - 0.25) < 1e-14
    s = _pair(0.0, 0.1)
    assert np.isfinite(s.ax).all()


def _five_bodies():
    pts = [(0.0, 0.0, 0.0, 1.0), (1.0, 0.0, 0.0, 2.0), (0.0, 1.0, 0.0, 1.5),
           (-1.0, -1.0, 0.5, 0.5), (0.5, 0.5, -0.5, 3.0)]
    s = O.State(5)
    for i, (x, y, z, m) in enumerate(pts):
        s.x[i], s.y[i], s.z[i], s.m[i] = x, y, z, m
    return s


def test_bh_theta_zero_matches_bf():
    s = _five_bodies()
    bf = s.copy()
    O.bf_accel(bf, 0.05)
    bh = s.copy()
    O.bh_accel(bh, 0.05, 0.0)
    assert np.abs(bh.acc() - bf.acc()).max() < 1e-12
    g = s.copy()
    O.bh_group_accel(g, 0.05, 0.0)
    assert np.abs(g.acc() - bf.acc()).max() < 1e-12


def test_bh_pair_magnitude():
    s = O.State(2)
    s.x[:] = [0.0, 2.0]
    s.m[:] = [1.0, 1.0]
    O.bh_accel(s, 0.0, 0.5)
    assert abs(s.ax[0] - 0.25) < 1e-14
    g = O.State(2)
    g.x[:] = [0.0, 2.0]
    g.m[:] = [1.0, 1.0]
    O.bh_group_accel(g, 0.0, 0.5)
    assert abs(g.ax[0] - 0.25) < 1e-14 and abs(g.ax[0] + g.ax[1]) < 1e-14


def test_reference_octree_com():
    s = O.State(2)
    s.x[:] = [-1.0, 1.0]
    s.m[:] = [1.0, 1.0]
    root = np.zeros(4)
    nn = O.lib().orc_ref_octree_summary(2, s.x, s.y, s.z, s.m, root)
    assert nn >= 3
    assert abs(root[0] - 2.0) < 1e-14 and np.abs(root[1:]).max() < 1e-14


# --- conservation invariants (crates/sim-core/tests/conservation.rs) -------

def test_momentum_and_com_conservation_plummer():
    for seed, check in ((77, "p"), (99, "com")):
        s = O.plummer(200, seed)
        eps = O.plummer_suggested_softening(200)
        rows = O.run(s, O.plummer_suggested_dt(), 1000, eps, 1000)
        last = rows[-1]
        if check == "p":
            assert np.linalg.norm(last[4:7]) < 1e-12
        else:
            assert np.linalg.norm(last[10:13]) < 1e-12


def test_two_body_angular_momentum():
    s = O.two_body(0.8)
    dt = O.two_body_period() / 10000.0
    eps = 1.0 * (1 - 0.8) * 0.001
    steps = int(round(50 * O.two_body_period() / dt))
    rows = O.run(s, dt, steps, eps, steps)
    L0 = np.linalg.norm(rows[0][7:10])
    L1 = np.linalg.norm(rows[-1][7:10])
    assert abs((L1 - L0) / L0) < 1e-8
    assert np.linalg.norm(rows[-1][4:7]) < 1e-14
