"""Host-side checks (no GPU): the NBODY snapshot format
(crates/sim-core/src/snapshot.rs:5-157), the cold-collapse generator
against the oracle, and the argument handling of gravis_headless
(crates/headless/src/main.rs:70-107), which exits before any device call.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gravis_amd", "gravis_headless")


@pytest.fixture(scope="module")
def g():
    import gravis_amd
    return gravis_amd


def test_snapshot_round_trip(g, tmp_path):
    p = g.scenarios.evrard(300, seed=3)
    p.vx[:] = np.linspace(-1, 1, 300)
    p.density[:] = np.arange(300) * 0.5
    p.ax[:] = 7.0
    p.pressure[:] = 3.0
    path = str(tmp_path / "snap.bin")
    g.snapshot.write(path, p, 1.25, 17, 0.01, 2.5e-3)
    raw = open(path, "rb").read()
    assert raw[:8] == b"NBODY003"
    assert len(raw) == 8 + 5 * 8 + 300 * (7 * 8 + 1 + 3 * 8)
    q, t, step, eps, dt, ver = g.snapshot.read(path)
    assert (t, step, eps, dt, ver) == (1.25, 17, 0.01, 2.5e-3, 3)
    for f in ("x", "y", "z", "vx", "vy", "vz", "mass", "internal_energy", "smoothing_length",
              "density"):
        assert np.array_equal(getattr(q, f), getattr(p, f)), f
    assert np.array_equal(q.particle_type, p.particle_type)
    # not stored: zero on read (snapshot.rs:110-137)
    for f in ("ax", "ay", "az", "pressure", "du_dt", "sound_speed", "alpha_visc"):
        assert not getattr(q, f).any(), f


def _write_v(path, ver, n, rng):
    a = [rng.normal(size=n) for _ in range(7)]
    head = (b"NBODY00%d" % ver) + struct.pack("<dQddQ", 0.5, 3, 0.02, 1e-3, n)
    body = b"".join(x.astype("<f8").tobytes() for x in a)
    types = None
    if ver >= 2:
        types = rng.integers(0, 5, n).astype(np.uint8)
        body += types.tobytes()
    open(path, "wb").write(head + body)
    return a, types


@pytest.mark.parametrize("ver", [1, 2])
def test_snapshot_reads_older_versions(g, tmp_path, ver):
    rng = np.random.default_rng(ver)
    path = str(tmp_path / f"v{ver}.bin")
    a, types = _write_v(path, ver, 50, rng)
    q, t, step, eps, dt, v = g.snapshot.read(path)
    assert (t, step, eps, dt, v) == (0.5, 3, 0.02, 1e-3, ver)
    for arr, f in zip(a, ("x", "y", "z", "vx", "vy", "vz", "mass")):
        assert np.array_equal(getattr(q, f), arr)
    if ver == 2:
        assert np.array_equal(q.particle_type, types)
    else:
        assert not q.particle_type.any()
    assert not q.internal_energy.any() and not q.smoothing_length.any()


def test_snapshot_rejects_bad_magic(g, tmp_path):
    path = str(tmp_path / "bad.bin")
    open(path, "wb").write(b"NBODY009" + bytes(40))
    with pytest.raises(g.GravisError):
        g.snapshot.read(path)


@pytest.mark.parametrize("gas", [False, True])
def test_cold_collapse_bit_exact_with_oracle(g, gas):
    for n, seed in ((1000, 42), (257, 9)):
        p = g.scenarios.cold_collapse(n, seed=seed, gas=gas)
        s = O.cold_collapse(n, seed=seed, gas=gas)
        for a, b in (("x", "x"), ("y", "y"), ("z", "z"), ("vx", "vx"), ("mass", "m"),
                     ("smoothing_length", "h"), ("internal_energy", "u")):
            assert np.array_equal(getattr(p, a), getattr(s, b)), (gas, n, a)
        assert np.array_equal(p.particle_type, s.type)
    # COM frame (particle.rs:189-226)
    assert abs((p.mass * p.x).sum()) < 1e-15


def _cli(*args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("args,msg", [
    (["-n", "0"], "Error: --particles must be > 0"),
    (["--eccentricity", "1"], "Error: --eccentricity must be in [0, 1), got 1"),
    (["--dt", "-0.5"], "Error: --dt must be positive, got -0.5"),
    (["--softening", "-1e-3"], "Error: --softening must be non-negative, got -0.001"),
    (["--backend", "cpu"], "the CPU backend is sim-core's own headless"),
    (["--scenario", "nope"], "Unknown scenario: nope"),
])
def test_cli_rejects_bad_arguments(args, msg):
    r = _cli(*args)
    assert r.returncode == 1, r.stderr
    assert msg in r.stderr


def test_cli_rejects_unknown_flag_and_algorithm():
    r = _cli("--frobnicate", "1")
    assert r.returncode == 2 and "unexpected argument" in r.stderr
    r = _cli("--algorithm", "fmm")
    assert r.returncode == 2 and "possible values: brute-force, barnes-hut" in r.stderr
