"""gravis_headless on the GPU: replays of the reference's own headless
artifacts and snapshot restarts.

  * m1_plummer_1k_diag.csv (headless --scenario plummer -n 1000 --steps 1000
    --diag-interval 100, book/src/ch01_gravity.md:225-228) and
    m1_kepler_diag.csv (--scenario kepler --steps 100000 --diag-interval
    10000): step and time columns identical as text, energies and angular
    momentum within 1e-9 relative (the GPU sums forces in another order than
    sim-core's i<j loop, and 1000 chaotic steps grow 1e-16 to ~1e-12),
    virial within 1e-6. The fixtures predate the thermal column; columns are
    matched by name.
  * Restart: 200 steps straight equal 100 steps + snapshot + restart + 100
    steps, digit for digit in the CSV.
"""
import csv
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gravis_amd", "gravis_headless")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def run(*args, timeout=300):
    print("gravis_headless", " ".join(args), flush=True)
    r = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr[-2000:]
    return r


def rows(path):
    with open(path) as f:
        return list(csv.DictReader(f))


def compare(ours, ref):
    assert len(ours) == len(ref)
    worst = 0.0
    for a, b in zip(ours, ref):
        assert a["step"] == b["step"]
        assert a["time"] == b["time"]
        for k in ("kinetic_energy", "potential_energy", "total_energy", "angular_momentum_mag"):
            x, y = float(a[k]), float(b[k])
            e = abs(x - y) / abs(y)
            worst = max(worst, e)
            assert e <= 1e-9, (a["step"], k, x, y)
        assert abs(float(a["virial_ratio"]) - float(b["virial_ratio"])) <= 1e-6
    return worst


def test_plummer_fixture_replay(tmp_path):
    out = str(tmp_path / "p.csv")
    r = run("--scenario", "plummer", "-n", "1000", "--steps", "1000", "--diag-interval", "100",
            "--diag-csv", out)
    worst = compare(rows(out), rows(os.path.join(GOLDEN, "m1_plummer_1k_diag.csv")))
    print(f"  worst relative difference {worst:.2e}", flush=True)
    err = r.stderr.splitlines()
    assert err[0] == "Scenario:   Plummer Sphere"
    assert err[1] == "Particles:  1000"
    assert err[2] == "Algorithm:  brute-force"
    assert err[4] == "dt:         1.000000e-2"
    assert any(l.startswith("Initial:    E=-1.4806687137e-1  2K/|U|=0.9949") for l in err)
    assert any(l.startswith("Step     1000  t=1.0000e1  E=") for l in err)
    assert "=== Final ===" in err


def test_kepler_fixture_replay(tmp_path):
    out = str(tmp_path / "k.csv")
    run("--scenario", "kepler", "--steps", "100000", "--diag-interval", "10000",
        "--diag-csv", out)
    worst = compare(rows(out), rows(os.path.join(GOLDEN, "m1_kepler_diag.csv")))
    print(f"  worst relative difference {worst:.2e}", flush=True)


@pytest.mark.parametrize("algo", ["brute-force", "barnes-hut"])
def test_snapshot_restart_is_seamless(tmp_path, algo):
    a = str(tmp_path / "a.csv")
    run("--scenario", "plummer", "-n", "3000", "--steps", "200", "--diag-interval", "50",
        "--algorithm", algo, "--diag-csv", a)
    snaps = str(tmp_path / "snaps")
    b = str(tmp_path / "b.csv")
    run("--scenario", "plummer", "-n", "3000", "--steps", "100", "--diag-interval", "50",
        "--algorithm", algo, "--snapshot-interval", "100", "--snapshot-dir", snaps)
    snap = os.path.join(snaps, "snap_00000100.bin")
    assert os.path.exists(snap)
    run("--restart", snap, "--steps", "100", "--diag-interval", "50", "--algorithm", algo,
        "--diag-csv", b)
    ra, rb = rows(a), rows(b)
    by_step = {r["step"]: r for r in ra}
    assert [r["step"] for r in rb] == ["100", "150", "200"]
    for r in rb:
        assert r == by_step[r["step"]], r["step"]


def test_evrard_collapse_through_the_cli(tmp_path):
    out = str(tmp_path / "e.csv")
    r = run("--scenario", "evrard-collapse", "-n", "3000", "--steps", "40", "--diag-interval",
            "10", "--algorithm", "barnes-hut", "--diag-csv", out)
    rs = rows(out)
    assert [x["step"] for x in rs] == ["0", "10", "20", "30", "40"]
    e0, e1 = float(rs[0]["total_energy"]), float(rs[-1]["total_energy"])
    assert abs(e1 - e0) / abs(e0) < 0.05
    assert float(rs[0]["thermal_energy"]) == pytest.approx(0.05, rel=1e-12)
    assert float(rs[-1]["kinetic_energy"]) > 0.0
    assert "Scenario:   Evrard Collapse" in r.stderr


def test_galaxy_collision_through_the_cli(tmp_path):
    out = str(tmp_path / "g.csv")
    r = run("--scenario", "galaxy-collision", "-n", "4001", "--steps", "50", "--diag-interval",
            "25", "--algorithm", "barnes-hut", "--diag-csv", out)
    err = r.stderr.splitlines()
    assert err[0] == "Scenario:   Galaxy Collision"
    assert err[1] == "Particles:  4000"
    rs = rows(out)
    assert [x["step"] for x in rs] == ["0", "25", "50"]
    e0, e1 = float(rs[0]["total_energy"]), float(rs[-1]["total_energy"])
    assert abs(e1 - e0) / abs(e0) < 1e-3
