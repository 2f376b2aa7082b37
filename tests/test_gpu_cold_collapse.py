"""Cold collapse on the HIP backend (ColdCollapse, crates/sim-core/src/
scenarios/cold_collapse.rs:56-123), both forms.

Collisionless (20,000 particles, R = 1, M = 1, seed 42, softening 0.3 times
the mean spacing, cold_collapse.rs:115-122):
  * Barnes-Hut theta = 0.5 forces, FP64 and mixed walks, against the
    oracle's f64 brute force on every particle: within the reference's own
    theta = 0.5 error (m3_accuracy.csv: RMS 4.942351e-3, max 3.588391e-2);
  * 100 KDK steps at the suggested dt (t_ff / 200, so half a free-fall
    time) with the FP64 brute force on 2,000 particles: the oracle's same 100
    steps agree to 1e-9 in every position, and energy changes by less than
    1e-4 relative.

With gas (500 particles, the reference's cold_collapse_gas.rs:9-46):
brute-force gravity at the suggested softening, sim-core's SPH; the first dt
is min(compute_timestep, t_ff / 500); 50 steps of step_with_sph with
dt = min(dt_next, 1.5 dt), positions, densities and internal energies
finite at every tenth step (the reference's check), and the GPU run against
the oracle's same 50 steps: every dt within 1e-9 relative, the final state
within 1e-7 of each field's largest magnitude.
"""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

PAIRS = (("x", "x"), ("y", "y"), ("z", "z"), ("vx", "vx"), ("vy", "vy"), ("vz", "vz"),
         ("ax", "ax"), ("ay", "ay"), ("az", "az"), ("mass", "m"), ("density", "rho"),
         ("pressure", "p"), ("internal_energy", "u"), ("du_dt", "dudt"),
         ("smoothing_length", "h"), ("sound_speed", "c"), ("alpha_visc", "alpha"))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def t_ff(radius=1.0, mass=1.0):
    return 0.5 * np.pi * np.sqrt(radius ** 3 / (2.0 * mass))


def softening(n, radius=1.0):
    return 0.3 * np.cbrt(4.0 / 3.0 * np.pi * radius ** 3 / n)


@pytest.fixture(scope="module")
def g():
    import gravis_amd
    return gravis_amd


def test_cold_collapse_barnes_hut_forces(g):
    n = 20000
    p = g.scenarios.cold_collapse(n, seed=42)
    s = O.cold_collapse(n, seed=42)
    assert np.array_equal(p.x, s.x) and np.array_equal(p.z, s.z)
    eps = softening(n)
    st = O.State(n)
    st.x, st.y, st.z, st.m = s.x.copy(), s.y.copy(), s.z.copy(), s.m.copy()
    ref, _ = O.bf_targets(st, eps, np.arange(n), nthreads=16)
    for prec in ("fp64", "mixed"):
        q = p.copy()
        g.BarnesHut(eps, 0.5, precision=prec).compute_accelerations(q)
        rel = np.sqrt(((q.acc() - ref) ** 2).sum(1) / (ref ** 2).sum(1))
        rms = float(np.sqrt((rel ** 2).mean()))
        say(f"cold collapse BH {prec}: rms {rms:.3e} max {rel.max():.3e}")
        assert rms <= 4.942351e-3 and rel.max() <= 3.588391e-2


def test_cold_collapse_kdk_matches_oracle(g):
    n = 2000
    eps, dt, steps = softening(n), t_ff() / 200.0, 100
    p = g.scenarios.cold_collapse(n, seed=7)
    s = O.cold_collapse(n, seed=7)
    st = O.State(n)
    for k in ("x", "y", "z", "vx", "vy", "vz", "m"):
        setattr(st, k, getattr(s, k).copy())
    bf = g.BruteForce(eps)
    bf.compute_accelerations(p)
    O.bf_accel(st, eps)
    d0 = g.diagnostics.compute(p, eps)
    say(f"cold collapse N={n}: {steps} FP64 KDK steps, dt {dt:.4e}")
    g.LeapfrogKDK().run(p, bf, dt, steps)
    O.kdk_steps(st, dt, steps, eps)
    d1 = g.diagnostics.compute(p, eps, time=steps * dt, step=steps)
    dx = max(np.abs(p.x - st.x).max(), np.abs(p.y - st.y).max(), np.abs(p.z - st.z).max())
    de = abs(d1.total_energy - d0.total_energy) / abs(d0.total_energy)
    say(f"  max |x_gpu - x_oracle| {dx:.2e}, |dE/E| {de:.2e}, virial {d1.virial_ratio:.3f}")
    assert dx <= 1e-9 and de <= 1e-4


def test_cold_collapse_gas_50_steps(g):
    import gravis_amd.sph as sph
    n = 500
    p = g.scenarios.cold_collapse(n, seed=42, gas=True)
    s = O.cold_collapse(n, seed=42, gas=True)
    for a, b in PAIRS:
        assert np.array_equal(getattr(p, a), getattr(s, b)), a
    eps = softening(n)
    gravity = g.BruteForce(eps)
    solver = sph.SphSolver()
    prm = O.SphParams.default()
    p.clear_accelerations()
    gravity.compute_accelerations(p)
    solver.compute(p)
    O.sph_bf_accel(s, eps)
    inter = O.sph_compute(s, prm)
    dt = min(O.sph_timestep(s, inter, prm), t_ff() / 500.0)
    dto = dt
    worst_dt = 0.0
    for step in range(50):
        nxt = sph.step_with_sph(p, gravity, solver, dt)
        nxo = O.sph_step(s, dto, gravity=1, eps=eps, params=prm)
        worst_dt = max(worst_dt, abs(nxt - nxo) / nxo)
        dt, dto = min(nxt, 1.5 * dt), min(nxo, 1.5 * dto)
        if step % 10 == 0:
            for f in ("x", "density", "internal_energy"):
                assert np.isfinite(getattr(p, f)).all(), (step, f)
    worst = max(np.abs(getattr(p, a) - getattr(s, b)).max() / max(np.abs(getattr(s, b)).max(),
                                                                  1e-300)
                for a, b in PAIRS)
    say(f"cold collapse gas: 50 steps, worst dt difference {worst_dt:.2e}, worst field "
        f"{worst:.2e}, rho max {p.density.max():.3e}")
    assert worst_dt <= 1e-9 and worst <= 1e-7
