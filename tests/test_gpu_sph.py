"""GPU parity of the SPH path (gravis_sph_compute / _step / _run) against the
CPU oracle of sim-core's SPH (oracle/gravis_oracle_sph.c, pinned in
tests/test_oracle_sph.py).

Both sides run the same FP64 operations in the reference's order; the lists
are the same sets of particles, gathered in Morton order on the GPU and in
the reference tree's octant order in the oracle, so sums differ only by
rounding order. Bars (measured values are printed):
  * neighbour-list lengths identical;
  * rho, h, omega, P, c within 1e-13 relative (measured <= 2e-14);
  * a, du/dt, div_v, dalpha_dt within 1e-12 of the largest magnitude of the
    field (pair sums with cancellation), max_mu within 1e-13 relative
    (measured <= 2e-15);
  * steps: state within 1e-12 after 3 steps (measured <= 4e-15), next dt
    within 1e-12.
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


@pytest.fixture(scope="module")
def G():
    import gravis_amd as g
    import gravis_amd.sph as sph
    return g, sph


def to_gpu_particles(g, s):
    p = g.Particles.zeros(s.n)
    for a, b in PAIRS:
        getattr(p, a)[:] = getattr(s, b)
    p.particle_type[:] = s.type
    return p


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def compare_compute(p, s, gi, oi, label):
    nn_ok = np.array_equal(gi.nneigh, oi["nneigh"])
    errs = {}
    for a, b in (("density", "rho"), ("smoothing_length", "h"), ("pressure", "p"),
                 ("sound_speed", "c")):
        m = s.type == 4
        x, y = getattr(p, a)[m], getattr(s, b)[m]
        errs[b] = (np.abs(x - y) / np.maximum(np.abs(y), 1e-300)).max()
    errs["omega"] = (np.abs(gi.omega - oi["omega"]) / np.abs(oi["omega"])).max()
    for a, b in (("ax", "ax"), ("ay", "ay"), ("az", "az"), ("du_dt", "dudt")):
        errs[b] = rel(getattr(p, a), getattr(s, b))
    errs["div_v"] = rel(gi.div_v, oi["div_v"])
    errs["dalpha"] = rel(gi.dalpha_dt, oi["dalpha_dt"])
    mm = oi["max_mu"] > 0
    errs["max_mu"] = (np.abs(gi.max_mu - oi["max_mu"])[mm] / oi["max_mu"][mm]).max() if mm.any() \
        else np.abs(gi.max_mu).max()
    say(f"  {label}: lists equal {nn_ok}, "
        + ", ".join(f"{k} {v:.1e}" for k, v in errs.items()))
    assert nn_ok, np.nonzero(gi.nneigh != oi["nneigh"])[0][:10]
    for k in ("rho", "h", "p", "c", "omega", "max_mu"):
        assert errs[k] <= 1e-13, (k, errs[k])
    for k in ("ax", "ay", "az", "dudt", "div_v", "dalpha"):
        assert errs[k] <= 1e-12, (k, errs[k])


def run_compute(G, s, **knobs):
    g, sph = G
    p = to_gpu_particles(g, s)
    solver = sph.SphSolver()
    for k, v in knobs.items():
        setattr(solver, k, v)
    gi = solver.compute(p, hooks=True)
    stats = solver.stats()
    prm = O.SphParams.default()
    for k, v in knobs.items():
        setattr(prm, k, v)
    oi = O.sph_compute(s, prm)
    return p, gi, oi, stats


def test_sph_compute_evrard(G):
    n = 4000
    say(f"sph compute: Evrard N={n}")
    s = O.evrard(n, seed=5)
    rng = np.random.default_rng(1)
    s.vx[:], s.vy[:], s.vz[:] = rng.normal(size=(3, n)) * 0.2  # viscosity active
    s.u[:] *= 1.0 + rng.random(n)                               # conductivity active
    p, gi, oi, st = run_compute(G, s)
    say(f"  stats {st}")
    compare_compute(p, s, gi, oi, "evrard")


def test_sph_compute_mixed_types_and_requeries(G):
    """Gas with h far too small (the iteration grows h past the padded radius,
    so lists are queried again) next to non-gas particles, which appear in
    gas lists but get no list of their own (neighbors.rs:62-78)."""
    rng = np.random.default_rng(2)
    n = 3000
    s = O.SphState(n)
    s.x[:], s.y[:], s.z[:] = rng.normal(size=(3, n))
    s.vx[:], s.vy[:], s.vz[:] = rng.normal(size=(3, n)) * 0.1
    s.m[:] = 1.0 / n
    gas = rng.random(n) < 0.7
    s.type[:] = np.where(gas, 4, 3)
    s.u[:] = np.where(gas, 0.05, 0.0)
    s.h[:] = np.where(gas, 0.02, 0.0)
    s.alpha[:] = np.where(gas, 1.0, 0.0)
    say("sph compute: mixed types, h0 too small")
    p, gi, oi, st = run_compute(G, s)
    say(f"  stats {st}")
    assert st["requeries"] > 0 and st["rounds"] > 1
    compare_compute(p, s, gi, oi, "mixed")
    assert np.all(gi.nneigh[~gas] == 0) and np.all(gi.div_v[~gas] == 0)


@pytest.mark.parametrize("leaves", [4, 0])
def test_sph_compute_depth_first_fallback(G, leaves):
    """A leaf capacity of 4 per wave sends nearly every group to the
    depth-first search, 0 sends all of them; results must not change."""
    g, sph = G
    n = 3000
    s = O.evrard(n, seed=9)
    say(f"sph compute: depth-first fallback (sph_query_leaves = {leaves})")
    ctx = g.default_context()
    ctx.set_tuning("sph_query_leaves", leaves)
    try:
        p, gi, oi, st = run_compute(G, s)
    finally:
        ctx.set_tuning("sph_query_leaves", 512)
    say(f"  stats {st}")
    assert st["dfs_groups"] > 0
    compare_compute(p, s, gi, oi, "dfs")
    p2 = to_gpu_particles(g, O.evrard(n, seed=9))
    gi2 = sph.SphSolver().compute(p2, hooks=True)
    assert np.array_equal(gi2.nneigh, gi.nneigh)
    assert np.array_equal(p2.density, p.density) and np.array_equal(p2.ax, p.ax)


def test_sph_compute_h_max_and_params(G):
    """h_max clamp (density.rs:71-75), no conductivity, other viscosity."""
    s = O.evrard(2000, seed=3)
    say("sph compute: h_max / alpha_cond = 0 / beta 1.5")
    p, gi, oi, st = run_compute(G, s, h_max=0.12, alpha_cond=0.0, beta_visc=1.5)
    assert np.all(p.smoothing_length <= 0.12)
    compare_compute(p, s, gi, oi, "params")


def compare_state(p, s, tol, label):
    worst = 0.0
    for a, b in PAIRS:
        e = rel(getattr(p, a), getattr(s, b))
        worst = max(worst, e)
        assert e <= tol, (label, a, e)
    say(f"  {label}: worst field error {worst:.1e}")


@pytest.mark.parametrize("gravity", ["none", "bf", "bh"])
def test_sph_step_matches_oracle(G, gravity):
    g, sph = G
    n = 2000
    s = O.evrard(n, seed=4)
    eps = O.evrard_softening(n)
    say(f"sph step: Evrard N={n}, gravity {gravity}")
    p = to_gpu_particles(g, s)
    solver = sph.SphSolver()
    grav = {"none": None, "bf": g.BruteForce(eps), "bh": g.BarnesHut(eps, 0.5)}[gravity]
    code = {"none": 0, "bf": 1, "bh": 3}[gravity]
    # initial forces as headless does (main.rs:151-154)
    if grav is not None:
        grav.compute_accelerations(p)
        if gravity == "bf":
            O.sph_bf_accel(s, eps)
        else:
            st = O.State(n)
            for f in ("x", "y", "z", "m"):
                getattr(st, f)[:] = getattr(s, f)
            O.bh_group_accel(st, eps, 0.5)
            s.ax[:], s.ay[:], s.az[:] = st.ax, st.ay, st.az
    solver.compute(p)
    O.sph_compute(s)
    compare_state(p, s, 1e-12, "initial")
    dt = 1e-3
    for k in range(3):
        dg = sph.step_with_sph(p, grav, solver, dt)
        do = O.sph_step(s, dt, gravity=code, eps=eps, theta=0.5)
        say(f"  step {k}: dt_next gpu {dg:.9e} oracle {do:.9e}")
        assert abs(dg - do) <= 1e-12 * do
        compare_state(p, s, 1e-12, f"step {k}")
        dt = do


def test_sph_run_follows_the_headless_dt_rule(G):
    g, sph = G
    n = 1500
    s = O.evrard(n, seed=6)
    eps = O.evrard_softening(n)
    say(f"sph run: Evrard N={n}, 8 steps, brute force")
    p = to_gpu_particles(g, s)
    solver = sph.SphSolver()
    grav = g.BruteForce(eps)
    grav.compute_accelerations(p)
    solver.compute(p)
    O.sph_bf_accel(s, eps)
    O.sph_compute(s)
    dt0 = 2e-3
    dt_g, t_g = sph.run(p, grav, solver, dt0, 8)
    dt, t = dt0, 0.0
    for _ in range(8):
        nxt = O.sph_step(s, dt, gravity=1, eps=eps)
        t += dt
        dt = min(nxt, 1.5 * dt)
    say(f"  dt gpu {dt_g:.9e} oracle {dt:.9e}; t gpu {t_g:.9e} oracle {t:.9e}")
    assert abs(dt_g - dt) <= 1e-12 * dt and abs(t_g - t) <= 1e-12 * t
    compare_state(p, s, 1e-12, "run")


def test_sph_evrard_conservation_on_gpu(G):
    """tests/evrard_collapse.rs:10-76 on the GPU path (N = 5000, BH gravity)."""
    g, sph = G
    n = 5000
    p = g.scenarios.evrard(n)
    eps = g.scenarios.evrard_softening(n)
    grav = g.BarnesHut(eps, 0.5)
    solver = sph.SphSolver()
    say(f"sph Evrard conservation N={n}, BH")
    grav.compute_accelerations(p)
    inter = solver.compute(p)
    s0 = O.evrard(n)  # oracle timestep on the same state
    for a, b in PAIRS:
        getattr(s0, b)[:] = getattr(p, a)
    dt = min(O.sph_timestep(s0, {"div_v": inter.div_v, "max_mu": inter.max_mu}), 1e-3)
    d0 = g.diagnostics.compute(p, eps)
    t, step = 0.0, 0
    while t < 0.5:
        if t + dt > 0.5:
            dt = 0.5 - t
        nxt = sph.step_with_sph(p, grav, solver, dt)
        t += dt
        dt = min(nxt, 2.0 * dt)
        step += 1
        assert step < 100_000
    d1 = g.diagnostics.compute(p, eps, t, step)
    de = abs(d1.total_energy - d0.total_energy) / abs(d0.total_energy)
    say(f"  {step} steps, dE/E = {de:.3e}, K {d0.kinetic_energy:.3e} -> {d1.kinetic_energy:.3e}, "
        f"Eth {d0.thermal_energy:.3e} -> {d1.thermal_energy:.3e}")
    assert de < 0.05
    assert d1.kinetic_energy > 2.0 * d0.kinetic_energy
    assert d1.thermal_energy > d0.thermal_energy
