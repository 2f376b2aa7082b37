"""Pins the CPU oracle of sim-core's SPH (oracle/gravis_oracle_sph.c).

No SPH fixture exists in the reference, so the oracle is pinned by the
reference's own unit and integration tests (sph/kernel.rs:110-230,
sph/neighbors.rs:120-200, sph/density.rs:160-260,
tests/evrard_collapse.rs:10-76) and by invariants a plausible slip would
break:
  * the kernel integrates to 1, dW/dh and grad W match central differences
    of W (a dropped 1/h or a wrong sign fails);
  * the tree ball query equals brute force;
  * pure hydrodynamics conserves momentum and total energy to the extent the
    neighbour lists are symmetric (a wrong sign or factor in the pressure,
    viscosity, heating or conductivity terms breaks one of them);
  * the Evrard collapse conserves energy to 5% up to t = 0.5 while K grows
    more than twofold and the thermal energy grows.
"""
import math

import numpy as np
import pytest

import oracle as O


# ---- kernel.rs:110-230 ------------------------------------------------------

def test_kernel_normalization_and_support():
    h, n = 1.0, 10_000
    dr = 2.0 * h / n
    total = sum(O.sph_w((i + 0.5) * dr, h) * 4 * math.pi * ((i + 0.5) * dr) ** 2 * dr
                for i in range(n))
    assert abs(total - 1.0) < 1e-3
    assert O.sph_w(2.0, 1.0) == 0.0 and O.sph_w(2.5, 1.0) == 0.0 and O.sph_w(1.99, 1.0) > 0
    prev = O.sph_w(0.0, 1.0)
    for i in range(1, 20):
        cur = O.sph_w(i * 0.1, 1.0)
        assert cur <= prev
        prev = cur


@pytest.mark.parametrize("r,h", [(0.7, 1.0), (0.05, 0.3), (1.9, 1.0), (0.4, 0.25)])
def test_dw_dh_is_the_h_derivative(r, h):
    e = 1e-6 * h
    fd = (O.sph_w(r, h + e) - O.sph_w(r, h - e)) / (2 * e)
    an = O.sph_dw_dh(r, h)
    assert abs(fd - an) <= 1e-5 * max(1.0, abs(an) * h ** 4)


@pytest.mark.parametrize("d,h", [((0.3, -0.2, 0.5), 0.8), ((0.2, 0.3, -0.1), 0.6),
                                 ((0.01, 0.0, 0.02), 0.05)])
def test_grad_w_is_the_spatial_gradient(d, h):
    dx, dy, dz = d
    r = math.sqrt(dx * dx + dy * dy + dz * dz)
    g = O.sph_grad_w(dx, dy, dz, r, h)
    e = 1e-7 * r
    for k in range(3):
        dp = list(d); dm = list(d)
        dp[k] += e; dm[k] -= e
        fd = (O.sph_w(math.sqrt(sum(v * v for v in dp)), h)
              - O.sph_w(math.sqrt(sum(v * v for v in dm)), h)) / (2 * e)
        assert abs(fd - g[k]) <= 1e-6 * np.abs(g).max()
    gr = O.sph_grad_w(-dx, -dy, -dz, r, h)
    assert np.abs(g + gr).max() < 1e-14
    assert abs(np.linalg.norm(g) - O.sph_abs_grad_w(r, h)) < 1e-14 * max(1, np.linalg.norm(g))
    assert (O.sph_grad_w(0, 0, 0, 0, 1.0) == 0).all()
    assert (O.sph_grad_w(3, 0, 0, 3, 1.0) == 0).all()


# ---- neighbours and density -------------------------------------------------

def _grid_gas(n_side, spacing, h, u, mass=1.0):
    n = n_side ** 3
    s = O.SphState(n)
    g = np.arange(n_side) * spacing
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    s.x[:], s.y[:], s.z[:] = X.ravel(), Y.ravel(), Z.ravel()
    s.m[:] = mass
    s.u[:] = u
    s.h[:] = h
    s.alpha[:] = 1.0
    s.type[:] = 4
    return s


def test_ball_query_matches_brute_force():
    """neighbors.rs:126-165 on a 5^3 grid, plus a random cloud."""
    rng = np.random.default_rng(3)
    for s, radius in ((_grid_gas(5, 0.5, 0.8, 1.0), 1.6), (None, 0.37)):
        if s is None:
            s = O.SphState(400)
            s.x[:], s.y[:], s.z[:] = rng.normal(size=(3, 400))
            s.m[:] = 1.0  # a zero-mass node counts as empty (octree.rs:124-126)
        for i in range(0, s.n, 7):
            d2 = (s.x - s.x[i]) ** 2 + (s.y - s.y[i]) ** 2 + (s.z - s.z[i]) ** 2
            bf = np.nonzero(d2 <= radius * radius)[0]
            tree = O.ball_query(s, i, radius)
            assert np.array_equal(np.sort(tree), bf), i


def test_uniform_density_recovery():
    """density.rs:181-256: interior density within 5% of m/s^3, omega ~ 1."""
    spacing = 0.5
    s = _grid_gas(8, spacing, 1.0, 1.0)
    inter = O.sph_compute(s)
    edge = 7 * spacing
    margin = 2 * s.h
    inside = ((s.x >= margin) & (s.y >= margin) & (s.z >= margin) & (s.x <= edge - margin)
              & (s.y <= edge - margin) & (s.z <= edge - margin))
    assert inside.sum() > 0
    expect = 1.0 / spacing ** 3
    assert np.mean(np.abs(s.rho[inside] - expect) / expect) < 0.05
    assert np.all(np.abs(inter["omega"][inside] - 1.0) < 0.3)
    # EOS (density.rs:111-119)
    g = 5.0 / 3.0
    assert np.allclose(s.p, (g - 1) * s.rho * s.u, rtol=1e-15)
    assert np.allclose(s.c, np.sqrt(g * (g - 1) * s.u), rtol=1e-15)


def _blob(seed):
    """Jittered 12^3 grid of gas with random velocities (viscosity on) and
    uneven u (conductivity on), pressure dominated."""
    rng = np.random.default_rng(seed)
    g = np.arange(12) * 0.1
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    n = X.size
    s = O.SphState(n)
    s.x[:] = X.ravel() + 0.02 * rng.normal(size=n)
    s.y[:] = Y.ravel() + 0.02 * rng.normal(size=n)
    s.z[:] = Z.ravel() + 0.02 * rng.normal(size=n)
    s.vx[:], s.vy[:], s.vz[:] = rng.normal(size=(3, n)) * 0.1
    s.m[:] = 1.0 / n
    s.u[:] = 1.0 + 0.5 * rng.random(n)
    s.h[:] = 0.12
    s.alpha[:] = 1.0
    s.type[:] = 4
    return s


def test_hydro_conserves_momentum_and_energy():
    """Sum m a = 0 and sum m (v.a + du/dt) = 0 up to the neighbour-list
    asymmetry (measured 2.6e-4 and 1.7e-5 here). Planting a sign flip of the
    P_j term, a 10% error on either P_i term, or a wrong sign or factor on the
    viscosity terms moves one of them to >= 1.2e-3."""
    s = _blob(5)
    inter = O.sph_compute(s)
    ma = s.m[:, None] * s.acc()
    assert np.all(np.abs(ma.sum(0)) <= 1e-3 * np.abs(ma).sum(0))
    pw = s.m * (s.vx * s.ax + s.vy * s.ay + s.vz * s.az)
    heat = s.m * s.dudt
    assert abs((pw + heat).sum()) <= 2e-4 * np.abs(pw).sum()
    assert (inter["max_mu"] > 0).any() and (inter["dalpha_dt"] != 0).any()


def _wendland_abs_grad(r, h):
    """|dW/dr| of W = 21/(16 pi h^3) (1 - q/2)^4 (1 + 2q), q = r/h, by hand:
    dW/dq = 21/(16 pi h^3) * (-5 q) (1 - q/2)^3."""
    q = r / h
    return 21 / (16 * math.pi * h ** 3) * 5 * q * (1 - q / 2) ** 3 / h


def test_two_particle_conduction_term():
    """At rest the only du/dt is the conductivity term of forces.rs:192-202,
    m alpha_c sqrt(|P_i - P_j| / rho_avg) (u_i - u_j) / (rho_i + rho_j) |grad W(h_avg)|,
    checked against the kernel derivative written out above; it is
    antisymmetric for equal masses."""
    s = O.SphState(2)
    s.x[:] = [0.0, 0.3]
    s.m[:] = 0.5
    s.u[:] = [2.0, 1.0]
    s.h[:] = 0.4
    s.alpha[:] = 1.0
    s.type[:] = 4
    O.sph_compute(s)
    r = 0.3
    havg = 0.5 * (s.h[0] + s.h[1])
    rs = s.rho[0] + s.rho[1]
    expect = 0.5 * 1.0 * math.sqrt(abs(s.p[0] - s.p[1]) / (0.5 * rs)) * (2.0 - 1.0) / rs \
        * _wendland_abs_grad(r, havg)
    assert expect > 0
    assert abs(s.dudt[0] - expect) <= 1e-12 * expect
    assert abs(s.dudt[0] + s.dudt[1]) <= 1e-15 * expect


def test_evrard_energy_conservation():
    """tests/evrard_collapse.rs:10-76 on the oracle (N = 1000, brute force)."""
    n = 1000
    s = O.evrard(n)
    eps = O.evrard_softening(n)
    O.sph_bf_accel(s, eps)
    inter = O.sph_compute(s)
    dt = min(O.sph_timestep(s, inter), 1e-3)
    d0 = O.sph_diagnostics(s, eps)
    t, step, t_end = 0.0, 0, 0.5
    while t < t_end:
        if t + dt > t_end:
            dt = t_end - t
        nxt = O.sph_step(s, dt, gravity=1, eps=eps)
        t += dt
        dt = min(nxt, dt * 2.0)
        step += 1
        assert step < 100_000
    d1 = O.sph_diagnostics(s, eps, time=t, step=step)
    assert abs(d1["E"] - d0["E"]) / abs(d0["E"]) < 0.05
    assert d1["K"] > 2.0 * d0["K"]
    assert d1["Eth"] > d0["Eth"]


def test_evrard_ics_follow_the_generator():
    """evrard_collapse.rs:45-57: all inside R, m = M/n, h = 1.5 spacing."""
    n = 2000
    s = O.evrard(n, seed=7)
    r2 = s.x ** 2 + s.y ** 2 + s.z ** 2
    assert r2.max() <= 1.0
    spacing = (4.0 / 3.0 * math.pi / n) ** (1 / 3)
    assert np.allclose(s.h, 1.5 * spacing, rtol=1e-15, atol=0) and np.all(s.m == 1.0 / n)
    assert np.all(s.u == 0.05) and np.all(s.alpha == 1.0) and np.all(s.type == 4)
    assert abs(O.evrard_softening(n) - 0.3 * spacing) < 1e-15
