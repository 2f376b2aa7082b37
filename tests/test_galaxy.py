"""Galaxy-collision initial conditions (GalaxyCollision,
crates/sim-core/src/scenarios/galaxy_collision.rs:20-450).

The reference publishes no per-particle values for this generator. Its own
tests (crates/sim-core/tests/galaxy.rs:90-210) pin counts, types,
determinism and the stability of an isolated pair; they are restated here.
The oracle's raw samples (before the COM shift) are then checked against
the profiles the generator claims to draw from. Each enclosed mass is found
by quadrature of its density, not from the generator's closed forms:
  * exponential disk: R Sigma(R) ~ x e^-x on [0, 6 R_d); sech^2 heights from
    U in [0.01, 0.99); per particle, v_R in [-0.15, 0.15] v_c,
    v_phi - v_c in [-0.075, 0.075] v_c and |v_z| <= 0.1 v_c, with
    v_c^2 = G M(<R) / R summed over disk, bulge and halo;
  * Plummer bulge: M(<r)/M uniform in [0.001, 1);
  * Hernquist halo: M(<r)/M uniform in [0.001, (R/(R+a))^2); isotropic
    directions; <|v|^2> / (v_c^2 / 2) = 3 E[-ln U], U in [0.001, 1).
The product generator (scenarios.cpp) is compared with the oracle bit for
bit. CPU only: both generators are host code.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate, stats

import oracle
import gravis_amd as g
from gravis_amd import _lib

RD, ZH, AB, AH, RMAX_D, RMAX_H = 1.0, 0.1, 0.3, 5.0, 6.0, 30.0
# component masses as the reference forms them: baryon_mass * fraction
M_DISK, M_BULGE, M_HALO = 1.0 * 0.7, 1.0 * (1.0 - 0.7), 1.0 * 10.0


def _cum_mass(density_r2, r):
    """M(<r) = int_0^r 4 pi s^2 rho(s) ds on sorted radii, by quadrature."""
    out = np.empty_like(r)
    prev, acc = 0.0, 0.0
    for k, rk in enumerate(r):
        acc += integrate.quad(density_r2, prev, rk, epsabs=0, epsrel=1e-12, limit=200)[0]
        out[k] = acc
        prev = rk
    return out


def _grid_mass():
    grid = np.concatenate([[1e-9], np.geomspace(1e-6, 60.0, 1500)])
    disk = _cum_mass(lambda s: 2 * math.pi * s * M_DISK / (2 * math.pi * RD * RD)
                     * math.exp(-s / RD), grid)
    bulge = _cum_mass(lambda s: 4 * math.pi * s * s * 3 * M_BULGE / (4 * math.pi * AB ** 3)
                      * (1 + s * s / (AB * AB)) ** -2.5, grid)
    halo = _cum_mass(lambda s: 4 * math.pi * s * s * M_HALO * AH
                     / (2 * math.pi * s * (s + AH) ** 3), grid)
    return grid, disk, bulge, halo


@pytest.fixture(scope="module")
def profiles():
    return _grid_mass()


def _interp(grid, y, r):
    return np.interp(np.log(r), np.log(grid), y)


@pytest.fixture(scope="module")
def raw():
    s, t = oracle.galaxy_collision(com_shift=False, n_per_galaxy=20000)
    return s, t


def _product(**kw):
    return g.scenarios.galaxy_collision(**kw)


def _oracle_matches_product(**kw):
    s, t = oracle.galaxy_collision(**kw)
    p = _product(**kw)
    for a, b in (("x", "x"), ("y", "y"), ("z", "z"), ("vx", "vx"), ("vy", "vy"),
                 ("vz", "vz"), ("m", "mass")):
        np.testing.assert_array_equal(getattr(s, a), getattr(p, b), err_msg=a)
    np.testing.assert_array_equal(t, p.particle_type)
    for f in ("ax", "ay", "az") + g.Particles.SPH_FIELDS:
        assert not getattr(p, f).any()


@pytest.mark.parametrize("kw", [dict(n_per_galaxy=1000), dict(n_per_galaxy=777, seed=7),
                                dict(n_per_galaxy=1000, separation=0.0, impact_parameter=0.0,
                                     approach_velocity=0.0),
                                dict(n_per_galaxy=300, disk_fraction=0.5, halo_mass_ratio=3.0,
                                     disk_scale_height=0.2, halo_max_radius=12.0)])
def test_product_generator_is_the_oracle_bit_for_bit(kw):
    _oracle_matches_product(**kw)


# ---- crates/sim-core/tests/galaxy.rs ----------------------------------------

def test_generates_correct_total_count():
    assert _product(n_per_galaxy=500).count == 1000


def test_has_all_particle_types():
    t = _product(n_per_galaxy=500).particle_type
    n_disk, n_bulge, n_halo = (int((t == k).sum()) for k in (1, 2, 3))
    assert n_disk > 0 and n_bulge > 0 and n_halo > 0
    assert n_disk + n_bulge + n_halo == t.size
    assert n_halo > n_disk
    # counts proportional to mass, rounded (galaxy_collision.rs:85-88):
    # 500 * 0.7 / 11 = 31.8 -> 32, 500 * 0.3 / 11 = 13.6 -> 14, rest halo
    assert (n_disk, n_bulge, n_halo) == (64, 28, 908)


def test_deterministic():
    a, b = _product(n_per_galaxy=200), _product(n_per_galaxy=200)
    for f in ("x", "vx", "particle_type"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f))
    c = _product(n_per_galaxy=200, seed=43)
    assert not np.array_equal(a.x, c.x)


def test_single_galaxy_stability_oracle():
    """galaxy.rs:154-210 on the oracle: two galaxies on top of each other,
    Barnes-Hut theta 0.5, suggested eps and dt, 500 KDK steps: virial ratio
    in [0.5, 2), |dE/E| < 5 %."""
    kw = dict(n_per_galaxy=1000, separation=0.0, impact_parameter=0.0, approach_velocity=0.0)
    s, _ = oracle.galaxy_collision(**kw)
    eps = g.scenarios.galaxy_suggested_softening(**kw)
    dt = g.scenarios.galaxy_suggested_dt(**kw)
    oracle.bh_accel(s, eps, 0.5)
    e0 = oracle.diagnostics(s, eps)["E"]
    oracle.kdk_steps(s, dt, 500, eps, algo=oracle.BH_REF, theta=0.5)
    d = oracle.diagnostics(s, eps, time=500 * dt, step=500)
    assert 0.5 <= d["virial"] < 2.0, d["virial"]
    assert abs((d["E"] - e0) / e0) < 0.05


# ---- profiles of the raw samples --------------------------------------------

def _split(s, t, galaxy):
    n = s.n // 2
    sl = slice(0, n) if galaxy == 0 else slice(n, 2 * n)
    sign = -1.0 if galaxy == 0 else 1.0
    # placement (galaxy_collision.rs:136-151): (-+7.5, -+1, 0), vx +-0.25
    x = s.x[sl] - sign * 7.5
    y = s.y[sl] - sign * 1.0
    vx = s.vx[sl] + sign * 0.25
    return dict(x=x, y=y, z=s.z[sl], vx=vx, vy=s.vy[sl], vz=s.vz[sl], m=s.m[sl], t=t[sl])


def test_counts_masses_and_placement(raw):
    s, t = raw
    for gal in (0, 1):
        d = _split(s, t, gal)
        k = d["t"]
        # 20000 * 0.7 / 11 = 1272.7, 20000 * 0.3 / 11 = 545.5 -> 545
        assert [int((k == c).sum()) for c in (1, 2, 3)] == [1273, 545, 18182]
        # components in generation order: disk, bulge, halo
        assert np.all(np.diff(k.astype(int)) >= 0)
        for c, mtot, cnt in ((1, M_DISK, 1273), (2, M_BULGE, 545), (3, M_HALO, 18182)):
            np.testing.assert_array_equal(d["m"][k == c], mtot / cnt)
        # the disk of each galaxy is centred on its placement point
        disk = k == 1
        assert abs(d["x"][disk].mean()) < 0.1 and abs(d["y"][disk].mean()) < 0.1
        assert abs(d["vx"][disk].mean()) < 0.1


def test_disk_radii_heights_and_rotation(raw, profiles):
    s, t = raw
    grid, md, mb, mh = profiles
    R_all, ratios = [], []
    for gal in (0, 1):
        d = _split(s, t, gal)
        k = d["t"] == 1
        x, y, z = d["x"][k], d["y"][k], d["z"][k]
        vx, vy, vz = d["vx"][k], d["vy"][k], d["vz"][k]
        R = np.hypot(x, y)
        assert R.max() < RMAX_D
        # heights: z = z0 atanh(2U - 1), U in [0.01, 0.99)
        assert np.abs(z).max() <= ZH * math.atanh(0.98) * (1 + 1e-12)
        cdf_z = lambda v: ((np.tanh(v / ZH) + 1) / 2 - 0.01) / 0.98
        assert stats.kstest(z, cdf_z).pvalue > 1e-3
        vR = (x * vx + y * vy) / R
        vphi = (x * vy - y * vx) / R
        vc = np.sqrt((_interp(grid, md, R) + _interp(grid, mb, R) + _interp(grid, mh, R)) / R)
        tol = 1e-6
        assert np.all(np.abs(vR) <= 0.15 * vc * (1 + tol))
        assert np.all(np.abs(vphi - vc) <= 0.075 * vc * (1 + tol))
        assert np.all(np.abs(vz) <= 0.1 * vc * (1 + tol))
        R_all.append(R)
        ratios.append(vphi / vc)
    R = np.concatenate(R_all)
    norm = integrate.quad(lambda u: u * math.exp(-u), 0, RMAX_D)[0]
    cdf_R = np.vectorize(lambda r: integrate.quad(lambda u: u * math.exp(-u), 0, r)[0] / norm)
    assert stats.kstest(R, cdf_R).pvalue > 1e-3
    # uniform dispersions: mean v_phi / v_c = 1 within 4 sigma
    r = np.concatenate(ratios)
    assert abs(r.mean() - 1.0) < 4 * 0.075 / math.sqrt(3) / math.sqrt(r.size)


def test_bulge_is_a_truncated_plummer_sphere(raw, profiles):
    s, t = raw
    grid, _, mb, _ = profiles
    rs = []
    for gal in (0, 1):
        d = _split(s, t, gal)
        k = d["t"] == 2
        rs.append(np.sqrt(d["x"][k] ** 2 + d["y"][k] ** 2 + d["z"][k] ** 2))
    r = np.concatenate(rs)
    frac = _interp(grid, mb, r) / M_BULGE
    assert frac.min() >= 0.001 * (1 - 1e-6)
    assert stats.kstest(frac, stats.uniform(0.001, 0.999).cdf).pvalue > 1e-3


def test_halo_is_a_truncated_hernquist_sphere(raw, profiles):
    s, t = raw
    grid, md, mb, mh = profiles
    rs, v2, cz = [], [], []
    for gal in (0, 1):
        d = _split(s, t, gal)
        k = d["t"] == 3
        x, y, z = d["x"][k], d["y"][k], d["z"][k]
        r = np.sqrt(x * x + y * y + z * z)
        rs.append(r)
        vc2 = (_interp(grid, md, r) + _interp(grid, mb, r) + _interp(grid, mh, r)) / r
        v2.append((d["vx"][k] ** 2 + d["vy"][k] ** 2 + d["vz"][k] ** 2) / (vc2 / 2))
        cz.append(z / r)
    r = np.concatenate(rs)
    assert r.max() < RMAX_H
    frac = _interp(grid, mh, r) / M_HALO
    u_max = float(_interp(grid, mh, np.array([RMAX_H]))[0]) / M_HALO
    assert stats.kstest(frac, stats.uniform(0.001, u_max - 0.001).cdf).pvalue > 1e-3
    cz = np.concatenate(cz)
    assert abs(cz.mean()) < 4 / math.sqrt(3 * cz.size)
    assert abs((cz ** 2).mean() - 1 / 3) < 4 * math.sqrt(4 / 45 / cz.size)
    # Box-Muller with u1 in [0.001, 1): E[g^2] = E[-ln u1] = 1 + a ln a / (1 - a)
    a = 0.001
    e_g2 = 1 + a * math.log(a) / (1 - a)
    v2 = np.concatenate(v2)
    assert abs(v2.mean() / (3 * e_g2) - 1) < 4 * math.sqrt(2 / 3) / math.sqrt(v2.size)


def test_suggested_dt_and_softening():
    # t_orb(R_d) / 500 with enclosed masses 0.26 M_disk, Plummer and
    # Hernquist at R_d (galaxy_collision.rs:159-185); quadrature cross-check
    # of the two spherical terms.
    mb = integrate.quad(lambda s: 4 * math.pi * s * s * 3 * M_BULGE / (4 * math.pi * AB ** 3)
                        * (1 + s * s / (AB * AB)) ** -2.5, 0, RD, epsrel=1e-13)[0]
    mh = integrate.quad(lambda s: 4 * math.pi * s * s * M_HALO * AH
                        / (2 * math.pi * s * (s + AH) ** 3), 0, RD, epsrel=1e-13)[0]
    m_enc = 0.26 * M_DISK + mb + mh
    expect = 2 * math.pi * RD / math.sqrt(m_enc / RD) / 500
    assert g.scenarios.galaxy_suggested_dt() == pytest.approx(expect, rel=1e-12)
    assert oracle.galaxy_suggested_dt() == g.scenarios.galaxy_suggested_dt()
    assert g.scenarios.galaxy_suggested_softening() == 0.05


def test_rejects_bad_arguments():
    L = _lib.load()
    prm = g.scenarios.galaxy_params(n_per_galaxy=10)
    p = g.Particles.zeros(19)
    st = g.sph_state(p)
    assert L.gravis_scenario_galaxy_collision(C.byref(prm), C.byref(st)) == _lib.INVALID_ARG
    prm = g.scenarios.galaxy_params(n_per_galaxy=0)
    p = g.Particles.zeros(0)
    st = g.sph_state(p)
    assert L.gravis_scenario_galaxy_collision(C.byref(prm), C.byref(st)) == _lib.INVALID_ARG
    with pytest.raises(TypeError):
        g.scenarios.galaxy_params(no_such_field=1)
