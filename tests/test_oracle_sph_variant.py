"""Pins the oracle's build-defined SPH variant (no reference counterpart):
cubic-spline M4 kernel, Newton-Raphson smoothing lengths to 1e-4, fixed
viscosity alpha, gravity softening eps = h, and the rho ~ 1/r generator.

Pins (each chosen so a plausible slip fails):
  * the kernel integrates to 1, dW/dh and grad W equal central differences;
  * the Newton h of every particle matches a bisection root of
    rho(h) = m (eta/h)^3 computed by brute force in numpy;
  * pure hydro conserves momentum and total energy (fixed alpha: the
    viscous heating still balances the kinetic loss);
  * eps = h brute force equals a numpy brute force; the group walk with
    per-target softening at theta = 0 equals it;
  * the 1/r generator's enclosed mass follows (r/R)^2 and is isotropic;
  * a collapse of the 1/r sphere conserves energy to 5% up to t = 0.3.
"""
import math

import numpy as np
import pytest

import oracle as O


def test_cubic_spline_normalization_and_derivatives():
    n, h = 20000, 1.0
    dr = 2.0 * h / n
    tot = sum(O.cs_w((i + 0.5) * dr, h) * 4 * math.pi * ((i + 0.5) * dr) ** 2 * dr
              for i in range(n))
    assert abs(tot - 1.0) < 1e-6
    assert O.cs_w(2.0, 1.0) == 0.0 and O.cs_w(1.999, 1.0) > 0.0
    # W is C1 at q = 1 and q = 2
    e = 1e-9
    assert abs(O.cs_w(1 - e, 1) - O.cs_w(1 + e, 1)) < 1e-8
    for r, h in ((0.3, 1.0), (0.9, 0.7), (1.5, 1.0), (0.05, 0.04)):
        d = 1e-6 * h
        fd = (O.cs_w(r, h + d) - O.cs_w(r, h - d)) / (2 * d)
        assert abs(fd - O.cs_dw_dh(r, h)) <= 1e-6 * max(1.0, abs(fd))
    for d3, h in (((0.3, -0.2, 0.5), 0.5), ((0.02, 0.01, -0.03), 0.03)):
        r = math.sqrt(sum(v * v for v in d3))
        g = O.cs_grad_w(*d3, r, h)
        for k in range(3):
            e = 1e-7 * r
            dp = list(d3); dm = list(d3)
            dp[k] += e; dm[k] -= e
            fd = (O.cs_w(math.sqrt(sum(v * v for v in dp)), h)
                  - O.cs_w(math.sqrt(sum(v * v for v in dm)), h)) / (2 * e)
            assert abs(fd - g[k]) <= 1e-6 * np.abs(g).max()


def _bisect_h(x, m, i, eta, lo, hi):
    """Root of rho_i(h) - m_i (eta/h)^3 by bisection, rho by brute force."""
    d = np.sqrt(((x - x[i]) ** 2).sum(1))

    def g(h):
        q = d / h
        f = np.where(q < 1, 1 - 1.5 * q ** 2 + 0.75 * q ** 3,
                     np.where(q < 2, 0.25 * (2 - q) ** 3, 0.0))
        return (m * f).sum() / (math.pi * h ** 3) - m[i] * (eta / h) ** 3

    assert g(lo) * g(hi) < 0
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        if g(lo) * g(mid) <= 0:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def test_newton_h_matches_bisection_root():
    s = O.evrard_1r(3000, seed=3)
    prm = O.SphParams.variant()
    O.sph_compute(s, prm)
    x = np.stack([s.x, s.y, s.z], 1)
    rng = np.random.default_rng(0)
    for i in rng.choice(s.n, 25, replace=False):
        hb = _bisect_h(x, s.m, int(i), prm.eta, 0.3 * s.h[i], 3.0 * s.h[i])
        assert abs(s.h[i] - hb) <= 2e-4 * hb, (i, s.h[i], hb)


def _blob(seed):
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


def test_variant_hydro_conserves_momentum_and_energy():
    s = _blob(5)
    inter = O.sph_compute(s, O.SphParams.variant())
    ma = s.m[:, None] * s.acc()
    assert np.all(np.abs(ma.sum(0)) <= 1e-3 * np.abs(ma).sum(0))
    pw = s.m * (s.vx * s.ax + s.vy * s.ay + s.vz * s.az)
    assert abs((pw + s.m * s.dudt).sum()) <= 2e-4 * np.abs(pw).sum()
    assert np.all(inter["dalpha_dt"] == 0.0)


def test_softening_h_gravity():
    rng = np.random.default_rng(7)
    n = 300
    s = O.SphState(n)
    s.x[:], s.y[:], s.z[:] = rng.normal(size=(3, n))
    s.m[:] = rng.random(n) + 0.5
    eps = 0.05 + 0.1 * rng.random(n)
    O.bf_accel_eps(s, eps)
    x = np.stack([s.x, s.y, s.z], 1)
    d = x[None, :, :] - x[:, None, :]
    r2 = (d ** 2).sum(2) + eps[:, None] ** 2
    np.fill_diagonal(r2, np.inf)
    a = (s.m[None, :, None] * d / r2[:, :, None] ** 1.5).sum(1)
    assert np.abs(s.acc() - a).max() <= 1e-12 * np.abs(a).max()
    t = O.SphState(n)
    for f in ("x", "y", "z", "m"):
        getattr(t, f)[:] = getattr(s, f)
    O.bh_group_accel_eps(t, eps, theta=0.0)
    assert np.abs(t.acc() - a).max() <= 1e-12 * np.abs(a).max()


def test_one_over_r_generator():
    n = 20000
    s = O.evrard_1r(n, seed=11)
    r = np.sqrt(s.x ** 2 + s.y ** 2 + s.z ** 2)
    assert r.max() <= 1.0
    # enclosed mass fraction (r/R)^2: Kolmogorov-Smirnov distance
    rs = np.sort(r)
    ks = np.abs(np.arange(1, n + 1) / n - rs ** 2).max()
    assert ks < 1.63 / math.sqrt(n)  # 1% level
    assert np.abs([s.x.mean(), s.y.mean(), s.z.mean()]).max() < 0.02
    rho = 1.0 / (2 * math.pi * np.maximum(r, 1e-6))
    assert np.allclose(s.h, 1.2 * np.cbrt((1.0 / n) / rho), rtol=1e-14)
    assert np.all(s.type == 4) and np.all(s.u == 0.05) and np.all(s.m == 1.0 / n)


def test_variant_collapse_conserves_energy():
    n = 2000
    s = O.evrard_1r(n, seed=42)
    prm = O.SphParams.variant()
    O.bf_accel_eps(s, s.h)
    inter = O.sph_compute(s, prm)
    dt = min(O.sph_timestep(s, inter, prm), 1e-3)

    def energy():
        x = np.stack([s.x, s.y, s.z], 1)
        d = x[None, :, :] - x[:, None, :]
        r2 = (d ** 2).sum(2) + s.h[:, None] ** 2
        np.fill_diagonal(r2, np.inf)
        # the eps = h force derives from a pair potential only when h_i = h_j;
        # use the symmetrised form for a diagnostic
        e2 = 0.5 * (s.h[:, None] ** 2 + s.h[None, :] ** 2)
        r2s = (d ** 2).sum(2) + e2
        np.fill_diagonal(r2s, np.inf)
        U = -0.5 * (s.m[:, None] * s.m[None, :] / np.sqrt(r2s)).sum()
        K = 0.5 * (s.m * (s.vx ** 2 + s.vy ** 2 + s.vz ** 2)).sum()
        return K + U + (s.m * s.u).sum(), K
    e0, k0 = energy()
    t = 0.0
    while t < 0.3:
        if t + dt > 0.3:
            dt = 0.3 - t
        nxt = O.sph_step(s, dt, gravity=1, params=prm)
        t += dt
        dt = min(nxt, 2 * dt)
    e1, k1 = energy()
    assert abs(e1 - e0) / abs(e0) < 0.05
    assert k1 > k0
