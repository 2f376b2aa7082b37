"""Pins the oracle's discrete adjoint of the KDK/brute-force map
(oracle/gravis_oracle_adjoint.c) by central finite differences of the
oracle's own forward map. The reference has no reverse mode, so FD is the
independent check; a planted sign error in the adjoint would fail it.

Loss: L = sum_i |x_i(T) - target_i|^2, so dL/dx_T = 2 (x_T - target).
"""
import numpy as np

import oracle as O


def _loss(s0, eps, dt, T, tgt):
    s = s0.copy()
    O.ens_forward(s, eps, dt, T)
    return float(((s.x - tgt[0]) ** 2 + (s.y - tgt[1]) ** 2 + (s.z - tgt[2]) ** 2).sum())


def test_gather_forward_matches_reference_kdk():
    """The ensemble forward is sim-core's KDK with the gather form of the
    force: it tracks the reference i<j form to rounding."""
    s = O.plummer(64, 3)
    a = s.copy()
    O.ens_forward(a, 0.05, 1e-3, 50)
    b = s.copy()
    O.bf_accel(b, 0.05)
    O.kdk_steps(b, 1e-3, 50, 0.05)
    assert np.abs(a.x - b.x).max() < 1e-12 and np.abs(a.vx - b.vx).max() < 1e-11


def test_adjoint_matches_central_fd():
    rng = np.random.default_rng(0)
    n, eps, dt, T = 24, 0.05, 1e-3, 60
    s = O.plummer(n, 5)
    s.m[:] = rng.uniform(0.5, 1.5, n) / n
    tgt = O.ens_targets(7, n)
    f = s.copy()
    O.ens_forward(f, eps, dt, T)
    gx, gy, gz = 2 * (f.x - tgt[0]), 2 * (f.y - tgt[1]), 2 * (f.z - tgt[2])
    adj = O.ens_adjoint(s, eps, dt, T, gx, gy, gz)
    h = 1e-6
    worst = 0.0
    for field in ("x", "y", "z", "vx", "vy", "vz", "m"):
        for i in rng.choice(n, 6, replace=False):
            sp, sm = s.copy(), s.copy()
            getattr(sp, field)[i] += h
            getattr(sm, field)[i] -= h
            fd = (_loss(sp, eps, dt, T, tgt) - _loss(sm, eps, dt, T, tgt)) / (2 * h)
            ad = adj[field][i]
            scale = max(abs(fd), 1e-3 * np.abs(adj[field]).max(), 1e-12)
            worst = max(worst, abs(ad - fd) / scale)
            assert abs(ad - fd) <= 1e-5 * scale, (field, i, ad, fd)
    assert worst < 1e-5


def test_adjoint_velocity_cotangent_path():
    """Cotangents on final velocities flow through the K2/D/K1 chain."""
    rng = np.random.default_rng(1)
    n, eps, dt, T = 10, 0.05, 1e-3, 20
    s = O.plummer(n, 8)
    gv = [rng.normal(size=n) for _ in range(3)]
    z = np.zeros(n)
    adj = O.ens_adjoint(s, eps, dt, T, z, z, z, gv)

    def lv(st):
        t = st.copy()
        O.ens_forward(t, eps, dt, T)
        return float((gv[0] * t.vx + gv[1] * t.vy + gv[2] * t.vz).sum())
    h = 1e-6
    for field in ("x", "vy", "m"):
        i = 3
        sp, sm = s.copy(), s.copy()
        getattr(sp, field)[i] += h
        getattr(sm, field)[i] -= h
        fd = (lv(sp) - lv(sm)) / (2 * h)
        assert abs(adj[field][i] - fd) <= 1e-6 * max(1.0, abs(fd)), (field, adj[field][i], fd)


def test_ensemble_counts_and_targets_are_seeded():
    c = O.ens_counts(42, 256)
    assert c.min() >= 512 and c.max() <= 8192
    assert (c == O.ens_counts(42, 256)).all()
    t = O.ens_targets(9, 1000)
    assert min(a.min() for a in t) >= -1.0 and max(a.max() for a in t) < 1.0
