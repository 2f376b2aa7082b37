"""GPU parity of the batched ensemble and its adjoint through the C ABI.

Bars:
  * FP64 ensemble forward == the oracle's per-system KDK (gather form) to
    1e-11 after 100 steps (summation order differs; no chaos at this length).
  * hipGraph replay is bit-identical to eager launches, replay after replay.
  * GPU adjoint (FP64) == the oracle adjoint (itself pinned by central FD,
    tests/test_oracle_adjoint.py) to 1e-8 relative.
  * GPU adjoint vs central finite differences (h = 1e-6, FP64 forward) on 64
    random coordinates per system: ||adj - fd|| / ||fd|| <= 1e-5 per system
    and every coordinate within 1e-5 of the probe vector's largest entry
    (a single small component is limited by ~1e-14 relative round-off in the
    FP64 loss, i.e. ~1e-8/h in the FD quotient).
  * Checkpoint spacing does not change the gradients (bit-identical).
  * Mixed-precision forward: RMS relative force error <= 1e-6 vs FP64.
"""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def E():
    import gravis_amd.ensemble as ens
    return ens


def make_ensemble(counts, seed0=42):
    """Concatenated Plummer spheres, system s seeded seed0 + s."""
    parts = [O.plummer(int(n), seed0 + s) for s, n in enumerate(counts)]
    st = {}
    for f, of in (("x", "x"), ("y", "y"), ("z", "z"), ("vx", "vx"), ("vy", "vy"), ("vz", "vz"),
                  ("m", "m")):
        st[f] = np.concatenate([getattr(p, of) for p in parts])
    n = len(st["x"])
    for f in ("ax", "ay", "az"):
        st[f] = np.zeros(n)
    return st, parts


def init_accel(E, st, counts, eps, precision="fp64"):
    import gravis_amd as g
    import ctypes as C
    ctx = g.default_context()
    off = E._offsets(counts)
    s = E._solver(eps, precision)
    import torch
    T = {k: torch.from_numpy(st[k]).cuda() for k in ("x", "y", "z", "m", "ax", "ay", "az")}
    torch.cuda.synchronize()
    ctx.check(ctx._L.gravis_ensemble_accel_device(
        ctx.handle, C.byref(s), len(counts), E._p(off),
        *[C.c_void_p(T[k].data_ptr()) for k in ("x", "y", "z", "m", "ax", "ay", "az")]))
    ctx.synchronize()
    for k in ("ax", "ay", "az"):
        st[k][:] = T[k].cpu().numpy()


def test_ensemble_forward_matches_oracle(E):
    counts = [100, 300, 513, 64]
    eps, dt, T = 0.05, 1e-3, 100
    say(f"ensemble forward {counts} T={T}")
    st, parts = make_ensemble(counts)
    init_accel(E, st, counts, eps)
    E.run(st, counts, eps, dt, T, "fp64", use_graph=True)
    off = E._offsets(counts)
    for s, p in enumerate(parts):
        O.ens_forward(p, eps, dt, T)
        lo, hi = int(off[s]), int(off[s + 1])
        for f in ("x", "y", "z", "vx", "vy", "vz"):
            d = np.abs(st[f][lo:hi] - getattr(p, f)).max()
            assert d <= 1e-11, (s, f, d)


def test_ensemble_graph_replay_is_bit_identical(E):
    counts = [700, 257, 1024]
    eps, dt = 0.05, 1e-3
    say("ensemble graph vs eager")
    st, _ = make_ensemble(counts, 7)
    init_accel(E, st, counts, eps, "mixed")
    a = {k: v.copy() for k, v in st.items()}
    b = {k: v.copy() for k, v in st.items()}
    for rep in range(3):  # replays with eager work between them
        E.run(a, counts, eps, dt, 5, "mixed", use_graph=True)
        E.run(b, counts, eps, dt, 5, "mixed", use_graph=False)
        for f in ("x", "vx", "ax", "vz"):
            assert np.array_equal(a[f], b[f]), (rep, f)


def test_ensemble_mixed_force_error(E):
    counts = [8192, 512, 3000]
    eps = 0.05
    say("ensemble mixed force error")
    st, _ = make_ensemble(counts, 3)
    s64 = {k: v.copy() for k, v in st.items()}
    init_accel(E, s64, counts, eps, "fp64")
    init_accel(E, st, counts, eps, "mixed")
    A = np.stack([s64["ax"], s64["ay"], s64["az"]], 1)
    B = np.stack([st["ax"], st["ay"], st["az"]], 1)
    rel = np.sqrt(((A - B) ** 2).sum(1) / (A ** 2).sum(1))
    say(f"  rms={np.sqrt((rel ** 2).mean()):.3e} max={rel.max():.3e}")
    assert np.sqrt((rel ** 2).mean()) <= 1e-6


@pytest.mark.parametrize("counts,eps", [([8192, 512, 3000, 1, 0, 65], 0.05),
                                         ([700, 257, 1024], 0.0)])
def test_ensemble_mixed_kdk_sorted_path(E, counts, eps):
    """Mixed KDK runs in per-system Morton order on a sorted copy and
    returns the state in the caller's order. After 3 steps the difference
    from the FP64 run, per component, is at most 1e-5 of the change over the
    run in RMS and 1e-3 of the largest change (a mix-up of the order would
    give errors of the size of the change itself), and the accelerations
    agree to RMS 1e-6 / max 1e-4 relative. Systems of 0 and 1 particles
    included; eps = 0 takes the masked self-pair path."""
    say(f"ensemble mixed kdk sorted path {counts} eps={eps}")
    st, _ = make_ensemble(counts, 11)
    init_accel(E, st, counts, eps, "fp64")
    s0 = {k: v.copy() for k, v in st.items()}
    s64 = {k: v.copy() for k, v in st.items()}
    E.run(s64, counts, eps, 1e-3, 3, "fp64", use_graph=True)
    E.run(st, counts, eps, 1e-3, 3, "mixed", use_graph=True)
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        chg = np.abs(s64[f] - s0[f])
        err = np.abs(st[f] - s64[f])
        rms_e, rms_c = np.sqrt((err ** 2).mean()), np.sqrt((chg ** 2).mean())
        say(f"  {f}: rms err / rms change {rms_e / rms_c:.2e}, max err / max change "
            f"{err.max() / chg.max():.2e}")
        assert rms_e <= 1e-5 * rms_c and err.max() <= 1e-3 * chg.max(), f
    A = np.stack([s64["ax"], s64["ay"], s64["az"]], 1)
    B = np.stack([st["ax"], st["ay"], st["az"]], 1)
    ok = (A ** 2).sum(1) > 0
    rel = np.sqrt(((A - B) ** 2).sum(1)[ok] / (A ** 2).sum(1)[ok])
    say(f"  rms={np.sqrt((rel ** 2).mean()):.3e} max={rel.max():.3e}")
    assert np.sqrt((rel ** 2).mean()) <= 1e-6 and rel.max() <= 1e-4


def _targets(counts):
    tx, ty, tz = [], [], []
    for s, n in enumerate(counts):
        t = O.ens_targets(10000 + s, int(n))
        tx.append(t[0]); ty.append(t[1]); tz.append(t[2])
    return np.concatenate(tx), np.concatenate(ty), np.concatenate(tz)


def test_adjoint_matches_oracle_and_is_checkpoint_invariant(E):
    counts = [40, 96, 33]
    eps, dt, T = 0.05, 1e-3, 40
    say(f"ensemble adjoint vs oracle {counts} T={T}")
    st, parts = make_ensemble(counts, 11)
    tgt = _targets(counts)
    res = E.adjoint(st, counts, eps, dt, T, targets=tgt, checkpoint_every=7)
    off = E._offsets(counts)
    for s, p in enumerate(parts):
        lo, hi = int(off[s]), int(off[s + 1])
        f = p.copy()
        O.ens_forward(f, eps, dt, T)
        gx = 2 * (f.x - tgt[0][lo:hi])
        gy = 2 * (f.y - tgt[1][lo:hi])
        gz = 2 * (f.z - tgt[2][lo:hi])
        loss = float((0.25 * (gx ** 2 + gy ** 2 + gz ** 2)).sum())
        assert abs(res["loss"][s] - loss) <= 1e-10 * loss
        ref = O.ens_adjoint(p, eps, dt, T, gx, gy, gz)
        for k in ("x", "y", "z", "vx", "vy", "vz", "m"):
            a, b = res[k][lo:hi], ref[k]
            assert np.abs(a - b).max() <= 1e-8 * np.abs(b).max(), (s, k)
    for ck in (1, 32, 1000):
        r2 = E.adjoint(st, counts, eps, dt, T, targets=tgt, checkpoint_every=ck)
        for k in ("x", "vy", "m"):
            assert np.array_equal(r2[k], res[k]), (ck, k)


def test_adjoint_matches_central_finite_differences(E):
    """64 random coordinates per system (positions, velocities, masses),
    central FD with h = 1e-6 on the FP64 forward, <= 1e-5 relative."""
    counts = [96, 160, 64, 200]
    eps, dt, T = 0.05, 1e-3, 200
    say(f"ensemble adjoint vs FD {counts} T={T}")
    st, _ = make_ensemble(counts, 21)
    tgt = _targets(counts)
    res = E.adjoint(st, counts, eps, dt, T, targets=tgt, checkpoint_every=32)
    off = E._offsets(counts)
    rng = np.random.default_rng(5)
    fields = ("x", "y", "z", "vx", "vy", "vz", "m")
    # build all perturbed copies as one big ensemble: each probe is a system copy
    probes = []
    for s in range(len(counts)):
        lo, hi = int(off[s]), int(off[s + 1])
        for q in range(64):
            f = fields[rng.integers(len(fields))]
            i = int(rng.integers(lo, hi))
            probes.append((s, f, i))
    h = 1e-6
    big_counts, big = [], {k: [] for k in fields}
    for s, f, i in probes:
        lo, hi = int(off[s]), int(off[s + 1])
        for sign in (+1, -1):
            for k in fields:
                v = st[k][lo:hi].copy()
                if k == f:
                    v[i - lo] += sign * h
                big[k].append(v)
            big_counts.append(hi - lo)
    bst = {k: np.concatenate(v) for k, v in big.items()}
    nb = len(bst["x"])
    for k in ("ax", "ay", "az"):
        bst[k] = np.zeros(nb)
    init_accel(E, bst, big_counts, eps)
    E.run(bst, big_counts, eps, dt, T, "fp64", use_graph=True)
    boff = E._offsets(big_counts)
    ad_all, fd_all = {}, {}
    for q, (s, f, i) in enumerate(probes):
        lo = int(off[s])
        L = []
        for k in range(2):
            b = 2 * q + k
            a0, a1 = int(boff[b]), int(boff[b + 1])
            t0 = tgt[0][lo:lo + (a1 - a0)], tgt[1][lo:lo + (a1 - a0)], tgt[2][lo:lo + (a1 - a0)]
            L.append(((bst["x"][a0:a1] - t0[0]) ** 2 + (bst["y"][a0:a1] - t0[1]) ** 2
                      + (bst["z"][a0:a1] - t0[2]) ** 2).sum())
        ad_all.setdefault(s, []).append(res[f][i])
        fd_all.setdefault(s, []).append((L[0] - L[1]) / (2 * h))
    worst = 0.0
    for s in ad_all:
        ad, fd = np.array(ad_all[s]), np.array(fd_all[s])
        rel = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
        coord = np.abs(ad - fd).max() / np.abs(fd).max()
        say(f"  system {s}: |adj-fd|/|fd| = {rel:.2e}, max coord err / max |fd| = {coord:.2e}")
        worst = max(worst, rel)
        assert rel <= 1e-5 and coord <= 1e-5, s
    say(f"  worst norm-wise relative |adjoint - FD| = {worst:.2e}")


def test_autograd_function_matches_adjoint(E):
    import torch
    counts = [50, 70]
    eps, dt, T = 0.05, 1e-3, 30
    say("EnsembleKDK autograd")
    st, _ = make_ensemble(counts, 31)
    tgt = _targets(counts)
    x0 = torch.tensor(np.stack([st["x"], st["y"], st["z"]]), device="cuda", requires_grad=True)
    v0 = torch.tensor(np.stack([st["vx"], st["vy"], st["vz"]]), device="cuda", requires_grad=True)
    m = torch.tensor(st["m"], device="cuda", requires_grad=True)
    xT, vT = E.EnsembleKDK.apply(x0, v0, m, counts, eps, dt, T)
    t = torch.tensor(np.stack(tgt), device="cuda")
    loss = ((xT - t) ** 2).sum()
    loss.backward()
    ref = E.adjoint(st, counts, eps, dt, T, targets=tgt)
    assert abs(loss.item() - ref["loss"].sum()) <= 1e-12 * ref["loss"].sum()
    gx = x0.grad.cpu().numpy()
    assert np.abs(gx[0] - ref["x"]).max() <= 1e-12 * np.abs(ref["x"]).max()
    assert np.abs(v0.grad.cpu().numpy()[2] - ref["vz"]).max() <= 1e-12 * np.abs(ref["vz"]).max()
    assert np.abs(m.grad.cpu().numpy() - ref["m"]).max() <= 1e-12 * np.abs(ref["m"]).max()


def test_host_ensemble_accel_matches_device(E):
    """gravis_ensemble_accel (host arrays, what a Rust host calls before
    gravis_ensemble_kdk) gives the device entry point's bits."""
    counts = [300, 1, 0, 517]
    for precision in ("fp64", "mixed"):
        st, _ = make_ensemble(counts, 5)
        dev = {k: v.copy() for k, v in st.items()}
        init_accel(E, dev, counts, 0.05, precision)
        E.accel(st, counts, 0.05, precision)
        for k in ("ax", "ay", "az"):
            assert np.array_equal(st[k], dev[k]), (precision, k)
