"""The batched differentiable ensemble at its configured layout: 256 Plummer
spheres (a = 1, M = 1, system s seeded 42 + s) with N_s drawn in
[512, 8192] from seed 42 (about 1.13 M particles), eps 0.05, dt 1e-3, 1,000
KDK steps, mixed forces, one hipGraph-captured step replayed; loss
sum_s sum_i |x_i(T) - target_i|^2 with targets U[-1, 1)^3 seeded 10000 + s;
adjoint with checkpoints every 32 steps.

Bars:
  * layout: the product's counts equal the oracle's;
  * forces of every system: mixed against FP64 on the GPU, RMS relative
    error <= 1e-6 per system (max <= 1e-4); FP64 on the GPU against the
    oracle's f64 brute force on four systems (first, last, smallest,
    largest), condition-normalised error |da| / sum_j |a_ij| <= 1e-12;
  * forward, 1,000 steps: the FP64 hipGraph run against the oracle's f64
    KDK on the smallest and the largest system, positions within 1e-8
    absolute; the mixed hipGraph run against the FP64 run, positions within
    1e-5 absolute (RMS over each system);
  * adjoint, 1,000 steps, checkpoints every 32, on four systems cut from the
    ensemble (smallest, largest, two random):
      - per system, the full 256-system mixed adjoint equals the four-system
        mixed adjoint bit for bit (systems do not interact);
      - the FP64-forward adjoint against central finite differences
        (h = 1e-6, FP64 forward) on 64 random coordinates per system:
        |adj - fd| / |fd| <= 1e-5;
      - the FP64-forward adjoint of the smallest system against the oracle's
        f64 adjoint: within 1e-8 of the largest component;
      - the mixed-forward adjoint against the FP64-forward adjoint:
        |g_mixed - g_fp64| / |g_fp64| <= 1e-4 per system and field.
"""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

EPS, DT, T, CK = 0.05, 1e-3, 1000, 32
FIELDS = ("x", "y", "z", "vx", "vy", "vz", "m")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def sub(st, off, systems):
    """The systems' slices of an ensemble state as a new ensemble."""
    out = {}
    for k, v in st.items():
        out[k] = np.concatenate([v[int(off[s]):int(off[s + 1])] for s in systems])
    return out


@pytest.fixture(scope="module")
def C5():
    import gravis_amd.ensemble as E
    cnts = E.counts(42, 256, 512, 8192)
    st = E.plummer_ensemble(cnts)
    off = E._offsets(cnts)
    tg = [np.concatenate(a) for a in zip(*[E.targets(10000 + s, int(c))
                                            for s, c in enumerate(cnts)])]
    order = np.argsort(cnts, kind="stable")
    rng = np.random.default_rng(77)
    pick = [int(order[0]), int(order[-1])]
    while len(pick) < 4:
        s = int(rng.integers(len(cnts)))
        if s not in pick:
            pick.append(s)
    say(f"C5: {len(cnts)} systems, N_total {int(cnts.sum())}, adjoint systems {pick} "
        f"(N = {[int(cnts[s]) for s in pick]})")
    O.ens_set_threads(16)
    yield dict(E=E, cnts=cnts, st=st, off=off, tg=tg, pick=pick)
    O.ens_set_threads(1)


def accel(E, st, cnts, precision):
    import ctypes as C
    import torch
    import gravis_amd as g
    ctx = g.default_context()
    s = E._solver(EPS, precision)
    off = E._offsets(cnts)
    Td = {k: torch.from_numpy(np.ascontiguousarray(st[k])).cuda()
          for k in ("x", "y", "z", "m")}
    A = [torch.zeros(len(st["x"]), dtype=torch.float64, device="cuda") for _ in range(3)]
    torch.cuda.synchronize()
    ctx.check(ctx._L.gravis_ensemble_accel_device(
        ctx.handle, C.byref(s), len(cnts), E._p(off),
        *[C.c_void_p(Td[k].data_ptr()) for k in ("x", "y", "z", "m")],
        *[C.c_void_p(a.data_ptr()) for a in A]))
    ctx.synchronize()
    return np.stack([a.cpu().numpy() for a in A], 1)


def test_c5_layout(C5):
    cnts = C5["cnts"]
    assert np.array_equal(cnts, O.ens_counts(42, 256, 512, 8192))
    assert cnts.min() >= 512 and cnts.max() <= 8192 and len(cnts) == 256


def test_c5_forces_every_system(C5):
    E, cnts, st, off = C5["E"], C5["cnts"], C5["st"], C5["off"]
    say("C5: FP64 and mixed forces, all systems")
    a64 = accel(E, st, cnts, "fp64")
    amx = accel(E, st, cnts, "mixed")
    worst_rms = worst_max = 0.0
    for s in range(len(cnts)):
        lo, hi = int(off[s]), int(off[s + 1])
        A, B = a64[lo:hi], amx[lo:hi]
        rel = np.sqrt(((A - B) ** 2).sum(1) / (A ** 2).sum(1))
        worst_rms = max(worst_rms, np.sqrt((rel ** 2).mean()))
        worst_max = max(worst_max, rel.max())
    say(f"C5: mixed vs FP64, worst per-system RMS {worst_rms:.3e}, worst max {worst_max:.3e}")
    assert worst_rms <= 1e-6 and worst_max <= 1e-4
    order = np.argsort(cnts, kind="stable")
    for s in sorted({0, len(cnts) - 1, int(order[0]), int(order[-1])}):
        lo, hi = int(off[s]), int(off[s + 1])
        p = O.State(hi - lo)
        for k, ok in (("x", "x"), ("y", "y"), ("z", "z"), ("m", "m")):
            getattr(p, ok)[:] = st[k][lo:hi]
        ref, absum = O.bf_targets(p, EPS, np.arange(hi - lo), nthreads=16)
        err = (np.sqrt(((a64[lo:hi] - ref) ** 2).sum(1)) / absum).max()
        say(f"C5: system {s} (N = {hi - lo}) FP64 vs oracle, condition-normalised {err:.2e}")
        assert err <= 1e-12, s


def test_c5_forward_1000_steps(C5):
    E, cnts, st, off = C5["E"], C5["cnts"], C5["st"], C5["off"]
    s64 = {k: v.copy() for k, v in st.items()}
    smx = {k: v.copy() for k, v in st.items()}
    for d, prec in ((s64, "fp64"), (smx, "mixed")):
        a = accel(E, d, cnts, prec)
        d["ax"][:], d["ay"][:], d["az"][:] = a[:, 0], a[:, 1], a[:, 2]
        say(f"C5: {prec} hipGraph forward, {T} steps")
        E.run(d, cnts, EPS, DT, T, prec, use_graph=True)
    worst = 0.0
    for s in range(len(cnts)):
        lo, hi = int(off[s]), int(off[s + 1])
        d2 = sum((smx[k][lo:hi] - s64[k][lo:hi]) ** 2 for k in ("x", "y", "z"))
        worst = max(worst, np.sqrt(d2.mean()))
    say(f"C5: mixed vs FP64 after {T} steps, worst per-system RMS position difference "
        f"{worst:.3e}")
    assert worst <= 1e-5
    order = np.argsort(cnts, kind="stable")
    for s in (int(order[0]), int(order[-1])):
        lo, hi = int(off[s]), int(off[s + 1])
        p = O.State(hi - lo)
        for k in ("x", "y", "z", "vx", "vy", "vz", "m"):
            getattr(p, k)[:] = st[k][lo:hi]
        say(f"C5: oracle forward, system {s} (N = {hi - lo}), {T} steps")
        O.ens_forward(p, EPS, DT, T)
        d = max(np.abs(s64[k][lo:hi] - getattr(p, k)).max() for k in ("x", "y", "z"))
        say(f"C5: system {s} FP64 GPU vs oracle after {T} steps: max |dx| {d:.2e}")
        assert d <= 1e-8, s


@pytest.fixture(scope="module")
def adjoints(C5):
    E, cnts, st, off, tg, pick = (C5[k] for k in ("E", "cnts", "st", "off", "tg", "pick"))
    sc = [int(cnts[s]) for s in pick]
    sst = sub(st, off, pick)
    stg = [np.concatenate([t[int(off[s]):int(off[s + 1])] for s in pick]) for t in tg]
    say(f"C5: full 256-system mixed adjoint, {T} steps, checkpoints every {CK}")
    full = E.adjoint(st, cnts, EPS, DT, T, targets=tg, checkpoint_every=CK, precision="mixed")
    say("C5: four-system adjoints (mixed, FP64)")
    mx = E.adjoint(sst, sc, EPS, DT, T, targets=stg, checkpoint_every=CK, precision="mixed")
    f64 = E.adjoint(sst, sc, EPS, DT, T, targets=stg, checkpoint_every=CK, precision="fp64")
    return dict(full=full, mx=mx, f64=f64, sc=sc, sst=sst, stg=stg)


def test_c5_adjoint_systems_independent(C5, adjoints):
    off, pick = C5["off"], C5["pick"]
    full, mx, sc = adjoints["full"], adjoints["mx"], adjoints["sc"]
    soff = np.concatenate([[0], np.cumsum(sc)])
    for q, s in enumerate(pick):
        lo, hi = int(off[s]), int(off[s + 1])
        for k in FIELDS:
            assert np.array_equal(full[k][lo:hi], mx[k][soff[q]:soff[q + 1]]), (s, k)
        assert full["loss"][s] == mx["loss"][q]


def test_c5_mixed_adjoint_vs_fp64_adjoint(C5, adjoints):
    mx, f64, sc = adjoints["mx"], adjoints["f64"], adjoints["sc"]
    soff = np.concatenate([[0], np.cumsum(sc)])
    worst = 0.0
    for q in range(len(sc)):
        a, b = soff[q], soff[q + 1]
        for k in FIELDS:
            e = np.linalg.norm(mx[k][a:b] - f64[k][a:b]) / np.linalg.norm(f64[k][a:b])
            worst = max(worst, e)
    say(f"C5: mixed-forward vs FP64-forward adjoint, worst |dg|/|g| {worst:.2e}")
    assert worst <= 1e-4


def test_c5_adjoint_vs_oracle_smallest(C5, adjoints):
    f64, sst, stg, sc = adjoints["f64"], adjoints["sst"], adjoints["stg"], adjoints["sc"]
    n0 = sc[0]
    p = O.State(n0)
    for k in FIELDS:
        getattr(p, k)[:] = sst[k][:n0]
    f = p.copy()
    say(f"C5: oracle forward + adjoint, N = {n0}, {T} steps")
    O.ens_forward(f, EPS, DT, T)
    g = [2 * (getattr(f, k) - stg[i][:n0]) for i, k in enumerate(("x", "y", "z"))]
    ref = O.ens_adjoint(p, EPS, DT, T, *g)
    worst = 0.0
    for k in FIELDS:
        e = np.abs(f64[k][:n0] - ref[k]).max() / np.abs(ref[k]).max()
        worst = max(worst, e)
        assert e <= 1e-8, k
    say(f"C5: GPU FP64 adjoint vs oracle adjoint, worst {worst:.2e}")


def test_c5_adjoint_vs_finite_differences(C5, adjoints):
    E = C5["E"]
    f64, sst, stg, sc = adjoints["f64"], adjoints["sst"], adjoints["stg"], adjoints["sc"]
    soff = np.concatenate([[0], np.cumsum(sc)]).astype(np.int64)
    rng = np.random.default_rng(5)
    probes = []
    for q in range(len(sc)):
        for _ in range(64):
            probes.append((q, FIELDS[rng.integers(len(FIELDS))],
                           int(rng.integers(soff[q], soff[q + 1]))))
    h = 1e-6
    big_counts, big = [], {k: [] for k in FIELDS}
    for q, f, i in probes:
        lo, hi = soff[q], soff[q + 1]
        for sign in (+1, -1):
            for k in FIELDS:
                v = sst[k][lo:hi].copy()
                if k == f:
                    v[i - lo] += sign * h
                big[k].append(v)
            big_counts.append(int(hi - lo))
    bst = {k: np.concatenate(v) for k, v in big.items()}
    a = accel(E, bst, big_counts, "fp64")
    bst["ax"], bst["ay"], bst["az"] = (np.ascontiguousarray(a[:, r]) for r in range(3))
    say(f"C5: FD forward, {len(big_counts)} perturbed systems, {T} steps (FP64)")
    E.run(bst, big_counts, EPS, DT, T, "fp64", use_graph=True)
    boff = E._offsets(big_counts)
    ad, fd = {}, {}
    for n, (q, f, i) in enumerate(probes):
        lo = soff[q]
        L = []
        for k in range(2):
            b = 2 * n + k
            a0, a1 = int(boff[b]), int(boff[b + 1])
            L.append(sum(((bst[c][a0:a1] - stg[r][lo:lo + (a1 - a0)]) ** 2).sum()
                          for r, c in enumerate(("x", "y", "z"))))
        ad.setdefault(q, []).append(f64[f][i])
        fd.setdefault(q, []).append((L[0] - L[1]) / (2 * h))
    worst = 0.0
    for q in ad:
        x, y = np.array(ad[q]), np.array(fd[q])
        rel = np.linalg.norm(x - y) / np.linalg.norm(y)
        say(f"C5: system {C5['pick'][q]} (N = {sc[q]}): |adj - fd| / |fd| = {rel:.2e}")
        worst = max(worst, rel)
    assert worst <= 1e-5
