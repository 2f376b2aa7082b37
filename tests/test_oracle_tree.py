"""Pins the oracle's linear-octree Barnes-Hut definition (Morton keys,
stable sort, leaf-16 tree, group walk). The reference has no counterpart
(its octree is a one-particle-per-leaf insertion tree, octree.rs:156-239), so
these are pinned by brute-force restatements on small inputs, closed forms
and invariants:

  * spread/interleave  -> naive per-bit interleave in Python
  * quantisation       -> exact integer cell of hand-placed points
  * stable sort        -> numpy's stable argsort
  * tree structure     -> partition / leaf-size / level invariants, and the
                          cell of every particle matches its key prefix
  * FP64 COM sums      -> numpy sums in the stated order, and exact for
                          dyadic data
  * group walk         -> theta = 0 equals brute force; accuracy at theta
                          0.3..1.0 within the reference's published error
                          (m3_accuracy.csv)
"""
import os

import numpy as np

import oracle as O


def _naive_key(qx, qy, qz):
    k = 0
    for b in range(21):
        k |= ((qx >> b) & 1) << (3 * b)
        k |= ((qy >> b) & 1) << (3 * b + 1)
        k |= ((qz >> b) & 1) << (3 * b + 2)
    return k


def test_morton_interleave_matches_naive():
    rng = np.random.default_rng(1)
    box = np.array([0.5, 0.5, 0.5, 0.5])  # cube [0,1]^3, cell = 2^-21
    q = rng.integers(0, 2 ** 21, size=(200, 3))
    pts = (q + 0.5) / 2.0 ** 21  # cell centres: exact quantisation
    keys = O.morton63(pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy(), box)
    for (qx, qy, qz), k in zip(q.tolist(), keys.tolist()):
        assert k == _naive_key(qx, qy, qz)
    assert int(keys.max()) < 2 ** 63


def test_morton_clamps_and_corners():
    box = np.array([0.0, 0.0, 0.0, 1.0])  # [-1, 1]^3
    x = np.array([-1.0, 1.0, 5.0, -7.0, 0.0])
    keys = O.morton63(x, np.zeros(5), np.zeros(5), box)
    assert keys[0] == _naive_key(0, 2 ** 20, 2 ** 20)
    assert keys[1] == _naive_key(2 ** 21 - 1, 2 ** 20, 2 ** 20)
    assert keys[2] == keys[1] and keys[3] == keys[0]
    assert keys[4] == _naive_key(2 ** 20, 2 ** 20, 2 ** 20)


def test_sort_is_stable_and_matches_numpy():
    rng = np.random.default_rng(2)
    keys = rng.integers(0, 50, size=5000).astype(np.uint64) << np.uint64(40)
    keys |= rng.integers(0, 3, size=5000).astype(np.uint64)
    sk, perm = O.sort_keys(keys)
    ref = np.argsort(keys, kind="stable")
    assert (perm.astype(np.int64) == ref).all()
    assert (sk == keys[ref]).all()


def _check_tree(s, t, leaf):
    n = s.n
    lv, fi, cn, ch, nc = t["level"], t["first"], t["cnt"], t["child"], t["nchild"]
    assert fi[0] == 0 and cn[0] == n and lv[0] == 0
    keys = t["keys"]
    leaves_cover = np.zeros(n, dtype=np.int64)
    for v in range(len(lv)):
        if nc[v] == 0:
            assert cn[v] <= leaf or lv[v] == 21
            leaves_cover[fi[v]:fi[v] + cn[v]] += 1
        else:
            assert cn[v] > leaf and lv[v] < 21
            kids = range(ch[v], ch[v] + nc[v])
            assert fi[ch[v]] == fi[v]
            assert sum(int(cn[c]) for c in kids) == cn[v]
            for c in kids:
                assert lv[c] == lv[v] + 1
                assert cn[c] > 0
            for a, b in zip(list(kids)[:-1], list(kids)[1:]):
                assert fi[a] + cn[a] == fi[b]
        # every particle of the node shares the node's key prefix
        sh = np.uint64(3 * (21 - int(lv[v])))
        pre = keys[fi[v]:fi[v] + cn[v]] >> sh
        assert (pre == pre[0]).all()
    assert (leaves_cover == 1).all()
    # BFS order: levels non-decreasing
    assert (np.diff(lv.astype(np.int64)) >= 0).all()


def test_linear_tree_invariants_plummer():
    s = O.plummer(3000, 5)
    t = O.linear_tree(s, 16)
    _check_tree(s, t, 16)
    perm = t["perm"].astype(np.int64)
    sm = s.m[perm]
    sx = s.x[perm]
    root = t["com"][0]
    assert abs(root[3] - sm.sum()) < 1e-13
    assert abs(root[0] - (sm * sx).sum() / sm.sum()) < 1e-13


def test_linear_tree_leaf_com_exact_on_dyadic_data():
    """Dyadic positions and masses make every FP64 sum exact, so each node's
    COM is exactly the mean of its particles."""
    rng = np.random.default_rng(3)
    n = 1000
    s = O.State(n)
    s.x[:] = rng.integers(-512, 512, n) / 256.0
    s.y[:] = rng.integers(-512, 512, n) / 256.0
    s.z[:] = rng.integers(-512, 512, n) / 256.0
    s.m[:] = rng.integers(1, 4, n) / 4.0
    t = O.linear_tree(s, 16)
    _check_tree(s, t, 16)
    perm = t["perm"].astype(np.int64)
    for v in range(len(t["level"])):
        idx = perm[t["first"][v]:t["first"][v] + t["cnt"][v]]
        M = s.m[idx].sum()
        assert t["com"][v][3] == M
        for d, arr in enumerate((s.x, s.y, s.z)):
            mx = (s.m[idx] * arr[idx]).sum()
            assert t["com"][v][d] == mx / M


def test_linear_tree_coincident_points_stop_at_level_21():
    s = O.State(40)
    s.x[:20] = 0.25
    s.x[20:] = -0.25
    s.m[:] = 1.0
    t = O.linear_tree(s, 16)
    _check_tree(s, t, 16)
    assert t["level"].max() == 21
    assert t["cnt"][t["nchild"] == 0].max() == 20


def _rms_max(ref_acc, acc):
    m2 = (ref_acc ** 2).sum(1)
    e2 = ((acc - ref_acc) ** 2).sum(1)
    ok = m2 >= 1e-30
    rel2 = e2[ok] / m2[ok]
    return float(np.sqrt(rel2.sum() / len(ref_acc))), float(np.sqrt(rel2.max()))


def test_group_walk_theta0_is_brute_force():
    s = O.plummer(700, 11)
    bf = s.copy()
    O.bf_accel(bf, 0.05)
    g = s.copy()
    O.bh_group_accel(g, 0.05, 0.0)
    scale = np.abs(bf.acc()).max()
    assert np.abs(g.acc() - bf.acc()).max() < 1e-12 * scale


def test_group_walk_within_reference_bh_error(golden_dir):
    import csv
    ref = {float(r[0]): (float(r[2]), float(r[3]))
           for r in list(csv.reader(open(os.path.join(golden_dir, "m3_accuracy.csv"))))[1:]}
    s = O.plummer(1000, 42)
    bf = s.copy()
    O.bf_accel(bf, 0.05)
    for theta, (rms_ref, max_ref) in ref.items():
        g = s.copy()
        O.bh_group_accel(g, 0.05, theta)
        rms, mx = _rms_max(bf.acc(), g.acc())
        assert rms <= rms_ref and mx <= max_ref, (theta, rms, mx)
        assert rms > 0.0


def test_group_walk_interaction_count_scales_n_log_n():
    counts = []
    for n in (4000, 16000):
        s = O.plummer(n, 42)
        counts.append(O.bh_group_accel(s, 0.05, 0.5) / n)
    # per-particle work grows far slower than N (brute force would be x4)
    assert counts[1] < 2.0 * counts[0]
