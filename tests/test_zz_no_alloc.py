"""gravis_accel_device_noalloc: the brute-force force call whose scratch
lives in the library's static device array (no device allocation), against
gravis_accel_device on the same inputs, through the C ABI. The entry point allocates nothing: every scratch
buffer (records, Morton-sort buffers, j-chunk partials) is carved out of
one static 20 MiB device array per GPU; n is limited to 65,536.

Two calls in a row on one context, the second with a smaller input (the
array keeps the first call's contents), for FP64 and for mixed precision:
  * FP64: |a_noalloc - a| / sum_j |a_ij| <= 1e-12 per particle (the j split
    differs, so the summation grouping does);
  * mixed: RMS |da|/|a| <= 1e-6, max <= 1e-4 against the same entry point's
    mixed result;
  * results are ADDED to ax/ay/az (accumulate contract, gravity.rs:6-10);
  * n above the stated limit (65,536) and Barnes-Hut are refused with a
    status and a message, and the context still works afterwards.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

NMAX = 65536


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def env():
    import torch
    import gravis_amd as g
    from gravis_amd import _lib
    L = _lib.load()
    f = L.gravis_accel_device_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.POINTER(_lib.Solver), C.c_size_t] + [C.c_void_p] * 7
    ctx = g.Context(0)  # its own stream; every call below synchronises the device
    yield dict(torch=torch, g=g, lib=_lib, L=L, ctx=ctx)
    ctx.close()


def run(env, entry, solver, s, base):
    """One force call through `entry` on device copies of s; a starts at
    `base` (the call adds to it)."""
    torch = env["torch"]
    T = {k: torch.from_numpy(getattr(s, k).copy()).cuda() for k in ("x", "y", "z", "m")}
    A = [torch.from_numpy(base[:, r].copy()).cuda() for r in range(3)]
    torch.cuda.synchronize()
    st = entry(env["ctx"].handle, C.byref(solver), s.n,
               *[C.c_void_p(T[k].data_ptr()) for k in ("x", "y", "z", "m")],
               *[C.c_void_p(a.data_ptr()) for a in A])
    torch.cuda.synchronize()
    return st, np.stack([a.cpu().numpy() for a in A], 1)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_matches_accel_device(env, precision):
    lib, L = env["lib"], env["L"]
    prec = lib.MIXED if precision == "mixed" else lib.FP64
    solver = lib.Solver(lib.BRUTE_FORCE, prec, 0.01, 0.0, 0, 0)
    for n, seed in ((NMAX, 42), (20011, 7)):
        s = O.plummer(n, seed)
        base = np.random.default_rng(seed).normal(size=(n, 3))
        say(f"noalloc {precision} n={n}")
        st, got = run(env, L.gravis_accel_device_noalloc, solver, s, base)
        assert st == lib.OK, L.gravis_last_error(env["ctx"].handle)
        st, ref = run(env, L.gravis_accel_device, solver, s, base)
        assert st == lib.OK
        got, ref = got - base, ref - base
        if precision == "fp64":
            _, absum = O.bf_targets(s, 0.01, np.arange(n), nthreads=16)
            err = (np.sqrt(((got - ref) ** 2).sum(1)) / absum).max()
            say(f"  condition-normalised difference {err:.2e}")
            assert err <= 1e-12
        else:
            rel = np.sqrt(((got - ref) ** 2).sum(1) / (ref ** 2).sum(1))
            rms = float(np.sqrt((rel ** 2).mean()))
            say(f"  rms {rms:.2e} max {rel.max():.2e}")
            assert rms <= 1e-6 and rel.max() <= 1e-4


def test_noalloc_refuses_beyond_its_limits(env):
    lib, L, ctx = env["lib"], env["L"], env["ctx"]
    s = O.plummer(NMAX + 1, 3)
    base = np.zeros((s.n, 3))
    st, _ = run(env, L.gravis_accel_device_noalloc,
                lib.Solver(lib.BRUTE_FORCE, lib.FP64, 0.01, 0.0, 0, 0), s, base)
    assert st == lib.INVALID_ARG
    assert b"65536" in L.gravis_last_error(ctx.handle)
    s = O.plummer(1000, 3)
    st, _ = run(env, L.gravis_accel_device_noalloc,
                lib.Solver(lib.BARNES_HUT, lib.FP64, 0.01, 0.5, 16, 64), s, np.zeros((1000, 3)))
    assert st == lib.UNSUPPORTED
    st, a = run(env, L.gravis_accel_device_noalloc,
                lib.Solver(lib.BRUTE_FORCE, lib.FP64, 0.01, 0.0, 0, 0), s, np.zeros((1000, 3)))
    assert st == lib.OK and np.isfinite(a).all() and np.abs(a).max() > 0
