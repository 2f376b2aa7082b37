"""Batched ensemble of independent N-body systems and its reverse mode
(C ABI: gravis_ensemble_kdk / gravis_ensemble_adjoint).

The reference has no ensemble and no reverse mode; the forward of every
system is sim-core's LeapfrogKDK (crates/sim-core/src/integrator.rs:41-77)
with brute-force gravity (gravity.rs:47-80). `EnsembleKDK` is the reverse-mode
API: a torch.autograd.Function whose backward runs the hand-written adjoint
kernels on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from . import default_context, Context

FIELDS = ("x", "y", "z", "vx", "vy", "vz")


def _solver(eps: float, precision: str) -> _lib.Solver:
    return _lib.Solver(_lib.BRUTE_FORCE, _lib.MIXED if precision == "mixed" else _lib.FP64,
                       float(eps), 0.0, 0, 0)


def _offsets(counts) -> np.ndarray:
    off = np.zeros(len(counts) + 1, dtype=np.uint64)
    off[1:] = np.cumsum(np.asarray(counts, dtype=np.uint64))
    return off


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def counts(seed: int = 42, nsys: int = 256, lo: int = 512, hi: int = 8192) -> np.ndarray:
    """Seeded system sizes N_s in [lo, hi] (gravis_scenario_ensemble_counts)."""
    c = np.zeros(nsys, dtype=np.uint32)
    _lib.check(None, _lib.load().gravis_scenario_ensemble_counts(seed, nsys, lo, hi, _p(c)))
    return c


def targets(seed: int, n: int):
    """Seeded loss targets U[-1, 1)^3 (gravis_scenario_targets)."""
    t = [np.zeros(n) for _ in range(3)]
    _lib.check(None, _lib.load().gravis_scenario_targets(seed, n, *[_p(a) for a in t]))
    return t


def plummer_ensemble(cnts, seed0: int = 42) -> dict:
    """Concatenated Plummer spheres (a = 1, M = 1), system s seeded seed0 + s."""
    from . import scenarios
    parts = [scenarios.plummer(int(n), seed0 + s) for s, n in enumerate(cnts)]
    st = {k: np.concatenate([getattr(p, k) for p in parts]) for k in FIELDS}
    st["m"] = np.concatenate([p.mass for p in parts])
    n = len(st["m"])
    for k in ("ax", "ay", "az"):
        st[k] = np.zeros(n)
    return st


def accel(state: dict, counts, eps: float, precision: str = "fp64",
          ctx: Context | None = None) -> None:
    """a = F(x) for every system into state["ax"/"ay"/"az"] (assigns;
    gravis_ensemble_accel)."""
    ctx = ctx or default_context()
    off = _offsets(counts)
    s = _solver(eps, precision)
    ctx.check(ctx._L.gravis_ensemble_accel(
        ctx.handle, C.byref(s), len(counts), _p(off),
        *[_p(state[k]) for k in ("x", "y", "z", "m", "ax", "ay", "az")]))


def run(state: dict, counts, eps: float, dt: float, nsteps: int, precision: str = "fp64",
        use_graph: bool = True, ctx: Context | None = None) -> None:
    """nsteps of KDK for every system; state holds numpy arrays x..vz, m,
    ax..az (a must be a(t) on entry; use accel() first). In place."""
    ctx = ctx or default_context()
    off = _offsets(counts)
    s = _solver(eps, precision)
    ctx.check(ctx._L.gravis_ensemble_kdk(
        ctx.handle, C.byref(s), len(counts), _p(off),
        *[_p(state[k]) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "ax", "ay", "az")],
        float(dt), int(nsteps), 1 if use_graph else 0))


def adjoint(state0: dict, counts, eps: float, dt: float, nsteps: int, targets=None,
            grad_pos=None, grad_vel=None, checkpoint_every: int = 32, precision: str = "fp64",
            ctx: Context | None = None) -> dict:
    """Cotangents of (x0, v0, m). Either targets=(tx, ty, tz) for the loss
    sum |x(T) - target|^2 (per-system loss returned under "loss"), or
    grad_pos=(gx, gy, gz) [+ grad_vel] cotangents of the final state."""
    ctx = ctx or default_context()
    off = _offsets(counts)
    n = int(off[-1])
    s = _solver(eps, precision)
    out = {k: np.zeros(n) for k in ("x", "y", "z", "vx", "vy", "vz", "m")}
    loss = np.zeros(len(counts))
    final = np.zeros(6 * n)
    keep = []

    def arr3(t):
        if t is None:
            return [None, None, None]
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in t]
        keep.extend(a)
        return [_p(v) for v in a]
    tp, gp, vp = arr3(targets), arr3(grad_pos), arr3(grad_vel)
    ctx.check(ctx._L.gravis_ensemble_adjoint(
        ctx.handle, C.byref(s), len(counts), _p(off),
        *[_p(np.ascontiguousarray(state0[k], dtype=np.float64)) for k in FIELDS + ("m",)],
        float(dt), int(nsteps), int(checkpoint_every), *tp, *gp, *vp, _p(loss), _p(final),
        *[_p(out[k]) for k in ("x", "y", "z", "vx", "vy", "vz", "m")]))
    out["loss"] = loss
    out["final"] = {k: final[i * n:(i + 1) * n].copy() for i, k in enumerate(FIELDS)}
    return out


try:
    import torch

    class EnsembleKDK(torch.autograd.Function):
        """x_T, v_T = EnsembleKDK.apply(x0, v0, m, counts, eps, dt, nsteps,
        precision, checkpoint_every). x0/v0 are (3, n) float64 CUDA tensors
        (rows x, y, z), m is (n,). backward runs the GPU adjoint."""

        @staticmethod
        def forward(fctx, x0, v0, m, counts, eps, dt, nsteps, precision="fp64",
                    checkpoint_every=32):
            g = default_context()
            for t in (x0, v0, m):
                if t.dtype != torch.float64 or not t.is_cuda:
                    raise TypeError("EnsembleKDK expects float64 CUDA tensors")
            x0c, v0c, mc = x0.contiguous(), v0.contiguous(), m.contiguous()
            n = mc.numel()
            off = _offsets(counts)
            st = torch.cat([x0c, v0c], 0).clone()
            acc = torch.zeros(3, n, dtype=torch.float64, device=x0.device)
            s = _solver(eps, precision)
            torch.cuda.synchronize()
            ptr = lambda t: C.c_void_p(t.data_ptr())
            # a(x0), then nsteps KDK
            g.check(g._L.gravis_ensemble_accel_device(
                g.handle, C.byref(s), len(counts), _p(off), *[ptr(st[r]) for r in range(3)],
                ptr(mc), *[ptr(acc[r]) for r in range(3)]))
            g.check(g._L.gravis_ensemble_kdk_device(
                g.handle, C.byref(s), len(counts), _p(off), *[ptr(st[r]) for r in range(6)],
                ptr(mc), *[ptr(acc[r]) for r in range(3)], float(dt), int(nsteps), 1))
            g.synchronize()
            fctx.save_for_backward(x0c, v0c, mc)
            fctx.cfg = (list(counts), eps, dt, nsteps, precision, checkpoint_every)
            return st[:3].clone(), st[3:].clone()

        @staticmethod
        def backward(fctx, gx, gv):
            x0, v0, m = fctx.saved_tensors
            counts, eps, dt, nsteps, precision, ck = fctx.cfg
            g = default_context()
            n = m.numel()
            off = _offsets(counts)
            s = _solver(eps, precision)
            gx = (gx if gx is not None else torch.zeros_like(x0)).contiguous()
            out = torch.zeros(7, n, dtype=torch.float64, device=x0.device)
            ptr = lambda t: C.c_void_p(t.data_ptr())
            gvp = [ptr(gv[r]) for r in range(3)] if gv is not None else [None] * 3
            if gv is not None:
                gv = gv.contiguous()
                gvp = [ptr(gv[r]) for r in range(3)]
            torch.cuda.synchronize()
            g.check(g._L.gravis_ensemble_adjoint_device(
                g.handle, C.byref(s), len(counts), _p(off), *[ptr(x0[r]) for r in range(3)],
                *[ptr(v0[r]) for r in range(3)], ptr(m), float(dt), int(nsteps), int(ck),
                None, None, None, *[ptr(gx[r]) for r in range(3)], *gvp, None, None,
                *[ptr(out[r]) for r in range(7)]))
            g.synchronize()
            return out[:3], out[3:6], out[6], None, None, None, None, None, None
except ImportError:  # torch absent: the numpy entry points above still work
    EnsembleKDK = None
