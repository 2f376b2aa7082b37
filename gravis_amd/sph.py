"""Smoothed particle hydrodynamics on the GPU: sim-core's SphSolver and
step_with_sph (crates/sim-core/src/sph/solver.rs:20-212) over the C ABI.

    sph = SphSolver()
    gravity = BarnesHut(eps, 0.5)
    p.clear_accelerations(); gravity.compute_accelerations(p); sph.compute(p)
    dt = step_with_sph(p, gravity, sph, dt)

Everything runs in the HIP library (gravis_sph_compute / gravis_sph_step /
gravis_sph_run); there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from . import Context, GravitySolver, Particles, default_context, sph_state

__all__ = ["SphSolver", "ForceIntermediates", "accelerations", "step_with_sph", "run"]


@dataclass
class ForceIntermediates:
    """ForceIntermediates (sph/forces.rs:45-53); omega and nneigh are the
    grad-h factors and final neighbour-list lengths (test hooks)."""
    div_v: np.ndarray
    max_mu: np.ndarray
    dalpha_dt: np.ndarray
    omega: np.ndarray | None = None
    nneigh: np.ndarray | None = None


class SphSolver:
    """SphSolver::new (solver.rs:38-53) with ForceParams::default
    (forces.rs:30-41); attributes may be changed before a call.
    variant=True selects the build's physics variant (cubic spline, Newton h
    to 1e-4, fixed alpha, no conduction, gravity eps = h, FP32 pair loops)."""

    def __init__(self, ctx: Context | None = None, variant: bool = False):
        self.ctx = ctx or default_context()
        d = _lib.SphParams()
        if variant:
            self.ctx._L.gravis_sph_variant_params(C.byref(d))
        else:
            self.ctx._L.gravis_sph_default_params(C.byref(d))
        for k, t in _lib.SphParams._fields_:
            v = getattr(d, k)
            setattr(self, k, float(v) if t is C.c_double else int(v))

    def params(self) -> _lib.SphParams:
        return _lib.SphParams(*[getattr(self, k) for k, _ in _lib.SphParams._fields_])

    def compute(self, p: Particles, hooks: bool = False) -> ForceIntermediates:
        """SphSolver::compute (solver.rs:61-80): density, h, EOS, then forces
        ADDED to p.ax/ay/az and p.du_dt."""
        n = p.count
        out = ForceIntermediates(np.zeros(n), np.zeros(n), np.zeros(n))
        om = nn = None
        if hooks:
            out.omega = np.zeros(n)
            out.nneigh = np.zeros(n, dtype=np.uint32)
            om = out.omega.ctypes.data_as(C.c_void_p)
            nn = out.nneigh.ctypes.data_as(C.c_void_p)
        prm = self.params()
        st = sph_state(p)
        self.ctx.check(self.ctx._L.gravis_sph_compute(
            self.ctx.handle, C.byref(prm), C.byref(st), out.div_v.ctypes.data_as(C.c_void_p),
            out.max_mu.ctypes.data_as(C.c_void_p), out.dalpha_dt.ctypes.data_as(C.c_void_p),
            om, nn))
        return out

    def stats(self) -> dict:
        a = (C.c_uint64 * 4)()
        self.ctx.check(self.ctx._L.gravis_sph_stats(self.ctx.handle, C.byref(a)))
        return dict(dfs_groups=a[0], requeries=a[1], rounds=a[2], pool_entries=a[3])


def _gravity_desc(gravity: GravitySolver | None):
    if gravity is None:
        return None
    return C.byref(gravity.desc())


def accelerations(p: Particles, gravity: GravitySolver | None, sph: SphSolver) -> None:
    """The gravity call of step_with_sph (solver.rs:180-181), accumulating
    into p.ax/ay/az; with sph.soft_h the softening of target i is h_i."""
    if gravity is None:
        return
    prm = sph.params()
    st = sph_state(p)
    sph.ctx.check(sph.ctx._L.gravis_sph_accel(sph.ctx.handle, C.byref(prm),
                                              C.byref(gravity.desc()), C.byref(st)))


def step_with_sph(p: Particles, gravity: GravitySolver | None, sph: SphSolver,
                  dt: float) -> float:
    """step_with_sph (solver.rs:145-212); returns the next dt."""
    nxt = C.c_double(0.0)
    prm = sph.params()
    st = sph_state(p)
    sph.ctx.check(sph.ctx._L.gravis_sph_step(sph.ctx.handle, C.byref(prm), _gravity_desc(gravity),
                                             C.byref(st), float(dt), C.byref(nxt)))
    return nxt.value


def run(p: Particles, gravity: GravitySolver | None, sph: SphSolver, dt: float, nsteps: int,
        time: float = 0.0) -> tuple[float, float]:
    """nsteps of headless's SPH loop (crates/headless/src/main.rs:178-187):
    dt = min(dt_next, 1.5 dt) after each step. Returns (next dt, time)."""
    d = C.c_double(float(dt))
    t = C.c_double(float(time))
    prm = sph.params()
    st = sph_state(p)
    sph.ctx.check(sph.ctx._L.gravis_sph_run(sph.ctx.handle, C.byref(prm), _gravity_desc(gravity),
                                            C.byref(st), int(nsteps), C.byref(d), C.byref(t)))
    return d.value, t.value
