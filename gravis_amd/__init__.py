"""gravis_amd -- MI355X (gfx950) force-and-integration engine for gravis.

Python mirror of sim-core's solver interface over the C ABI
(include/gravis.h, libgravis_hip.so). Names, argument meaning and error
behaviour follow the reference:

  Particles                 crates/sim-core/src/particle.rs:21-227
  GravitySolver             crates/sim-core/src/gravity.rs:6-10 (accumulate)
  NoGravity, BruteForce     crates/sim-core/src/gravity.rs:13-80
  BarnesHut                 crates/sim-core/src/barnes_hut.rs:21-192
  LeapfrogKDK               crates/sim-core/src/integrator.rs:5-77
  diagnostics.compute/_fast crates/sim-core/src/diagnostics.rs:37-48

Every force, integration and diagnostic runs on the GPU through the HIP
library; nothing here computes physics on the CPU, and a missing library
raises instead of falling back.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import GravisError

__all__ = ["scenarios", "Particles", "Context", "GravitySolver", "NoGravity", "BruteForce", "BarnesHut",
           "LeapfrogKDK", "diagnostics", "GravisError", "default_context", "render_pack",
           "GPU_PARTICLE"]


class Context:
    """Owns a gravis_ctx (device buffers, stream, tree arena)."""

    def __init__(self, device: int = 0):
        self._L = _lib.load()
        h = C.c_void_p()
        st = self._L.gravis_create(device, C.byref(h))
        if st != _lib.OK:
            raise GravisError(st, f"gravis_create(device={device}) failed: "
                              + self._L.gravis_status_string(st).decode())
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            self._L.gravis_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, st):
        _lib.check(self.handle, st)

    def synchronize(self):
        self.check(self._L.gravis_synchronize(self.handle))

    def set_stream(self, stream_ptr: int | None):
        self.check(self._L.gravis_set_stream(self.handle, C.c_void_p(stream_ptr or 0)))

    def set_profiling(self, on: bool):
        self.check(self._L.gravis_set_profiling(self.handle, 1 if on else 0))

    def set_tuning(self, key: str, value: int):
        self.check(self._L.gravis_set_tuning(self.handle, key.encode(), int(value)))

    def tuning(self) -> dict:
        """Current launch shapes (gravis_get_tuning)."""
        t = _lib.Tuning()
        self.check(self._L.gravis_get_tuning(self.handle, C.byref(t)))
        return {k: getattr(t, k) for k, _ in _lib.Tuning._fields_ if k != "best_ms"}

    def autotune(self, gravity: "GravitySolver", p: "Particles") -> dict:
        """Time the candidate launch shapes of `gravity` on p's positions and
        masses and keep the fastest in this context (gravis_autotune)."""
        t = _lib.Tuning()
        d = gravity.desc()
        self.check(self._L.gravis_autotune(self.handle, C.byref(d), p.count,
                                           *[_ptr(getattr(p, f)) for f in ("x", "y", "z",
                                                                           "mass")],
                                           C.byref(t)))
        return {k: getattr(t, k) for k, _ in _lib.Tuning._fields_}

    def bh_stats(self) -> dict:
        """Work counters of the last Barnes-Hut call."""
        out = (C.c_uint64 * 5)()
        self.check(self._L.gravis_bh_stats(self.handle, C.byref(out)))
        return dict(zip(("node_tests", "accepted_nodes", "leaf_particles", "pairs", "nodes"),
                        list(out)))

    def kernel_time(self, name: str):
        ms = C.c_double()
        n = C.c_uint64()
        self.check(self._L.gravis_kernel_time(self.handle, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value


_DEFAULT = None


def default_context() -> Context:
    global _DEFAULT
    if _DEFAULT is None:
        _DEFAULT = Context(0)
    return _DEFAULT


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.c_void_p)


def _u8(a: np.ndarray):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(C.c_void_p)


class Particles:
    """Structure-of-arrays particle storage (particle.rs:21-64), f64, with
    the SPH fields and particle types (particle.rs:2-10; 4 = gas)."""

    FIELDS = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "mass")
    SPH_FIELDS = ("density", "pressure", "internal_energy", "du_dt", "smoothing_length",
                  "sound_speed", "alpha_visc")
    GAS = 4

    def __init__(self, capacity: int = 0):
        self.count = 0
        for f in self.FIELDS + self.SPH_FIELDS:
            setattr(self, f, np.zeros(0, dtype=np.float64))
        self.particle_type = np.zeros(0, dtype=np.uint8)

    @classmethod
    def from_arrays(cls, x, y, z, vx, vy, vz, mass):
        p = cls()
        n = len(x)
        p.count = n
        for f, v in zip(("x", "y", "z", "vx", "vy", "vz", "mass"), (x, y, z, vx, vy, vz, mass)):
            setattr(p, f, np.ascontiguousarray(v, dtype=np.float64).copy())
        for f in ("ax", "ay", "az") + cls.SPH_FIELDS:
            setattr(p, f, np.zeros(n))
        p.particle_type = np.zeros(n, dtype=np.uint8)
        return p

    @classmethod
    def zeros(cls, n: int):
        p = cls()
        p.count = n
        for f in cls.FIELDS + cls.SPH_FIELDS:
            setattr(p, f, np.zeros(n))
        p.particle_type = np.zeros(n, dtype=np.uint8)
        return p

    def _append(self, values: dict, ptype: int) -> int:
        for f in self.FIELDS + self.SPH_FIELDS:
            setattr(self, f, np.append(getattr(self, f), float(values.get(f, 0.0))))
        self.particle_type = np.append(self.particle_type, np.uint8(ptype))
        self.count += 1
        return self.count - 1

    def add(self, x, y, z, vx, vy, vz, mass) -> int:
        """particle.rs:94-96: append one particle, a = 0; returns its index."""
        return self._append(dict(x=x, y=y, z=z, vx=vx, vy=vy, vz=vz, mass=mass), 0)

    def add_gas(self, x, y, z, vx, vy, vz, mass, internal_energy, smoothing_length) -> int:
        """particle.rs:136-154: a gas particle with alpha_visc = 1."""
        return self._append(dict(x=x, y=y, z=z, vx=vx, vy=vy, vz=vz, mass=mass,
                                 internal_energy=internal_energy,
                                 smoothing_length=smoothing_length, alpha_visc=1.0), self.GAS)

    def is_gas(self, i: int) -> bool:
        return int(self.particle_type[i]) == self.GAS

    def has_gas(self) -> bool:
        return bool((self.particle_type == self.GAS).any())

    def clear_accelerations(self):
        """particle.rs:173-179."""
        self.ax[:] = 0.0
        self.ay[:] = 0.0
        self.az[:] = 0.0

    def clear_sph_rates(self):
        """particle.rs:182-186."""
        self.du_dt[:] = 0.0

    def copy(self) -> "Particles":
        p = Particles()
        p.count = self.count
        for f in self.FIELDS + self.SPH_FIELDS:
            setattr(p, f, getattr(self, f).copy())
        p.particle_type = self.particle_type.copy()
        return p

    def acc(self) -> np.ndarray:
        return np.stack([self.ax, self.ay, self.az], axis=1)

    def pos(self) -> np.ndarray:
        return np.stack([self.x, self.y, self.z], axis=1)


class GravitySolver:
    """gravity.rs:6-10. compute_accelerations ACCUMULATES into p.ax/ay/az;
    the caller zeroes them first."""

    algorithm = _lib.NO_GRAVITY

    def __init__(self, softening: float = 0.0, theta: float = 0.0, precision: str = "fp64",
                 leaf_size: int = 16, group_size: int = 64, ctx: Context | None = None):
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.softening = float(softening)
        self.theta = float(theta)
        self.precision = precision
        self.leaf_size = leaf_size
        self.group_size = group_size
        self.ctx = ctx or default_context()

    def desc(self) -> _lib.Solver:
        return _lib.Solver(self.algorithm, _lib.MIXED if self.precision == "mixed" else _lib.FP64,
                           self.softening, self.theta, self.leaf_size, self.group_size)

    def compute_accelerations(self, p: Particles) -> None:
        d = self.desc()
        L = self.ctx._L
        self.ctx.check(L.gravis_accel(self.ctx.handle, C.byref(d), p.count, _ptr(p.x), _ptr(p.y),
                                      _ptr(p.z), _ptr(p.mass), _ptr(p.ax), _ptr(p.ay),
                                      _ptr(p.az)))

    def compute_accelerations_device(self, n: int, x, y, z, m, ax, ay, az) -> None:
        """Device-pointer form (ints from torch's data_ptr()); asynchronous on
        the context stream."""
        d = self.desc()
        L = self.ctx._L
        self.ctx.check(L.gravis_accel_device(self.ctx.handle, C.byref(d), n,
                                             *[C.c_void_p(v) for v in (x, y, z, m, ax, ay, az)]))


class NoGravity(GravitySolver):
    """gravity.rs:13-19."""
    algorithm = _lib.NO_GRAVITY

    def __init__(self, ctx: Context | None = None):
        super().__init__(0.0, 0.0, "fp64", ctx=ctx)


class BruteForce(GravitySolver):
    """gravity.rs:36-80. precision="mixed" selects FP32 pair forces on FP64
    tile-relative separations with compensated accumulation."""
    algorithm = _lib.BRUTE_FORCE

    def __init__(self, softening: float, precision: str = "fp64", ctx: Context | None = None):
        super().__init__(softening, 0.0, precision, ctx=ctx)


class BarnesHut(GravitySolver):
    """barnes_hut.rs:21-40, linear octree with leaf_size particles per leaf
    and wavefront groups of group_size targets."""
    algorithm = _lib.BARNES_HUT

    def __init__(self, softening: float, theta: float, leaf_size: int = 16,
                 group_size: int = 64, precision: str = "fp64", ctx: Context | None = None):
        super().__init__(softening, theta, precision, leaf_size, group_size, ctx)


class LeapfrogKDK:
    """integrator.rs:41-77. step() is one Integrator::step; run() fuses
    nsteps on the device (state stays in HBM between steps)."""

    def step(self, p: Particles, gravity: GravitySolver, dt: float) -> None:
        self.run(p, gravity, dt, 1)

    def run(self, p: Particles, gravity: GravitySolver, dt: float, nsteps: int) -> None:
        d = gravity.desc()
        ctx = gravity.ctx
        ctx.check(ctx._L.gravis_kdk(ctx.handle, C.byref(d), p.count,
                                    *[_ptr(getattr(p, f)) for f in ("x", "y", "z", "vx", "vy",
                                                                    "vz", "mass", "ax", "ay",
                                                                    "az")],
                                    float(dt), int(nsteps)))


class _Diagnostics:
    """diagnostics.rs:37-48 (FP64 on the GPU)."""

    @staticmethod
    def _run(p: Particles, softening: float, with_pot: bool, time: float, step: int,
             ctx: Context | None):
        ctx = ctx or default_context()
        out = _lib.Diag()
        gas = p.has_gas()
        ctx.check(ctx._L.gravis_diagnostics(ctx.handle, p.count, float(softening),
                                            1 if with_pot else 0,
                                            *[_ptr(getattr(p, f)) for f in ("x", "y", "z", "vx",
                                                                            "vy", "vz", "mass")],
                                            _ptr(p.internal_energy) if gas else None,
                                            _u8(p.particle_type) if gas else None,
                                            float(time), int(step), C.byref(out)))
        return out

    def compute(self, p: Particles, softening: float, time: float = 0.0, step: int = 0,
                ctx: Context | None = None):
        return self._run(p, softening, True, time, step, ctx)

    def compute_fast(self, p: Particles, time: float = 0.0, step: int = 0,
                     ctx: Context | None = None):
        return self._run(p, 0.0, False, time, step, ctx)


diagnostics = _Diagnostics()

# GpuParticle (crates/render-core/src/gpu_types.rs:5-9): 32 bytes
GPU_PARTICLE = np.dtype([("position", np.float32, 3), ("mass", np.float32),
                         ("color", np.float32, 4)])


def render_pack(p: Particles, ctx: Context | None = None) -> np.ndarray:
    """The render view of p: one GpuParticle per particle (positions, masses
    and internal energies rounded to f32 as the sim thread does,
    crates/native-app/src/sim_thread.rs:250-258; colours by particle_color,
    crates/render-core/src/color.rs:9-52), packed on the GPU
    (gravis_render_pack). Returns a structured array of GPU_PARTICLE."""
    ctx = ctx or default_context()
    out = np.zeros(p.count, dtype=GPU_PARTICLE)
    ctx.check(ctx._L.gravis_render_pack(ctx.handle, p.count,
                                        *[_ptr(getattr(p, f)) for f in ("x", "y", "z", "mass")],
                                        _ptr(p.internal_energy), _u8(p.particle_type),
                                        out.ctypes.data_as(C.c_void_p)))
    return out

# Particles field -> gravis_sph_state member
_SPH_MAP = dict(x="x", y="y", z="z", vx="vx", vy="vy", vz="vz", ax="ax", ay="ay", az="az",
                m="mass", rho="density", p="pressure", u="internal_energy", dudt="du_dt",
                h="smoothing_length", c="sound_speed", alpha="alpha_visc")


class _Snapshot:
    """sim-core snapshots (crates/sim-core/src/snapshot.rs:5-157)."""

    @staticmethod
    def write(path: str, p: Particles, time: float, step: int, softening: float,
              dt: float) -> None:
        """Snapshot::capture + write_to (NBODY003)."""
        _lib.check(None, _lib.load().gravis_snapshot_write(path.encode(), time, step, softening,
                                                            dt, C.byref(sph_state(p))))

    @staticmethod
    def read(path: str):
        """Snapshot::read_from: returns (particles, time, step, softening, dt,
        version); accelerations are zero (recompute before stepping)."""
        L = _lib.load()
        v = C.c_int32()
        t, so, dt = C.c_double(), C.c_double(), C.c_double()
        st_, cnt = C.c_uint64(), C.c_uint64()
        _lib.check(None, L.gravis_snapshot_info(path.encode(), C.byref(v), C.byref(t),
                                                C.byref(st_), C.byref(so), C.byref(dt),
                                                C.byref(cnt)))
        p = Particles.zeros(cnt.value)
        _lib.check(None, L.gravis_snapshot_read(path.encode(), C.byref(sph_state(p))))
        return p, t.value, st_.value, so.value, dt.value, v.value


snapshot = _Snapshot()


def sph_state(p: Particles) -> _lib.SphState:
    """gravis_sph_state over the host arrays of p (borrowed, not copied)."""
    st = _lib.SphState()
    st.n = p.count
    for k in _lib.SPH_FIELDS:
        setattr(st, k, _ptr(getattr(p, _SPH_MAP[k])).value)
    st.type = _u8(p.particle_type).value
    return st


def morton_sort(x, y, z, ctx: Context | None = None):
    """Device Morton keys (sorted) + stable permutation + root cube."""
    ctx = ctx or default_context()
    n = len(x)
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    keys = np.zeros(n, dtype=np.uint64)
    perm = np.zeros(n, dtype=np.uint32)
    box = np.zeros(4)
    ctx.check(ctx._L.gravis_morton_sort(ctx.handle, n, _ptr(x), _ptr(y), _ptr(z),
                                        keys.ctypes.data_as(C.c_void_p),
                                        perm.ctypes.data_as(C.c_void_p), _ptr(box)))
    return keys, perm, box


class _Scenarios:
    """sim-core scenario generators (crates/sim-core/src/scenarios/), same
    seeds give the reference's particles."""

    @staticmethod
    def plummer(n: int, seed: int = 42, scale_radius: float = 1.0,
                total_mass: float = 1.0) -> Particles:
        """PlummerSphere { n, seed, .. }.generate() (plummer_sphere.rs:106-133)."""
        L = _lib.load()
        a = [np.zeros(n) for _ in range(7)]
        _lib.check(None, L.gravis_scenario_plummer(n, scale_radius, total_mass, seed,
                                                   *[_ptr(v) for v in a]))
        return Particles.from_arrays(*a)

    @staticmethod
    def plummer_suggested_dt(scale_radius: float = 1.0, total_mass: float = 1.0) -> float:
        return _lib.load().gravis_plummer_suggested_dt(scale_radius, total_mass)

    @staticmethod
    def plummer_suggested_softening(n: int, scale_radius: float = 1.0) -> float:
        return _lib.load().gravis_plummer_suggested_softening(n, scale_radius)

    @staticmethod
    def two_body(eccentricity: float = 0.5, mass_ratio: float = 0.5,
                 semi_major_axis: float = 1.0) -> Particles:
        """TwoBody::generate (two_body.rs:49-92)."""
        L = _lib.load()
        a = [np.zeros(2) for _ in range(7)]
        _lib.check(None, L.gravis_scenario_two_body(eccentricity, mass_ratio, semi_major_axis,
                                                    *[_ptr(v) for v in a]))
        return Particles.from_arrays(*a)

    @staticmethod
    def plummer_pair(n_each: int, seed_a: int = 42, seed_b: int = 43, separation: float = 10.0,
                     impact: float = 0.0, scale_radius: float = 1.0,
                     total_mass: float = 1.0) -> Particles:
        """Two Plummer spheres on a parabolic approach (layout of
        galaxy_collision.rs:134-154)."""
        L = _lib.load()
        a = [np.zeros(2 * n_each) for _ in range(7)]
        _lib.check(None, L.gravis_scenario_plummer_pair(n_each, scale_radius, total_mass, seed_a,
                                                        seed_b, separation, impact,
                                                        *[_ptr(v) for v in a]))
        return Particles.from_arrays(*a)

    @staticmethod
    def two_body_period(semi_major_axis: float = 1.0) -> float:
        return _lib.load().gravis_two_body_period(semi_major_axis)

    @staticmethod
    def evrard(n: int, seed: int = 42, total_mass: float = 1.0, radius: float = 1.0,
               u0: float = 0.05) -> Particles:
        """EvrardCollapse::generate (evrard_collapse.rs:45-57): n gas
        particles uniform in a sphere, h = 1.5 (V/n)^(1/3), at rest."""
        L = _lib.load()
        p = Particles.zeros(n)
        st = sph_state(p)
        _lib.check(None, L.gravis_scenario_evrard(n, total_mass, radius, u0, seed, C.byref(st)))
        return p

    @staticmethod
    def evrard_1r(n: int, seed: int = 42, total_mass: float = 1.0, radius: float = 1.0,
                  u0: float = 0.05, eta: float = 1.2) -> Particles:
        """Gas sphere with rho ~ 1/r (no reference counterpart): r = R sqrt(U),
        isotropic, at rest, h = eta (m / rho(r))^(1/3)."""
        L = _lib.load()
        p = Particles.zeros(n)
        st = sph_state(p)
        _lib.check(None, L.gravis_scenario_evrard_1r(n, total_mass, radius, u0, eta, seed,
                                                     C.byref(st)))
        return p

    @staticmethod
    def cold_collapse(n: int, seed: int = 42, radius: float = 1.0, total_mass: float = 1.0,
                      gas: bool = False) -> Particles:
        """ColdCollapse::generate (cold_collapse.rs:56-102)."""
        L = _lib.load()
        p = Particles.zeros(n)
        st = sph_state(p)
        _lib.check(None, L.gravis_scenario_cold_collapse(n, radius, total_mass, seed,
                                                         1 if gas else 0, C.byref(st)))
        return p

    @staticmethod
    def galaxy_params(**kw) -> _lib.GalaxyParams:
        """GalaxyCollision::default() (galaxy_collision.rs:51-68) with the
        given fields replaced."""
        g = _lib.GalaxyParams()
        _lib.load().gravis_galaxy_default_params(C.byref(g))
        for k, v in kw.items():
            if not hasattr(g, k):
                raise TypeError(f"GalaxyCollision has no field {k!r}")
            setattr(g, k, v)
        return g

    @staticmethod
    def galaxy_collision(**kw) -> Particles:
        """GalaxyCollision::generate (galaxy_collision.rs:75-157); keyword
        arguments are GalaxyCollision fields. 2 * n_per_galaxy particles of
        types 1 (disk), 2 (bulge), 3 (halo)."""
        L = _lib.load()
        g = _Scenarios.galaxy_params(**kw)
        p = Particles.zeros(2 * g.n_per_galaxy)
        st = sph_state(p)
        _lib.check(None, L.gravis_scenario_galaxy_collision(C.byref(g), C.byref(st)))
        return p

    @staticmethod
    def galaxy_suggested_dt(**kw) -> float:
        return _lib.load().gravis_galaxy_suggested_dt(C.byref(_Scenarios.galaxy_params(**kw)))

    @staticmethod
    def galaxy_suggested_softening(**kw) -> float:
        return _lib.load().gravis_galaxy_suggested_softening(
            C.byref(_Scenarios.galaxy_params(**kw)))

    @staticmethod
    def evrard_softening(n: int, radius: float = 1.0) -> float:
        """evrard_collapse.rs:67-70: 0.3 (V/n)^(1/3)."""
        return _lib.load().gravis_evrard_softening(n, radius)


scenarios = _Scenarios()
