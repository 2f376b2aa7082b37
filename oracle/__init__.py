"""CPU oracle for the gravis force-and-integration path (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker or the timed CPU baseline. The product
(gravis_amd) never imports it.

The C restatement lives in gravis_oracle.c (reference citations per function
there); this module is a thin ctypes layer over liboracle.so returning numpy
arrays.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

BF, BH_REF, BH_LINEAR = 0, 1, 2

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_sz, _d, _u64, _u32, _i = C.c_size_t, C.c_double, C.c_uint64, C.c_uint32, C.c_int


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    path = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(path):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "liboracle.so"],
                       check=True, stdout=subprocess.DEVNULL)
    return path


def lib():
    global _LIB
    if _LIB is None:
        L = C.CDLL(build())
        sig = {
            "orc_chacha20_block": (None, [_u32p, _u64, _u64, _u32p]),
            "orc_rng_stream_u64": (None, [_u64, _sz, _u64p]),
            "orc_rng_seed_key": (None, [_u64, _u32p]),
            "orc_plummer": (None, [_sz, _d, _d, _u64] + [_dp] * 7),
            "orc_plummer_suggested_dt": (_d, [_d, _d]),
            "orc_plummer_suggested_softening": (_d, [_sz, _d]),
            "orc_two_body": (None, [_d, _d, _d] + [_dp] * 7),
            "orc_two_body_period": (_d, [_d]),
            "orc_shift_to_com": (None, [_sz] + [_dp] * 7),
            "orc_bf_accel": (None, [_sz, _d] + [_dp] * 7),
            "orc_bf_targets": (None, [_sz, _d] + [_dp] * 4 + [_sz, _i64p] + [_dp] * 4 + [_i]),
            "orc_bh_accel": (None, [_sz, _d, _d] + [_dp] * 7 + [_i]),
            "orc_ref_octree_summary": (_sz, [_sz] + [_dp] * 5),
            "orc_root_box": (None, [_sz, _dp, _dp, _dp, _dp]),
            "orc_morton63": (None, [_sz, _dp, _dp, _dp, _dp, _u64p]),
            "orc_sort_keys": (None, [_sz, _u64p, _u64p, _u32p]),
            "orc_linear_tree": (_sz, [_sz] + [_dp] * 4 + [_u32, _sz, _u64p, _u32p] + [_u32p] * 5
                                + [_dp, _dp]),
            "orc_bh_group_accel": (_u64, [_sz, _d, _d, _u32, _u32, _u32, _i] + [_dp] * 7 + [_i]),
            "orc_kdk_steps": (None, [_sz, _i, _d, _d] + [_dp] * 10 + [_d, _u64, _i]),
            "orc_run": (_sz, [_sz, _i, _d, _d] + [_dp] * 7 + [_d, _u64, _u64, _dp, _sz, _i]),
            "orc_galaxy_collision": (None, [C.c_void_p, _i] + [_dp] * 7 + [C.c_void_p]),
            "orc_galaxy_suggested_dt": (_d, [C.c_void_p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        f = L.orc_diagnostics
        f.restype = None
        f.argtypes = [_sz, _d, _i] + [_dp] * 7 + [C.c_void_p, C.c_void_p, _d, _u64, _dp]
        _LIB = L
    return _LIB


def _z(n):
    return np.zeros(n, dtype=np.float64)


class State:
    """SoA particle state mirroring sim-core's Particles (particle.rs:21-64),
    gravity fields only."""

    def __init__(self, n: int):
        self.n = n
        for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "m"):
            setattr(self, k, _z(n))

    def copy(self) -> "State":
        s = State(self.n)
        for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "m"):
            setattr(s, k, getattr(self, k).copy())
        return s

    def pos(self):
        return np.stack([self.x, self.y, self.z], axis=1)

    def acc(self):
        return np.stack([self.ax, self.ay, self.az], axis=1)


def chacha20_block(key, counter: int, stream: int = 0) -> np.ndarray:
    out = np.zeros(16, dtype=np.uint32)
    lib().orc_chacha20_block(np.ascontiguousarray(key, dtype=np.uint32), counter, stream, out)
    return out


def rng_stream(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint64)
    lib().orc_rng_stream_u64(seed, n, out)
    return out


def rng_seed_key(seed: int) -> np.ndarray:
    out = np.zeros(8, dtype=np.uint32)
    lib().orc_rng_seed_key(seed, out)
    return out


def plummer(n: int, seed: int = 42, a: float = 1.0, total_mass: float = 1.0) -> State:
    s = State(n)
    lib().orc_plummer(n, a, total_mass, seed, s.x, s.y, s.z, s.vx, s.vy, s.vz, s.m)
    return s


def plummer_suggested_dt(a: float = 1.0, total_mass: float = 1.0) -> float:
    return lib().orc_plummer_suggested_dt(a, total_mass)


def plummer_suggested_softening(n: int, a: float = 1.0) -> float:
    return lib().orc_plummer_suggested_softening(n, a)


class GalaxyParams(C.Structure):
    """GalaxyCollision's fields and Default (galaxy_collision.rs:20-68)."""
    _fields_ = [("n_per_galaxy", C.c_uint64), ("seed", C.c_uint64)] + [
        (k, C.c_double) for k in ("disk_fraction", "halo_mass_ratio", "baryon_mass",
                                  "disk_scale_radius", "disk_scale_height",
                                  "bulge_scale_radius", "halo_scale_radius", "disk_max_radius",
                                  "halo_max_radius", "separation", "impact_parameter",
                                  "approach_velocity")]

    @classmethod
    def default(cls, **kw):
        p = cls(50000, 42, 0.7, 10.0, 1.0, 1.0, 0.1, 0.3, 5.0, 6.0, 30.0, 15.0, 2.0, 0.5)
        for k, v in kw.items():
            setattr(p, k, v)
        return p


def galaxy_collision(com_shift: bool = True, **kw):
    """GalaxyCollision::generate (galaxy_collision.rs:75-157). Keyword
    arguments override Default fields; com_shift=False returns the samples
    before shift_to_com_frame. Returns (State, particle types)."""
    p = GalaxyParams.default(**kw)
    n = 2 * p.n_per_galaxy
    s = State(n)
    t = np.zeros(n, dtype=np.uint8)
    lib().orc_galaxy_collision(C.byref(p), 1 if com_shift else 0, s.x, s.y, s.z, s.vx, s.vy, s.vz, s.m,
                               t.ctypes.data_as(C.c_void_p))
    return s, t


def galaxy_suggested_dt(**kw) -> float:
    return lib().orc_galaxy_suggested_dt(C.byref(GalaxyParams.default(**kw)))


def two_body(e: float = 0.5, mass_ratio: float = 0.5, a: float = 1.0) -> State:
    s = State(2)
    lib().orc_two_body(e, mass_ratio, a, s.x, s.y, s.z, s.vx, s.vy, s.vz, s.m)
    return s


def two_body_period(a: float = 1.0) -> float:
    return lib().orc_two_body_period(a)


def bf_accel(s: State, eps: float) -> None:
    """Accumulate brute-force accelerations into s.ax/ay/az."""
    lib().orc_bf_accel(s.n, eps, s.x, s.y, s.z, s.m, s.ax, s.ay, s.az)


def bf_targets(s: State, eps: float, targets, nthreads: int = 0):
    t = np.ascontiguousarray(targets, dtype=np.int64)
    ax, ay, az, ab = _z(len(t)), _z(len(t)), _z(len(t)), _z(len(t))
    lib().orc_bf_targets(s.n, eps, s.x, s.y, s.z, s.m, len(t), t, ax, ay, az, ab, nthreads)
    return np.stack([ax, ay, az], axis=1), ab


def bh_accel(s: State, eps: float, theta: float, nthreads: int = 0) -> None:
    lib().orc_bh_accel(s.n, eps, theta, s.x, s.y, s.z, s.m, s.ax, s.ay, s.az, nthreads)


FRONTIER_CAP = 4096  # the device's group-walk frontier capacity (kGroupFrontierCap)


def bh_group_accel(s: State, eps: float, theta: float, leaf: int = 16, group: int = 64,
                   nthreads: int = 0, frontier_cap: int = FRONTIER_CAP, split: bool = True) -> int:
    """The device's group walk restated. split=True: the target groups of its
    Barnes-Hut gravity path (runs of consecutive small children cut into
    runs of `group`); False: those of its SPH step (whole children packed)."""
    return lib().orc_bh_group_accel(s.n, eps, theta, leaf, group, frontier_cap, int(split),
                                    s.x, s.y, s.z, s.m, s.ax, s.ay, s.az, nthreads)


def root_box(x, y, z) -> np.ndarray:
    out = _z(4)
    lib().orc_root_box(len(x), np.ascontiguousarray(x), np.ascontiguousarray(y),
                       np.ascontiguousarray(z), out)
    return out


def morton63(x, y, z, box) -> np.ndarray:
    keys = np.zeros(len(x), dtype=np.uint64)
    lib().orc_morton63(len(x), np.ascontiguousarray(x), np.ascontiguousarray(y),
                       np.ascontiguousarray(z), np.ascontiguousarray(box, dtype=np.float64), keys)
    return keys


def sort_keys(keys):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    sk = np.zeros_like(keys)
    perm = np.zeros(len(keys), dtype=np.uint32)
    lib().orc_sort_keys(len(keys), keys, sk, perm)
    return sk, perm


def linear_tree(s: State, leaf: int = 16):
    """Return dict with the leaf-`leaf` linear octree of s (see gravis_oracle.c)."""
    n = s.n
    cap = 2 * n + 64
    out = {k: np.zeros(cap, dtype=np.uint32) for k in ("level", "first", "cnt", "child", "nchild")}
    com = np.zeros(4 * cap, dtype=np.float64)
    sk = np.zeros(n, dtype=np.uint64)
    perm = np.zeros(n, dtype=np.uint32)
    box = _z(4)
    nn = lib().orc_linear_tree(n, s.x, s.y, s.z, s.m, leaf, cap, sk, perm, out["level"],
                               out["first"], out["cnt"], out["child"], out["nchild"], com, box)
    assert nn > 0
    res = {k: v[:nn].copy() for k, v in out.items()}
    res["com"] = com[: 4 * nn].reshape(nn, 4).copy()
    res["keys"], res["perm"], res["box"] = sk, perm, box
    return res


def kdk_steps(s: State, dt: float, nsteps: int, eps: float, algo: int = BF,
              theta: float = 0.5, nthreads: int = 0) -> None:
    lib().orc_kdk_steps(s.n, algo, eps, theta, s.x, s.y, s.z, s.vx, s.vy, s.vz, s.m,
                        s.ax, s.ay, s.az, dt, nsteps, nthreads)


def run(s: State, dt: float, nsteps: int, eps: float, diag_interval: int, algo: int = BF,
        theta: float = 0.5, nthreads: int = 0) -> np.ndarray:
    """headless-style run: returns rows of diagnostics (see diagnostics())."""
    max_rows = 2 + (nsteps // diag_interval if diag_interval else 0)
    rows = np.zeros(19 * max_rows, dtype=np.float64)
    nr = lib().orc_run(s.n, algo, eps, theta, s.x, s.y, s.z, s.vx, s.vy, s.vz, s.m, dt,
                       nsteps, diag_interval, rows, max_rows, nthreads)
    return rows[: 19 * nr].reshape(nr, 19)


DIAG_FIELDS = ("K", "U", "Eth", "E", "px", "py", "pz", "Lx", "Ly", "Lz", "cx", "cy", "cz",
               "cvx", "cvy", "cvz", "virial", "time", "step")


def diagnostics(s: State, eps: float, with_potential: bool = True, time: float = 0.0,
                step: int = 0) -> dict:
    out = _z(19)
    lib().orc_diagnostics(s.n, eps, 1 if with_potential else 0, s.x, s.y, s.z, s.vx, s.vy,
                          s.vz, s.m, None, None, time, step, out)
    return dict(zip(DIAG_FIELDS, out.tolist()))


# ---- SPH (gravis_oracle_sph.c) ---------------------------------------------

SPH_FIELDS = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "m", "rho", "p", "u", "dudt",
              "h", "c", "alpha")


class _SphStateC(C.Structure):
    _fields_ = [("n", C.c_size_t)] + [(f, C.POINTER(C.c_double)) for f in SPH_FIELDS] + [
        ("type", C.POINTER(C.c_uint8))]


class SphParams(C.Structure):
    """SphSolver::new + ForceParams::default (solver.rs:40-53, forces.rs:30-41)
    plus the build's variant knobs (0 = sim-core's choice; see
    gravis_oracle_sph.c)."""
    _fields_ = [(k, C.c_double) for k in ("gamma", "eta", "h_max", "cfl", "force_factor",
                                          "beta_visc", "alpha_min", "alpha_max", "alpha_cond",
                                          "sigma_visc", "h_tol")] + [
        (k, C.c_int32) for k in ("kernel", "h_solver", "max_h_iter", "alpha_switch",
                                 "pair_fp32", "soft_h")]

    @classmethod
    def default(cls):
        p = cls()
        _sph_lib().orc_sph_default_params(C.byref(p))
        return p

    @classmethod
    def variant(cls):
        """Cubic spline, Newton h to 1e-4, fixed alpha, no conduction, eps = h."""
        p = cls()
        _sph_lib().orc_sph_variant_params(C.byref(p))
        return p


class SphState:
    """Particles with the SPH fields (particle.rs:47-63); type 4 = gas."""

    def __init__(self, n: int):
        self.n = n
        for f in SPH_FIELDS:
            setattr(self, f, np.zeros(n))
        self.type = np.zeros(n, dtype=np.uint8)

    def copy(self):
        s = SphState(self.n)
        for f in SPH_FIELDS + ("type",):
            setattr(s, f, getattr(self, f).copy())
        return s

    def c_struct(self):
        st = _SphStateC()
        st.n = self.n
        for f in SPH_FIELDS:
            setattr(st, f, getattr(self, f).ctypes.data_as(C.POINTER(C.c_double)))
        st.type = self.type.ctypes.data_as(C.POINTER(C.c_uint8))
        return st

    def acc(self):
        return np.stack([self.ax, self.ay, self.az], axis=1)


_SPH = None


def _sph_lib():
    global _SPH
    if _SPH is None:
        L = lib()
        sp = C.POINTER(_SphStateC)
        pp = C.POINTER(SphParams)
        spec = {
            "orc_sph_default_params": (None, [pp]),
            "orc_sph_variant_params": (None, [pp]),
            "orc_cs_w": (_d, [_d, _d]),
            "orc_cs_dw_dh": (_d, [_d, _d]),
            "orc_cs_grad_w": (None, [_d, _d, _d, _d, _d, _dp]),
            "orc_evrard_1r": (None, [_sz, _d, _d, _d, _d, _u64, sp]),
            "orc_cold_collapse": (None, [_sz, _d, _d, _u64, _i, sp]),
            "orc_cold_collapse_ex": (None, [_sz, _d, _d, _u64, _i, _i, sp]),
            "orc_bf_accel_eps": (None, [_sz] + [_dp] * 8),
            "orc_bh_group_accel_eps": (_u64, [_sz, _dp, _d, _u32, _u32, _u32] + [_dp] * 7 + [_i]),
            "orc_w": (_d, [_d, _d]),
            "orc_dw_dh": (_d, [_d, _d]),
            "orc_grad_w": (None, [_d, _d, _d, _d, _d, _dp]),
            "orc_abs_grad_w": (_d, [_d, _d]),
            "orc_ball_query": (_sz, [sp, _sz, _d, _u32p, _sz]),
            "orc_sph_compute": (None, [sp, pp, _dp, _dp, _dp, _dp, _u32p]),
            "orc_sph_compute_targets": (None, [sp, pp, _sz, _i64p, _dp, _dp, _dp, _dp, _u32p,
                                               _i]),
            "orc_bf_targets_eps": (None, [_sz, _dp] + [_dp] * 4 + [_sz, _i64p] + [_dp] * 3
                                   + [_i]),
            "orc_sph_timestep": (_d, [sp, pp, _dp, _dp]),
            "orc_sph_step": (_d, [sp, pp, _i, _d, _d, _d]),
            "orc_evrard": (None, [_sz, _d, _d, _d, _u64, sp]),
            "orc_evrard_softening": (_d, [_sz, _d]),
        }
        for name, (res, args) in spec.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _SPH = L
    return _SPH


def sph_diagnostics(s: SphState, eps: float, with_potential: bool = True, time: float = 0.0,
                    step: int = 0) -> dict:
    """diagnostics::compute with the gas thermal energy (diagnostics.rs:37-148)."""
    out = _z(19)
    lib().orc_diagnostics(s.n, eps, 1 if with_potential else 0, s.x, s.y, s.z, s.vx, s.vy,
                          s.vz, s.m, s.u.ctypes.data_as(C.c_void_p),
                          s.type.ctypes.data_as(C.c_void_p), time, step, out)
    return dict(zip(DIAG_FIELDS, out.tolist()))


def sph_bf_accel(s: SphState, eps: float) -> None:
    """BruteForce::compute_accelerations on an SPH state (accumulates)."""
    lib().orc_bf_accel(s.n, eps, s.x, s.y, s.z, s.m, s.ax, s.ay, s.az)


def sph_w(r: float, h: float) -> float:
    return _sph_lib().orc_w(r, h)


def sph_dw_dh(r: float, h: float) -> float:
    return _sph_lib().orc_dw_dh(r, h)


def sph_grad_w(dx: float, dy: float, dz: float, r: float, h: float) -> np.ndarray:
    g = np.zeros(3)
    _sph_lib().orc_grad_w(dx, dy, dz, r, h, g)
    return g


def sph_abs_grad_w(r: float, h: float) -> float:
    return _sph_lib().orc_abs_grad_w(r, h)


def ball_query(s: SphState, i: int, radius: float) -> np.ndarray:
    """Octree::query_ball (octree.rs:245-308) from particle i, in tree order."""
    st = s.c_struct()
    out = np.zeros(s.n, dtype=np.uint32)
    k = _sph_lib().orc_ball_query(C.byref(st), i, radius, out, s.n)
    return out[:k].copy()


def evrard(n: int, seed: int = 42, total_mass: float = 1.0, radius: float = 1.0,
           u0: float = 0.05) -> SphState:
    s = SphState(n)
    st = s.c_struct()
    _sph_lib().orc_evrard(n, total_mass, radius, u0, seed, C.byref(st))
    return s


def evrard_1r(n: int, seed: int = 42, total_mass: float = 1.0, radius: float = 1.0,
              u0: float = 0.05, eta: float = 1.2) -> SphState:
    """Variant ICs: rho ~ 1/r gas sphere (see gravis_oracle_sph.c)."""
    s = SphState(n)
    st = s.c_struct()
    _sph_lib().orc_evrard_1r(n, total_mass, radius, u0, eta, seed, C.byref(st))
    return s


def cold_collapse(n: int, seed: int = 42, radius: float = 1.0, total_mass: float = 1.0,
                  gas: bool = False, com_shift: bool = True) -> SphState:
    """ColdCollapse::generate (cold_collapse.rs:56-102); com_shift=False
    returns the samples before shift_to_com_frame."""
    s = SphState(n)
    st = s.c_struct()
    _sph_lib().orc_cold_collapse_ex(n, radius, total_mass, seed, 1 if gas else 0,
                                    1 if com_shift else 0, C.byref(st))
    return s


def cs_w(r: float, h: float) -> float:
    return _sph_lib().orc_cs_w(r, h)


def cs_dw_dh(r: float, h: float) -> float:
    return _sph_lib().orc_cs_dw_dh(r, h)


def cs_grad_w(dx, dy, dz, r, h) -> np.ndarray:
    g = np.zeros(3)
    _sph_lib().orc_cs_grad_w(dx, dy, dz, r, h, g)
    return g


def bf_accel_eps(s: SphState, eps_i: np.ndarray) -> None:
    """Brute force with per-target softening eps_i (accumulates)."""
    _sph_lib().orc_bf_accel_eps(s.n, np.ascontiguousarray(eps_i, dtype=np.float64), s.x, s.y,
                                s.z, s.m, s.ax, s.ay, s.az)


def bh_group_accel_eps(s: SphState, eps_i: np.ndarray, theta: float = 0.5) -> None:
    """Linear-octree group walk with per-target softening (accumulates)."""
    _sph_lib().orc_bh_group_accel_eps(s.n, np.ascontiguousarray(eps_i, dtype=np.float64),
                                      theta, 16, 64, FRONTIER_CAP, s.x, s.y, s.z, s.m, s.ax,
                                      s.ay, s.az, 0)


def evrard_softening(n: int, radius: float = 1.0) -> float:
    return _sph_lib().orc_evrard_softening(n, radius)


def sph_compute(s: SphState, params: SphParams | None = None):
    """SphSolver::compute: accumulates SPH a and du/dt, updates rho/h/P/c;
    returns dict(div_v, max_mu, dalpha_dt, omega, nneigh)."""
    params = params or SphParams.default()
    n = s.n
    out = {k: np.zeros(n) for k in ("div_v", "max_mu", "dalpha_dt", "omega")}
    nn = np.zeros(n, dtype=np.uint32)
    st = s.c_struct()
    _sph_lib().orc_sph_compute(C.byref(st), C.byref(params), out["div_v"], out["max_mu"],
                               out["dalpha_dt"], out["omega"], nn)
    out["nneigh"] = nn
    return out


def sph_compute_targets(s: SphState, targets, params: SphParams | None = None,
                        nthreads: int = 0):
    """SphSolver::compute restricted to `targets` (indices): the targets'
    and their neighbours' density iteration, the targets' forces. Updates
    rho/h/P/c of the particles visited and accumulates the targets' a and
    du/dt in s; returns per-target dict(div_v, max_mu, dalpha_dt, omega,
    nneigh). Each value equals what sph_compute gives for that particle."""
    params = params or SphParams.default()
    t = np.ascontiguousarray(targets, dtype=np.int64)
    k = len(t)
    out = {q: np.zeros(k) for q in ("div_v", "max_mu", "dalpha_dt", "omega")}
    nn = np.zeros(k, dtype=np.uint32)
    st = s.c_struct()
    _sph_lib().orc_sph_compute_targets(C.byref(st), C.byref(params), k, t, out["div_v"],
                                       out["max_mu"], out["dalpha_dt"], out["omega"], nn,
                                       nthreads)
    out["nneigh"] = nn
    return out


def bf_targets_eps(s, eps_t, targets, nthreads: int = 0):
    """Brute-force accelerations at `targets` with softening eps_t[k] for
    target k (sources in index order); returns (k, 3)."""
    t = np.ascontiguousarray(targets, dtype=np.int64)
    a = [np.zeros(len(t)) for _ in range(3)]
    _sph_lib().orc_bf_targets_eps(s.n, np.ascontiguousarray(eps_t, dtype=np.float64), s.x,
                                  s.y, s.z, s.m, len(t), t, *a, nthreads)
    return np.stack(a, 1)


def sph_timestep(s: SphState, inter: dict, params: SphParams | None = None) -> float:
    params = params or SphParams.default()
    st = s.c_struct()
    return _sph_lib().orc_sph_timestep(C.byref(st), C.byref(params), inter["div_v"],
                                       inter["max_mu"])


def sph_step(s: SphState, dt: float, gravity: int = 1, eps: float = 0.0, theta: float = 0.5,
             params: SphParams | None = None) -> float:
    """step_with_sph; gravity 0 none, 1 brute force, 2 reference Barnes-Hut,
    3 linear-octree group Barnes-Hut (leaf 16, group 64)."""
    params = params or SphParams.default()
    st = s.c_struct()
    return _sph_lib().orc_sph_step(C.byref(st), C.byref(params), gravity, eps, theta, dt)


# ---- ensemble + adjoint (gravis_oracle_adjoint.c) -------------------------

def _adj_lib():
    L = lib()
    if not getattr(L, "_adj_ready", False):
        spec = {
            "orc_ens_forward": (None, [_sz, _d, _d, _u64] + [_dp] * 7),
            "orc_ens_adjoint": (None, [_sz, _d, _d, _u64] + [_dp] * 10 + [C.c_void_p] * 3
                                + [_dp] * 7),
            "orc_ens_counts": (None, [_u64, _sz, _u32, _u32, _u32p]),
            "orc_ens_targets": (None, [_u64, _sz, _dp, _dp, _dp]),
            "orc_ens_set_threads": (None, [_i]),
        }
        for name, (res, args) in spec.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        L._adj_ready = True
    return L


def ens_set_threads(n: int) -> None:
    """Threads for ens_forward / ens_adjoint (default 1; results do not
    depend on it)."""
    _adj_lib().orc_ens_set_threads(int(n))


def ens_forward(s: State, eps: float, dt: float, nsteps: int) -> None:
    """KDK with gather-form brute force (a(x0) first); state updated in place."""
    _adj_lib().orc_ens_forward(s.n, eps, dt, nsteps, s.x, s.y, s.z, s.vx, s.vy, s.vz, s.m)


def ens_adjoint(s: State, eps: float, dt: float, nsteps: int, gx, gy, gz, gv=None):
    """Cotangents of (x0, v0, m) for cotangents (gx, gy, gz) of x_T (and gv
    of v_T). Returns dict of arrays x, y, z, vx, vy, vz, m."""
    n = s.n
    out = {k: _z(n) for k in ("x", "y", "z", "vx", "vy", "vz", "m")}
    gv_ptrs = [None, None, None]
    keep = []
    if gv is not None:
        keep = [np.ascontiguousarray(a, dtype=np.float64) for a in gv]
        gv_ptrs = [a.ctypes.data_as(C.c_void_p) for a in keep]
    _adj_lib().orc_ens_adjoint(n, eps, dt, nsteps, s.x, s.y, s.z, s.vx, s.vy, s.vz, s.m,
                               np.ascontiguousarray(gx, dtype=np.float64),
                               np.ascontiguousarray(gy, dtype=np.float64),
                               np.ascontiguousarray(gz, dtype=np.float64), *gv_ptrs,
                               out["x"], out["y"], out["z"], out["vx"], out["vy"], out["vz"],
                               out["m"])
    return out


def ens_counts(seed: int, nsys: int, lo: int = 512, hi: int = 8192) -> np.ndarray:
    c = np.zeros(nsys, dtype=np.uint32)
    _adj_lib().orc_ens_counts(seed, nsys, lo, hi, c)
    return c


def ens_targets(seed: int, n: int):
    tx, ty, tz = _z(n), _z(n), _z(n)
    _adj_lib().orc_ens_targets(seed, n, tx, ty, tz)
    return tx, ty, tz
