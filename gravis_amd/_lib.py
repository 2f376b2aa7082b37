"""ctypes binding of libgravis_hip.so (the C ABI declared in include/gravis.h).

Loading fails loudly: there is no CPU fallback anywhere in this package. If
the shared library is missing, build it with `python -c "import
__graft_entry__ as g; g.build()"` (or `make -C gravis_amd/csrc`).
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgravis_hip.so")

# gravis_status
OK, INVALID_ARG, HIP_ERROR, OUT_OF_MEMORY, UNSUPPORTED, INTERNAL = range(6)
# gravis_algorithm / gravis_precision
BRUTE_FORCE, BARNES_HUT, NO_GRAVITY = 0, 1, 2
FP64, MIXED = 0, 1

EXPORTS = (
    "gravis_version", "gravis_status_string", "gravis_create", "gravis_destroy",
    "gravis_last_error", "gravis_set_stream", "gravis_synchronize", "gravis_set_profiling",
    "gravis_kernel_time", "gravis_accel", "gravis_accel_device", "gravis_kdk",
    "gravis_kdk_device", "gravis_diagnostics", "gravis_diagnostics_device",
    "gravis_morton_sort", "gravis_scenario_plummer", "gravis_plummer_suggested_dt",
    "gravis_plummer_suggested_softening", "gravis_scenario_two_body", "gravis_two_body_period",
    "gravis_scenario_plummer_pair", "gravis_bh_stats", "gravis_ensemble_kdk",
    "gravis_ensemble_kdk_device", "gravis_ensemble_accel_device", "gravis_ensemble_adjoint",
    "gravis_ensemble_adjoint_device", "gravis_set_tuning", "gravis_scenario_ensemble_counts",
    "gravis_scenario_targets", "gravis_sph_default_params", "gravis_sph_compute",
    "gravis_sph_compute_device", "gravis_sph_step", "gravis_sph_run", "gravis_sph_run_device",
    "gravis_sph_stats", "gravis_scenario_evrard", "gravis_evrard_softening",
    "gravis_scenario_evrard_1r", "gravis_sph_variant_params", "gravis_sph_accel",
    "gravis_sph_accel_device", "gravis_snapshot_write", "gravis_snapshot_info",
    "gravis_snapshot_read", "gravis_scenario_cold_collapse", "gravis_galaxy_default_params",
    "gravis_scenario_galaxy_collision", "gravis_galaxy_suggested_dt",
    "gravis_galaxy_suggested_softening", "gravis_get_tuning", "gravis_autotune",
    "gravis_accel_device_noalloc", "gravis_render_pack", "gravis_render_pack_device",
    "gravis_ensemble_accel",
)


class Solver(C.Structure):
    _fields_ = [("algorithm", C.c_int32), ("precision", C.c_int32), ("softening", C.c_double),
                ("theta", C.c_double), ("leaf_size", C.c_uint32), ("group_size", C.c_uint32)]


class Diag(C.Structure):
    _fields_ = [("time", C.c_double), ("step", C.c_uint64), ("kinetic_energy", C.c_double),
                ("potential_energy", C.c_double), ("total_energy", C.c_double),
                ("momentum", C.c_double * 3), ("angular_momentum", C.c_double * 3),
                ("center_of_mass", C.c_double * 3), ("center_of_mass_velocity", C.c_double * 3),
                ("virial_ratio", C.c_double), ("thermal_energy", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("bf_ipt", C.c_int32), ("bf_tile", C.c_int32), ("bf_waves_per_cu", C.c_int32),
                ("bh_group_size", C.c_uint32), ("bf_matrix", C.c_int32), ("best_ms", C.c_double)]


class SphParams(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("gamma", "eta", "h_max", "cfl", "force_factor",
                                           "beta_visc", "alpha_min", "alpha_max", "alpha_cond",
                                           "sigma_visc", "h_tol")] + [
        (k, C.c_int32) for k in ("kernel", "h_solver", "max_h_iter", "alpha_switch",
                                 "pair_fp32", "soft_h")]


class GalaxyParams(C.Structure):
    _fields_ = [("n_per_galaxy", C.c_uint64), ("seed", C.c_uint64)] + [
        (k, C.c_double) for k in ("disk_fraction", "halo_mass_ratio", "baryon_mass",
                                  "disk_scale_radius", "disk_scale_height",
                                  "bulge_scale_radius", "halo_scale_radius", "disk_max_radius",
                                  "halo_max_radius", "separation", "impact_parameter",
                                  "approach_velocity")]


SPH_FIELDS = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "m", "rho", "p", "u", "dudt",
              "h", "c", "alpha")


class SphState(C.Structure):
    _fields_ = ([("n", C.c_size_t)] + [(k, C.c_void_p) for k in SPH_FIELDS]
                + [("type", C.c_void_p)])


_LIB = None


class GravisError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"gravis status {status}: {message}")
        self.status = status


def load() -> C.CDLL:
    """Load the HIP library (raises if it is not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    # PyTorch-ROCm ships its own libamdhip64.so.7 with the same SONAME as
    # /opt/rocm's. Whichever is loaded first serves the whole process, and
    # torch cannot initialise on a newer runtime loaded before it, so torch
    # (when installed) is imported first and this library binds to its
    # runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run make -C gravis_amd/csrc "
                          "(this package has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp, dp, sz, d, i32 = C.c_void_p, C.POINTER(C.c_double), C.c_size_t, C.c_double, C.c_int
    st = C.c_int
    spec = {
        "gravis_version": (C.c_char_p, []),
        "gravis_status_string": (C.c_char_p, [st]),
        "gravis_create": (st, [i32, C.POINTER(vp)]),
        "gravis_destroy": (None, [vp]),
        "gravis_last_error": (C.c_char_p, [vp]),
        "gravis_set_stream": (st, [vp, vp]),
        "gravis_synchronize": (st, [vp]),
        "gravis_set_profiling": (st, [vp, i32]),
        "gravis_kernel_time": (st, [vp, C.c_char_p, dp, C.POINTER(C.c_uint64)]),
        "gravis_accel": (st, [vp, C.POINTER(Solver), sz] + [vp] * 7),
        "gravis_accel_device": (st, [vp, C.POINTER(Solver), sz] + [vp] * 7),
        "gravis_accel_device_noalloc": (st, [vp, C.POINTER(Solver), sz] + [vp] * 7),
        "gravis_render_pack": (st, [vp, sz] + [vp] * 7),
        "gravis_render_pack_device": (st, [vp, sz] + [vp] * 7),
        "gravis_kdk": (st, [vp, C.POINTER(Solver), sz] + [vp] * 10 + [d, C.c_uint64]),
        "gravis_kdk_device": (st, [vp, C.POINTER(Solver), sz] + [vp] * 10 + [d, C.c_uint64]),
        "gravis_diagnostics": (st, [vp, sz, d, i32] + [vp] * 9 + [d, C.c_uint64,
                                                                  C.POINTER(Diag)]),
        "gravis_diagnostics_device": (st, [vp, sz, d, i32] + [vp] * 9 + [d, C.c_uint64,
                                                                         C.POINTER(Diag)]),
        "gravis_morton_sort": (st, [vp, sz, vp, vp, vp, vp, vp, vp]),
        "gravis_scenario_plummer": (st, [sz, d, d, C.c_uint64] + [vp] * 7),
        "gravis_plummer_suggested_dt": (d, [d, d]),
        "gravis_plummer_suggested_softening": (d, [sz, d]),
        "gravis_scenario_two_body": (st, [d, d, d] + [vp] * 7),
        "gravis_two_body_period": (d, [d]),
        "gravis_scenario_plummer_pair": (st, [sz, d, d, C.c_uint64, C.c_uint64, d, d] + [vp] * 7),
        "gravis_bh_stats": (st, [vp, C.POINTER(C.c_uint64 * 5)]),
        "gravis_set_tuning": (st, [vp, C.c_char_p, C.c_int64]),
        "gravis_scenario_ensemble_counts": (st, [C.c_uint64, i32, C.c_uint32, C.c_uint32, vp]),
        "gravis_scenario_targets": (st, [C.c_uint64, sz, vp, vp, vp]),
        "gravis_ensemble_kdk": (st, [vp, C.POINTER(Solver), i32, vp] + [vp] * 10
                                + [d, C.c_uint64, i32]),
        "gravis_ensemble_accel_device": (st, [vp, C.POINTER(Solver), i32, vp] + [vp] * 7),
        "gravis_ensemble_accel": (st, [vp, C.POINTER(Solver), i32, vp] + [vp] * 7),
        "gravis_ensemble_kdk_device": (st, [vp, C.POINTER(Solver), i32, vp] + [vp] * 10
                                       + [d, C.c_uint64, i32]),
        "gravis_ensemble_adjoint": (st, [vp, C.POINTER(Solver), i32, vp] + [vp] * 7
                                    + [d, C.c_uint64, C.c_uint32] + [vp] * 9 + [vp] * 9),
        "gravis_ensemble_adjoint_device": (st, [vp, C.POINTER(Solver), i32, vp] + [vp] * 7
                                           + [d, C.c_uint64, C.c_uint32] + [vp] * 9
                                           + [vp] * 9),
        "gravis_sph_default_params": (None, [C.POINTER(SphParams)]),
        "gravis_sph_variant_params": (None, [C.POINTER(SphParams)]),
        "gravis_sph_compute": (st, [vp, C.POINTER(SphParams), C.POINTER(SphState)] + [vp] * 5),
        "gravis_sph_compute_device": (st, [vp, C.POINTER(SphParams), C.POINTER(SphState)]
                                      + [vp] * 5),
        "gravis_sph_step": (st, [vp, C.POINTER(SphParams), vp, C.POINTER(SphState), d, dp]),
        "gravis_sph_run": (st, [vp, C.POINTER(SphParams), vp, C.POINTER(SphState), C.c_uint64,
                                dp, dp]),
        "gravis_sph_run_device": (st, [vp, C.POINTER(SphParams), vp, C.POINTER(SphState),
                                       C.c_uint64, dp, dp]),
        "gravis_sph_stats": (st, [vp, C.POINTER(C.c_uint64 * 4)]),
        "gravis_sph_accel": (st, [vp, C.POINTER(SphParams), vp, C.POINTER(SphState)]),
        "gravis_sph_accel_device": (st, [vp, C.POINTER(SphParams), vp, C.POINTER(SphState)]),
        "gravis_scenario_evrard": (st, [sz, d, d, d, C.c_uint64, C.POINTER(SphState)]),
        "gravis_evrard_softening": (d, [sz, d]),
        "gravis_scenario_evrard_1r": (st, [sz, d, d, d, d, C.c_uint64, C.POINTER(SphState)]),
        "gravis_scenario_cold_collapse": (st, [sz, d, d, C.c_uint64, i32, C.POINTER(SphState)]),
        "gravis_snapshot_write": (st, [C.c_char_p, d, C.c_uint64, d, d, C.POINTER(SphState)]),
        "gravis_snapshot_info": (st, [C.c_char_p, C.POINTER(i32), dp, C.POINTER(C.c_uint64), dp,
                                      dp, C.POINTER(C.c_uint64)]),
        "gravis_snapshot_read": (st, [C.c_char_p, C.POINTER(SphState)]),
        "gravis_galaxy_default_params": (None, [C.POINTER(GalaxyParams)]),
        "gravis_get_tuning": (st, [vp, C.POINTER(Tuning)]),
        "gravis_autotune": (st, [vp, C.POINTER(Solver), sz] + [vp] * 4 + [C.POINTER(Tuning)]),
        "gravis_scenario_galaxy_collision": (st, [C.POINTER(GalaxyParams),
                                                  C.POINTER(SphState)]),
        "gravis_galaxy_suggested_dt": (d, [C.POINTER(GalaxyParams)]),
        "gravis_galaxy_suggested_softening": (d, [C.POINTER(GalaxyParams)]),
    }
    for name, (res, args) in spec.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def check(ctx, status: int) -> None:
    if status != OK:
        L = load()
        msg = L.gravis_last_error(ctx).decode() if ctx else ""
        raise GravisError(status, msg or L.gravis_status_string(status).decode())
