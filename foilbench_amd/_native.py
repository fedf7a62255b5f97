"""ctypes binding of libfoilbench_amd.so (the C ABI declared in include/foilbench_amd.h).

The library is loaded from this package directory. There is no CPU fallback: if the
library is missing, or no GPU is visible, the solvers raise instead of computing anything.
"""

from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libfoilbench_amd.so"

FB_OK = 0
FB_F64, FB_F32, FB_F32_MIXED, FB_F16_MIXED = 0, 1, 2, 3
PRECISION_NAMES = {FB_F64: "f64", FB_F32: "f32", FB_F32_MIXED: "f32-mixed", FB_F16_MIXED: "f16-mixed"}


class NativeError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, status: int, message: str) -> None:
        super().__init__(message)
        self.status = status


class fb_domain(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("periodic_x", C.c_int32), ("periodic_y", C.c_int32),
        ("x0", C.c_double), ("x1", C.c_double), ("y0", C.c_double), ("y1", C.c_double),
    ]


class fb_foil(C.Structure):
    _fields_ = [
        ("chord", C.c_double), ("pivot_x", C.c_double), ("pivot_y", C.c_double),
        ("half_scale", C.c_double), ("camber_fwd", C.c_double), ("camber_aft", C.c_double),
        ("two_p", C.c_double), ("one_minus_two_p", C.c_double), ("p", C.c_double),
        ("cambered", C.c_int32), ("pad_", C.c_int32),
    ]


class fb_pose(C.Structure):
    _fields_ = [("angle_degrees", C.c_double), ("cos_a", C.c_double), ("sin_a", C.c_double)]


class fb_lbm_params(C.Structure):
    _fields_ = [
        ("omega_plus", C.c_double), ("omega_minus", C.c_double), ("u_lattice", C.c_double),
        ("target_ux", C.c_double), ("target_uy", C.c_double), ("wall_omega_rad", C.c_double),
        ("wall_scale", C.c_double), ("velocity_scale", C.c_double),
        ("feq_boundary", C.c_double * 9), ("feq_uncovered", C.c_double * 9),
    ]


class fb_lbm_stats(C.Structure):
    _fields_ = [
        ("min_population", C.c_double), ("max_speed", C.c_double), ("density_excursion", C.c_double),
        ("nonfinite_populations", C.c_int64), ("bad_fluid_density", C.c_int64),
        ("nonfinite_velocity", C.c_int64),
    ]


class fb_sf_params(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("viscosity", C.c_double), ("tolerance", C.c_double),
        ("max_iterations", C.c_int32), ("maccormack", C.c_int32),
        ("freestream_x", C.c_double), ("freestream_y", C.c_double), ("wall_omega_rad", C.c_double),
        ("cfl_limit", C.c_double), ("channel_walls", C.c_int32), ("mixed_poisson", C.c_int32),
        ("warm_pressure", C.c_int32), ("face_transport", C.c_int32),
    ]


class fb_sf_report(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("viscosity_iterations", C.c_int32), ("viscosity_residual", C.c_double),
        ("viscosity_converged", C.c_int32), ("pressure_iterations", C.c_int32), ("pressure_residual", C.c_double),
        ("pressure_info", C.c_int32), ("pressure_converged", C.c_int32), ("pressure_refinements", C.c_int32),
        ("pad_", C.c_int32), ("projection_cfl", C.c_double),
    ]


class fb_mac_stats(C.Structure):
    _fields_ = [
        ("max_face_speed", C.c_double), ("max_wall_speed", C.c_double), ("max_cell_speed", C.c_double),
        ("divergence_linf", C.c_double), ("solid_leakage", C.c_double), ("nonfinite_faces", C.c_int64),
        ("nonfinite_wall", C.c_int64),
    ]


SF_STATUS = {
    0: None,
    1: ("convergence_failure", "viscosity", "implicit viscosity did not converge"),
    2: ("nonfinite_state", "projection", "projection received non-finite velocity"),
    3: ("excessive_velocity", "projection", "projection CFL exceeds its limit"),
    4: ("nonfinite_state", "projection", "pressure RHS contains non-finite values"),
    5: ("projection_failure", "projection", "pressure RHS exceeds the safe CG dot-product range"),
    6: ("nonfinite_state", "projection", "pressure CG produced a non-finite iterate"),
    7: ("projection_failure", "projection", "pressure CG failed: floating-point breakdown"),
    8: ("projection_failure", "projection", "pressure CG did not converge"),
    9: ("nonfinite_state", "projection", "projection produced non-finite velocity"),
}

class fb_pic_params(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("viscosity", C.c_double), ("tolerance", C.c_double),
        ("max_iterations", C.c_int32), ("channel_walls", C.c_int32),
        ("freestream_x", C.c_double), ("freestream_y", C.c_double), ("wall_omega_rad", C.c_double),
        ("blend", C.c_double), ("wall_travel", C.c_double), ("maximum_radius", C.c_double),
        ("mixed_poisson", C.c_int32), ("pad_", C.c_int32),
    ]


class fb_pic_report(C.Structure):
    _fields_ = [("grid", fb_sf_report), ("unsupported_faces", C.c_int64), ("swept_hits", C.c_int64),
                ("escaped", C.c_int64)]


class fb_pic_stats(C.Structure):
    _fields_ = [
        ("max_particle_speed", C.c_double), ("max_grid_speed", C.c_double), ("particles_inside", C.c_int64),
        ("nonfinite", C.c_int64), ("fluid_cells", C.c_int64), ("empty_cells", C.c_int64),
        ("underfilled_cells", C.c_int64), ("particle_count", C.c_int64),
    ]


FB_DIAG_HISTOGRAM = 1024


class fb_diag_params(C.Structure):
    _fields_ = [
        ("freestream_u", C.c_double), ("wake_threshold", C.c_double), ("wake_col0", C.c_int32),
        ("down_col0", C.c_int32), ("angular_velocity_rad", C.c_double), ("diagonal_link", C.c_double),
        ("centers_x", C.POINTER(C.c_double)), ("centers_y", C.POINTER(C.c_double)),
    ]


class fb_diag(C.Structure):
    _fields_ = [
        ("kinetic_energy", C.c_double), ("enstrophy", C.c_double), ("divergence_l2", C.c_double),
        ("wake_width", C.c_double), ("recirculation_area", C.c_double), ("density_mean", C.c_double),
        ("cut_link_adjacent_normal_speed", C.c_double), ("nonfinite", C.c_int64), ("fluid_cells", C.c_int64),
        ("empty_cells", C.c_int64), ("underfilled_cells", C.c_int64), ("max_particles", C.c_int64),
    ]


class fb_server_options(C.Structure):
    _fields_ = [("sf_workers", C.c_int32), ("lbm_chunk", C.c_int32), ("use_graphs", C.c_int32),
                ("sf_batch", C.c_int32), ("sf_gather_us", C.c_double)]


FB_SERVE_SF = 0
FB_SERVE_LBM = 1
FB_TUNE_DST_ROWS = 0
FB_TUNE_LBM_LIST_BLOCKS = 1
FB_TUNE_LBM_FUSE = 2
FB_TUNE_LBM_ROWS = 3
FB_TUNE_DST_PPT = 4
FB_TUNE_LBM_NT = 5


class fb_server_result(C.Structure):
    _fields_ = [
        ("id", C.c_int64), ("kind", C.c_int32), ("status", C.c_int32), ("sf_status", C.c_int32),
        ("steps_done", C.c_int32), ("submitted_ms", C.c_double), ("started_ms", C.c_double),
        ("finished_ms", C.c_double), ("pressure_iterations", C.c_int64), ("lbm_stats", fb_lbm_stats),
    ]


_P = C.c_void_p
_SIGNATURES: dict[str, tuple] = {
    "fb_last_error": (C.c_char_p, []),
    "fb_version": (C.c_int, []),
    "fb_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "fb_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "fb_ctx_destroy": (C.c_int, [_P]),
    "fb_ctx_synchronize": (C.c_int, [_P]),
    "fb_copy_benchmark": (C.c_int, [_P, C.c_int64, C.c_int, C.POINTER(C.c_double)]),
    "fb_naca_cell_sdf": (C.c_int, [_P, C.POINTER(fb_domain), C.POINTER(fb_foil), C.POINTER(fb_pose), _P, _P]),
    "fb_naca_points": (C.c_int, [_P, C.POINTER(fb_foil), C.POINTER(fb_pose), C.c_int64, _P, _P, _P]),
    "fb_lbm_create": (C.c_int, [_P, C.POINTER(fb_domain), C.POINTER(fb_foil), C.c_int, C.c_int, _P, _P,
                                C.POINTER(_P)]),
    "fb_lbm_destroy": (C.c_int, [_P]),
    "fb_lbm_set_state": (C.c_int, [_P, _P, _P]),
    "fb_lbm_get_state": (C.c_int, [_P, _P, _P]),
    "fb_lbm_set_equilibrium": (C.c_int, [_P, _P, _P]),
    "fb_lbm_set_pose": (C.c_int, [_P, C.POINTER(fb_pose)]),
    "fb_lbm_update_pose": (C.c_int, [_P, C.POINTER(fb_pose), C.POINTER(fb_lbm_params)]),
    "fb_lbm_get_geometry": (C.c_int, [_P, _P, _P]),
    "fb_lbm_checkpoint": (C.c_int, [_P]),
    "fb_lbm_rollback": (C.c_int, [_P]),
    "fb_lbm_rescale": (C.c_int, [_P, C.c_double]),
    "fb_lbm_advance": (C.c_int, [_P, C.POINTER(fb_lbm_params), C.c_int, C.POINTER(fb_pose),
                                 C.POINTER(fb_lbm_stats)]),
    "fb_lbm_stats_now": (C.c_int, [_P, C.POINTER(fb_lbm_params), C.POINTER(fb_lbm_stats)]),
    "fb_lbm_macroscopic": (C.c_int, [_P, C.c_double, _P, _P]),
    "fb_lbm_run_fixed": (C.c_int, [_P, C.POINTER(fb_lbm_params), C.c_int]),
    "fb_lbm_time_fixed": (C.c_int, [_P, C.POINTER(fb_lbm_params), C.c_int, C.POINTER(C.c_double)]),
    "fb_lbm_forces": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "fb_mac_forces": (C.c_int, [_P, C.c_double, C.c_double, C.POINTER(C.c_double)]),
    "fb_mac_skew_rate": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "fb_mac_set_pressure": (C.c_int, [_P, _P]),
    "fb_lbm_run_batch": (C.c_int, [C.POINTER(_P), C.POINTER(fb_lbm_params), C.c_int, C.c_int, C.c_int,
                                   C.POINTER(fb_lbm_stats)]),
    "fb_mem_info": (C.c_int, [C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "fb_tune_set": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "fb_tune_get": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]),
    "fb_poisson_time_precond": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                          C.c_int, C.POINTER(C.c_double)]),
    "fb_server_create": (C.c_int, [C.c_int, C.POINTER(fb_server_options), C.POINTER(_P)]),
    "fb_server_destroy": (C.c_int, [_P]),
    "fb_server_submit_sf": (C.c_int, [_P, C.c_int64, _P, C.POINTER(fb_sf_params), C.POINTER(fb_pose), C.c_int]),
    "fb_server_submit_lbm": (C.c_int, [_P, C.c_int64, _P, C.POINTER(fb_lbm_params), C.c_int]),
    "fb_server_poll": (C.c_int, [_P, C.POINTER(fb_server_result), C.c_int, C.POINTER(C.c_int)]),
    "fb_server_load": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "fb_server_clock": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "fb_mac_create": (C.c_int, [_P, C.POINTER(fb_domain), C.POINTER(fb_foil), C.c_int, C.POINTER(_P)]),
    "fb_mac_destroy": (C.c_int, [_P]),
    "fb_mac_set_faces": (C.c_int, [_P, _P, _P]),
    "fb_mac_get_faces": (C.c_int, [_P, _P, _P]),
    "fb_mac_set_pose": (C.c_int, [_P, C.POINTER(fb_pose)]),
    "fb_mac_get_geometry": (C.c_int, [_P, _P, _P]),
    "fb_mac_get_pressure": (C.c_int, [_P, _P]),
    "fb_mac_checkpoint": (C.c_int, [_P]),
    "fb_mac_rollback": (C.c_int, [_P]),
    "fb_sf_substep": (C.c_int, [_P, C.POINTER(fb_sf_params), C.POINTER(fb_pose), C.POINTER(fb_sf_report)]),
    "fb_mac_project": (C.c_int, [_P, C.POINTER(fb_sf_params), C.POINTER(fb_sf_report)]),
    "fb_mac_get_stats": (C.c_int, [_P, C.c_double, C.POINTER(fb_mac_stats)]),
    "fb_pic_create": (C.c_int, [_P, C.POINTER(fb_domain), C.POINTER(fb_foil), C.c_int, C.POINTER(_P)]),
    "fb_pic_destroy": (C.c_int, [_P]),
    "fb_pic_mac": (C.c_int, [_P, C.POINTER(_P)]),
    "fb_pic_set_pose": (C.c_int, [_P, C.POINTER(fb_pose)]),
    "fb_pic_set_grid": (C.c_int, [_P, _P, _P]),
    "fb_pic_get_grid": (C.c_int, [_P, _P, _P]),
    "fb_pic_set_rng": (C.c_int, [_P, C.c_uint64, C.c_uint64]),
    "fb_pic_get_rng": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "fb_pic_seed": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int64)]),
    "fb_pic_count": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "fb_pic_set_particles": (C.c_int, [_P, C.c_int64, _P, _P]),
    "fb_pic_get_particles": (C.c_int, [_P, _P, _P]),
    "fb_pic_particle_to_grid": (C.c_int, [_P, C.c_int, C.POINTER(C.c_double)]),
    "fb_pic_regrid": (C.c_int, [_P, _P, C.c_int64]),
    "fb_pic_resolve_collisions": (C.c_int, [_P, C.POINTER(fb_pose), C.c_double]),
    "fb_pic_substep": (C.c_int, [_P, C.POINTER(fb_pic_params), C.POINTER(fb_pose), _P, C.POINTER(fb_pic_report)]),
    "fb_pic_get_stats": (C.c_int, [_P, C.POINTER(fb_pose), C.POINTER(fb_pic_stats)]),
    "fb_pic_checkpoint": (C.c_int, [_P]),
    "fb_pic_rollback": (C.c_int, [_P]),
    "fb_pic_particle_to_faces": (C.c_int, [_P, C.c_int, C.c_int64, _P, _P, _P, _P, C.c_int, C.c_int, C.c_double,
                                           C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, _P, _P,
                                           C.POINTER(C.c_double)]),
    "fb_pic_particle_to_faces_noalloc": (C.c_int, [_P, C.c_int, C.c_int64, _P, _P, _P, _P, C.c_int, C.c_int,
                                                   C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                                   _P, _P, _P]),
    "fb_pic_faces_to_particle": (C.c_int, [_P, C.c_int, C.c_int64, _P, _P, _P, C.c_int, C.c_int, C.c_double,
                                           C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, _P]),
    "fb_mac_diagnostics": (C.c_int, [_P, C.POINTER(fb_diag_params), C.POINTER(fb_diag)]),
    "fb_lbm_diagnostics": (C.c_int, [_P, C.c_double, C.c_int, C.POINTER(fb_diag_params), C.POINTER(fb_diag)]),
    "fb_pic_diagnostics": (C.c_int, [_P, C.POINTER(fb_diag_params), C.POINTER(fb_diag), _P, _P]),
    "fb_mac_sample_velocity": (C.c_int, [_P, C.c_int64, _P, _P]),
    "fb_lbm_sample_velocity": (C.c_int, [_P, C.c_double, C.c_int, C.c_int64, _P, _P]),
    "fb_pic_sample_velocity": (C.c_int, [_P, C.c_int64, _P, _P]),
    "fb_pic_particle_to_cells": (C.c_int, [_P, C.c_int, C.c_int64, _P, _P, C.c_int, C.c_int, C.c_double,
                                           C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                           C.c_int, _P]),
    "fb_pic_cells_to_particle": (C.c_int, [_P, C.c_int, C.c_int64, _P, _P, C.c_int, C.c_int, C.c_double,
                                           C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, _P]),
}

_lock = threading.Lock()
_lib: C.CDLL | None = None


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


def load(require_gpu: bool = True) -> C.CDLL:
    """Load the library once; raise if it is missing (there is no fallback)."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build it with `python -m foilbench_amd.build` "
                    "(the HIP library is required; there is no CPU fallback)"
                )
            lib = C.CDLL(str(LIB_PATH))
            for name, (res, args) in _SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
        lib = _lib
    if require_gpu:
        n = C.c_int(0)
        rc = lib.fb_device_count(C.byref(n))
        if rc != FB_OK or n.value < 1:
            raise RuntimeError("no HIP device is visible; the foilbench_amd solvers run only on the GPU")
    return lib


def check(rc: int) -> None:
    if rc != FB_OK:
        msg = _lib.fb_last_error().decode() if _lib is not None else "unknown error"
        raise NativeError(rc, msg or f"status {rc}")


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    if not a.flags.c_contiguous:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """One HIP device and stream (fb_ctx)."""

    _default: "Context | None" = None

    def __init__(self, device: int = 0) -> None:
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.fb_ctx_create(device, C.byref(h)))
        self.handle = h
        self.device = device

    @classmethod
    def default(cls) -> "Context":
        if cls._default is None:
            cls._default = Context(0)
        return cls._default

    def synchronize(self) -> None:
        check(self.lib.fb_ctx_synchronize(self.handle))

    def __del__(self) -> None:  # pragma: no cover - interpreter shutdown order
        try:
            if self.handle:
                self.lib.fb_ctx_destroy(self.handle)
                self.handle = None
        except Exception:
            pass
