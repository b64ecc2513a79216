"""ctypes binding of include/sf_seepage.h (the C-ABI of libsf_seepage.so).

Only plain pointers and sizes cross the boundary; numpy arrays are passed by
address.  Every failing call raises SFError carrying sf_last_error().
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "_lib" / "libsf_seepage.so"

SF_OK = 0
SF_E_INVALID = -1
SF_E_CAPACITY = -2
SF_E_HIP = -3
SF_E_STATE = -4
SF_E_IO = -5
SF_E_NOTFOUND = -6

SF_SPH_SOLVER = 0
SF_WCSPH_SOLVER = 1
SF_TYPE_SF = 0
SF_TYPE_MULTI_SF = 1

SF_SORT_RADIX = 0
SF_SORT_INCREMENTAL = 1

FIELDS = {
    "pos": (0, "f3"), "vel": (1, "f3"), "acc": (2, "f3"), "col": (3, "f3"), "label": (4, "u64"),
    "mass": (5, "f1"), "radius": (6, "f1"), "density": (7, "f1"), "pressure": (8, "f1"),
    "voidage": (9, "f1"), "saturation": (10, "f1"), "max_saturation": (11, "f1"),
    "avg_flow_vel": (12, "f3"), "avg_drag_force": (13, "f3"), "adhesion_force": (14, "f3"),
    "avg_adhesion_force": (15, "f3"), "cda0asat": (16, "f3"), "amcamcp": (17, "f2"),
    "particle2cell": (18, "u32"), "sort_perm": (19, "u32"), "sorted_keys": (20, "u32"),
    "cell_start": (21, "u32c"), "nbr_count": (22, "u32"), "nbr_candidates": (23, "u32"),
    "b_pos": (24, "bf3"), "b_label": (25, "bu64"), "b_volume": (26, "bf1"), "b_cell_start": (27, "u32c"),
}


PASSES = ["hash", "sort", "gather", "cell_start", "density", "avg_flow", "voidage_adhesion", "sand_momentum",
          "water_momentum", "advect"]


class SFError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class sf_float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]

    @classmethod
    def of(cls, v):
        return cls(float(v[0]), float(v[1]), float(v[2]))

    def tolist(self):
        return [self.x, self.y, self.z]


class sf_int3(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("z", C.c_int32)]

    def tolist(self):
        return [self.x, self.y, self.z]


class sf_params(C.Structure):
    """Mirror of CudaSeepageflowParams (cuda_seepageflow_params.h:25-69)."""
    _fields_ = [(n, C.c_float) for n in (
        "sph_mass", "sph_density", "sph_particle_radius", "sph_kernel_radius",
        "sph_stiff", "sph_visc", "sph_nu", "sph_bnu",
        "dem_mass", "dem_density", "dem_particle_radius", "dem_kernel_radius",
        "dem_young", "dem_poisson", "dem_tan_friction_angle", "dem_damping",
        "sf_c0", "sf_csat", "sf_cmc", "sf_cmc_p", "sf_cd",
        "sf_a0", "sf_asat", "sf_amc", "sf_amc_p")] + [
        ("sf_dry_sand_color", sf_float3), ("sf_wet_sand_color", sf_float3),
        ("gravity", sf_float3), ("dt", C.c_float),
        ("solver_type", C.c_int32), ("sf_type", C.c_int32)]


class sf_boundary_params(C.Structure):
    """Mirror of CudaBoundaryParams (cuda_boundary_params.h:19-27)."""
    _fields_ = [("kernel_radius", C.c_float), ("lowest_point", sf_float3), ("highest_point", sf_float3),
                ("world_size", sf_float3), ("world_center", sf_float3), ("grid_size", sf_int3)]


class sf_solver_options(C.Structure):
    """Mirror of the CudaWCSphSFSolver constructor arguments (cuda_wcsph_sf_solver.cuh:22-32)."""
    _fields_ = [("negative_scale", C.c_float), ("time_step_limit_scale", C.c_float),
                ("speed_of_sound", C.c_float)]


SF_EMITTER_SQUARE, SF_EMITTER_CIRCLE, SF_EMITTER_RECTANGLE = 0, 1, 2


class sf_emitter_params(C.Structure):
    """CudaSphEmitterParams (cuda_sph_params.h:56-70) as sf_set_emitter takes it."""
    _fields_ = [("enable", C.c_int), ("run", C.c_int), ("emit_pos", sf_float3), ("emit_vel", sf_float3),
                ("emit_col", sf_float3), ("emit_type", C.c_int), ("emit_radius", C.c_float),
                ("emit_width", C.c_float), ("emit_height", C.c_float)]


class sf_volume_data(C.Structure):
    """SeepageflowVolumeData / SeepageflowMultiVolumeData (cuda_volume_emitter.cuh:43-63) as caller arrays."""
    _fields_ = [("pos3", C.c_void_p), ("col3", C.c_void_p), ("label", C.c_void_p), ("mass", C.c_void_p),
                ("radius", C.c_void_p), ("cda0asat3", C.c_void_p), ("amcamcp2", C.c_void_p), ("size", C.c_size_t),
                ("capacity", C.c_size_t), ("sand_min_radius", C.c_float)]


class sf_scene_params(C.Structure):
    """Per-scene overrides of a batched handle (names as in sf_params)."""
    _fields_ = [(n, C.c_float) for n in ("dem_tan_friction_angle", "sf_c0", "sf_cmc", "sf_cmc_p", "sf_csat",
                                         "sf_cd")]


_lib = None

_VP = C.c_void_p
_SIG = {
    "sf_abi_version": (C.c_int, []),
    "sf_last_error": (C.c_char_p, []),
    "sf_validate_params": (C.c_int, [C.POINTER(sf_params), C.POINTER(sf_boundary_params)]),
    "sf_grid_size": (C.c_int, [sf_float3, sf_float3, C.c_float, C.POINTER(sf_int3)]),
    "sf_build_world_boundary": (C.c_int, [sf_float3, sf_float3, C.c_float, C.c_size_t, _VP, _VP,
                                          C.POINTER(C.c_size_t)]),
    "sf_write_bgeo": (C.c_int, [C.c_char_p, C.c_size_t, _VP, _VP, _VP, _VP]),
    "sf_read_bgeo": (C.c_int, [C.c_char_p, C.c_int, _VP, C.c_size_t, C.POINTER(C.c_size_t)]),
    "sf_bgeo_to_vtk": (C.c_int, [C.c_char_p, C.c_char_p]),
    "sf_build_seepage_shape_volume": (C.c_int, [C.POINTER(sf_volume_data), _VP, C.c_size_t, sf_float3, C.c_float,
                                                C.c_int, C.c_float]),
    "sf_build_seepage_shape_multi_volume": (C.c_int, [C.POINTER(sf_volume_data), _VP, C.c_size_t, sf_float3,
                                                      C.c_float, sf_float3, C.c_float, C.c_float, C.c_int,
                                                      C.c_float]),
    "sf_create": (C.c_int, [C.POINTER(sf_params), C.POINTER(sf_boundary_params), C.POINTER(sf_solver_options),
                            C.c_size_t, C.c_int, C.POINTER(_VP)]),
    "sf_destroy": (C.c_int, [_VP]),
    "sf_set_params": (C.c_int, [_VP, C.POINTER(sf_params)]),
    "sf_set_stream": (C.c_int, [_VP, _VP]),
    "sf_synchronize": (C.c_int, [_VP]),
    "sf_upload_particles": (C.c_int, [_VP, C.c_size_t, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "sf_upload_boundary": (C.c_int, [_VP, C.c_size_t, _VP, _VP]),
    "sf_build_neighbors": (C.c_int, [_VP]),
    "sf_step": (C.c_int, [_VP, C.c_float]),
    "sf_step_n": (C.c_int, [_VP, C.c_float, C.c_int]),
    "sf_update_system": (C.c_int, [_VP, C.c_float, C.POINTER(C.c_float)]),
    "sf_step_profiled": (C.c_int, [_VP, C.c_float, C.c_int, _VP, C.POINTER(C.c_float)]),
    "sf_num_sub_steps": (C.c_int, [_VP, C.POINTER(C.c_size_t)]),
    "sf_update_frame": (C.c_int, [_VP, C.c_float, C.c_int, C.POINTER(C.c_size_t), C.POINTER(C.c_float)]),
    "sf_size": (C.c_int, [_VP, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "sf_download": (C.c_int, [_VP, C.c_int, _VP, C.c_size_t]),
    "sf_device_buffer": (C.c_int, [_VP, C.c_int, C.POINTER(_VP), C.POINTER(C.c_size_t)]),
    "sf_set_field": (C.c_int, [_VP, C.c_int, _VP, C.c_size_t]),
    "sf_export_bgeo": (C.c_int, [_VP, C.c_char_p, C.c_char_p]),
    "sf_particle_view": (C.c_int, [_VP, C.c_int, C.POINTER(_VP)]),
    "sf_grid_build": (C.c_int, [_VP, C.c_size_t, sf_float3, C.c_float, sf_int3, _VP, _VP, _VP, _VP]),
    "sf_grid_build_noalloc": (C.c_int, [_VP, C.c_size_t, sf_float3, C.c_float, sf_int3, _VP, _VP, _VP, _VP]),
    "sf_update_solver": (C.c_int, [_VP, C.c_float, C.POINTER(C.c_float)]),
    "sf_export_bgeo_device": (C.c_int, [C.c_char_p, C.c_char_p, _VP, _VP, _VP, _VP, C.c_size_t]),
    "sf_batch_create": (C.c_int, [C.POINTER(sf_params), C.POINTER(sf_boundary_params), C.POINTER(sf_solver_options),
                                  C.c_size_t, _VP, _VP, C.c_int, C.POINTER(_VP)]),
    "sf_batch_set_scene_params": (C.c_int, [_VP, _VP]),
    "sf_batch_scene_offsets": (C.c_int, [_VP, C.POINTER(C.c_size_t), _VP, C.c_size_t]),
    "sf_step_graph": (C.c_int, [_VP, C.c_float, C.c_int]),
    "sf_set_launch_config": (C.c_int, [_VP, C.c_int, C.c_int, C.c_int]),
    "sf_get_launch_config": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "sf_autotune": (C.c_int, [_VP, C.c_int, _VP, _VP]),
    "sf_set_sort_strategy": (C.c_int, [_VP, C.c_int]),
    "sf_set_prefix_scan_threshold": (C.c_int, [C.c_int]),
    "sf_set_interface_lanes": (C.c_int, [_VP, C.c_int]),
    "sf_set_emitter": (C.c_int, [_VP, C.POINTER(sf_emitter_params)]),
    "sf_get_emitter": (C.c_int, [_VP, C.POINTER(C.c_int), C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "sf_step_timed": (C.c_int, [_VP, C.c_float, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "sf_step_timed_every": (C.c_int, [_VP, C.c_float, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float),
                                      C.POINTER(C.c_float)]),
    "sf_diff_forward": (C.c_int, [_VP, C.c_float, C.c_int, C.c_int]),
    "sf_diff_reset": (C.c_int, [_VP, _VP, C.c_int]),
    "sf_diff_get": (C.c_int, [_VP, C.c_int, _VP, C.c_int]),
    "sf_diff_backward": (C.c_int, [_VP, _VP, _VP, _VP, _VP, _VP, _VP, C.c_int]),
}


def lib() -> C.CDLL:
    """Load libsf_seepage.so; fails loudly when it has not been built."""
    global _lib
    if _lib is None:
        path = Path(os.environ.get("SF_SEEPAGE_LIB", LIB_PATH))
        if not path.exists():
            raise ImportError(f"{path} is missing: run `python -m sph_seepage_flows_amd.build` "
                              "(or __graft_entry__.build()) first")
        L = C.CDLL(str(path))
        for name, (res, args) in _SIG.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != SF_OK:
        msg = lib().sf_last_error()
        raise SFError(rc, msg.decode() if msg else "")


def ptr(a) -> int | None:
    if a is None:
        return None
    assert isinstance(a, np.ndarray) and a.flags.c_contiguous
    return a.ctypes.data
