"""sph_seepage_flows_amd -- MI355X (gfx950) coupled SPH-DEM seepage step.

The compute lives in libsf_seepage.so (hand-written HIP kernels behind the
C-ABI of include/sf_seepage.h).  This package is the thin Python host side:
ctypes bindings, parameter presets, deterministic scene generators and the
SFSystem mirror of the reference's CudaSFSystem API.
"""
from ._abi import (SF_SPH_SOLVER, SF_TYPE_MULTI_SF, SF_TYPE_SF, SF_WCSPH_SOLVER, SFError, lib, sf_boundary_params,
                   sf_params, sf_solver_options)
from .params import Physics, make_params
from .scenes import Scene, drop_in_1m, mixed, small_block, world_boundary
from .system import BatchSystem, SFSystem, write_bgeo

__all__ = [
    "SFError", "SFSystem", "BatchSystem", "Scene", "Physics", "make_params", "small_block", "drop_in_1m", "mixed",
    "world_boundary", "write_bgeo", "lib", "sf_params", "sf_boundary_params", "sf_solver_options",
    "SF_SPH_SOLVER", "SF_WCSPH_SOLVER", "SF_TYPE_SF", "SF_TYPE_MULTI_SF",
]
