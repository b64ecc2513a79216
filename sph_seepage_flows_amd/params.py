"""Parameter presets -> sf_params / sf_boundary_params / sf_solver_options.

Field meanings follow CudaSeepageflowParams (cuda_seepageflow_params.h:25-69)
and the derivations of the reference driver (seepage_flows/src/seepageflow/main.cpp:39-104):
  sph_mass   = 0.8 * (2 r)^3 * rho0                          (main.cpp:44-45)
  sph_nu     = 2 * visc * h * c,  sph_bnu = 0.2 * h * c      (main.cpp:47-51)
  dem_mass   = rho_s * 4/3 pi r_s^3                          (main.cpp:61)
  grid_size  = int3((hi - lo) / h)                           (main.cpp:104)
  sph_stiff  = 0.001 * c^2 * rho0 / 7 with c = 100             (main.cpp:48,53)
`c0` is the solver's speed of sound used by the CFL limit
(CudaWCSphSFSolver::SetSpeedOfSound, cuda_wcsph_sf_solver.cuh:22-32).  A Tait
stiffness of c0^2 rho0 / gamma = 228,571 Pa (c0 = 40) drives the reference's
boundary-particle walls unstable within 50 steps of the small block scene
(max speed 280 m/s), so the presets keep the reference driver's stiffness.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np

from ._abi import (SF_TYPE_MULTI_SF, SF_TYPE_SF, SF_WCSPH_SOLVER, lib, sf_boundary_params, sf_float3, sf_int3,
                   sf_params, sf_solver_options)

DRY_SAND = (0.88, 0.79552, 0.5984)   # main.cpp:79
WET_SAND = (0.38, 0.29, 0.14)        # main.cpp:80
WATER_COLOR = (127 / 255, 205 / 255, 255 / 255)  # main.cpp:90


@dataclass
class Physics:
    lo: tuple = (0.0, 0.0, 0.0)
    hi: tuple = (1.0, 1.0, 1.0)
    r_water: float = 0.01
    r_sand: float = 0.01
    rho0: float = 1000.0
    rho_sand: float = 2650.0
    c0: float = 40.0          # CFL speed of sound
    gamma: float = 7.0
    stiff_sound_speed: float = 100.0  # sound_speed of main.cpp:48 (stiffness and viscosity scale)
    stiff_factor: float = 0.001       # main.cpp:53
    visc: float = 0.05        # main.cpp:47
    boundary_friction: float = 0.2  # main.cpp:50
    young: float = 1e6
    poisson: float = 0.3
    friction: float = 0.5     # tan of the friction angle
    damping: float = 0.4      # main.cpp:66
    sf_c0: float = 0.7        # main.cpp:68-77
    sf_csat: float = 0.0
    sf_cmc: float = 1.0
    sf_cmc_p: float = 0.01
    sf_cd: float = 0.5
    sf_a0: float = 0.0
    sf_asat: float = 0.8
    sf_amc: float = 1.5
    sf_amc_p: float = 0.5
    gravity: tuple = (0.0, -9.81, 0.0)
    dt: float = 1e-4
    solver_type: int = SF_WCSPH_SOLVER
    sf_type: int = SF_TYPE_SF
    negative_scale: float = 0.0
    time_step_limit_scale: float = 3.0

    @property
    def h(self) -> float:
        return float(np.float32(4.0) * np.float32(self.r_water))


def f32(x) -> float:
    return float(np.float32(x))


def make_params(ph: Physics):
    """Builds the three C structs for a Physics preset (FP32 as the reference does)."""
    h = np.float32(ph.h)
    rw = np.float32(ph.r_water)
    rs = np.float32(ph.r_sand)
    rho0 = np.float32(ph.rho0)
    diam = np.float32(2.0) * rw
    p = sf_params()
    p.sph_mass = f32(np.float32(0.8) * diam * diam * diam * rho0)
    p.sph_density = f32(rho0)
    p.sph_particle_radius = f32(rw)
    p.sph_kernel_radius = f32(h)
    cs = np.float32(ph.stiff_sound_speed)
    p.sph_stiff = f32(np.float32(ph.stiff_factor) * cs * cs * rho0 / np.float32(ph.gamma))
    p.sph_visc = f32(ph.visc)
    p.sph_nu = f32((np.float32(ph.visc) + np.float32(ph.visc)) * h * cs)
    p.sph_bnu = f32(np.float32(ph.boundary_friction) * h * cs)
    p.dem_density = f32(ph.rho_sand)
    p.dem_particle_radius = f32(rs)
    p.dem_mass = f32(np.float32(ph.rho_sand) * (np.float32(4.0 / 3.0) * np.float32(3.1415926) * (rs * rs * rs)))
    p.dem_kernel_radius = f32(np.float32(4.0) * rs)
    p.dem_young = f32(ph.young)
    p.dem_poisson = f32(ph.poisson)
    p.dem_tan_friction_angle = f32(ph.friction)
    p.dem_damping = f32(ph.damping)
    for n in ("sf_c0", "sf_csat", "sf_cmc", "sf_cmc_p", "sf_cd", "sf_a0", "sf_asat", "sf_amc", "sf_amc_p"):
        setattr(p, n, f32(getattr(ph, n)))
    p.sf_dry_sand_color = sf_float3.of(DRY_SAND)
    p.sf_wet_sand_color = sf_float3.of(WET_SAND)
    p.gravity = sf_float3.of(ph.gravity)
    p.dt = f32(ph.dt)
    p.solver_type = ph.solver_type
    p.sf_type = ph.sf_type

    b = sf_boundary_params()
    b.kernel_radius = f32(h)
    b.lowest_point = sf_float3.of(ph.lo)
    b.highest_point = sf_float3.of(ph.hi)
    lo = np.array(ph.lo, np.float32)
    hi = np.array(ph.hi, np.float32)
    b.world_size = sf_float3.of(hi - lo)
    b.world_center = sf_float3.of((hi + lo) / np.float32(2.0))
    g = sf_int3()
    from ._abi import check
    check(lib().sf_grid_size(b.lowest_point, b.highest_point, b.kernel_radius, g))
    b.grid_size = g

    o = sf_solver_options()
    o.negative_scale = f32(ph.negative_scale)
    o.time_step_limit_scale = f32(ph.time_step_limit_scale)
    o.speed_of_sound = f32(ph.c0)
    return p, b, o


def grid_size_py(lo, hi, h) -> tuple[int, int, int]:
    """make_int3((hi - lo) / h) evaluated in FP32 without the library."""
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    q = (hi - lo) / np.float32(h)
    return tuple(int(v) for v in q)
