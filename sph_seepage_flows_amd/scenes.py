"""Deterministic synthetic scenes (numpy, FP32) for the seepage step.

Lattices are filled in (i, j, k) order with i (x) outermost, as
CudaVolumeEmitter::BuildSphVolume does (cuda_volume_emitter.cpp:15-34), at a
spacing of 2r from `origin`, and truncated to the requested count.  Jitter is
per component U(-a, a) from a counter-based SplitMix64 of (seed, stream,
particle, component), so the same scene comes out on every host and thread
count.  (The reference's own jitter draws from std::random_device and is not
reproducible, cuda_volume_emitter.cpp:41-54.)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import params as P
from ._abi import SF_TYPE_MULTI_SF, check, lib, ptr, sf_float3

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    x = x.astype(np.uint64)
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        z = x
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, stream: int, count: int, comps: int) -> np.ndarray:
    """U[0,1) float32 array (count, comps) from (seed, stream, index, component)."""
    idx = np.arange(count * comps, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) * np.uint64(0xD1B54A32D192ED03)) ^ (np.uint64(stream) << np.uint64(40)) ^ idx
    u = splitmix64(key) >> np.uint64(40)  # 24 random bits
    return (u.astype(np.float32) * np.float32(1.0 / (1 << 24))).reshape(count, comps)


def normal(seed: int, stream: int, count: int, comps: int) -> np.ndarray:
    u1 = uniform(seed, stream, count, comps).astype(np.float64)
    u2 = uniform(seed, stream + 1000003, count, comps).astype(np.float64)
    return (np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2 * np.pi * u2)).astype(np.float32)


def lattice(origin, spacing, counts, n, jitter=0.0, seed=0, stream=0) -> np.ndarray:
    nx, ny, nz = counts
    if nx * ny * nz < n:
        raise ValueError(f"lattice {counts} holds {nx * ny * nz} < {n} points")
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ijk = np.stack([i.ravel(), j.ravel(), k.ravel()], 1)[:n].astype(np.float32)
    o = np.asarray(origin, np.float32)
    p = o + ijk * np.float32(spacing)
    if jitter:
        p = p + (uniform(seed, stream, n, 3) * np.float32(2.0) - np.float32(1.0)) * np.float32(jitter)
    return p.astype(np.float32)


@dataclass
class Scene:
    name: str
    physics: P.Physics
    pos: np.ndarray
    label: np.ndarray
    mass: np.ndarray
    radius: np.ndarray
    col: np.ndarray
    vel: np.ndarray | None = None
    cda0asat: np.ndarray | None = None
    amcamcp: np.ndarray | None = None
    bpos: np.ndarray | None = None
    blabel: np.ndarray | None = None
    meta: dict = field(default_factory=dict)
    max_sat: np.ndarray | None = None  # initial maxSaturation history (None: zero)
    # sf_params fields set after the Physics derivations (as a driver that
    # changes CUDA_SEEPAGEFLOW_PARAMS after computing it does)
    overrides: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    @property
    def nb(self) -> int:
        return 0 if self.bpos is None else int(self.bpos.shape[0])

    def params(self):
        p, b, o = P.make_params(self.physics)
        for k, v in self.overrides.items():
            setattr(p, k, v)
        return p, b, o


def world_boundary(lo, hi, r) -> tuple[np.ndarray, np.ndarray]:
    """CudaBoundaryEmitter::BuildWorldBoundary through the C-ABI host helper."""
    L = lib()
    import ctypes as C
    n = C.c_size_t(0)
    check(L.sf_build_world_boundary(sf_float3.of(lo), sf_float3.of(hi), float(np.float32(r)), 0, None, None, n))
    pos = np.zeros((n.value, 3), np.float32)
    lab = np.zeros(n.value, np.uint64)
    check(L.sf_build_world_boundary(sf_float3.of(lo), sf_float3.of(hi), float(np.float32(r)), n.value, ptr(pos),
                                    ptr(lab), n))
    return pos, lab


def _assemble(name, ph, blocks, vel=None, boundary=True, extra_b=None, meta=None):
    """blocks: list of (positions, label)."""
    p, prm = ph, P.make_params(ph)[0]
    pos = np.concatenate([b[0] for b in blocks]).astype(np.float32)
    label = np.concatenate([np.full(len(b[0]), b[1], np.uint64) for b in blocks])
    water = label == 0
    mass = np.where(water, np.float32(prm.sph_mass), np.float32(prm.dem_mass)).astype(np.float32)
    radius = np.where(water, np.float32(prm.sph_particle_radius), np.float32(prm.dem_particle_radius)).astype(np.float32)
    col = np.where(water[:, None], np.array(P.WATER_COLOR, np.float32), np.array(P.DRY_SAND, np.float32)).astype(np.float32)
    cda = np.tile(np.array([p.sf_cd, p.sf_a0, p.sf_asat], np.float32), (len(pos), 1))
    amc = np.tile(np.array([p.sf_amc, p.sf_amc_p], np.float32), (len(pos), 1))
    bpos = blab = None
    if boundary:
        bpos, blab = world_boundary(p.lo, p.hi, p.r_water)
        if extra_b is not None:
            bpos = np.concatenate([bpos, extra_b[0]]).astype(np.float32)
            blab = np.concatenate([blab, extra_b[1]]).astype(np.uint64)
    return Scene(name, ph, np.ascontiguousarray(pos), label, mass, radius, np.ascontiguousarray(col),
                 None if vel is None else np.ascontiguousarray(vel, np.float32), cda, amc, bpos, blab, meta or {})


def small_block(seed: int = 0, sf_type: int = 0) -> Scene:
    """Plain method, small: 16,384 water + 8,192 sand in a 1 m box (r = 0.01,
    h = 4r, dt = 1e-4).  Water: 15 x 25 x 44 lattice at x < 0.3, y < 0.5
    (no jitter).  Sand: 15 x 13 x 44 lattice at x in [0.5, 0.8], jitter
    U(-0.1r, 0.1r) from `seed`."""
    ph = P.Physics(sf_type=sf_type)
    r = ph.r_water
    water = lattice((r, r, r), 2 * r, (15, 25, 44), 16384)
    rs = ph.r_sand
    sand = lattice((0.5 + rs, rs, rs), 2 * rs, (15, 13, 44), 8192, jitter=0.1 * rs, seed=seed, stream=1)
    return _assemble("small_block_24k", ph, [(water, 0), (sand, 1)], meta={"seed": seed})


def drop_in_1m(seed: int = 1, n_water: int = 600_000, n_sand: int = 400_000) -> Scene:
    """Plain GPU drop-in at scale: 600k water + 400k sand in a 2 x 1 x 1 m box.
    The stated volumes hold only 60k + 20k grains at r = 0.01, so the radii
    shrink until the lattices fit: water column 0.6 x 0.8 x 1 m at 2r = 0.0092
    (65 x 86 x 108), sand pile 0.4 x 0.4 x 1 m at x in [0.8, 1.2] with
    2r = 0.0072 (55 x 55 x 138); h = 4 r_water = 0.0184.  Both lattices get
    U(-0.1r, 0.1r) jitter from `seed`."""
    rw, rs = 0.0046, 0.0036
    ph = P.Physics(hi=(2.0, 1.0, 1.0), r_water=rw, r_sand=rs)
    water = lattice((rw, rw, rw), 2 * rw, (65, 86, 108), n_water, jitter=0.1 * rw, seed=seed, stream=0)
    sand = lattice((0.8 + rs, rs, rs), 2 * rs, (55, 55, 138), n_sand, jitter=0.1 * rs, seed=seed, stream=1)
    return _assemble("drop_in_1m", ph, [(water, 0), (sand, 1)], meta={"seed": seed})


def mixed(seed: int = 7, sf_type: int = 0, n_water: int = 1440, n_sand: int = 1152, collider: bool = True,
          vel_sigma: float = 0.2, box: float = 0.5) -> Scene:
    """Parity scene that exercises every coupling term in one step: a water
    block and a sand block interpenetrate (the sand lattice is offset by r),
    both touch the walls, all particles carry random velocities, and a small
    sphere of label-1 collider boundary particles sits in the overlap."""
    ph = P.Physics(hi=(box, box, box), sf_type=sf_type)
    r = ph.r_water
    water = lattice((0.05, 2 * r, 2 * r), 2 * r, (12, 10, 12), n_water, jitter=0.2 * r, seed=seed, stream=0)
    sand = lattice((0.2 + r, r, r), 2 * r, (12, 8, 12), n_sand, jitter=0.1 * r, seed=seed, stream=1)
    vel = normal(seed, 7, n_water + n_sand, 3) * np.float32(vel_sigma)
    extra = None
    if collider:
        # Fibonacci sphere of 120 label-1 boundary particles, radius 0.04, centre in the overlap
        k = np.arange(120, dtype=np.float64) + 0.5
        phi = np.arccos(1 - 2 * k / 120)
        th = np.pi * (1 + 5 ** 0.5) * k
        c = np.array([0.27, 0.12, 0.2])
        sp = c + 0.04 * np.stack([np.cos(th) * np.sin(phi), np.cos(phi), np.sin(th) * np.sin(phi)], 1)
        extra = (sp.astype(np.float32), np.ones(120, np.uint64))
    sc = _assemble(f"mixed_{n_water + n_sand}", ph, [(water, 0), (sand, 1)], vel=vel, extra_b=extra,
                   meta={"seed": seed})
    if sf_type == SF_TYPE_MULTI_SF:
        u = uniform(seed, 11, sc.n, 1)[:, 0]
        sc.cda0asat[:, 0] = np.float32(0.2) + u  # per-particle drag coefficient
    # nonzero saturation history U(0, 0.5) on the grains, so the colour / maxSat
    # update runs both ways (grains wetter and drier than their history)
    ms = np.zeros(sc.n, np.float32)
    ms[n_water:] = uniform(seed, 12, n_sand, 1)[:, 0] * np.float32(0.5)
    sc.max_sat = ms
    return sc


def _block(n, x0, r, max_rows=50, depth=44, cols=15):
    """Lattice extents for n particles of the small-block layout: `cols` x
    rows x `depth` at spacing 2r from (x0 + r, r, r); rows grow up to the box
    height (`max_rows`), beyond that columns grow along x."""
    rows = max(1, math.ceil(n / (cols * depth)))
    if rows > max_rows:
        rows = max_rows
        cols = math.ceil(n / (rows * depth))
    return (cols, rows, depth)


def ensemble(seed: int = 3, n_scenes: int = 256, n_min: int = 5_000, n_max: int = 60_000) -> list[Scene]:
    """Batched ensemble: `n_scenes` water-block + sand-pile scenes of the
    small-block kind (1 m box, r = 0.01, h = 4r), each with its own total count
    log-uniform in [n_min, n_max], water fraction U(0.3, 0.7), friction
    U(0.3, 0.7), capillary coefficient sigma U(0, 0.05) and drag coefficient D
    log-uniform in [1e3, 1e4], all from `seed`.  Capillary sigma enters the
    Bezier controls as c0 = 0.7 sigma, cmc = sigma (cmc_p = 0.01, csat = 0, the
    shape of the reference driver's curve, main.cpp:68-72); the drag enters as
    cd = g rho0 / D (the reference's drag is g rho0 eps^2 / (cd eps^3 / (1 - eps)^2),
    cuda_sph_sf_solver_gpu.cuh:858-862).  Water blocks start at x = 0, sand piles at
    x = 0.5; sand gets U(-0.1r, 0.1r) jitter."""
    u = uniform(seed, 0, n_scenes, 5).astype(np.float64)
    out = []
    for q in range(n_scenes):
        n = int(round(math.exp(math.log(n_min) + u[q, 0] * (math.log(n_max) - math.log(n_min)))))
        nw = int(round(n * (0.3 + 0.4 * u[q, 1])))
        ns = n - nw
        sigma = 0.05 * u[q, 3]
        D = math.exp(math.log(1e3) + u[q, 4] * (math.log(1e4) - math.log(1e3)))
        ph = P.Physics(friction=0.3 + 0.4 * u[q, 2], sf_c0=0.7 * sigma, sf_cmc=sigma, sf_cmc_p=0.01, sf_csat=0.0,
                       sf_cd=9.81 * 1000.0 / D)
        r, rs = ph.r_water, ph.r_sand
        water = lattice((r, r, r), 2 * r, _block(nw, 0.0, r), nw)
        sand = lattice((0.5 + rs, rs, rs), 2 * rs, _block(ns, 0.5, rs), ns, jitter=0.1 * rs, seed=seed,
                       stream=100 + q)
        out.append(_assemble(f"ensemble_{q}", ph, [(water, 0), (sand, 1)],
                             meta={"seed": seed, "scene": q, "sigma": sigma, "drag_D": D}))
    return out


def dam(seed: int = 2, n_sand: int = 2_500_000, n_water: int = 1_500_000, r: float = 0.005,
        base: float = 2.0, crest: float = 0.5, height: float = 0.6, depth: float = 0.55) -> Scene:
    """Seepage embankment: a trapezoidal sand dam (base 2 m, crest 0.5 m,
    height 0.6 m) with the water reservoir upstream at 0.55 m depth, both
    extruded along z far enough to hold the counts (2.5M sand + 1.5M water at
    r = 0.005: cross-sections 0.75 m^2 and about 0.45 m^2 give about 3.34 m).
    Lattices at spacing 2r are filled slab by slab along z (z outermost, so the
    truncation to the exact count trims the last slab) and within a slab in
    (x, y) order.  Sand: friction 0.6, capillary sigma 0.02 (c0 = 0.7 sigma,
    cmc = sigma), drag D = 5e3 (cd = g rho0 / D), U(-0.1r, 0.1r) jitter and an
    initial maxSaturation U(0, 0.1) per grain from `seed`.  The reservoir
    keeps one lattice spacing of clearance from the upstream face.  Box: x from
    0 to 0.3 m past the downstream toe, y to 1 m."""
    sp = 2 * r
    slope = (base - crest) / 2 / height  # horizontal run per unit height of each face

    def sand_slab(x0):
        xs = x0 + r + sp * np.arange(int((base) / sp) + 1)
        ys = r + sp * np.arange(int(height / sp) + 1)
        X, Y = np.meshgrid(xs, ys, indexing="ij")
        keep = (Y <= height) & (X >= x0 + slope * Y) & (X <= x0 + base - slope * Y)
        return np.stack([X[keep], Y[keep]], 1)

    def water_slab(x0):
        xs = r + sp * np.arange(int((x0 + slope * depth) / sp) + 1)
        ys = r + sp * np.arange(int(depth / sp) + 1)
        X, Y = np.meshgrid(xs, ys, indexing="ij")
        keep = (Y <= depth) & (X <= x0 + slope * Y - sp)
        return np.stack([X[keep], Y[keep]], 1)

    slab_s = sand_slab(0.0)
    n_slabs = math.ceil(n_sand / len(slab_s))
    x0 = 0.0
    while len(water_slab(x0)) * n_slabs < n_water:
        x0 += sp
    slab_s = sand_slab(x0)
    n_slabs = math.ceil(n_sand / len(slab_s))
    slab_w = water_slab(x0)
    n_slabs = max(n_slabs, math.ceil(n_water / len(slab_w)))

    def extrude(slab, n):
        k = np.arange(math.ceil(n / len(slab)))
        z = r + sp * k
        P = np.empty((len(k), len(slab), 3), np.float64)
        P[:, :, :2] = slab[None]
        P[:, :, 2] = z[:, None]
        return P.reshape(-1, 3)[:n].astype(np.float32)

    hi_x = x0 + base + 0.3
    h = 4 * r
    hi_z = math.ceil((n_slabs * sp + sp) / h) * h
    ph = P.Physics(hi=(float(np.float32(hi_x)), 1.0, float(np.float32(hi_z))), r_water=r, r_sand=r, friction=0.6,
                   sf_c0=0.7 * 0.02, sf_cmc=0.02, sf_cmc_p=0.01, sf_csat=0.0, sf_cd=9.81 * 1000.0 / 5e3)
    sand = extrude(slab_s, n_sand)
    sand = sand + (uniform(seed, 1, n_sand, 3) * np.float32(2.0) - np.float32(1.0)) * np.float32(0.1 * r)
    water = extrude(slab_w, n_water)
    sc = _assemble("dam_4m", ph, [(water, 0), (sand.astype(np.float32), 1)],
                   meta={"seed": seed, "dam_x0": x0, "slabs": n_slabs})
    ms = np.zeros(sc.n, np.float32)
    ms[n_water:] = uniform(seed, 3, n_sand, 1)[:, 0] * np.float32(0.1)
    sc.max_sat = ms
    return sc


def column(seed: int = 4, n_sand: int = 48_000, n_water: int = 16_000, width: float = 0.2, height: float = 0.4,
           pour_speed: float = 1.0, vel_sigma: float = 0.01, drag_D: float = 5e3, sigma: float = 0.02,
           friction: float = 0.5) -> Scene:
    """Differentiable seepage column: `n_sand` grains filling a width x height
    x width column (the spacing 2r that packs 48k grains into 0.2 x 0.4 x 0.2 m
    is 0.00693 m, so r = 0.003466), with `n_water` water particles (same r) in
    a block one spacing above it, poured downwards at `pour_speed`.  Both
    lattices are filled layer by layer from the bottom (y outermost), so the
    truncation to the exact count trims the top layer.  Sand positions get
    U(-0.1r, 0.1r) jitter (no pair starts exactly touching, where the liquid
    bridge switches on) and every particle an initial velocity N(0, vel_sigma)
    per component, both from `seed`.  Physics
    controls as in the ensemble: drag D (cd = g rho0 / D), capillary sigma
    (c0 = 0.7 sigma, cmc = sigma), friction = tan of the friction angle."""
    sp = (width * width * height / n_sand) ** (1.0 / 3.0)
    r = sp / 2
    nxz = int(width / sp)

    def layers(n, y0):
        ny = math.ceil(n / (nxz * nxz))
        j, i, k = np.meshgrid(np.arange(ny), np.arange(nxz), np.arange(nxz), indexing="ij")
        p = np.stack([r + i.ravel() * sp, y0 + r + j.ravel() * sp, r + k.ravel() * sp], 1)[:n]
        return p.astype(np.float32), ny

    sand, nys = layers(n_sand, 0.0)
    sand = (sand + (uniform(seed, 1, n_sand, 3) * np.float32(2.0) - np.float32(1.0)) * np.float32(0.1 * r)).astype(
        np.float32)
    water, nyw = layers(n_water, (nys + 1) * sp)
    top = (nys + 1 + nyw + 2) * sp
    h = 4 * r
    hi_y = math.ceil(top / h) * h
    ph = P.Physics(hi=(width, float(np.float32(hi_y)), width), r_water=r, r_sand=r, friction=friction,
                   sf_c0=0.7 * sigma, sf_cmc=sigma, sf_cmc_p=0.01, sf_csat=0.0, sf_cd=9.81 * 1000.0 / drag_D)
    vel = normal(seed, 5, n_water + n_sand, 3) * np.float32(vel_sigma)
    vel[:n_water, 1] -= np.float32(pour_speed)
    return _assemble(f"column_{n_water + n_sand}", ph, [(water, 0), (sand, 1)], vel=vel,
                     meta={"seed": seed, "spacing": sp, "drag_D": drag_D, "sigma": sigma, "friction": friction})


SCENES = {"small_block": small_block, "drop_in_1m": drop_in_1m, "mixed": mixed, "dam": dam, "column": column}


def main_emitter(ph: P.Physics) -> dict:
    """The reference driver's inflow emitter (main.cpp:85-95) for the box of
    `ph`: a disc of radius 0.17 at centre + (size.x / 10, size.y / 5,
    size.z / 2.3) shooting along -z at 5 m/s; keyword arguments of
    SFSystem.set_emitter."""
    from ._abi import SF_EMITTER_CIRCLE
    lo, hi = np.array(ph.lo, np.float32), np.array(ph.hi, np.float32)
    size, centre = hi - lo, (hi + lo) / np.float32(2.0)
    pos = centre + size / np.array([10.0, 5.0, 2.3], np.float32)
    col = np.array([127.0, 205.0, 255.0], np.float32) / np.float32(255.0)
    return dict(pos=tuple(float(v) for v in pos), vel=(0.0, 0.0, -5.0), col=tuple(float(v) for v in col),
                kind=SF_EMITTER_CIRCLE, radius=0.17, width=0.22, height=0.18)


def shape_scene(volume, name: str = "shape", lo=(0.0, 0.0, 0.0), hi=(1.5, 1.5, 2.5)) -> Scene:
    """The reference driver's bunny scene (main.cpp:25-221) around sand built
    from bgeo shapes (`volume`, a scene_io.VolumeData of multi-material
    shapes): 1.5 x 1.5 x 2.5 m box, water r = 0.01 (h = 0.04, no water at
    start: it comes from `main_emitter`), sand density 2700, E = 1e5,
    MULTI_SF, gravity -9.8, solver speed of sound 100 (the solver's default),
    dem_particle_radius = the smallest shape radius and
    dt = 0.5 R / sqrt(E / rho_s) with that R (main.cpp:171-172); dem_mass
    keeps its value for r = 0.005 as main.cpp:61 computes it first."""
    ph = P.Physics(lo=tuple(lo), hi=tuple(hi), r_water=0.01, r_sand=0.005, rho_sand=2700.0, young=1e5,
                   gravity=(0.0, -9.8, 0.0), c0=100.0, sf_type=SF_TYPE_MULTI_SF)
    R = np.float32(volume.sand_min_radius)
    dt = np.float32(0.5) * R / np.sqrt(np.float32(ph.young) / np.float32(ph.rho_sand))
    ph.dt = float(dt)
    if len(volume.cda0asat) != volume.n:
        raise ValueError("shape_scene takes a multi-material volume (build_shape_multi_volume)")
    bpos, blab = world_boundary(ph.lo, ph.hi, ph.r_water)
    return Scene(name, ph, np.ascontiguousarray(volume.pos), volume.label.copy(), volume.mass.copy(),
                 volume.radius.copy(), np.ascontiguousarray(volume.col), None, volume.cda0asat.copy(),
                 volume.amcamcp.copy(), bpos, blab, meta={"emitter": main_emitter(ph), "max_num": 500_000},
                 overrides={"dem_particle_radius": float(R)})

