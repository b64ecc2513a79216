"""ctypes wrapper of the CPU oracle (oracle/sf_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker / reported baseline and
never as the thing measured or shipped.  The product library does not load
or link it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libsf_oracle.so"

FIELD_IDS = {
    "pos": (0, 3), "vel": (1, 3), "acc": (2, 3), "col": (3, 3), "label": (4, "u64"), "mass": (5, 1),
    "radius": (6, 1), "density": (7, 1), "pressure": (8, 1), "voidage": (9, 1), "saturation": (10, 1),
    "max_saturation": (11, 1), "avg_flow_vel": (12, 3), "avg_drag_force": (13, 3), "adhesion_force": (14, 3),
    "avg_adhesion_force": (15, 3), "cda0asat": (16, 3), "particle2cell": (18, "u32"), "sort_perm": (19, "u32"),
    "sorted_keys": (20, "u32"), "cell_start": (21, "u32c"), "nbr_count": (22, "u32"),
    "nbr_candidates": (23, "u32"), "b_pos": (24, "b3"), "b_label": (25, "bu64"), "b_volume": (26, "b1"),
    "b_cell_start": (27, "u32c"),
}

_lib = None


def build(force: bool = False) -> Path:
    if force or not LIB.exists() or LIB.stat().st_mtime < (HERE / "sf_oracle.cpp").stat().st_mtime:
        subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        VP, D, F, SZ = C.c_void_p, C.c_double, C.c_float, C.c_size_t
        sig = {
            "orc_last_error": (C.c_char_p, []),
            "orc_threads": (C.c_int, [C.c_int]),
            "orc_poly6": (D, [D, D]),
            "orc_spiky_grad": (None, [D, VP, VP]),
            "orc_adhesion": (D, [D, D]),
            "orc_bezier": (D, [D, D, D, D, D]),
            "orc_dem_force": (None, [VP, VP, D, D, D, D, VP]),
            "orc_capillary_force": (None, [VP, D, D, VP, VP]),
            "orc_capillary_rupture": (D, [D]),
            "orc_world_boundary_force": (None, [VP, VP, D, D, D, D, D, VP, VP, VP]),
            "orc_hash": (None, [VP, SZ, VP, F, VP, VP]),
            "orc_world_boundary": (SZ, [VP, VP, F, VP, VP]),
            "orc_create": (VP, [C.c_int, VP, VP, VP]),
            "orc_destroy": (None, [VP]),
            "orc_set_params": (None, [VP, VP]),
            "orc_upload_particles": (None, [VP, SZ, VP, VP, VP, VP, VP, VP, VP, VP]),
            "orc_upload_boundary": (None, [VP, SZ, VP, VP]),
            "orc_build_neighbors": (None, [VP]),
            "orc_step": (None, [VP, F, C.c_int]),
            "orc_update_system": (None, [VP, F]),
            "orc_num_sub_steps": (SZ, [VP]),
            "orc_set_emitter": (None, [VP, C.c_int, C.c_int, VP, VP, VP, C.c_int, F, F, F, SZ]),
            "orc_emitter_state": (None, [VP, VP, VP, VP]),
            "orc_size": (SZ, [VP]),
            "orc_get": (C.c_int, [VP, C.c_int, VP]),
            "orc_set_state": (None, [VP, VP, VP]),
            "orc_set_maxsat": (None, [VP, VP]),
            "orc_get_ids": (None, [VP, VP]),
            "orc_tangent_window": (C.c_int, [VP, VP, VP, SZ, VP, VP, VP, VP, VP, VP, SZ, VP, VP, F, C.c_int,
                                             C.c_int, VP, VP, VP, VP, VP, VP, VP]),
        }
        for k, (r, a) in sig.items():
            fn = getattr(L, k)
            fn.restype = r
            fn.argtypes = a
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags.c_contiguous
    return a.ctypes.data


def threads(k: int = 0) -> int:
    return lib().orc_threads(int(k))


class OracleSystem:
    """Same call sequence as SFSystem: upload, build_neighbors, step, update_system, get."""

    def __init__(self, scene, precision: int = 64, params=None):
        p, b, o = params if params is not None else scene.params()
        self._keep = (p, b, o)
        self.C = int(b.grid_size.x) * int(b.grid_size.y) * int(b.grid_size.z)
        self._h = lib().orc_create(int(precision), C.byref(p), C.byref(b), C.byref(o))
        f = lambda a, dt=np.float32: None if a is None else np.ascontiguousarray(a, dt)
        self._arr = [f(scene.pos), f(scene.vel), f(scene.col), np.ascontiguousarray(scene.label, np.uint64),
                     f(scene.mass), f(scene.radius), f(scene.cda0asat), f(scene.amcamcp)]
        lib().orc_upload_particles(self._h, scene.n, *[_p(a) for a in self._arr])
        if getattr(scene, "max_sat", None) is not None:
            self._ms = np.ascontiguousarray(scene.max_sat, np.float32)
            lib().orc_set_maxsat(self._h, _p(self._ms))
        self.nb = 0
        if scene.bpos is not None:
            bp = np.ascontiguousarray(scene.bpos, np.float32)
            bl = np.ascontiguousarray(scene.blabel, np.uint64)
            lib().orc_upload_boundary(self._h, len(bp), _p(bp), _p(bl))
            self.nb = len(bp)

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n(self) -> int:
        return int(lib().orc_size(self._h))

    def set_params(self, p):
        self._keep = (p,) + self._keep[1:]
        lib().orc_set_params(self._h, C.byref(p))

    def build_neighbors(self):
        lib().orc_build_neighbors(self._h)

    def step(self, dt: float, n: int = 1):
        lib().orc_step(self._h, float(dt), int(n))

    def update_system(self, render_interval: float):
        lib().orc_update_system(self._h, float(render_interval))

    def num_sub_steps(self) -> int:
        return int(lib().orc_num_sub_steps(self._h))

    def size(self) -> int:
        return int(lib().orc_size(self._h))

    def set_emitter(self, cap: int, pos, vel, col, kind: int = 0, radius: float = 0.0, width: float = 0.0,
                    height: float = 0.0, enable: bool = True, run: bool = True):
        """Inflow emitter (kind 0 square, 1 circle, 2 rectangle); cap = the GPU
        handle's max particle count (emission needs n + samples < cap)."""
        f3 = [np.ascontiguousarray(a, np.float32) for a in (pos, vel, col)]
        lib().orc_set_emitter(self._h, int(enable), int(run), *[_p(a) for a in f3], int(kind), float(radius),
                              float(width), float(height), int(cap))
        self._keep_em = f3

    def emitter_state(self):
        e, p, m = C.c_int(), C.c_size_t(), C.c_size_t()
        lib().orc_emitter_state(self._h, C.byref(e), C.byref(p), C.byref(m))
        return bool(e.value), p.value, m.value

    def set_state(self, pos, vel):
        pos = np.ascontiguousarray(pos, np.float64)
        vel = np.ascontiguousarray(vel, np.float64)
        lib().orc_set_state(self._h, _p(pos), _p(vel))

    def ids(self) -> np.ndarray:
        """Upload slot of each particle, in the current (sorted) order."""
        out = np.zeros(self.n, np.uint32)
        lib().orc_get_ids(self._h, _p(out))
        return out

    def get_by_id(self, name: str) -> np.ndarray:
        """A per-particle field in upload order."""
        v = self.get(name)
        out = np.empty_like(v)
        out[self.ids()] = v
        return out

    def get(self, name: str) -> np.ndarray:
        fid, kind = FIELD_IDS[name]
        n = self.n
        if kind in ("b3", "b1", "bu64"):
            cnt, kind = self.nb, {"b3": 3, "b1": 1, "bu64": "u64"}[kind]
        else:
            cnt = n
        if kind == 3:
            out = np.zeros((cnt, 3), np.float64)
        elif kind == 1:
            out = np.zeros(cnt, np.float64)
        elif kind == "u64":
            out = np.zeros(cnt, np.uint64)
        elif kind == "u32":
            out = np.zeros(cnt, np.uint32)
        elif kind == "u32c":
            out = np.zeros(self.C + 1, np.uint32)
        rc = lib().orc_get(self._h, fid, _p(out) if out.size else None)
        if rc:
            raise RuntimeError(lib().orc_last_error().decode())
        return out


def tangent_window(scene, dt: float, steps: int, vdot=None, pdot=None, params=None, xdot=None):
    """FP64 forward-mode tangent of `steps` fixed-dt steps from the scene's
    initial state.  vdot: (ndir, n, 3) initial-velocity directions; pdot:
    (ndir, 4) directions in (sf_cd, sf_c0, sf_cmc, dem_tan_friction_angle);
    xdot: (ndir, n, 3) initial-position directions;
    ndir <= 4.  Returns (pos (n,3), sat (n), posdot (ndir,n,3), satdot (ndir,n))
    in upload order; sat is the last step's saturation."""
    p, b, o = params if params is not None else scene.params()
    n = scene.n
    given = [a for a in (vdot, pdot, xdot) if a is not None]
    ndir = int(given[0].shape[0]) if given else 0
    f = lambda a, dt_=np.float32: None if a is None else np.ascontiguousarray(a, dt_)
    pos, vel, lab = f(scene.pos), f(scene.vel), np.ascontiguousarray(scene.label, np.uint64)
    mass, rad, ms = f(scene.mass), f(scene.radius), f(getattr(scene, "max_sat", None))
    bp = f(scene.bpos) if scene.bpos is not None else None
    bl = np.ascontiguousarray(scene.blabel, np.uint64) if scene.blabel is not None else None
    vd = None if vdot is None else np.ascontiguousarray(vdot, np.float64)
    pd = None if pdot is None else np.ascontiguousarray(pdot, np.float64)
    xd = None if xdot is None else np.ascontiguousarray(xdot, np.float64)
    out_p = np.zeros((n, 3)); out_s = np.zeros(n)
    pdt = np.zeros((max(ndir, 1), n, 3)); sdt = np.zeros((max(ndir, 1), n))
    rc = lib().orc_tangent_window(C.byref(p), C.byref(b), C.byref(o), n, _p(pos), _p(vel), _p(lab), _p(mass),
                                  _p(rad), _p(ms), 0 if bp is None else len(bp), _p(bp), _p(bl), float(dt),
                                  int(steps), ndir, _p(vd), _p(xd), _p(pd), _p(out_p), _p(out_s), _p(pdt),
                                  _p(sdt))
    if rc:
        raise RuntimeError(lib().orc_last_error().decode())
    return out_p, out_s, pdt[:ndir], sdt[:ndir]


# ---- scalar restatements (double) for the pinning tests ----
def poly6(h, r):
    return lib().orc_poly6(float(h), float(r))


def spiky_grad(h, r3):
    r = np.ascontiguousarray(r3, np.float64)
    o = np.zeros(3)
    lib().orc_spiky_grad(float(h), _p(r), _p(o))
    return o


def adhesion(h, r):
    return lib().orc_adhesion(float(h), float(r))


def bezier(py0, py1, px1, py2, sr):
    return lib().orc_bezier(float(py0), float(py1), float(px1), float(py2), float(sr))


def dem_force(d3, v3, rij, kn, ks, tanphi):
    d = np.ascontiguousarray(d3, np.float64)
    v = np.ascontiguousarray(v3, np.float64)
    o = np.zeros(3)
    lib().orc_dem_force(_p(d), _p(v), float(rij), float(kn), float(ks), float(tanphi), _p(o))
    return o


def capillary_force(d3, R, sr, bez4):
    d = np.ascontiguousarray(d3, np.float64)
    b = np.ascontiguousarray(bez4, np.float64)
    o = np.zeros(3)
    lib().orc_capillary_force(_p(d), float(R), float(sr), _p(b), _p(o))
    return o


def capillary_rupture(R):
    return lib().orc_capillary_rupture(float(R))


def world_boundary_force(p3, v3, R, Rb, young, poisson, tanphi, lo3, hi3):
    a = [np.ascontiguousarray(x, np.float64) for x in (p3, v3, lo3, hi3)]
    o = np.zeros(3)
    lib().orc_world_boundary_force(_p(a[0]), _p(a[1]), float(R), float(Rb), float(young), float(poisson),
                                   float(tanphi), _p(a[2]), _p(a[3]), _p(o))
    return o


def hash_keys(pos, lo, h, grid):
    pos = np.ascontiguousarray(pos, np.float32)
    lo = np.ascontiguousarray(lo, np.float32)
    g = np.ascontiguousarray(grid, np.int32)
    out = np.zeros(len(pos), np.uint32)
    lib().orc_hash(_p(pos), len(pos), _p(lo), float(np.float32(h)), _p(g), _p(out))
    return out


def world_boundary(lo, hi, r):
    lo = np.ascontiguousarray(lo, np.float32)
    hi = np.ascontiguousarray(hi, np.float32)
    n = lib().orc_world_boundary(_p(lo), _p(hi), float(np.float32(r)), None, None)
    pos = np.zeros((n, 3), np.float32)
    lab = np.zeros(n, np.uint64)
    lib().orc_world_boundary(_p(lo), _p(hi), float(np.float32(r)), _p(pos), _p(lab))
    return pos, lab
