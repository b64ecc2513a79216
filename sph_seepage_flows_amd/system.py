"""Python mirror of the reference's per-frame system API over the C-ABI.

    sys = SFSystem.from_scene(scene)             # CudaSFParticles + searchers + solver + CudaSFSystem
    ms  = sys.UpdateSystem(render_interval)      # CudaBaseSystem::UpdateSystem (cuda_base_system.cu:29-56)
    k   = sys.GetNumOfSubTimeSteps()             # cuda_base_system.cuh:38
    sys.step(dt)                                 # one sub-step with a fixed dt
    pos = sys.get("pos")                         # CudaSFParticles::GetPosPtr() contents, as numpy
    sys.export_bgeo(folder, name)                # ExportBgeoFileCUDA (kiri_utils.h:364-448)

The arithmetic all happens in libsf_seepage.so on the GPU; this class only
moves host arrays across the boundary.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import FIELDS, check, lib, ptr


class SFSystem:
    def __init__(self, params, bparams, opts=None, max_num: int = 0, device: int = -1):
        self._p, self._b, self._o = params, bparams, opts
        h = C.c_void_p()
        check(lib().sf_create(C.byref(params), C.byref(bparams), C.byref(opts) if opts is not None else None,
                              int(max_num), int(device), C.byref(h)))
        self._h = h

    # ---- construction helpers ----
    @classmethod
    def from_scene(cls, scene, max_num: int | None = None, device: int = -1):
        p, b, o = scene.params()
        s = cls(p, b, o, max_num=max(scene.n, 1) if max_num is None else max_num, device=device)
        s.upload_particles(scene)
        if getattr(scene, "max_sat", None) is not None:
            s.set_field("max_saturation", scene.max_sat)
        if scene.bpos is not None:
            s.upload_boundary(scene.bpos, scene.blabel)
        return s

    def upload_particles(self, scene) -> None:
        f = lambda a, dt=np.float32: None if a is None else np.ascontiguousarray(a, dt)
        pos, vel, col = f(scene.pos), f(scene.vel), f(scene.col)
        lab = np.ascontiguousarray(scene.label, np.uint64)
        mass, rad = f(scene.mass), f(scene.radius)
        cda, amc = f(scene.cda0asat), f(scene.amcamcp)
        check(lib().sf_upload_particles(self._h, scene.n, ptr(pos), ptr(vel), ptr(col), ptr(lab), ptr(mass),
                                        ptr(rad), ptr(cda), ptr(amc)))

    def upload_boundary(self, bpos, blabel=None) -> None:
        bp = np.ascontiguousarray(bpos, np.float32)
        bl = None if blabel is None else np.ascontiguousarray(blabel, np.uint64)
        check(lib().sf_upload_boundary(self._h, len(bp), ptr(bp), ptr(bl)))

    # ---- lifetime ----
    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().sf_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self) -> int:
        return self._h.value

    # ---- stepping (reference names) ----
    def BuildGNSearcher(self) -> None:
        check(lib().sf_build_neighbors(self._h))

    def UpdateSystem(self, render_interval: float) -> float:
        ms = C.c_float(0.0)
        check(lib().sf_update_system(self._h, float(render_interval), C.byref(ms)))
        return ms.value

    def update_frame(self, render_interval: float, adaptive: bool = True) -> tuple[int, float]:
        """One frame of the reference driver's sub-step loop in the library
        (sf_update_frame): (sub-steps run, GPU ms)."""
        k, ms = C.c_size_t(0), C.c_float(0.0)
        check(lib().sf_update_frame(self._h, float(render_interval), int(adaptive), C.byref(k), C.byref(ms)))
        return k.value, ms.value

    def UpdateSolver(self, render_interval: float) -> float:
        """CudaWCSphSFSolver::UpdateSolver on the last BuildGNSearcher's
        structures (sf_update_solver): passes, sub-step count, advect; no emission."""
        ms = C.c_float(0.0)
        check(lib().sf_update_solver(self._h, float(render_interval), C.byref(ms)))
        return ms.value

    def GetNumOfSubTimeSteps(self) -> int:
        k = C.c_size_t(0)
        check(lib().sf_num_sub_steps(self._h, C.byref(k)))
        return k.value

    def step(self, dt: float, n: int = 1) -> None:
        check(lib().sf_step_n(self._h, float(dt), int(n)))

    def set_sort_strategy(self, incremental: bool) -> None:
        """Incremental stable sort (default) or the LSD radix sort every step."""
        check(lib().sf_set_sort_strategy(self._h, _abi.SF_SORT_INCREMENTAL if incremental else _abi.SF_SORT_RADIX))

    def set_emitter(self, pos, vel, col, kind: int = _abi.SF_EMITTER_SQUARE, radius: float = 0.0,
                    width: float = 0.0, height: float = 0.0, enable: bool = True, run: bool = True) -> None:
        """Inflow emitter (sf_set_emitter): emits water in UpdateSystem."""
        e = _abi.sf_emitter_params(int(enable), int(run), _abi.sf_float3.of(pos), _abi.sf_float3.of(vel),
                                   _abi.sf_float3.of(col), int(kind), float(radius), float(width), float(height))
        check(lib().sf_set_emitter(self._h, C.byref(e)))

    def emitter_state(self) -> tuple[bool, int, int]:
        """(enabled, samples per emission, particles emitted so far)."""
        e, p, m = C.c_int(), C.c_size_t(), C.c_size_t()
        check(lib().sf_get_emitter(self._h, C.byref(e), C.byref(p), C.byref(m)))
        return bool(e.value), p.value, m.value

    def set_interface_lanes(self, lanes: int) -> None:
        """Lanes per particle (1, 2 or 4; default 4) of the avg-flow and
        voidage/adhesion passes (sf_set_interface_lanes).  More lanes shorten
        the walks where the phases mix; results change at rounding level."""
        check(lib().sf_set_interface_lanes(self._h, int(lanes)))

    def step_graph(self, dt: float, n: int) -> None:
        """n fixed-dt steps replayed from a captured HIP graph (same results as step)."""
        check(lib().sf_step_graph(self._h, float(dt), int(n)))

    NEIGHBOUR_PASSES = ("density", "avg_flow", "voidage_adhesion", "sand_momentum", "water_momentum")

    def autotune(self, reps: int = 10) -> dict:
        """Times every launch shape of the neighbour passes on the current
        particles and keeps the fastest. Returns {pass: (nb, cap, best_ms, default_ms)}."""
        best = np.zeros(5, np.float32)
        dflt = np.zeros(5, np.float32)
        check(lib().sf_autotune(self._h, int(reps), ptr(best), ptr(dflt)))
        out = {}
        for q, name in enumerate(self.NEIGHBOUR_PASSES):
            nb, cap = self.launch_config(name)
            out[name] = (nb, cap, float(best[q]), float(dflt[q]))
        return out

    def launch_config(self, name: str) -> tuple[int, int]:
        nb, cap = C.c_int(), C.c_int()
        check(lib().sf_get_launch_config(self._h, _abi.PASSES.index(name), C.byref(nb), C.byref(cap)))
        return nb.value, cap.value

    def set_launch_config(self, name: str, nb: int, cap: int = 16) -> None:
        check(lib().sf_set_launch_config(self._h, _abi.PASSES.index(name), int(nb), int(cap)))

    def step_profiled(self, dt: float, n: int, per_pass: bool = True):
        """n fixed-dt steps timed with HIP events on the handle's stream.
        Returns (total_ms, {pass: mean ms per step} or None)."""
        pm = np.zeros(len(_abi.PASSES), np.float32) if per_pass else None
        tot = C.c_float(0.0)
        check(lib().sf_step_profiled(self._h, float(dt), int(n), ptr(pm), C.byref(tot)))
        return tot.value, (dict(zip(_abi.PASSES, map(float, pm))) if per_pass else None)

    def step_timed(self, dt: float, n: int, pass_name: str, every: int = 1):
        """n fixed-dt steps with HIP events around one pass only, on every
        `every`-th step (sf_step_timed_every).  Returns (total_ms, mean ms of
        that pass over the steps it was timed on)."""
        pm, tot = C.c_float(0.0), C.c_float(0.0)
        check(lib().sf_step_timed_every(self._h, float(dt), int(n), _abi.PASSES.index(pass_name), int(every),
                                        C.byref(pm), C.byref(tot)))
        return tot.value, pm.value

    def set_params(self, params) -> None:
        check(lib().sf_set_params(self._h, C.byref(params)))
        self._p = params

    def synchronize(self) -> None:
        check(lib().sf_synchronize(self._h))

    def sizes(self) -> tuple[int, int, int]:
        n, nb, c = C.c_size_t(), C.c_size_t(), C.c_size_t()
        check(lib().sf_size(self._h, C.byref(n), C.byref(nb), C.byref(c)))
        return n.value, nb.value, c.value

    # ---- data ----
    def get(self, name: str) -> np.ndarray:
        fid, kind = FIELDS[name]
        n, nb, c = self.sizes()
        cnt = nb if kind.startswith("b") else n
        kind = kind.lstrip("b")
        if kind == "f3":
            out = np.zeros((cnt, 3), np.float32)
        elif kind == "f2":
            out = np.zeros((cnt, 2), np.float32)
        elif kind == "f1":
            out = np.zeros(cnt, np.float32)
        elif kind == "u64":
            out = np.zeros(cnt, np.uint64)
        elif kind == "u32":
            out = np.zeros(cnt, np.uint32)
        elif kind == "u32c":
            if name == "b_cell_start":  # the boundary grid is one scene's grid
                g = self._b.grid_size
                c = g.x * g.y * g.z
            out = np.zeros(c + 1, np.uint32)
        else:
            raise KeyError(name)
        check(lib().sf_download(self._h, fid, ptr(out) if out.size else None, out.nbytes))
        return out

    def set_field(self, name: str, values) -> None:
        """Overwrites vel / col / max_saturation / cda0asat / amcamcp in the current slot order."""
        fid, _ = FIELDS[name]
        a = np.ascontiguousarray(values, np.float32)
        check(lib().sf_set_field(self._h, fid, ptr(a), a.nbytes))

    def device_buffer(self, name: str) -> tuple[int, int]:
        fid, _ = FIELDS[name]
        p, s = C.c_void_p(), C.c_size_t()
        check(lib().sf_device_buffer(self._h, fid, C.byref(p), C.byref(s)))
        return p.value or 0, s.value

    def view(self, name: str) -> int:
        """Device address of the field's reference-layout array (sf_particle_view):
        what CudaSFParticles::Get*Ptr() / GetCellStart() return."""
        fid, _ = FIELDS[name]
        p = C.c_void_p()
        check(lib().sf_particle_view(self._h, fid, C.byref(p)))
        return p.value or 0

    def export_bgeo(self, folder: str, name: str) -> str:
        check(lib().sf_export_bgeo(self._h, folder.encode(), name.encode()))
        return f"{folder}/{name}.bgeo"


def export_bgeo_device(folder: str, name: str, pos_dev: int, col_dev: int, radius_dev: int, label_dev: int,
                       n: int) -> str:
    """ExportBgeoFileCUDA(folder, name, pos, col, radius, label, n) over device
    arrays in the reference's layout (sf_export_bgeo_device)."""
    check(lib().sf_export_bgeo_device(folder.encode(), name.encode(), pos_dev, col_dev, radius_dev, label_dev, int(n)))
    return f"{folder}/{name}.bgeo"


def write_bgeo(path: str, pos, col, radius, label) -> None:
    pos = np.ascontiguousarray(pos, np.float32)
    col = np.ascontiguousarray(col, np.float32)
    radius = np.ascontiguousarray(radius, np.float32)
    label = np.ascontiguousarray(label, np.uint64)
    check(lib().sf_write_bgeo(path.encode(), len(pos), ptr(pos), ptr(col), ptr(radius), ptr(label)))


class BatchSystem(SFSystem):
    """Many independent scenes in one handle (sf_batch_create): every pass is
    one launch over all scenes.  Scenes share the box, grid, boundary and all
    parameters except friction, the capillary curve and the drag coefficient
    (sf_scene_params); scene s keeps the slot range offsets[s]:offsets[s+1]."""

    def __init__(self, params, bparams, opts, counts, scene_params, device: int = -1):
        self._p, self._b, self._o = params, bparams, opts
        counts = np.ascontiguousarray(counts, np.uint64)
        sp = (_abi.sf_scene_params * len(counts))(*scene_params)
        h = C.c_void_p()
        check(lib().sf_batch_create(C.byref(params), C.byref(bparams), C.byref(opts) if opts is not None else None,
                                    len(counts), ptr(counts), C.addressof(sp), int(device), C.byref(h)))
        self._h = h
        self._sp = sp

    @staticmethod
    def scene_params_of(physics) -> "_abi.sf_scene_params":
        p = _abi.sf_scene_params()
        p.dem_tan_friction_angle = float(np.float32(physics.friction))
        for n in ("sf_c0", "sf_cmc", "sf_cmc_p", "sf_csat", "sf_cd"):
            setattr(p, n, float(np.float32(getattr(physics, n))))
        return p

    @classmethod
    def from_scenes(cls, scenes, device: int = -1):
        """scenes: Scene objects with one shared box / grid / radii / boundary;
        each scene's own friction, capillary controls and drag go to its table row."""
        p, b, o = scenes[0].params()
        s = cls(p, b, o, [sc.n for sc in scenes], [cls.scene_params_of(sc.physics) for sc in scenes], device)
        cat = lambda name, dt=np.float32: None if getattr(scenes[0], name) is None else np.ascontiguousarray(
            np.concatenate([getattr(sc, name) for sc in scenes]), dt)
        pos, vel, col = cat("pos"), cat("vel"), cat("col")
        if vel is None and any(sc.vel is not None for sc in scenes):
            vel = np.concatenate([sc.vel if sc.vel is not None else np.zeros((sc.n, 3), np.float32)
                                  for sc in scenes]).astype(np.float32)
        lab = cat("label", np.uint64)
        mass, rad, cda, amc = cat("mass"), cat("radius"), cat("cda0asat"), cat("amcamcp")
        check(lib().sf_upload_particles(s._h, len(pos), ptr(pos), ptr(vel), ptr(col), ptr(lab), ptr(mass),
                                        ptr(rad), ptr(cda), ptr(amc)))
        if any(getattr(sc, "max_sat", None) is not None for sc in scenes):
            s.set_field("max_saturation", np.concatenate(
                [sc.max_sat if sc.max_sat is not None else np.zeros(sc.n, np.float32) for sc in scenes]))
        if scenes[0].bpos is not None:
            s.upload_boundary(scenes[0].bpos, scenes[0].blabel)
        return s

    def set_scene_params(self, scene_params) -> None:
        sp = (_abi.sf_scene_params * len(scene_params))(*scene_params)
        check(lib().sf_batch_set_scene_params(self._h, C.addressof(sp)))
        self._sp = sp

    def offsets(self) -> np.ndarray:
        ns = C.c_size_t()
        check(lib().sf_batch_scene_offsets(self._h, C.byref(ns), None, 0))
        off = np.zeros(ns.value + 1, np.uint64)
        check(lib().sf_batch_scene_offsets(self._h, None, ptr(off), len(off)))
        return off.astype(np.int64)

    def get_scene(self, name: str, s: int) -> np.ndarray:
        off = self.offsets()
        return self.get(name)[off[s]:off[s + 1]]
