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

    def GetNumOfSubTimeSteps(self) -> int:
        k = C.c_size_t(0)
        check(lib().sf_num_sub_steps(self._h, C.byref(k)))
        return k.value

    def step(self, dt: float, n: int = 1) -> None:
        check(lib().sf_step_n(self._h, float(dt), int(n)))

    def step_profiled(self, dt: float, n: int, per_pass: bool = True):
        """n fixed-dt steps timed with HIP events on the handle's stream.
        Returns (total_ms, {pass: mean ms per step} or None)."""
        pm = np.zeros(len(_abi.PASSES), np.float32) if per_pass else None
        tot = C.c_float(0.0)
        check(lib().sf_step_profiled(self._h, float(dt), int(n), ptr(pm), C.byref(tot)))
        return tot.value, (dict(zip(_abi.PASSES, map(float, pm))) if per_pass else None)

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
            out = np.zeros(c + 1, np.uint32)
        else:
            raise KeyError(name)
        check(lib().sf_download(self._h, fid, ptr(out) if out.size else None, out.nbytes))
        return out

    def device_buffer(self, name: str) -> tuple[int, int]:
        fid, _ = FIELDS[name]
        p, s = C.c_void_p(), C.c_size_t()
        check(lib().sf_device_buffer(self._h, fid, C.byref(p), C.byref(s)))
        return p.value or 0, s.value

    def export_bgeo(self, folder: str, name: str) -> str:
        check(lib().sf_export_bgeo(self._h, folder.encode(), name.encode()))
        return f"{folder}/{name}.bgeo"


def write_bgeo(path: str, pos, col, radius, label) -> None:
    pos = np.ascontiguousarray(pos, np.float32)
    col = np.ascontiguousarray(col, np.float32)
    radius = np.ascontiguousarray(radius, np.float32)
    label = np.ascontiguousarray(label, np.uint64)
    check(lib().sf_write_bgeo(path.encode(), len(pos), ptr(pos), ptr(col), ptr(radius), ptr(label)))
