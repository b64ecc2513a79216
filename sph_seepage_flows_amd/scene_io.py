"""Scene input: bgeo shapes and the sand volumes built from them.

`read_bgeo` is ReadBgeoFileForGPU (seepage_flows/include/kiri_utils.h:252-308)
and `read_multi_bgeo` is ReadMultiBgeoFilesForGPU (kiri_utils.h:310-361), both
through the library's bgeo V5 reader (`sf_read_bgeo`).  `VolumeData` holds
what SeepageflowVolumeData / SeepageflowMultiVolumeData hold
(cuda_volume_emitter.cuh:43-63), and its two builders are
CudaVolumeEmitter::BuildSeepageflowShapeVolume / ...ShapeMultiVolume
(cuda_volume_emitter.cpp:109-175) through `sf_build_seepage_shape_volume` /
`sf_build_seepage_shape_multi_volume`.  `bgeo_to_vtk` is Partio2VTK
(kiri_utils.h:54-250) through `sf_bgeo_to_vtk`.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._abi import check, lib, ptr, sf_float3, sf_volume_data


def read_bgeo(path: str | os.PathLike, flip_yz: bool = False) -> np.ndarray:
    """(n, 4) float32 rows (x, y, z, pscale) of a bgeo V5 file (plain or
    gzip); pscale is 0.01 when the file has none; flip_yz gives (x, z, y)."""
    p = os.fsencode(path)
    n = C.c_size_t(0)
    check(lib().sf_read_bgeo(p, int(flip_yz), None, 0, n))
    out = np.zeros((n.value, 4), np.float32)
    check(lib().sf_read_bgeo(p, int(flip_yz), ptr(out), n.value, n))
    return out


def read_multi_bgeo(paths, flip_yz: bool = False) -> np.ndarray:
    """The points of several files, concatenated in the given order."""
    parts = [read_bgeo(p, flip_yz) for p in paths]
    return np.concatenate(parts) if parts else np.zeros((0, 4), np.float32)


def bgeo_to_vtk(bgeo_path: str | os.PathLike, vtk_path: str | os.PathLike) -> None:
    """Partio2VTK (kiri_utils.h:54-250): a binary legacy VTK unstructured grid
    of the file's particles and point attributes (sf_bgeo_to_vtk)."""
    check(lib().sf_bgeo_to_vtk(os.fsencode(bgeo_path), os.fsencode(vtk_path)))


def pbr_path(root: str | os.PathLike, folder: str, name: str) -> str:
    """<root>bgeo/<folder>/<name>.bgeo, the path ReadBgeoFileForGPU builds
    from DB_PBR_PATH (kiri_utils.h:258-261)."""
    return os.path.join(os.fspath(root), "bgeo", folder, name + ".bgeo")


class VolumeData:
    """Sand particles built from shapes: pos, col, label, mass, radius and
    (multi-material) cda0asat / amcamcp, plus sand_min_radius."""

    def __init__(self):
        self.pos = np.zeros((0, 3), np.float32)
        self.col = np.zeros((0, 3), np.float32)
        self.label = np.zeros(0, np.uint64)
        self.mass = np.zeros(0, np.float32)
        self.radius = np.zeros(0, np.float32)
        self.cda0asat = np.zeros((0, 3), np.float32)
        self.amcamcp = np.zeros((0, 2), np.float32)
        self.sand_min_radius = np.float32(0.0)

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    def _grown(self, extra: int, multi: bool):
        m = self.n + extra
        grow = lambda a, shape, dt=np.float32: np.ascontiguousarray(
            np.concatenate([a, np.zeros((extra,) + shape, dt)]))
        arrays = dict(pos=grow(self.pos, (3,)), col=grow(self.col, (3,)), label=grow(self.label, (), np.uint64),
                      mass=grow(self.mass, ()), radius=grow(self.radius, ()))
        if multi:
            arrays["cda0asat"] = grow(self.cda0asat, (3,))
            arrays["amcamcp"] = grow(self.amcamcp, (2,))
        d = sf_volume_data()
        d.pos3, d.col3 = ptr(arrays["pos"]), ptr(arrays["col"])
        d.label, d.mass, d.radius = ptr(arrays["label"]), ptr(arrays["mass"]), ptr(arrays["radius"])
        if multi:
            d.cda0asat3, d.amcamcp2 = ptr(arrays["cda0asat"]), ptr(arrays["amcamcp"])
        d.size, d.capacity = self.n, m
        d.sand_min_radius = float(self.sand_min_radius)
        return d, arrays

    def _commit(self, d, arrays):
        for k, v in arrays.items():
            setattr(self, k, v)
        self.sand_min_radius = np.float32(d.sand_min_radius)

    def build_shape_volume(self, shape, color, sand_density, offset_y=False, world_lowest_y=0.0):
        """BuildSeepageflowShapeVolume: appends `shape` ((n, 4) rows) as sand."""
        s = np.ascontiguousarray(shape, np.float32)
        d, arrays = self._grown(len(s), multi=False)
        check(lib().sf_build_seepage_shape_volume(C.byref(d), ptr(s), len(s), sf_float3.of(color),
                                                  float(np.float32(sand_density)), int(offset_y),
                                                  float(np.float32(world_lowest_y))))
        self._commit(d, arrays)

    def build_shape_multi_volume(self, shape, color, sand_density, cda0asat, amcamcp, offset_y=False,
                                 world_lowest_y=0.0):
        """BuildSeepageflowShapeMultiVolume: as build_shape_volume, with the
        shape's (cd, a0, asat) and (amc, amcp) on each of its particles."""
        s = np.ascontiguousarray(shape, np.float32)
        if self.n and len(self.cda0asat) != self.n:
            raise ValueError("a volume built by build_shape_volume has no per-particle coefficients")
        d, arrays = self._grown(len(s), multi=True)
        amc = np.asarray(amcamcp, np.float32)
        check(lib().sf_build_seepage_shape_multi_volume(C.byref(d), ptr(s), len(s), sf_float3.of(color),
                                                        float(np.float32(sand_density)), sf_float3.of(cda0asat),
                                                        float(amc[0]), float(amc[1]), int(offset_y),
                                                        float(np.float32(world_lowest_y))))
        self._commit(d, arrays)
