"""Raw device memory for the tests through the HIP runtime the library is
bound to: allocation and copies of the device arrays the library reads or
hands out by address.  torch ships a HIP runtime of its own under the same
soname; whichever of the two the process loaded first serves both, so the
runtime is looked up among the mapped libraries after the library is loaded."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_rt = None


def _library_runtime_path() -> str:
    from sph_seepage_flows_amd._abi import lib

    lib()  # the library and its HIP runtime are mapped now
    with open("/proc/self/maps") as f:
        for line in f:
            path = line.split()[-1] if len(line.split()) >= 6 else ""
            if os.path.basename(path).startswith("libamdhip64.so"):
                return path
    return "/opt/rocm/lib/libamdhip64.so"


def runtime() -> C.CDLL:
    global _rt
    if _rt is None:
        _rt = C.CDLL(_library_runtime_path())
        _rt.hipMemcpy.restype = C.c_int
        _rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _rt.hipMalloc.restype = C.c_int
        _rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _rt.hipFree.restype = C.c_int
        _rt.hipFree.argtypes = [C.c_void_p]
        _rt.hipDeviceSynchronize.restype = C.c_int
        _rt.hipStreamCreate.restype = C.c_int
        _rt.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        _rt.hipStreamDestroy.restype = C.c_int
        _rt.hipStreamDestroy.argtypes = [C.c_void_p]
        _rt.hipStreamSynchronize.restype = C.c_int
        _rt.hipStreamSynchronize.argtypes = [C.c_void_p]
    return _rt


def d2h(addr: int, shape, dtype) -> np.ndarray:
    out = np.zeros(shape, dtype)
    if out.nbytes:
        rc = runtime().hipMemcpy(out.ctypes.data, C.c_void_p(addr), out.nbytes, 2)  # hipMemcpyDeviceToHost
        assert rc == 0, f"hipMemcpy D2H failed ({rc})"
    return out


def h2d(addr: int, a: np.ndarray) -> None:
    a = np.ascontiguousarray(a)
    if a.nbytes:
        rc = runtime().hipMemcpy(C.c_void_p(addr), a.ctypes.data, a.nbytes, 1)  # hipMemcpyHostToDevice
        assert rc == 0, f"hipMemcpy H2D failed ({rc})"


class DeviceArray:
    """A device allocation of the library's runtime, freed on close / deletion."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        rc = runtime().hipMalloc(C.byref(p), max(self.nbytes, 1))
        assert rc == 0, f"hipMalloc failed ({rc})"
        self.ptr = p.value

    @classmethod
    def of(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        h2d(d.ptr, a)
        return d

    def get(self, shape, dtype) -> np.ndarray:
        return d2h(self.ptr, shape, dtype)

    def close(self) -> None:
        if self.ptr:
            runtime().hipFree(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
