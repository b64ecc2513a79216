"""Raw device-memory copies for the tests (HIP runtime through ctypes): reads
and writes the device arrays the library hands out by address."""
from __future__ import annotations

import ctypes as C

import numpy as np

_rt = None


def runtime() -> C.CDLL:
    global _rt
    if _rt is None:
        for name in ("libamdhip64.so", "libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _rt = C.CDLL(name)
                break
            except OSError:
                continue
        if _rt is None:
            raise OSError("HIP runtime library not found")
        _rt.hipMemcpy.restype = C.c_int
        _rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
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
