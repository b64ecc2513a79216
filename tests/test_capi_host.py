"""CPU-only checks of the C-ABI library: it loads, exports every symbol the
header declares, and its host-only helpers behave (no GPU calls)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import sph_seepage_flows_amd as sf
from sph_seepage_flows_amd import _abi
from sph_seepage_flows_amd.params import Physics, grid_size_py, make_params

ROOT = Path(__file__).resolve().parents[1]


def header_symbols():
    txt = (ROOT / "include" / "sf_seepage.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(sf_\w+)\s*\(", txt, flags=re.M)))


def test_library_exports_every_header_symbol():
    L = C.CDLL(str(_abi.LIB_PATH))
    syms = header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert set(syms) <= set(_abi._SIG), "ctypes signatures must cover the header"


def test_abi_version_and_error_text():
    L = sf.lib()
    assert L.sf_abi_version() == 2
    p, b, o = make_params(Physics())
    bad = _abi.sf_params.from_buffer_copy(p)
    bad.sph_density = -1.0
    rc = L.sf_validate_params(C.byref(bad), C.byref(b))
    assert rc == _abi.SF_E_INVALID
    assert b"sph_density" in L.sf_last_error()
    assert L.sf_validate_params(C.byref(p), C.byref(b)) == 0


def test_grid_size_matches_fp32_truncation():
    for lo, hi, h in [((0, 0, 0), (1, 1, 1), 0.04), ((0, 0, 0), (2, 1, 1), 0.0184), ((-0.3, 0.1, 0), (1.2, 1.5, 2.5), 0.04)]:
        g = _abi.sf_int3()
        _abi.check(sf.lib().sf_grid_size(_abi.sf_float3.of(lo), _abi.sf_float3.of(hi), float(np.float32(h)), g))
        assert tuple(g.tolist()) == grid_size_py(lo, hi, h)
    # extents an int cannot hold, negative or NaN: an error, not a truncation
    for lo, hi, h in [((0, 0, 0), (1, 1, 1), 1e-12), ((0, 0, 0), (-1, 1, 1), 0.04), ((0, 0, 0), (np.nan, 1, 1), 0.04)]:
        g = _abi.sf_int3()
        rc = sf.lib().sf_grid_size(_abi.sf_float3.of(lo), _abi.sf_float3.of(hi), float(np.float32(h)), g)
        assert rc == _abi.SF_E_INVALID
    b = make_params(Physics())[1]
    bad = _abi.sf_boundary_params.from_buffer_copy(b)
    bad.grid_size.x += 1
    assert sf.lib().sf_validate_params(C.byref(make_params(Physics())[0]), C.byref(bad)) == _abi.SF_E_INVALID


def test_world_boundary_matches_oracle(oracle_mod):
    for lo, hi, r in [((0, 0, 0), (1, 1, 1), 0.01), ((0, 0, 0), (2, 1, 1), 0.0046), ((0, 0, 0), (1.5, 1.5, 2.5), 0.01)]:
        p, l = sf.world_boundary(lo, hi, r)
        q, m = oracle_mod.world_boundary(lo, hi, r)
        np.testing.assert_array_equal(p, q)
        np.testing.assert_array_equal(l, m)
    p, _ = sf.world_boundary((0, 0, 0), (1, 1, 1), 0.01)
    assert len(p) == 6 * 51 * 51  # 50 spacings per side -> 51 x 51 per plane, edges duplicated


def test_world_boundary_size_limits():
    """A box of 1e6 spacings per side would hold ~6e12 boundary particles:
    refused up front rather than walked (a count-only call walks too)."""
    n = C.c_size_t(0)
    rc = sf.lib().sf_build_world_boundary(_abi.sf_float3.of((0, 0, 0)), _abi.sf_float3.of((1e4, 1e4, 1e4)), 0.005,
                                          0, None, None, C.byref(n))
    assert rc == _abi.SF_E_INVALID
    rc = sf.lib().sf_build_world_boundary(_abi.sf_float3.of((0, 0, 0)), _abi.sf_float3.of((1, 1, 1)), 0.01, 0, None,
                                          None, C.byref(n))
    assert rc == 0 and n.value == 6 * 51 * 51


def test_world_boundary_capacity_error():
    n = C.c_size_t(0)
    pos = np.zeros((10, 3), np.float32)
    lab = np.zeros(10, np.uint64)
    rc = sf.lib().sf_build_world_boundary(_abi.sf_float3.of((0, 0, 0)), _abi.sf_float3.of((1, 1, 1)), 0.01, 10,
                                          _abi.ptr(pos), _abi.ptr(lab), C.byref(n))
    assert rc == _abi.SF_E_CAPACITY
