"""sf_grid_build: the uniform-grid searcher (CudaGNBaseSearcher::BuildGNSearcher,
cuda_neighbor_searcher.cu:36-51) on caller device arrays.  Keys, the stable
permutation and cellStart are compared bit for bit with the oracle's FP32
hash restatement, numpy's stable argsort and a lower-bound count, for grids
whose radix sort takes one, two and three passes (the work pair and the
caller's pair swap roles with the pass count), and on the edge cases."""
import numpy as np
import pytest

from sph_seepage_flows_amd import _abi, scenes
from sph_seepage_flows_amd._abi import check, lib, sf_float3, sf_int3

from _hip import DeviceArray

pytestmark = pytest.mark.gpu


def grid_build_ref(oracle_mod, pos, lo, h, g):
    keys = oracle_mod.hash_keys(pos, lo, h, g)
    perm = np.argsort(keys, kind="stable").astype(np.uint32)
    sk = keys[perm]
    C_ = int(g[0]) * int(g[1]) * int(g[2])
    cs = np.searchsorted(sk, np.arange(C_ + 1, dtype=np.uint64), side="left").astype(np.uint32)
    return sk, perm, cs


def run_grid(pos, lo, h, g, fn="sf_grid_build"):
    """one grid build through the C-ABI on device arrays of the library's runtime"""
    n = len(pos)
    C_ = int(g[0]) * int(g[1]) * int(g[2])
    p = DeviceArray.of(np.ascontiguousarray(pos, np.float32).reshape(-1, 3))
    sk = DeviceArray.of(np.full(max(n, 1), 0x7F7F7F7F, np.uint32))
    pm = DeviceArray.of(np.full(max(n, 1), 0x7F7F7F7F, np.uint32))
    cs = DeviceArray.of(np.full(C_ + 1, 0x7F7F7F7F, np.uint32))
    check(getattr(lib(), fn)(p.ptr if n else None, n, sf_float3.of(lo), float(np.float32(h)),
                             sf_int3(*[int(v) for v in g]), sk.ptr if n else None, pm.ptr if n else None, cs.ptr,
                             None))
    return sk.get(n, np.uint32), pm.get(n, np.uint32), cs.get(C_ + 1, np.uint32)


@pytest.mark.parametrize("name", ["small_block", "drop_in_1m", "mixed"])
def test_grid_build_matches_oracle(name, oracle_mod):
    sc = {"small_block": scenes.small_block, "drop_in_1m": scenes.drop_in_1m, "mixed": scenes.mixed}[name]()
    p, b, _ = sc.params()
    lo = np.array(sc.physics.lo, np.float32)
    g = (b.grid_size.x, b.grid_size.y, b.grid_size.z)
    sk, pm, cs = run_grid(sc.pos, lo, b.kernel_radius, g)
    rsk, rpm, rcs = grid_build_ref(oracle_mod, sc.pos, lo, b.kernel_radius, g)
    print(f"[grid] {name}: n={sc.n} cells={len(rcs) - 1}", flush=True)
    np.testing.assert_array_equal(sk, rsk)
    np.testing.assert_array_equal(pm, rpm)
    np.testing.assert_array_equal(cs, rcs)


def test_grid_build_edge_cases(oracle_mod):
    rng = np.random.default_rng(3)
    lo = np.zeros(3, np.float32)
    # one radix pass (216 cells), particles outside the box, on faces and non-finite
    pts = rng.uniform(-0.2, 1.2, (5000, 3)).astype(np.float32)
    pts[:7] = [[np.nan, 0.5, 0.5], [np.inf, 0.5, 0.5], [-np.inf, 0.5, 0.5], [1.0, 1.0, 1.0], [0.0, 0.0, 0.0],
               [0.5, np.nan, -1.0], [0.99999994, 0.5, 0.5]]
    for g, h, pos in (((6, 6, 6), 1.0 / 6, pts), ((25, 25, 25), 0.04, pts), ((1, 1, 1), 1.0, pts[:100]),
                      ((25, 25, 25), 0.04, np.tile(np.float32([[0.5, 0.5, 0.5]]), (3000, 1))),
                      ((25, 25, 25), 0.04, pts[:1])):
        sk, pm, cs = run_grid(pos, lo, h, g)
        rsk, rpm, rcs = grid_build_ref(oracle_mod, pos, lo, h, g)
        np.testing.assert_array_equal(sk, rsk)
        np.testing.assert_array_equal(pm, rpm)
        np.testing.assert_array_equal(cs, rcs)
    # no particles: every cell starts at 0
    sk, pm, cs = run_grid(np.zeros((0, 3), np.float32), lo, 0.04, (25, 25, 25))
    assert len(cs) == 25 ** 3 + 1 and not cs.any()


def test_grid_build_errors():
    cs = DeviceArray(16 * 4)
    L = lib()
    for g, h in (((0, 2, 2), 0.1), ((2, 2, 2), 0.0), ((2, 2, 2), float("nan"))):
        rc = L.sf_grid_build(None, 0, sf_float3.of((0, 0, 0)), h, sf_int3(*g), None, None, cs.ptr, None)
        assert rc == _abi.SF_E_INVALID
    rc = L.sf_grid_build(None, 5, sf_float3.of((0, 0, 0)), 0.1, sf_int3(2, 2, 2), None, None, cs.ptr, None)
    assert rc == _abi.SF_E_INVALID
