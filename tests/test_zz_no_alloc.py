"""sf_grid_build_noalloc: the grid build with its work area in a
library-owned device array instead of an allocation.  Its outputs equal
sf_grid_build's bit for bit on two calls in a row, the second with a smaller
input (the array keeps the first call's contents), at the 1M-particle size
and on small inputs; inputs above its maximum are refused."""
import numpy as np
import pytest

from sph_seepage_flows_amd import _abi, scenes
from sph_seepage_flows_amd._abi import check, lib, sf_float3, sf_int3

from _hip import DeviceArray, runtime

pytestmark = pytest.mark.gpu

MAX_N = 1 << 20


def build(fn, pos, lo, h, g):
    """one grid build through the C-ABI on device arrays of the library's runtime"""
    n = len(pos)
    C_ = int(g[0]) * int(g[1]) * int(g[2])
    p = DeviceArray.of(np.ascontiguousarray(pos, np.float32))
    sk, pm = DeviceArray.of(np.full(max(n, 1), 0xFFFFFFFF, np.uint32)), DeviceArray.of(np.full(max(n, 1), 0xFFFFFFFF, np.uint32))
    cs = DeviceArray.of(np.full(C_ + 1, 0xFFFFFFFF, np.uint32))
    check(getattr(lib(), fn)(p.ptr if n else None, n, sf_float3.of(lo), float(np.float32(h)),
                             sf_int3(*[int(v) for v in g]), sk.ptr if n else None, pm.ptr if n else None, cs.ptr,
                             None))
    assert runtime().hipDeviceSynchronize() == 0
    return sk.get(n, np.uint32), pm.get(n, np.uint32), cs.get(C_ + 1, np.uint32)


def test_noalloc_equals_allocating_build_on_two_calls():
    big = scenes.drop_in_1m()
    small = scenes.small_block()
    pb, bb, _ = big.params()
    ps, bs, _ = small.params()
    calls = [("drop_in_1m", big.pos, big.physics.lo, bb), ("small_block", small.pos, small.physics.lo, bs),
             ("small_block_part", small.pos[:777], small.physics.lo, bs)]
    for name, pos, lo, b in calls:
        g = (b.grid_size.x, b.grid_size.y, b.grid_size.z)
        print(f"[noalloc] {name}: n={len(pos)}", flush=True)
        ref = build("sf_grid_build", pos, np.float32(lo), b.kernel_radius, g)
        got = build("sf_grid_build_noalloc", pos, np.float32(lo), b.kernel_radius, g)
        for a, r, what in zip(got, ref, ("sorted keys", "permutation", "cellStart")):
            np.testing.assert_array_equal(a, r, err_msg=f"{name}: {what}")


def test_noalloc_refuses_more_than_its_maximum():
    cs = DeviceArray(9 * 4)
    x = DeviceArray(3 * 4 * (MAX_N + 1))
    k = DeviceArray(4 * (MAX_N + 1))
    rc = lib().sf_grid_build_noalloc(x.ptr, MAX_N + 1, sf_float3.of((0, 0, 0)), 0.5, sf_int3(2, 2, 2), k.ptr, k.ptr,
                                     cs.ptr, None)
    assert rc == _abi.SF_E_CAPACITY
    assert "1048576" in lib().sf_last_error().decode()
