"""GPU tests of the C-ABI surface: bgeo export of device state, device buffer
access, parameter changes, capacity / label errors, the profiled step."""
import ctypes as C
import struct

import numpy as np
import pytest

import sph_seepage_flows_amd as sf
from sph_seepage_flows_amd import SFError, SFSystem, _abi, scenes

pytestmark = pytest.mark.gpu


def test_export_bgeo_roundtrip(tmp_path):
    sc = scenes.mixed()
    g = SFSystem.from_scene(sc)
    g.step(1e-4, 3)
    path = g.export_bgeo(str(tmp_path / "frames"), "0001")
    data = open(path, "rb").read()
    n = sc.n
    assert data[:5] == b"BgeoV" and struct.unpack(">i", data[9:13])[0] == n
    rec = np.frombuffer(data[98:98 + 36 * n], dtype=">f4").reshape(n, 9)  # after the 41-byte header + 57 bytes of attribute headers
    np.testing.assert_array_equal(rec[:, :3].astype(np.float32), g.get("pos"))
    np.testing.assert_array_equal(rec[:, 4:7].astype(np.float32), g.get("col"))
    np.testing.assert_array_equal(rec[:, 7].astype(np.float32), g.get("radius"))
    lab = np.frombuffer(data[98:98 + 36 * n], dtype=">i4").reshape(n, 9)[:, 8]
    np.testing.assert_array_equal(lab.astype(np.uint64), g.get("label"))
    assert data[-2:] == b"\x00\xff"


def test_device_buffers_and_sizes():
    sc = scenes.mixed()
    g = SFSystem.from_scene(sc)
    g.step(1e-4)
    n, nb, cells = g.sizes()
    assert (n, nb) == (sc.n, sc.nb) and cells == 12 ** 3
    for name, stride in (("pos", 16), ("density", 8), ("cell_start", 4), ("acc", 16), ("b_pos", 16)):
        p, s = g.device_buffer(name)
        assert p != 0 and s == stride


def test_set_params_changes_gravity():
    sc = scenes.mixed(collider=False)
    g = SFSystem.from_scene(sc)
    p, b, o = sc.params()
    p.gravity = _abi.sf_float3.of((0.0, 0.0, -3.0))
    g.set_params(p)
    g.step(1e-4)
    # an isolated far-away check is not available in this scene; the mean
    # acceleration of all particles must now point along -z, not -y
    a = g.get("acc").mean(0)
    assert abs(a[2] + 3.0) < abs(a[1] + 9.81)


def test_errors_are_reported_not_crashing():
    sc = scenes.mixed()
    p, b, o = sc.params()
    g = SFSystem(p, b, o, max_num=10)
    with pytest.raises(SFError) as e:
        g.upload_particles(sc)
    assert e.value.code == _abi.SF_E_CAPACITY
    bad = scenes.mixed()
    bad.label = bad.label.copy()
    bad.label[3] = 2
    g2 = SFSystem(p, b, o, max_num=bad.n)
    with pytest.raises(SFError) as e:
        g2.upload_particles(bad)
    assert e.value.code == _abi.SF_E_INVALID and "label" in str(e.value)
    with pytest.raises(SFError):
        g2.step(1e-4)  # nothing uploaded yet
    g3 = SFSystem.from_scene(sc)
    with pytest.raises(SFError):
        g3.get("sort_perm")  # the searcher has not run
    with pytest.raises(SFError):
        g3.step(float("nan"))


def test_profiled_step_matches_plain_step():
    sc = scenes.small_block()
    a = SFSystem.from_scene(sc)
    b = SFSystem.from_scene(sc)
    a.step(1e-4, 5)
    tot, passes = b.step_profiled(1e-4, 5)
    np.testing.assert_array_equal(a.get("pos"), b.get("pos"))
    assert tot > 0 and set(passes) == set(_abi.PASSES) and all(v >= 0 for v in passes.values())
    c = SFSystem.from_scene(sc)
    tot_c, ms = c.step_timed(1e-4, 5, "water_momentum")
    np.testing.assert_array_equal(a.get("pos"), c.get("pos"))
    assert 0 < ms < tot_c
    for name in ("hash", "advect"):  # first and last pass
        d = SFSystem.from_scene(sc)
        assert d.step_timed(1e-4, 2, name)[1] > 0
    assert _abi.lib().sf_step_timed(c._h, 1e-4, 1, len(_abi.PASSES), None, None) == _abi.SF_E_INVALID
    # events on every 3rd of 7 steps only (steps 0, 3, 6): same trajectory
    e = SFSystem.from_scene(sc)
    tot_e, ms_e = e.step_timed(1e-4, 7, "water_momentum", every=3)
    a.step(1e-4, 2)
    np.testing.assert_array_equal(a.get("pos"), e.get("pos"))
    assert 0 < ms_e < tot_e
    for every in (0, 8):
        assert _abi.lib().sf_step_timed_every(e._h, 1e-4, 7, 0, every, None, None) == _abi.SF_E_INVALID


def test_graph_after_an_eager_radix_step_matches_plain_steps():
    """One plain step (the first build: a radix sort), then graph replays: the
    replays must read the previous keys from the buffers the incremental sort
    rewrites, not from the radix sort's."""
    sc = scenes.small_block()
    a = SFSystem.from_scene(sc)
    b = SFSystem.from_scene(sc)
    a.step(1e-4, 8)
    b.step(1e-4, 1)
    b.step_graph(1e-4, 7)
    for f in ("pos", "vel", "col", "sort_perm", "sorted_keys"):
        np.testing.assert_array_equal(a.get(f), b.get(f), err_msg=f)
    # an odd count and a second call reuse the graph
    a.step(1e-4, 5)
    b.step_graph(1e-4, 5)
    np.testing.assert_array_equal(a.get("pos"), b.get("pos"))


def test_positions_written_between_steps_are_hashed_again():
    """A step's advect writes the next build's cell keys; positions written
    through the device buffer afterwards must be hashed again, also by a
    captured graph.  Reference: the same run with the radix sort."""
    sc = scenes.small_block()
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemcpy.restype = C.c_int
    runs = {}
    for mode in ("incremental", "incremental_graph", "radix"):
        g = SFSystem.from_scene(sc)
        if mode == "radix":
            g.set_sort_strategy(False)
        run = g.step_graph if mode == "incremental_graph" else g.step
        g.step(1e-4, 3)
        run(1e-4, 4)
        pos, mass, lab = g.get("pos"), g.get("mass"), g.get("label")
        pm = np.zeros((sc.n, 4), np.float32)
        pm[:, :3] = pos
        pm[:, 0] += np.float32(0.05)  # more than a cell (h = 0.04): every key changes
        pm[:, 3] = np.where(lab == 1, -mass, mass)  # the label is the sign of the mass
        ptr, stride = g.device_buffer("pos")
        assert stride == 16
        assert hip.hipMemcpy(ptr, pm.ctypes.data, pm.nbytes, 1) == 0  # host to device
        np.testing.assert_array_equal(g.get("pos"), pm[:, :3])
        run(1e-4, 4)
        runs[mode] = {f: g.get(f) for f in ("pos", "vel", "sort_perm", "sorted_keys", "cell_start")}
    for mode in ("incremental", "incremental_graph"):
        for f, v in runs["radix"].items():
            np.testing.assert_array_equal(runs[mode][f], v, err_msg=f"{mode} {f}")


def test_caller_streams_give_the_plain_results():
    """Steps queued on caller streams (sf_set_stream), switched between calls
    and back to the handle's own stream (NULL), plain and graph-replayed, give
    the plain handle's states bit for bit."""
    from _hip import runtime
    rt = runtime()
    streams = [C.c_void_p(), C.c_void_p()]
    for s in streams:
        assert rt.hipStreamCreate(C.byref(s)) == 0
    lib = _abi.lib()
    try:
        sc = scenes.mixed()
        a = SFSystem.from_scene(sc)
        b = SFSystem.from_scene(sc)
        a.step(1e-4, 12)
        for s in (streams[0], streams[1], None):
            assert lib.sf_set_stream(b._h, s) == 0
            b.step(1e-4, 2)
            b.step_graph(1e-4, 2)
        assert lib.sf_synchronize(b._h) == 0
        for f in ("pos", "vel", "density", "saturation", "sort_perm"):
            np.testing.assert_array_equal(a.get(f), b.get(f), err_msg=f)
    finally:
        for s in streams:
            rt.hipStreamSynchronize(s)
            rt.hipStreamDestroy(s)
