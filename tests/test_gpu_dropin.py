"""The main.cpp-style C++ driver on the KIRI drop-in layer
(sph_seepage_flows_amd/cpp/seepage_main.cpp) against the Python host on the
same library: identical scene construction and bit-identical frames of the
adaptive sub-step loop (main.cpp:235-247)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from sph_seepage_flows_amd import SFSystem, scenes

pytestmark = pytest.mark.gpu
DRIVER = Path(__file__).resolve().parents[1] / "sph_seepage_flows_amd" / "_lib" / "seepage_main"


def read_bgeo(path):
    data = Path(path).read_bytes()
    n = int.from_bytes(data[9:13], "big")
    rec = np.frombuffer(data[98:98 + 36 * n], dtype=">f4").reshape(n, 9)
    lab = np.frombuffer(data[98:98 + 36 * n], dtype=">i4").reshape(n, 9)[:, 8]
    return rec[:, :3].astype(np.float32), rec[:, 4:7].astype(np.float32), rec[:, 7].astype(np.float32), lab


def test_cpp_driver_matches_python_host(tmp_path):
    assert DRIVER.exists(), "build() produces the driver"
    out = tmp_path / "frames"
    r = subprocess.run([str(DRIVER), "--frames", "1", "--out", str(out), "--export-initial"], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    sc = scenes.small_block()
    p0, c0, rad0, l0 = read_bgeo(out / "0000.bgeo")
    np.testing.assert_array_equal(p0, sc.pos)
    np.testing.assert_array_equal(c0, sc.col)
    np.testing.assert_array_equal(rad0, sc.radius)
    np.testing.assert_array_equal(l0.astype(np.uint64), sc.label)
    # same frame through the Python host (float32 remaining-time arithmetic as in main.cpp)
    g = SFSystem.from_scene(sc)
    remaining = np.float32(1.0 / 30.0)
    sub = 0
    while remaining > np.float32(1e-6):
        g.UpdateSystem(float(remaining))
        remaining = np.float32(remaining - remaining / np.float32(g.GetNumOfSubTimeSteps()))
        sub += 1
    p1, c1, rad1, l1 = read_bgeo(out / "0001.bgeo")
    print(f"sub-steps in frame 1: {sub}", flush=True)
    np.testing.assert_array_equal(p1, g.get("pos"))
    np.testing.assert_array_equal(c1, g.get("col"))
    np.testing.assert_array_equal(l1.astype(np.uint64), g.get("label"))


def test_cpp_driver_with_emitter_matches_python_host(tmp_path):
    """main.cpp's inflow emitter (circle r = 0.17, v = (0, 0, -5)) switched on:
    the driver's frame equals the Python host's with the same emitter,
    emitted particles included."""
    from sph_seepage_flows_amd import _abi
    out = tmp_path / "frames"
    r = subprocess.run([str(DRIVER), "--frames", "1", "--out", str(out), "--emit"], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    sc = scenes.small_block()
    g = SFSystem.from_scene(sc, max_num=sc.n + 20000)
    f = np.float32
    g.set_emitter(pos=(f(0.5) + f(1) / f(10), f(0.5) + f(1) / f(5), f(0.5) + f(1) / f(2.3)), vel=(0.0, 0.0, -5.0),
                  col=(f(127) / f(255), f(205) / f(255), f(255) / f(255)), kind=_abi.SF_EMITTER_CIRCLE,
                  radius=0.17)
    remaining = np.float32(1.0 / 30.0)
    while remaining > np.float32(1e-6):
        g.UpdateSystem(float(remaining))
        remaining = np.float32(remaining - remaining / np.float32(g.GetNumOfSubTimeSteps()))
    enabled, k, emitted = g.emitter_state()
    print(f"emitter: {k} samples, {emitted} emitted, n = {g.sizes()[0]}", flush=True)
    assert emitted >= k > 0 and g.sizes()[0] == sc.n + emitted
    p1, c1, rad1, l1 = read_bgeo(out / "0001.bgeo")
    np.testing.assert_array_equal(p1, g.get("pos"))
    np.testing.assert_array_equal(c1, g.get("col"))
    np.testing.assert_array_equal(l1.astype(np.uint64), g.get("label"))


def _ball(centre, radius, spacing):
    k = np.arange(-radius, radius + 1e-9, spacing)
    g = np.stack(np.meshgrid(k, k, k, indexing="ij"), -1).reshape(-1, 3)
    return (g[(g ** 2).sum(1) <= radius * radius] + np.asarray(centre)).astype(np.float32)


def test_cpp_driver_shape_scene_matches_python_host(tmp_path):
    """main.cpp's own scene set-up (shape objects read from bgeo, multi-material
    volume on the floor, circle emitter from the first frame): the driver's
    initial particles and first frame equal the Python host's
    (scene_io + scenes.shape_scene) bit for bit."""
    from sph_seepage_flows_amd import scene_io, write_bgeo
    res = tmp_path / "res"
    # a sand ball in place of the bunny asset, and a block of smaller grains
    ball = _ball((0.75, 0.4, 1.25), 0.12, 0.01)
    i, j, k = np.meshgrid(np.arange(20), np.arange(6), np.arange(20), indexing="ij")
    block = (np.stack([i, j, k], -1).reshape(-1, 3) * np.float32(0.008) + np.float32([0.2, 0.025, 0.3])).astype(
        np.float32)
    objs = [("bunny", "bunny", ball, 0.005, 0.5), ("dam", "dam2", block, 0.004, 0.015)]
    for folder, name, pts, ps, _ in objs:
        path = Path(scene_io.pbr_path(res, folder, name))
        path.parent.mkdir(parents=True)
        write_bgeo(str(path), pts, np.zeros_like(pts), np.full(len(pts), ps, np.float32), np.ones(len(pts), np.uint64))
    out = tmp_path / "frames"
    r = subprocess.run([str(DRIVER), "--frames", "1", "--out", str(out), "--export-initial", "--resources", str(res),
                        "--shape", "bunny/bunny", "--shape", "dam/dam2:0.015"], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    v = scene_io.VolumeData()
    dry = (0.88, 0.79552, 0.5984)
    for folder, name, _, _, cd in objs:
        shape = scene_io.read_bgeo(scene_io.pbr_path(res, folder, name))
        v.build_shape_multi_volume(shape, dry, 2700.0, (cd, 0.0, 0.8), (1.5, 0.5), offset_y=True, world_lowest_y=0.0)
    sc = scenes.shape_scene(v)
    assert sc.n == len(ball) + len(block) and v.sand_min_radius == np.float32(0.004)
    p0, c0, rad0, l0 = read_bgeo(out / "0000.bgeo")
    np.testing.assert_array_equal(p0, sc.pos)
    np.testing.assert_array_equal(c0, sc.col)
    np.testing.assert_array_equal(rad0, sc.radius)
    np.testing.assert_array_equal(l0.astype(np.uint64), sc.label)
    g = SFSystem.from_scene(sc, max_num=sc.meta["max_num"])
    g.set_emitter(**sc.meta["emitter"])
    remaining = np.float32(1.0 / 30.0)
    sub = 0
    while remaining > np.float32(1e-6):
        g.UpdateSystem(float(remaining))
        remaining = np.float32(remaining - remaining / np.float32(g.GetNumOfSubTimeSteps()))
        sub += 1
    enabled, k, emitted = g.emitter_state()
    print(f"shape scene: {sc.n} sand, {sub} sub-steps, {emitted} emitted", flush=True)
    assert emitted > 0
    p1, c1, rad1, l1 = read_bgeo(out / "0001.bgeo")
    np.testing.assert_array_equal(p1, g.get("pos"))
    np.testing.assert_array_equal(c1, g.get("col"))
    np.testing.assert_array_equal(l1.astype(np.uint64), g.get("label"))


def test_cpp_driver_device_frame_loop_is_identical(tmp_path):
    """--device-frame (CudaSFSystem::UpdateFrame, sf_update_frame) writes the
    same frames, byte for byte, as the UpdateSystem loop, emitter on."""
    outs = []
    for extra in ([], ["--device-frame"]):
        out = tmp_path / ("dev" if extra else "host")
        r = subprocess.run([str(DRIVER), "--frames", "2", "--out", str(out), "--emit", *extra], capture_output=True,
                           text=True, timeout=300)
        print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
        assert r.returncode == 0
        outs.append(out)
    for f in ("0001.bgeo", "0002.bgeo"):
        assert (outs[0] / f).read_bytes() == (outs[1] / f).read_bytes(), f


def test_cpp_driver_solver_api_is_identical(tmp_path):
    """--solver-api steps with the reference's OnUpdateSolver sequence
    (searcher->BuildGNSearcher(particles); solver->UpdateSolver(particles,
    boundaries, searcher->GetCellStart(), bsearcher->GetCellStart(), ...),
    cuda_sf_system.cu:66-84) and writes the same frames, byte for byte, as
    the UpdateSystem loop; the frames go out through main.cpp's own export
    call on the particles' Get*Ptr() arrays (main.cpp:264-275)."""
    outs = []
    for extra in ([], ["--solver-api"]):
        out = tmp_path / ("api" if extra else "sys")
        r = subprocess.run([str(DRIVER), "--frames", "2", "--out", str(out), *extra], capture_output=True, text=True,
                           timeout=300)
        print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
        assert r.returncode == 0 and "failed" not in r.stdout
        outs.append(out)
    for f in ("0001.bgeo", "0002.bgeo"):
        assert (outs[0] / f).read_bytes() == (outs[1] / f).read_bytes(), f
