"""Scene input on the host (no GPU): the bgeo V5 point reader behind
ReadBgeoFileForGPU (kiri_utils.h:252-308, over Partio::readBGEO,
extlib/include/partio/io/BGEO.cpp:196-309) and the sand shape builders
BuildSeepageflowShape(Multi)Volume (cuda_volume_emitter.cpp:109-175).

The reader is checked against files assembled here byte by byte from the
classic bgeo layout with known contents (every attribute kind partio reads
past, a particle-system primitive and detail attributes after the points,
gzip), against the library's own writer, and on each error partio refuses.
The builders are checked against what the mathematics fixes: the mass of a
sphere of the given density, the floor offset, the smallest radius, order
and labels."""
import gzip
import struct

import numpy as np
import pytest

from sph_seepage_flows_amd import SFError, _abi, scene_io, scenes, write_bgeo

DRY = (0.88, 0.79552, 0.5984)


def bgeo_bytes(pos, attrs, version=5, magic=b"Bgeo", prims=0, detail=()):
    """attrs: list of (name, houdini type, size, per-point values (n, size) or
    None, extra) -- type 4 (indexed string) takes its string table as extra."""
    n = len(pos)
    b = bytearray(magic + b"V")
    b += struct.pack(">iiiii", version, n, prims, 0, 0)
    b += struct.pack(">iiii", len(attrs), 0, 0, len(detail))
    for name, typ, size, _, extra in attrs:
        b += struct.pack(">H", len(name)) + name.encode() + struct.pack(">Hi", size, typ)
        if typ == 4:
            b += struct.pack(">i", len(extra))
            for s in extra:
                b += struct.pack(">H", len(s)) + s.encode()
        else:
            b += b"\x7f\x7f\x7f\x7f" * size  # default values (ignored)
    for i in range(n):
        b += struct.pack(">ffff", *pos[i], 1.0)
        for name, typ, size, vals, _ in attrs:
            fmt = ">" + ("i" if typ in (1, 4) else "f") * size
            b += struct.pack(fmt, *vals[i])
    for _ in range(prims):  # particle-system primitive: point count + 16-bit indices
        b += struct.pack(">ii", 0x8000, n) + struct.pack(">" + "H" * n, *range(n))
    for name, typ, size, vals, _ in detail:
        b += struct.pack(">H", len(name)) + name.encode() + struct.pack(">Hi", size, typ) + b"\0" * 4 * size
    for name, typ, size, vals, _ in detail:
        b += struct.pack(">" + "f" * size, *vals)
    b += b"\x00\xff"
    return bytes(b)


def _points(n, seed=0):
    rng = np.random.default_rng(seed)
    return (rng.random((n, 3)) * 2 - 0.5).astype(np.float32)


def test_reader_known_file_every_attribute_kind(tmp_path):
    n = 23
    pos = _points(n)
    rng = np.random.default_rng(1)
    ps = (rng.random((n, 1)) * 0.01 + 0.001).astype(np.float32)
    attrs = [("Cd", 5, 3, rng.random((n, 3)).astype(np.float32), None),
             ("id", 1, 1, np.arange(n)[:, None] * 7 - 3, None),
             ("name", 4, 1, (np.arange(n) % 2)[:, None], ["sand", "rock"]),
             ("pscale", 0, 1, ps, None),
             ("v", 0, 3, rng.random((n, 3)).astype(np.float32), None),
             ("pscale", 0, 1, ps * 3, None)]  # a second "pscale": the first one is the one found
    f = tmp_path / "shape.bgeo"
    f.write_bytes(bgeo_bytes(pos, attrs, prims=1, detail=[("time", 0, 1, [2.5], None)]))
    got = scene_io.read_bgeo(f)
    assert got.shape == (n, 4) and got.dtype == np.float32
    np.testing.assert_array_equal(got[:, :3], pos)
    np.testing.assert_array_equal(got[:, 3], ps[:, 0])
    flip = scene_io.read_bgeo(f, flip_yz=True)
    np.testing.assert_array_equal(flip[:, 0], pos[:, 0])
    np.testing.assert_array_equal(flip[:, 1], pos[:, 2])
    np.testing.assert_array_equal(flip[:, 2], pos[:, 1])
    np.testing.assert_array_equal(flip[:, 3], ps[:, 0])
    # gzip-compressed copy reads the same
    g = tmp_path / "shape_gz.bgeo"
    g.write_bytes(gzip.compress(f.read_bytes()))
    np.testing.assert_array_equal(scene_io.read_bgeo(g), got)


def test_reader_without_pscale_uses_0_01(tmp_path):
    n = 9
    pos = _points(n, 3)
    f = tmp_path / "p.bgeo"
    f.write_bytes(bgeo_bytes(pos, [("Cd", 5, 3, np.ones((n, 3), np.float32), None)]))
    got = scene_io.read_bgeo(f)
    np.testing.assert_array_equal(got[:, :3], pos)
    assert (got[:, 3] == np.float32(0.01)).all()


def test_reader_reads_the_library_writer(tmp_path):
    n = 100
    pos = _points(n, 5)
    rad = np.linspace(0.002, 0.006, n).astype(np.float32)
    f = tmp_path / "w.bgeo"
    write_bgeo(str(f), pos, np.zeros((n, 3), np.float32), rad, np.ones(n, np.uint64))
    got = scene_io.read_bgeo(f)
    np.testing.assert_array_equal(got[:, :3], pos)
    np.testing.assert_array_equal(got[:, 3], rad)
    # ReadMultiBgeoFilesForGPU concatenates in order
    both = scene_io.read_multi_bgeo([f, f], flip_yz=True)
    assert both.shape == (2 * n, 4)
    np.testing.assert_array_equal(both[n:, 1], pos[:, 2])


def test_reader_empty_and_count_only(tmp_path):
    f = tmp_path / "e.bgeo"
    f.write_bytes(bgeo_bytes(np.zeros((0, 3), np.float32), []))
    assert scene_io.read_bgeo(f).shape == (0, 4)
    g = tmp_path / "c.bgeo"
    g.write_bytes(bgeo_bytes(_points(4), []))
    import ctypes as C
    cnt = C.c_size_t(0)
    assert _abi.lib().sf_read_bgeo(str(g).encode(), 0, None, 0, cnt) == _abi.SF_OK and cnt.value == 4
    out = np.zeros((3, 4), np.float32)
    assert _abi.lib().sf_read_bgeo(str(g).encode(), 0, out.ctypes.data, 3, cnt) == _abi.SF_E_CAPACITY


@pytest.mark.parametrize("case", ["json", "magic", "version", "string_attr", "unknown_attr", "truncated",
                                  "truncated_table", "missing"])
def test_reader_errors(tmp_path, case):
    pos = _points(5)
    f = tmp_path / "bad.bgeo"
    if case == "json":
        f.write_bytes(b"\x7fNSJ" + b"\0" * 64)
    elif case == "magic":
        f.write_bytes(bgeo_bytes(pos, [], magic=b"Bgeq"))
    elif case == "version":
        f.write_bytes(bgeo_bytes(pos, [], version=4))
    elif case == "string_attr":
        f.write_bytes(bgeo_bytes(pos, [("s", 2, 1, np.zeros((5, 1)), None)]))
    elif case == "unknown_attr":
        f.write_bytes(bgeo_bytes(pos, [("s", 9, 1, np.zeros((5, 1)), None)]))
    elif case == "truncated":
        f.write_bytes(bgeo_bytes(pos, [])[:41 + 16 * 4 + 3])
    elif case == "truncated_table":
        f.write_bytes(bgeo_bytes(pos, [("pscale", 0, 1, np.ones((5, 1), np.float32), None)])[:45])
    else:
        f = tmp_path / "absent.bgeo"
    with pytest.raises(SFError) as e:
        scene_io.read_bgeo(f)
    assert e.value.code == _abi.SF_E_IO


def _shape(n, seed, r_lo=0.003, r_hi=0.006, y0=0.4):
    rng = np.random.default_rng(seed)
    s = np.zeros((n, 4), np.float32)
    s[:, :3] = (rng.random((n, 3)) * 0.3 + np.array([0.5, y0, 0.5])).astype(np.float32)
    s[:, 3] = (rng.random(n) * (r_hi - r_lo) + r_lo).astype(np.float32)
    return s


def _ulp_close(a, b64, ulps):
    a = np.asarray(a, np.float32)
    return np.abs(a.astype(np.float64) - b64) <= ulps * np.spacing(np.abs(a)).astype(np.float64)


def test_shape_volume_mass_labels_and_min_radius():
    s = _shape(200, 0)
    v = scene_io.VolumeData()
    v.build_shape_volume(s, DRY, 2700.0)
    assert v.n == 200 and (v.label == 1).all()
    np.testing.assert_array_equal(v.pos, s[:, :3])  # no offset: positions as read
    np.testing.assert_array_equal(v.radius, s[:, 3])
    np.testing.assert_array_equal(v.col, np.tile(np.array(DRY, np.float32), (200, 1)))
    # mass of a sphere: rho (4/3) pi r^3, up to FP32 rounding of the product chain
    m = 2700.0 * 4.0 / 3.0 * np.pi * s[:, 3].astype(np.float64) ** 3
    assert _ulp_close(v.mass, m, 4).all()
    assert v.sand_min_radius == s[:, 3].min()
    # the single-material builder restarts the smallest radius on each call
    v.build_shape_volume(_shape(10, 1, r_lo=0.007, r_hi=0.008), DRY, 2700.0)
    assert v.n == 210 and v.sand_min_radius >= np.float32(0.007)


def test_multi_volume_offset_to_the_floor():
    a, b = _shape(150, 2, y0=0.6), _shape(80, 3, r_lo=0.002, r_hi=0.005, y0=0.9)
    v = scene_io.VolumeData()
    v.build_shape_multi_volume(a, DRY, 2700.0, (0.5, 0.0, 0.8), (1.5, 0.5), offset_y=True, world_lowest_y=0.1)
    R = a[:, 3].min()
    assert v.sand_min_radius == R
    # x, z untouched; the lowest point of the shape now sits 5 R above the floor
    np.testing.assert_array_equal(v.pos[:, [0, 2]], a[:, [0, 2]])
    low = v.pos[np.argmin(a[:, 1]), 1]
    assert abs(float(low) - (0.1 + 5 * float(R))) < 1e-6
    # y shift is one constant for every particle
    shift = a[:, 1].astype(np.float64) - v.pos[:, 1]
    assert np.ptp(shift) < 2e-7
    np.testing.assert_array_equal(v.cda0asat, np.tile(np.array([0.5, 0.0, 0.8], np.float32), (150, 1)))
    np.testing.assert_array_equal(v.amcamcp, np.tile(np.array([1.5, 0.5], np.float32), (150, 1)))
    first = v.pos.copy()
    # second object: the smallest radius carries over, and every particle held
    # (the first object too) moves by the second shape's offset
    v.build_shape_multi_volume(b, DRY, 2700.0, (0.015, 0.0, 0.8), (1.5, 0.5), offset_y=True, world_lowest_y=0.1)
    R2 = min(R, b[:, 3].min())
    assert v.n == 230 and v.sand_min_radius == R2
    off = np.float32(b[:, 1].min()) - (np.float32(0.1) + np.float32(R2) * np.float32(5.0))
    np.testing.assert_array_equal(v.pos[:150, 1], first[:, 1] - off)
    np.testing.assert_array_equal(v.pos[150:, 1], b[:, 1] - off)
    assert abs(float(v.pos[150:, 1].min()) - (0.1 + 5 * float(R2))) < 1e-6
    np.testing.assert_array_equal(v.cda0asat[150:, 0], np.float32(0.015))


def test_volume_errors():
    import ctypes as C
    d = _abi.sf_volume_data()
    d.size, d.capacity = 0, 3
    s = _shape(4, 0)
    L = _abi.lib()
    rc = L.sf_build_seepage_shape_volume(C.byref(d), s.ctypes.data, 4, _abi.sf_float3.of(DRY), 2700.0, 0, 0.0)
    assert rc == _abi.SF_E_CAPACITY
    d.capacity = 4
    assert L.sf_build_seepage_shape_volume(C.byref(d), s.ctypes.data, 4, _abi.sf_float3.of(DRY), 2700.0, 0,
                                           0.0) == _abi.SF_E_INVALID  # null arrays
    v = scene_io.VolumeData()
    v.build_shape_volume(s, DRY, 2700.0)
    with pytest.raises(ValueError):
        v.build_shape_multi_volume(s, DRY, 2700.0, (0.5, 0, 0.8), (1.5, 0.5))


def test_shape_scene_follows_the_reference_driver(tmp_path):
    s = _shape(300, 4, r_lo=0.004, r_hi=0.006)
    f = tmp_path / "bgeo" / "bunny" / "bunny.bgeo"
    f.parent.mkdir(parents=True)
    write_bgeo(str(f), s[:, :3], np.zeros((300, 3), np.float32), s[:, 3], np.ones(300, np.uint64))
    assert scene_io.pbr_path(str(tmp_path) + "/", "bunny", "bunny") == str(f)
    v = scene_io.VolumeData()
    v.build_shape_multi_volume(scene_io.read_bgeo(f), DRY, 2700.0, (0.5, 0.0, 0.8), (1.5, 0.5), True, 0.0)
    sc = scenes.shape_scene(v)
    p, b, o = sc.params()
    R = np.float32(s[:, 3].min())
    assert p.dem_particle_radius == R and p.sf_type == _abi.SF_TYPE_MULTI_SF
    # dt = 0.5 R / sqrt(E / rho_s): the Hertz-like contact time scale of main.cpp:172
    assert abs(p.dt - 0.5 * float(R) / np.sqrt(1e5 / 2700.0)) < 1e-6 * p.dt
    assert (b.grid_size.x, b.grid_size.y, b.grid_size.z) == (37, 37, 62)
    assert o.speed_of_sound == 100.0 and p.gravity.y == np.float32(-9.8)
    e = sc.meta["emitter"]
    np.testing.assert_allclose(e["pos"], (0.75 + 0.15, 0.75 + 0.3, 1.25 + 2.5 / 2.3), rtol=1e-6)
    assert sc.n == 300 and sc.nb > 0 and (sc.label == 1).all()


def parse_vtk(data: bytes):
    """Independent reader of the binary legacy VTK layout (big-endian)."""
    out, at = {}, 0

    def line():
        nonlocal at
        e = data.index(b"\n", at)
        s = data[at:e].decode()
        at = e + 1
        return s

    assert line() == "# vtk DataFile Version 4.1"
    out["title"] = line()
    assert line() == "BINARY" and line() == "DATASET UNSTRUCTURED_GRID"
    h = line().split()
    assert h[0] == "POINTS" and h[2] == "float"
    n = int(h[1])
    out["points"] = np.frombuffer(data[at:at + 12 * n], ">f4").reshape(n, 3)
    at += 12 * n + 1
    h = line().split()
    assert h[0] == "CELLS" and int(h[1]) == n and int(h[2]) == 2 * n
    out["cells"] = np.frombuffer(data[at:at + 8 * n], ">i4").reshape(n, 2)
    at += 8 * n + 1
    assert line() == f"CELL_TYPES {n}"
    out["types"] = np.frombuffer(data[at:at + 4 * n], ">i4")
    at += 4 * n + 1
    assert line() == f"POINT_DATA {n}"
    h = line().split()
    assert h[:2] == ["FIELD", "FieldData"]
    fields = {}
    for _ in range(int(h[2])):
        name, comps, cnt, typ = line().split()
        assert int(cnt) == n
        k = int(comps) * n
        fields[name] = np.frombuffer(data[at:at + 4 * k], ">f4" if typ == "float" else ">i4").reshape(n, int(comps))
        at += 4 * k + 1
    assert at == len(data)
    out["fields"] = fields
    return out


def test_bgeo_to_vtk_export_layout(tmp_path):
    """A frame as the exporter writes it (Cd, pscale, label): points, vertex
    cells over the particle index, and the three attributes with all their
    components."""
    n = 17
    pos = _points(n, 8)
    rng = np.random.default_rng(9)
    col = rng.random((n, 3)).astype(np.float32)
    rad = (rng.random(n) * 0.01).astype(np.float32)
    lab = rng.integers(0, 2, n).astype(np.uint64)
    src = tmp_path / "0001.bgeo"
    write_bgeo(str(src), pos, col, rad, lab)
    dst = tmp_path / "0001.vtk"
    scene_io.bgeo_to_vtk(src, dst)
    v = parse_vtk(dst.read_bytes())
    assert v["title"] == "Seepage Flow Framework Particles Data"
    np.testing.assert_array_equal(v["points"], pos)
    np.testing.assert_array_equal(v["cells"], np.stack([np.ones(n), np.arange(n)], 1))
    assert (v["types"] == 1).all()
    assert list(v["fields"]) == ["Cd", "pscale", "label"]
    np.testing.assert_array_equal(v["fields"]["Cd"], col)
    np.testing.assert_array_equal(v["fields"]["pscale"][:, 0], rad)
    np.testing.assert_array_equal(v["fields"]["label"][:, 0], lab.astype(np.int32))


def test_bgeo_to_vtk_ids_vectors_and_skipped_kinds(tmp_path):
    n = 6
    pos = _points(n, 10)
    ids = (np.arange(n) * 3 + 100)[:, None]
    vel = np.arange(3 * n, dtype=np.float32).reshape(n, 3)
    attrs = [("id", 1, 1, ids, None), ("v", 5, 3, vel, None), ("name", 4, 1, np.zeros((n, 1), int), ["a"]),
             ("my attr", 0, 1, np.full((n, 1), 2.5, np.float32), None)]
    src = tmp_path / "s.bgeo"
    src.write_bytes(bgeo_bytes(pos, attrs))
    dst = tmp_path / "s.vtk"
    scene_io.bgeo_to_vtk(src, dst)
    v = parse_vtk(dst.read_bytes())
    np.testing.assert_array_equal(v["cells"][:, 1], ids[:, 0])  # cells point at the "id" values
    assert list(v["fields"]) == ["v", "my_attr"]  # no id, no indexed string; whitespace -> '_'
    np.testing.assert_array_equal(v["fields"]["v"], vel)
    assert (v["fields"]["my_attr"] == np.float32(2.5)).all()
    # an empty file writes nothing
    e = tmp_path / "e.bgeo"
    e.write_bytes(bgeo_bytes(np.zeros((0, 3), np.float32), []))
    scene_io.bgeo_to_vtk(e, tmp_path / "e.vtk")
    assert not (tmp_path / "e.vtk").exists()
