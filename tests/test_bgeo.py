"""bgeo V5 writer: byte-exact against a buffer assembled here from the layout of
Partio::writeBGEO (extlib/include/partio/io/BGEO.cpp:311-476) with the
attributes ExportBgeoFileCUDA registers (kiri_utils.h:390-394)."""
import struct

import numpy as np

from sph_seepage_flows_amd import write_bgeo


def expected_bytes(pos, col, rad, lab):
    n = len(pos)
    b = bytearray()
    b += struct.pack(">i", (ord("B") << 24) | (ord("g") << 16) | (ord("e") << 8) | ord("o"))
    b += b"V"
    b += struct.pack(">i", 5)            # version
    b += struct.pack(">iii", n, 0, 0)    # nPoints, nPrims, nPointGroups
    b += struct.pack(">iiiii", 0, 3, 0, 0, 0)  # nPrimGroups, nPointAttrib, nVertexAttrib, nPrimAttrib, nAttrib
    for name, count, typ in (("Cd", 3, 0), ("pscale", 1, 0), ("label", 1, 1)):
        b += struct.pack(">h", len(name)) + name.encode()
        b += struct.pack(">Hi", count, typ) + struct.pack(">" + "i" * count, *([0] * count))
    for i in range(n):
        b += struct.pack(">ffff", *pos[i], 1.0)
        b += struct.pack(">fff", *col[i])
        b += struct.pack(">f", rad[i])
        b += struct.pack(">i", int(lab[i]))
    b += b"\x00\xff"
    return bytes(b)


def test_bgeo_bytes(tmp_path):
    rng = np.random.default_rng(0)
    n = 37
    pos = rng.random((n, 3), dtype=np.float32)
    col = rng.random((n, 3), dtype=np.float32)
    rad = rng.random(n, dtype=np.float32) * np.float32(0.01)
    lab = rng.integers(0, 2, n).astype(np.uint64)
    path = tmp_path / "0001.bgeo"
    write_bgeo(str(path), pos, col, rad, lab)
    got = path.read_bytes()
    assert got == expected_bytes(pos, col, rad, lab)
    assert got[:5] == b"BgeoV"
    assert len(got) == 5 + 4 * 9 + (2 + 2 + 2 + 4 + 12) + (2 + 6 + 2 + 4 + 4) + (2 + 5 + 2 + 4 + 4) + 36 * n + 2


def test_bgeo_empty(tmp_path):
    path = tmp_path / "empty.bgeo"
    z = np.zeros((0, 3), np.float32)
    write_bgeo(str(path), z, z, np.zeros(0, np.float32), np.zeros(0, np.uint64))
    assert path.read_bytes() == expected_bytes(z, z, np.zeros(0), np.zeros(0))
