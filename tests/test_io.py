"""elsa's MHD and EDF files and the geometry file (elsa/io/tests/test_EDFHandler.cpp,
test_MHDHandler.cpp, test_IO.cpp)."""
import os

import numpy as np
import pytest

from elsa_amd import io


def test_edf_header_and_round_trip(tmp_path):
    """test_EDFHandler.cpp:29-70 (1D) and :74-115 (2D): the header lines and a round trip."""
    for shape, lines in (((5,), ["Dim_1 = 5;", "Size = 20;", "Spacing = 1;"]),
                         ((10, 5), ["Dim_1 = 5;", "Dim_2 = 10;", "Size = 200;", "Spacing = 1 1;"])):
        data = np.arange(np.prod(shape), dtype=np.float32).reshape(shape)
        text = io.EDF.header_text(data)
        parts = [p for p in text[1:].split("\n") if p.strip() and p.strip() != "}"]
        assert parts[:4] == ["HeaderID = EH:000001:000000:000000;", "Image = 1;", "ByteOrder = LowByteFirst;",
                             "DataType = FloatValue;"]
        assert parts[4:4 + len(lines)] == lines
        assert len(text) % 1024 == 0
        p = tmp_path / "a.edf"
        io.EDF.write(data, str(p))
        back, sp = io.EDF.read(str(p))
        np.testing.assert_array_equal(back, data)
        assert list(sp) == [1.0] * len(shape)


def test_mhd_round_trip_and_header(tmp_path):
    data = np.random.default_rng(0).random((4, 6, 5)).astype(np.float32)
    meta = tmp_path / "v.mhd"
    io.MHD.write(data, str(meta), "v.raw", spacing=[1.0, 2.0, 0.5])
    lines = meta.read_text().splitlines()
    assert lines == ["NDims = 3", "DimSize = 5 6 4", "ElementSpacing = 1 2 0.5", "ElementType = MET_FLOAT",
                     "ElementByteOrderMSB = False", "ElementDataFile = v.raw"]
    back, sp = io.MHD.read(str(meta))
    np.testing.assert_array_equal(back, data)
    np.testing.assert_array_equal(sp, [1.0, 2.0, 0.5])
    u16 = (np.arange(30, dtype=np.uint16) * 7).reshape(5, 6)
    io.write(u16, str(tmp_path / "u.mhd"))
    np.testing.assert_array_equal(io.read(str(tmp_path / "u.mhd"))[0], u16)


def test_readers_reject_like_elsa(tmp_path):
    """test_IO.cpp:11-40 and the header checks of the two handlers."""
    with pytest.raises(ValueError):
        io.read("hello.something")
    with pytest.raises(IOError):
        io.read(str(tmp_path / "hello.edf"))
    (tmp_path / "b.mhd").write_text("NDims = 2\nDimSize = 2 2\nElementType = MET_FLOAT\n"
                                    "ElementByteOrderMSB = True\nElementDataFile = b.raw\n")
    with pytest.raises(ValueError, match="little endian"):
        io.MHD.read(str(tmp_path / "b.mhd"))
    (tmp_path / "c.mhd").write_text("NDims = 2\nDimSize = 2 2 2\nElementType = MET_FLOAT\nElementDataFile = c.raw\n")
    with pytest.raises(ValueError, match="dimension size mismatch"):
        io.MHD.read(str(tmp_path / "c.mhd"))
    i8 = np.zeros((2, 2), np.int8)
    io.EDF.write(i8, str(tmp_path / "i8.edf"))  # elsa writes int8 but reads only u16/f32/f64
    with pytest.raises(ValueError):
        io.EDF.read(str(tmp_path / "i8.edf"))
    bad = io.EDF.header_text(np.zeros(4, np.float32)).replace("Size = 16;", "Size = 12;")
    (tmp_path / "s.edf").write_bytes(bad.encode() + np.zeros(4, np.float32).tobytes())
    with pytest.raises(ValueError, match="size inconsistency"):
        io.EDF.read(str(tmp_path / "s.edf"))


def test_geometry_file_round_trip(tmp_path):
    import oracle

    g = oracle.circle_geometry(3, (16, 12, 10), 7, 360, 800.0, 40.0, det=(20, 18))
    path = str(tmp_path / "traj.geom")
    io.write_geometry(path, g.vol, g.det, (g.P, g.Pinv, g.C, g.ext, g.sdd))
    vol, det, arrays = io.read_geometry(path)
    assert tuple(vol) == g.vol and tuple(det) == g.det
    for a, b in zip(arrays, (g.P, g.Pinv, g.C, g.ext, g.sdd)):
        np.testing.assert_array_equal(a, np.asarray(b, np.float32).reshape(a.shape))
    with open(path, "r+b") as f:
        f.write(b"X")
    with pytest.raises(ValueError):
        io.read_geometry(path)
