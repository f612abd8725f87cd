"""Grid file formats (.uni, .raw) of the reference's Grid<T>::save / load
(fileio/iogrids.cpp), host side: byte layout of the MNT3 header at the
offsets of the reference's packed struct, round trips, the checks and
error texts of readGridUni, and files written and read by the reference's own
Python .uni module (golden files under tests/golden)."""
import gzip
import os
import struct

import numpy as np
import pytest

from mantaflow_amd import gridio


def test_uni_header_layout(tmp_path):
    p = str(tmp_path / "a.uni")
    a = np.arange(5 * 4 * 3, dtype=np.float32) * 0.5
    gridio.write_uni(p, a, (5, 4, 3), gridio.TYPE_REAL | gridio.TYPE_LEVELSET, timestamp=1234567)
    raw = gzip.open(p, "rb").read()
    # "MNT3" + int dimX, dimY, dimZ, gridType, elementType, bytesPerElement
    # + char info[252] + int dimT + uint64 timestamp = 4 + 288 bytes, then data
    assert raw[:4] == b"MNT3" and len(raw) == 4 + 288 + a.nbytes
    assert struct.unpack_from("<6i", raw, 4) == (5, 4, 3, 17, 1, 4)
    assert struct.unpack_from("<i", raw, 4 + 24 + 252)[0] == 0
    assert struct.unpack_from("<Q", raw, 4 + 24 + 252 + 4)[0] == 1234567
    assert raw[4 + 24:4 + 24 + 8] == b"mantaflo"
    assert np.array_equal(np.frombuffer(raw[4 + 288:], "<f4"), a)


@pytest.mark.parametrize("gtype,ncomp,dt", [(gridio.TYPE_REAL, 1, np.float32), (gridio.TYPE_INT, 1, np.int32),
                                            (gridio.TYPE_VEC3, 3, np.float32),
                                            (gridio.TYPE_VEC3 | gridio.TYPE_MAC, 3, np.float32),
                                            (gridio.TYPE_INT | gridio.TYPE_FLAGS, 1, np.int32)])
def test_uni_round_trip(tmp_path, gtype, ncomp, dt):
    rng = np.random.default_rng(3)
    n = 7 * 6 * 5
    a = (rng.standard_normal((n, ncomp)) * 100).astype(dt)
    if ncomp == 1:
        a = a[:, 0]
    p = str(tmp_path / "g.uni")
    gridio.write_uni(p, a, (7, 6, 5), gtype)
    h, b = gridio.read_uni(p, (7, 6, 5), gtype)
    assert h["elementType"] == (0 if dt == np.int32 else (2 if ncomp == 3 else 1))
    assert h["bytesPerElement"] == 4 * ncomp and h["dims"] == (7, 6, 5)
    assert b.dtype == np.dtype("<i4" if dt == np.int32 else "<f4") and np.array_equal(b, a)


def test_uni_read_checks(tmp_path):
    p = str(tmp_path / "r.uni")
    gridio.write_uni(p, np.zeros(24, np.float32), (4, 3, 2), gridio.TYPE_REAL)
    # real <> levelset and vec3 <> mac are the same type for reading
    gridio.read_uni(p, (4, 3, 2), gridio.TYPE_REAL | gridio.TYPE_LEVELSET)
    with pytest.raises(gridio.GridIOError, match="grid dim doesn't match"):
        gridio.read_uni(p, (4, 3, 3), gridio.TYPE_REAL)
    with pytest.raises(gridio.GridIOError, match="grid type doesn't match"):
        gridio.read_uni(p, (4, 3, 2), gridio.TYPE_INT)
    q = str(tmp_path / "bad.uni")
    with gzip.open(q, "wb") as f:
        f.write(b"MNT9" + bytes(300))
    with pytest.raises(gridio.GridIOError, match="Unknown header 'MNT9'"):
        gridio.read_uni(q)
    assert gridio.unify_grid_type(gridio.TYPE_MAC) == gridio.TYPE_MAC | gridio.TYPE_VEC3
    with pytest.raises(gridio.GridIOError, match="unknown element type"):
        gridio.write_uni(q, np.zeros(1), (1, 1, 1), gridio.TYPE_FLAGS)


def test_raw_round_trip_and_length_check(tmp_path):
    p = str(tmp_path / "g.raw")
    a = np.arange(60, dtype=np.float32) - 7.25
    gridio.write_raw(p, a, False)
    assert gzip.open(p, "rb").read() == a.astype("<f4").tobytes()
    assert np.array_equal(gridio.read_raw(p, 60, False), a)
    with pytest.raises(gridio.GridIOError, match="stream length does not match"):
        gridio.read_raw(p, 61, False)


# golden files written by the reference's own Python .uni writer
# (tensorflow/tools/uniio.py, tests/golden/make_uni_golden.py), with what its
# reader returned for them
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GOLDEN_CASES = [("real_5x4x3", (5, 4, 3), 1, 1), ("vec3_4x3x2", (4, 3, 2), 12, 3), ("level_6x5x1", (6, 5, 1), 17, 1)]


@pytest.mark.parametrize("name,dims,gtype,nc", GOLDEN_CASES)
def test_uni_reads_reference_writer_files(name, dims, gtype, nc):
    ref = np.load(os.path.join(GOLDEN, "uniio_read.npz"))
    h, a = gridio.read_uni(os.path.join(GOLDEN, name + ".uni"), dims=dims, grid_type=gtype)
    hv = ref[name + "_header"]
    assert h["dims"] == tuple(int(v) for v in hv[:3]) == dims
    assert (h["gridType"], h["elementType"], h["bytesPerElement"], h["dimT"], h["timestamp"]) == \
        tuple(int(v) for v in hv[3:])
    want = ref[name + "_content"].reshape(-1, nc) if nc == 3 else ref[name + "_content"].reshape(-1)
    assert a.dtype == np.float32 and np.array_equal(a, want)
    assert np.array_equal(a.reshape(-1), ref[name + "_data"].reshape(-1))


@pytest.mark.parametrize("name,dims,gtype,nc", GOLDEN_CASES)
def test_uni_writes_reference_reader_bytes(tmp_path, name, dims, gtype, nc):
    """This package's writer and the reference's writer give the same
    uncompressed stream (id, header fields, cell data); only the free-text
    info field differs."""
    import gzip
    ref = np.load(os.path.join(GOLDEN, "uniio_read.npz"))
    data = ref[name + "_data"].reshape(-1, nc) if nc == 3 else ref[name + "_data"].reshape(-1)
    ts = int(ref[name + "_header"][7])
    p = str(tmp_path / "x.uni")
    gridio.write_uni(p, data, dims, gtype, timestamp=ts)
    with gzip.open(p, "rb") as f:
        mine = f.read()
    with gzip.open(os.path.join(GOLDEN, name + ".uni"), "rb") as f:
        theirs = f.read()
    assert len(mine) == len(theirs) == 4 + 288 + 4 * nc * dims[0] * dims[1] * dims[2]
    assert mine[:28] == theirs[:28]          # id + dimX..bytesPerElement
    assert mine[280:] == theirs[280:]        # dimT, timestamp, data
