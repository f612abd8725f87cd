"""Grid file formats of the reference's Grid<T>::save / load for device grids
(host-side: the grid is downloaded, written, read and uploaded in the
reference's host layout, x fastest, Vec3/MAC as xyz per cell).

.uni (fileio/iogrids.cpp:386-437 write, :440-509 read, header :38-44): a gzip
stream of the 4-byte id "MNT3", a packed 288-byte header
    int dimX, dimY, dimZ, gridType, elementType, bytesPerElement;
    char info[252]; int dimT; unsigned long long timestamp;
(little endian, no padding) and the cell data as float32 / int32.  The
element type is 0 for int, 1 for real, 2 for Vec3 (the first of the type
bits Int, Real, Vec3 that is set); reading checks the dimensions and the
grid type up to real<>levelset and vec3<>mac (unifyGridType, :213-221).

.raw (iogrids.cpp:255-291): the cell data alone, gzip-compressed.

The element layout is this module's own restatement of the reference's
struct and write order (the format has to match byte for byte).
"""
import gzip
import struct
import time

import numpy as np

TYPE_REAL, TYPE_INT, TYPE_VEC3, TYPE_MAC, TYPE_LEVELSET, TYPE_FLAGS = 1, 2, 4, 8, 16, 32
UNI_ID = b"MNT3"
UNI_HEADER = struct.Struct("<6i252siQ")  # 288 bytes
INFO = b"mantaflow_amd (MI355X grid path), float32"


class GridIOError(RuntimeError):
    pass


def element_type(grid_type):
    """0 int, 1 real, 2 vec3 (writeGridUni's order of tests)."""
    if grid_type & TYPE_INT:
        return 0
    if grid_type & TYPE_REAL:
        return 1
    if grid_type & TYPE_VEC3:
        return 2
    raise GridIOError("writeGridUni: unknown element type")


def unify_grid_type(t):
    if t & TYPE_REAL:
        t |= TYPE_LEVELSET
    if t & TYPE_LEVELSET:
        t |= TYPE_REAL
    if t & TYPE_VEC3:
        t |= TYPE_MAC
    if t & TYPE_MAC:
        t |= TYPE_VEC3
    return t


def _data_bytes(arr, et):
    dt = np.dtype("<i4") if et == 0 else np.dtype("<f4")
    return np.ascontiguousarray(arr, dtype=dt).tobytes()


def write_uni(path, arr, dims, grid_type, timestamp=None):
    """arr: cells in host order (n or n x 3); dims = (nx, ny, nz)."""
    et = element_type(grid_type)
    nx, ny, nz = (int(v) for v in dims)
    ncomp = 3 if et == 2 else 1
    if np.asarray(arr).size != nx * ny * nz * ncomp:
        raise GridIOError("writeGridUni: array size does not match the grid size")
    if timestamp is None:
        timestamp = int(time.time() * 1e6)
    head = UNI_HEADER.pack(nx, ny, nz, int(grid_type), et, 4 * ncomp, INFO, 0, int(timestamp))
    with gzip.open(path, "wb", compresslevel=1) as f:
        f.write(UNI_ID)
        f.write(head)
        f.write(_data_bytes(arr, et))
    return 1


def read_uni_header(f):
    ident = f.read(4)
    if ident != UNI_ID:
        raise GridIOError("readGridUni: Unknown header '%s' " % ident.decode("latin-1"))
    raw = f.read(UNI_HEADER.size)
    if len(raw) != UNI_HEADER.size:
        raise GridIOError("can't read file, no header present")
    nx, ny, nz, gt, et, bpe, info, dimt, ts = UNI_HEADER.unpack(raw)
    return {"dims": (nx, ny, nz), "gridType": gt, "elementType": et, "bytesPerElement": bpe,
            "info": info.split(b"\0", 1)[0].decode("latin-1"), "dimT": dimt, "timestamp": ts}


def read_uni(path, dims=None, grid_type=None):
    """Returns (header, array); with dims / grid_type given, checks them as
    readGridUni does."""
    with gzip.open(path, "rb") as f:
        h = read_uni_header(f)
        nx, ny, nz = h["dims"]
        if dims is not None and tuple(int(v) for v in dims) != (nx, ny, nz):
            raise GridIOError("grid dim doesn't match, %s vs %s" % ((nx, ny, nz), tuple(dims)))
        if grid_type is not None and unify_grid_type(h["gridType"]) != unify_grid_type(int(grid_type)):
            raise GridIOError("grid type doesn't match %d vs %d" % (h["gridType"], grid_type))
        et, bpe = h["elementType"], h["bytesPerElement"]
        ncomp = 3 if et == 2 else 1
        if bpe != 4 * ncomp:
            raise GridIOError("grid element size doesn't match %d vs %d" % (bpe, 4 * ncomp))
        n = nx * ny * nz
        raw = f.read(n * bpe)
        if len(raw) != n * bpe:
            raise GridIOError("can't read uni file, stream length does not match, %d vs %d" % (n * bpe, len(raw)))
    a = np.frombuffer(raw, dtype="<i4" if et == 0 else "<f4")
    return h, (a.reshape(n, 3) if ncomp == 3 else a).copy()


def write_raw(path, arr, is_int):
    with gzip.open(path, "wb", compresslevel=1) as f:
        f.write(_data_bytes(arr, 0 if is_int else 1))
    return 1


def read_raw(path, n_values, is_int):
    nbytes = 4 * n_values
    with gzip.open(path, "rb") as f:
        raw = f.read(nbytes + 1)
    if len(raw) != nbytes:
        raise GridIOError("can't read raw file, stream length does not match, %d vs %d" % (nbytes, len(raw)))
    return np.frombuffer(raw, dtype="<i4" if is_int else "<f4").copy()
