"""Text wire formats of the path (reference: utils/fileio.py:13-112).

``.vbr``  seven lines ``name=[a,b,...]`` in the order val, indx, bindx,
          rpntr, cpntr, bpntrb, bpntre (utils/fileio.py:13-35).
``.vbrc`` eleven lines: val, csr_val, indptr, indices, indx, bindx, rpntr,
          cpntr, bpntrb, bpntre, ublocks (utils/convert_real_to_vbr.py:84-95,
          parsed by utils/fileio.py:66-93).

Values are written with ``str()`` of each element exactly as the reference
does, so files round-trip text-identically for the same input arrays.  The
readers accept the same files; unlike the reference's ``read_vbr`` they also
accept empty arrays.  Readers return numpy arrays (float64 values, int64
indices) instead of Python lists.

The readers and writers run in the native library (``csrc/vbr_text.cpp``):
the file is memory-mapped and every line is parsed by all allowed CPU
threads with correctly rounded conversion, so the arrays equal Python's
float() / int() of each token bit for bit; the writer formats doubles as
str() does.  Without the built library they fall back to numpy parsing and
Python formatting (same results, slower).

Binary container (``save_vbr_bin``/``load_vbr_bin``, ``write_vbrc_bin``/
``read_vbrc_bin``): a 64-byte header (magic ``SABLEVB1``, kind 0 = .vbr
arrays / 1 = .vbrc arrays, array count), one 48-byte descriptor per array
(name, element type, count, file offset) and the raw little-endian arrays at
64-byte aligned offsets, read back by memory map without parsing.  Binary
``.npz`` variants (``save_vbr_npz``/``load_vbr_npz``) are also kept; they are
loaded with ``allow_pickle=False``.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib
import struct

import numpy as np

from .vbr import VBR

FILEPATH = pathlib.Path(__file__).resolve().parent
BASE_PATH = os.path.join(FILEPATH, "..")


def _fmt(arr):
    a = np.asarray(arr)
    if a.dtype == np.float64 or a.dtype.kind in "iu":
        return ",".join(map(str, a.tolist()))
    return ",".join(map(str, a))


def _write_line(f, name, arr):
    f.write(f"{name}=[{_fmt(arr)}]\n")


def _native():
    """The native library, or None when it is not built."""
    try:
        from . import _lib
    except (ImportError, OSError):
        return None
    return _lib


def _native_kind(a):
    a = np.asarray(a)
    if a.dtype == np.float64:
        return 0
    if a.dtype.kind in "iu":
        return 1
    return None


def _write_lines(path, named, nthreads=0):
    """One `name=[...]` line per (name, array), natively when every array is
    float64 or integer, else with Python's str() (float32 values keep their
    own str(), as the reference would print them)."""
    L = _native()
    arrs = [np.asarray(a) for _, a in named]
    kinds = [_native_kind(a) for a in arrs]
    if L is None or any(k is None for k in kinds):
        with open(path, "w") as f:
            for (name, _), a in zip(named, arrs):
                _write_line(f, name, a)
        return
    arrs = [np.ascontiguousarray(a, dtype=np.float64 if k == 0 else np.int64) for a, k in zip(arrs, kinds)]
    n = len(arrs)
    names = (C.c_char_p * n)(*[nm.encode() for nm, _ in named])
    kk = (L.I32 * n)(*kinds)
    ptrs = (L.P * n)(*[a.ctypes.data if len(a) else None for a in arrs])
    counts = (L.I64 * n)(*[len(a) for a in arrs])
    if L.lib.sable_text_write(os.fsencode(path), n, names, kk, ptrs, counts, 0, nthreads):
        raise OSError(L.lib.sable_text_last_error().decode())


def write_vbr_matrix(filename: str, vbr_matrix: VBR, dir_name: str, nthreads: int = 0):
    """utils/fileio.py:13-35."""
    assert isinstance(vbr_matrix, VBR)
    assert isinstance(filename, str)
    os.makedirs(dir_name, exist_ok=True)
    _write_lines(os.path.join(dir_name, filename + ".vbr"),
                 [(name, getattr(vbr_matrix, name)) for name in
                  ("val", "indx", "bindx", "rpntr", "cpntr", "bpntrb", "bpntre")], nthreads)


def write_vbrc(path, val2, csr_val, indptr, indices, indx2, bindx, rpntr,
               cpntr, bpntrb, bpntre, ublocks, nthreads: int = 0):
    """The .vbrc writer of utils/convert_real_to_vbr.py:84-95."""
    _write_lines(path, [("val", val2), ("csr_val", csr_val), ("indptr", indptr), ("indices", indices),
                        ("indx", indx2), ("bindx", bindx), ("rpntr", rpntr), ("cpntr", cpntr),
                        ("bpntrb", bpntrb), ("bpntre", bpntre), ("ublocks", ublocks)], nthreads)


def _read_lines_native(L, filename, kinds, nthreads):
    n = len(kinds)
    counts = (L.I64 * n)()
    h = L.P()
    if L.lib.sable_text_open(os.fsencode(filename), n, nthreads, C.byref(h), counts):
        raise ValueError(L.lib.sable_text_last_error().decode())
    try:
        out = []
        for i, k in enumerate(kinds):
            a = np.empty(counts[i], dtype=k)
            fn = L.lib.sable_text_read_f64 if k == np.float64 else L.lib.sable_text_read_i64
            if fn(h, i, a.ctypes.data if len(a) else None):
                raise ValueError(L.lib.sable_text_last_error().decode())
            out.append(a)
    finally:
        L.lib.sable_text_close(h)
    return out


def _read_lines(filename, kinds, nthreads=0):
    """The first len(kinds) lines of a text file as arrays (np.float64 /
    np.int64 per kinds)."""
    L = _native()
    if L is not None:
        return _read_lines_native(L, filename, kinds, nthreads)
    with open(filename) as f:
        lines = [f.readline() for _ in range(len(kinds))]
    if any(not ln for ln in lines):
        raise ValueError(f"{filename}: fewer than {len(kinds)} lines")
    return [_parse(ln, k) for ln, k in zip(lines, kinds)]


def _parse(line, kind):
    body = line.split("=", 1)[1].strip()
    if not (body.startswith("[") and body.endswith("]")):
        raise ValueError(f"malformed line: {line[:40]!r}")
    body = body[1:-1]
    if body == "":
        return np.zeros(0, dtype=kind)
    return np.array(body.split(","), dtype=kind)


def read_vbr(filename, nthreads: int = 0):
    """utils/fileio.py:55-64: (val, indx, bindx, rpntr, cpntr, bpntrb, bpntre)."""
    return tuple(_read_lines(filename, [np.float64] + [np.int64] * 6, nthreads))


def read_vbrc(filename, nthreads: int = 0):
    """utils/fileio.py:66-93.  Returns (val, indx, bindx, rpntr, cpntr,
    bpntrb, bpntre, ublocks, indptr, indices, csr_val) in the reference's
    return order."""
    (val, csr_val, indptr, indices, indx, bindx, rpntr, cpntr, bpntrb, bpntre,
     ublocks) = _read_lines(filename, [np.float64, np.float64] + [np.int64] * 9, nthreads)
    return val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, ublocks, indptr, indices, csr_val


def write_dense_vector(val: float, size: int, dir_name=None):
    """utils/fileio.py:37-44 (file keyed by length, as the reference)."""
    dir_name = dir_name or os.path.join(BASE_PATH, "Generated_dense_tensors")
    os.makedirs(dir_name, exist_ok=True)
    with open(os.path.join(dir_name, f"generated_vector_{size}.vector"), "w") as f:
        f.write(",".join(map(str, [val] * size)) + "\n")


def write_dense_matrix(val: float, m: int, n: int, dir_name=None):
    """utils/fileio.py:46-53."""
    dir_name = dir_name or os.path.join(BASE_PATH, "Generated_dense_tensors")
    os.makedirs(dir_name, exist_ok=True)
    with open(os.path.join(dir_name, f"generated_matrix_{m}x{n}.matrix"), "w") as f:
        f.write(",".join(map(str, [val] * (n * m))) + "\n")


def read_dense_values(filename):
    """Parse a ``.vector``/``.matrix`` file (comma separated, one line)."""
    with open(filename) as f:
        line = f.readline().strip().rstrip(",")
    return np.array(line.split(","), dtype=np.float64) if line else np.zeros(0)


def write_mm_file(filename, M):
    """utils/fileio.py:105-112 (dense array -> MatrixMarket coordinate)."""
    M = np.asarray(M)
    nz = np.argwhere(M != 0)
    with open(filename, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{M.shape[0]} {M.shape[1]} {len(nz)}\n")
        for i, j in nz:
            f.write(f"{i + 1} {j + 1} {M[i][j]}\n")


def save_vbr_npz(path, vbr: VBR):
    np.savez(path, val=np.asarray(vbr.val), indx=np.asarray(vbr.indx),
             bindx=np.asarray(vbr.bindx), rpntr=np.asarray(vbr.rpntr),
             cpntr=np.asarray(vbr.cpntr), bpntrb=np.asarray(vbr.bpntrb),
             bpntre=np.asarray(vbr.bpntre))


def load_vbr_npz(path) -> VBR:
    with np.load(path, allow_pickle=False) as z:
        return VBR(z["val"], z["indx"], z["bindx"], z["rpntr"], z["cpntr"],
                   z["bpntrb"], z["bpntre"])


# ---- binary container ------------------------------------------------------

_BIN_MAGIC = b"SABLEVB1"
_BIN_TYPES = {1: np.dtype("<f8"), 2: np.dtype("<i8"), 3: np.dtype("<f4"), 4: np.dtype("<i4")}
_BIN_CODES = {v: k for k, v in _BIN_TYPES.items()}
_VBR_NAMES = ("val", "indx", "bindx", "rpntr", "cpntr", "bpntrb", "bpntre")
_VBRC_NAMES = ("val", "csr_val", "indptr", "indices", "indx", "bindx", "rpntr", "cpntr", "bpntrb", "bpntre",
               "ublocks")


def save_arrays_bin(path, kind: int, named):
    """Writes (name, array) pairs into the binary container.  float64,
    int64, float32 and int32 arrays keep their type; other integer arrays
    are stored as int64, other floating arrays as float64."""
    arrs = []
    for name, a in named:
        a = np.asarray(a)
        dt = a.dtype.newbyteorder("<")
        if dt not in _BIN_CODES:
            dt = np.dtype("<i8") if a.dtype.kind in "iub" else np.dtype("<f8")
        if len(name.encode()) > 16:
            raise ValueError(f"array name {name!r} longer than 16 bytes")
        arrs.append((name, np.ascontiguousarray(a.ravel(), dtype=dt)))
    off = 64 + 48 * len(arrs)
    descs = []
    for name, a in arrs:
        off = (off + 63) // 64 * 64
        descs.append(struct.pack("<16sIIQQQ", name.encode(), _BIN_CODES[a.dtype], 0, len(a), off, a.nbytes))
        off += a.nbytes
    with open(path, "wb") as f:
        f.write(struct.pack("<8sII48x", _BIN_MAGIC, kind, len(arrs)))
        for d in descs:
            f.write(d)
        for (name, a), d in zip(arrs, descs):
            f.seek(struct.unpack("<16sIIQQQ", d)[4])
            a.tofile(f)


def load_arrays_bin(path, mmap: bool = True):
    """Reads the binary container: (kind, {name: array}).  With mmap the
    arrays are read-only views of the mapped file."""
    with open(path, "rb") as f:
        head = f.read(64)
        if len(head) < 64 or head[:8] != _BIN_MAGIC:
            raise ValueError(f"{path}: not a SABLEVB1 container")
        kind, n = struct.unpack("<II", head[8:16])
        descs = [struct.unpack("<16sIIQQQ", f.read(48)) for _ in range(n)]
    size = os.path.getsize(path)
    out = {}
    for name, code, _, count, off, nbytes in descs:
        dt = _BIN_TYPES.get(code)
        if dt is None or nbytes != count * dt.itemsize or (count and off + nbytes > size):
            raise ValueError(f"{path}: bad descriptor for {name!r}")
        key = name.rstrip(b"\0").decode()
        if count == 0:
            out[key] = np.zeros(0, dtype=dt)
        elif mmap:
            out[key] = np.memmap(path, dtype=dt, mode="r", offset=off, shape=(count,))
        else:
            out[key] = np.fromfile(path, dtype=dt, count=count, offset=off)
    return kind, out


def save_vbr_bin(path, vbr: VBR):
    save_arrays_bin(path, 0, [(n, getattr(vbr, n)) for n in _VBR_NAMES])


def load_vbr_bin(path, mmap: bool = True) -> VBR:
    kind, a = load_arrays_bin(path, mmap)
    if kind != 0 or any(n not in a for n in _VBR_NAMES):
        raise ValueError(f"{path}: not a VBR container")
    return VBR(*(a[n] for n in _VBR_NAMES))


def write_vbrc_bin(path, val2, csr_val, indptr, indices, indx2, bindx, rpntr, cpntr, bpntrb, bpntre, ublocks):
    save_arrays_bin(path, 1, list(zip(_VBRC_NAMES, (val2, csr_val, indptr, indices, indx2, bindx, rpntr, cpntr,
                                                    bpntrb, bpntre, ublocks))))


def read_vbrc_bin(path, mmap: bool = True):
    """Same return order as read_vbrc."""
    kind, a = load_arrays_bin(path, mmap)
    if kind != 1 or any(n not in a for n in _VBRC_NAMES):
        raise ValueError(f"{path}: not a .vbrc container")
    return (a["val"], a["indx"], a["bindx"], a["rpntr"], a["cpntr"], a["bpntrb"], a["bpntre"], a["ublocks"],
            a["indptr"], a["indices"], a["csr_val"])
