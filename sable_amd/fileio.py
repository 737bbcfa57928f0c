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

Binary ``.npz`` variants (``save_vbr_npz``/``load_vbr_npz``) hold the same
arrays without text parsing; they are loaded with ``allow_pickle=False``.
"""
from __future__ import annotations

import os
import pathlib

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


def write_vbr_matrix(filename: str, vbr_matrix: VBR, dir_name: str):
    """utils/fileio.py:13-35."""
    assert isinstance(vbr_matrix, VBR)
    assert isinstance(filename, str)
    os.makedirs(dir_name, exist_ok=True)
    with open(os.path.join(dir_name, filename + ".vbr"), "w") as f:
        for name in ("val", "indx", "bindx", "rpntr", "cpntr", "bpntrb", "bpntre"):
            _write_line(f, name, getattr(vbr_matrix, name))


def write_vbrc(path, val2, csr_val, indptr, indices, indx2, bindx, rpntr,
               cpntr, bpntrb, bpntre, ublocks):
    """The .vbrc writer of utils/convert_real_to_vbr.py:84-95."""
    with open(path, "w") as f:
        _write_line(f, "val", val2)
        _write_line(f, "csr_val", csr_val)
        _write_line(f, "indptr", indptr)
        _write_line(f, "indices", indices)
        _write_line(f, "indx", indx2)
        _write_line(f, "bindx", bindx)
        _write_line(f, "rpntr", rpntr)
        _write_line(f, "cpntr", cpntr)
        _write_line(f, "bpntrb", bpntrb)
        _write_line(f, "bpntre", bpntre)
        _write_line(f, "ublocks", ublocks)


def _parse(line, kind):
    body = line.split("=", 1)[1].strip()
    if not (body.startswith("[") and body.endswith("]")):
        raise ValueError(f"malformed line: {line[:40]!r}")
    body = body[1:-1]
    if body == "":
        return np.zeros(0, dtype=kind)
    return np.array(body.split(","), dtype=kind)


def read_vbr(filename):
    """utils/fileio.py:55-64: (val, indx, bindx, rpntr, cpntr, bpntrb, bpntre)."""
    with open(filename) as f:
        lines = [f.readline() for _ in range(7)]
    kinds = [np.float64] + [np.int64] * 6
    return tuple(_parse(l, k) for l, k in zip(lines, kinds))


def read_vbrc(filename):
    """utils/fileio.py:66-93.  Returns (val, indx, bindx, rpntr, cpntr,
    bpntrb, bpntre, ublocks, indptr, indices, csr_val) in the reference's
    return order."""
    with open(filename) as f:
        lines = [f.readline() for _ in range(11)]
    val = _parse(lines[0], np.float64)
    csr_val = _parse(lines[1], np.float64)
    indptr = _parse(lines[2], np.int64)
    indices = _parse(lines[3], np.int64)
    indx = _parse(lines[4], np.int64)
    bindx = _parse(lines[5], np.int64)
    rpntr = _parse(lines[6], np.int64)
    cpntr = _parse(lines[7], np.int64)
    bpntrb = _parse(lines[8], np.int64)
    bpntre = _parse(lines[9], np.int64)
    ublocks = _parse(lines[10], np.int64)
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
