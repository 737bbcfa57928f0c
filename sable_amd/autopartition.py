"""Block partitions from adjacent-line similarity (reference:
src/autopartition.py).

``cut_indices2(A, cut_threshold, similarity)`` returns ``(col_indices,
row_indices)`` exactly like src/autopartition.py:34-68: columns first, then
rows; a run of structurally empty lines is skipped and cut right after its
end; otherwise a cut goes between lines i and i+1 when
``similarity(line_i, line_{i+1}) < cut_threshold``.

``similarity2`` (src/autopartition.py:73-74) is max(a.b, a[1:].b[:-1],
a[:-1].b[1:]) / max(nnz(a), nnz(b)); 0/0 gives NaN, which never cuts.
``similarity2_numba`` (:76-94) returns 0 when either line has no nonzero
value (so such a pair always cuts).  Both are provided under the reference's
names (numba is not needed).

Instead of densifying two lines per step (the reference's per-line
``toarray()``), the three dot products of every adjacent pair are computed at
once with sparse elementwise products (each summed in ascending index
order), and only the final scan loops in Python.  ``device=True`` runs both
on the GPU (``sable_axis_cuts``: one wavefront per adjacent pair, and the
scan in its closed form per position), with the same result bit for bit.  A
user-supplied similarity function falls back to the per-pair loop.
"""
from __future__ import annotations

import gc
import pathlib

import numpy as np
import scipy.io
import scipy.sparse

from .convert import convert_sparse_to_vbr, convert_vbr_to_compressed


def similarity2(a, b):
    """src/autopartition.py:73-74 (dense 1-D arrays)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(np.count_nonzero(a), np.count_nonzero(b))
    num = max(a.dot(b), a[1:].dot(b[:-1]), a[:-1].dot(b[1:]))
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.float64(num) / den


def similarity2_numba(a, b):
    """src/autopartition.py:76-94 semantics: 0 when either line is all zero."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    ca, cb = np.count_nonzero(a), np.count_nonzero(b)
    if ca == 0 or cb == 0:
        return 0.0
    if len(a) > 1 and len(b) > 1:
        m = max(a.dot(b), a[1:].dot(b[:-1]), a[:-1].dot(b[1:]))
    else:
        m = a.dot(b)
    return m / max(ca, cb)


def similarity1(a, b):
    """src/autopartition.py:70-71."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return (a.dot(b) + a[1:].dot(b[:-1]) + a[:-1].dot(b[1:])) / (3 * max(np.count_nonzero(a), np.count_nonzero(b)))


def _colsum(P):
    """Column sums of a sparse matrix, each summed in ascending row order (one
    rounding per addition: np.bincount accumulates in index order), the
    order the GPU partitioner uses too."""
    P = scipy.sparse.csc_matrix(P)
    P.sort_indices()
    cols = np.repeat(np.arange(P.shape[1]), np.diff(P.indptr))
    return np.bincount(cols, weights=P.data, minlength=P.shape[1]).astype(np.float64)


def _pair_similarities(M, kind):
    """Similarity of every adjacent column pair (i, i+1) of the CSC matrix M.
    Each dot product is the sequential ascending-row sum of the entry
    products (exact for integer values, the reference's tests/Franz8.mtx
    included)."""
    n = M.shape[1]
    if n < 2:
        return np.zeros(0)
    L = M[:, :-1]
    R = M[:, 1:]
    d0 = _colsum(L.multiply(R))
    if M.shape[0] > 1:
        d1 = _colsum(L[1:, :].multiply(R[:-1, :]))
        d2 = _colsum(L[:-1, :].multiply(R[1:, :]))
        num = np.maximum(d0, np.maximum(d1, d2))
    else:
        num = d0
    nzv = M.copy()
    nzv.data = (nzv.data != 0).astype(np.float64)
    cnt = np.asarray(nzv.sum(axis=0)).ravel()
    den = np.maximum(cnt[:-1], cnt[1:])
    with np.errstate(invalid="ignore", divide="ignore"):
        sim = num / den
    if kind == "numba":
        sim = np.where((cnt[:-1] == 0) | (cnt[1:] == 0), 0.0, sim)
    elif kind == "s1":
        d = np.asarray(d0, dtype=np.float64)
        if M.shape[0] > 1:
            d = d + d1 + d2
        with np.errstate(invalid="ignore", divide="ignore"):
            sim = d / (3 * den)
    return sim


def _scan(nnz, sim, threshold):
    """The sequential cut rule of src/autopartition.py:38-50 over one axis."""
    n = len(nnz)
    cuts = [0]
    i = 0
    while i < n - 1:
        if nnz[i] == 0:
            while i < n - 1 and nnz[i + 1] == 0:
                i += 1
            cuts.append(i + 1)
        elif sim[i] < threshold:
            cuts.append(i + 1)
        i += 1
    if cuts[-1] != n:
        cuts.append(n)
    return cuts


def _axis_cuts(M, threshold, similarity):
    """Cuts along the columns of the CSC matrix M."""
    M = M.tocsc()
    M.sum_duplicates()
    nnz = np.diff(M.indptr)
    kind = {similarity2: "s2", similarity2_numba: "numba", similarity1: "s1"}.get(similarity)
    if kind is not None:
        sim = _pair_similarities(M.astype(np.float64), kind)
    else:
        sim = np.empty(max(M.shape[1] - 1, 0))
        for i in range(M.shape[1] - 1):
            if nnz[i] == 0:
                continue
            sim[i] = similarity(M[:, i].toarray().ravel(), M[:, i + 1].toarray().ravel())
    return _scan(nnz, sim, threshold)


_GPU_KIND = {similarity2: 0, similarity2_numba: 1, similarity1: 2}


def _axis_cuts_gpu(M, threshold, similarity, return_sim=False):
    """Cuts along the columns of M on the GPU (sable_axis_cuts): the same
    similarities as the host path, bit for bit, and the same scan."""
    from . import runtime
    from ._lib import lib, check
    kind = _GPU_KIND.get(similarity)
    if kind is None:
        raise ValueError("the GPU partitioner takes similarity2, similarity2_numba or similarity1")
    M = scipy.sparse.csc_matrix(M).astype(np.float64)
    M.sum_duplicates()
    M.sort_indices()
    n = M.shape[1]
    ip = np.ascontiguousarray(M.indptr, dtype=np.int64)
    ix = np.ascontiguousarray(M.indices, dtype=np.int32)
    v = np.ascontiguousarray(M.data, dtype=np.float64)
    cut = np.zeros(n + 1, dtype=np.uint8)
    sim = np.zeros(max(n - 1, 1), dtype=np.float64)
    runtime.init()
    check(lib.sable_axis_cuts(ip.ctypes.data, ix.ctypes.data, v.ctypes.data, n, M.shape[0], kind,
                              float(threshold), cut.ctypes.data, sim.ctypes.data, None))
    cuts = np.flatnonzero(cut).tolist()
    return (cuts, sim[:max(n - 1, 0)]) if return_sim else cuts


def cut_indices2(A, cut_threshold, similarity, device=False):
    """src/autopartition.py:34-68.  Returns (col_indices, row_indices).
    device=True computes both axes on the GPU (same result)."""
    A = scipy.sparse.csc_matrix(A)
    axis = _axis_cuts_gpu if device else _axis_cuts
    cols = axis(A, cut_threshold, similarity)
    rows = axis(A.T.tocsc(), cut_threshold, similarity)
    return cols, rows


def cut_indices2_fast(A, cut_threshold, similarity):
    """src/autopartition.py:96-177 (same cuts as cut_indices2)."""
    return cut_indices2(A, cut_threshold, similarity)


def cut_indices1(A, cut_threshold, similarity):
    """src/autopartition.py:16-32: cut at i+1 iff similarity < threshold (no
    empty-run rule)."""
    A = scipy.sparse.csc_matrix(A)
    out = []
    for M in (A, A.T.tocsc()):
        kind = {similarity2: "s2", similarity2_numba: "numba", similarity1: "s1"}.get(similarity)
        if kind is None:
            sim = np.array([similarity(M[:, i].toarray().ravel(), M[:, i + 1].toarray().ravel())
                            for i in range(M.shape[1] - 1)])
        else:
            sim = _pair_similarities(M.astype(np.float64), kind)
        out.append([0] + [i + 1 for i in range(M.shape[1] - 1) if sim[i] < cut_threshold] + [M.shape[1]])
    return out[0], out[1]


def my_convert_dense_to_vbrc(file_info, cut_threshold, cut_indices, similarity, *, router=None,
                             allow_empty_rows=False, device=False):
    """src/autopartition.py:196-205: .mtx -> partition -> .vbr -> routed
    .vbrc.  Returns (val, rpntr, cpntr, indx, bindx, bpntrb, bpntre, ublocks,
    indptr, indices, csr_val) like the reference.  device=True runs the
    partition (when cut_indices is cut_indices2), the block materialisation
    and the compression on the GPU, with the same results."""
    src_path, dest_path = file_info
    mtx = scipy.io.mmread(src_path)
    A = scipy.sparse.csc_matrix(mtx, copy=False)
    del mtx
    gc.collect()
    kw = {"device": True} if device and cut_indices is cut_indices2 else {}
    cpntr, rpntr = cut_indices(A, cut_threshold, similarity, **kw)
    stem = pathlib.Path(src_path).resolve().stem
    dst = pathlib.Path(dest_path).resolve().parent
    val, indx, bindx, bpntrb, bpntre = convert_sparse_to_vbr(A, rpntr, cpntr, stem, dst,
                                                             allow_empty_rows=allow_empty_rows, device=device)
    val, indx, bindx, bpntrb, bpntre, ublocks, indptr, indices, csr_val = convert_vbr_to_compressed(
        val, rpntr, cpntr, indx, bindx, bpntrb, bpntre, stem, dst, router=router, device=device)
    return val, rpntr, cpntr, indx, bindx, bpntrb, bpntre, ublocks, indptr, indices, csr_val
