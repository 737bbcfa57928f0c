"""VBR inspector: block extraction, per-block routing, CSR assembly.

Vectorised restatement of utils/convert_real_to_vbr.py with the reference's
names, argument order and return tuples:

* ``convert_sparse_to_vbr(mat, rpntr, cpntr, fname, dst_dir)`` -> (val, indx,
  bindx, bpntrb, bpntre)   (utils/convert_real_to_vbr.py:98-199)
* ``convert_vbr_to_compressed(val, rpntr, cpntr, indx, bindx, bpntrb, bpntre,
  fname, dst_dir, density=None)`` -> (val2, indx2, bindx, bpntrb, bpntre,
  ublocks, indptr, indices, csr_val)   (utils/convert_real_to_vbr.py:18-96)

Results are bit-identical to the reference's loops (the same IEEE double
density statistic, the same strict ``>`` comparison, the same CSR padding
rule and ordering) but run as numpy array operations, so a 10^8-value matrix
is inspected in seconds instead of the reference's per-element Python loop.

``density=None`` in the reference loads a scikit-learn model
(``models/density_threshold_spmv.pkl``, :19-20,58); here it asks the learned
block router (``sable_amd.router``) when one is given via ``router=``, and
otherwise falls back to the fixed threshold ``DEFAULT_DENSITY``.
"""
from __future__ import annotations

import logging
import os

import numpy as np
import scipy.sparse

from .fileio import write_vbrc, _write_line
from .vbr import validate_vbr, block_table

logger = logging.getLogger(__name__)

# Mixed-matrix default threshold of the reference's synthetic flows
# (bench_sparse.py:16, synthesize_matrices/gen.py:72-87).
DEFAULT_DENSITY = 15

ROUTE_DENSE = 0
ROUTE_SPARSE = 1
ROUTE_DENSE_MFMA = 2    # SpMM: bf16 hi+lo split on MFMA


def block_nonzero_counts(val, indx):
    """Per block: number of stored values ``!= 0.0`` (the routing count of
    utils/convert_real_to_vbr.py:41-47) and the block size."""
    val = np.asarray(val)
    ix = np.asarray(indx, dtype=np.int64)
    sizes = np.diff(ix)
    nz = (val != 0)
    csum = np.concatenate([[0], np.cumsum(nz, dtype=np.int64)])
    counts = csum[ix[1:]] - csum[ix[:-1]]
    return counts, sizes


def block_density_percent(counts, sizes):
    """(dense/(dense+sparse))*100 in IEEE double, elementwise — the same two
    correctly rounded operations as utils/convert_real_to_vbr.py:50."""
    return (counts.astype(np.float64) / sizes.astype(np.float64)) * 100.0


def threshold_routes(val, indx, density):
    """Route per block under a fixed threshold: dense iff density% > T."""
    counts, sizes = block_nonzero_counts(val, indx)
    pct = block_density_percent(counts, sizes)
    return np.where(pct > density, ROUTE_DENSE, ROUTE_SPARSE).astype(np.uint8)


def compress_with_routes(val, rpntr, cpntr, indx, bindx, bpntrb, bpntre, routes):
    """Build the compressed (.vbrc) arrays for a given per-block route.

    Returns (val2, indx2, ublocks, indptr, indices, csr_val) exactly as
    utils/convert_real_to_vbr.py:57-82 would for the same routing decisions.
    """
    val = np.asarray(val)
    ix = np.asarray(indx, dtype=np.int64)
    routes = np.asarray(routes)
    nb = len(ix) - 1
    sizes = np.diff(ix)
    nrows, ncols = int(rpntr[-1]), int(cpntr[-1])
    dense_b = routes != ROUTE_SPARSE      # both dense routes keep the block
    elem_dense = np.repeat(dense_b, sizes)
    val2 = val[elem_dense]
    indx2 = np.concatenate([[0], np.cumsum(sizes[dense_b])]).astype(np.int64)
    ublocks = np.nonzero(~dense_b)[0].astype(np.int64)
    if len(ublocks) == 0 or nb == 0:
        empty_f = np.zeros(0, dtype=np.float64)
        empty_i = np.zeros(0, dtype=np.int64)
        return val2, indx2, ublocks, empty_i, empty_i, empty_f
    brow, bcol, r0, h, c0, w = block_table(rpntr, cpntr, bindx, bpntrb, bpntre)
    sel = np.nonzero((~elem_dense) & (val != 0))[0]
    blk = np.repeat(np.arange(nb), sizes)[sel]
    local = sel - ix[blk]
    ii = r0[blk] + local % h[blk]
    jj = c0[blk] + local // h[blk]
    vv = val[sel]
    if len(ii) == 0:
        # every sparse-routed block is all zeros: the reference's COO list is
        # empty and it writes no CSR (:81-83)
        empty_f = np.zeros(0, dtype=np.float64)
        empty_i = np.zeros(0, dtype=np.int64)
        return val2, indx2, ublocks, empty_i, empty_i, empty_f
    if not (np.any(ii == nrows - 1) and np.any(jj == ncols - 1)):
        ii = np.concatenate([ii, [nrows - 1]])
        jj = np.concatenate([jj, [ncols - 1]])
        vv = np.concatenate([vv.astype(np.float64), [0.0]])
    csr = scipy.sparse.coo_array((vv, (ii, jj))).tocsr()
    assert len(csr.indptr) == nrows + 1
    return val2, indx2, ublocks, csr.indptr, csr.indices, csr.data


def compress_with_routes_device(val, rpntr, cpntr, indx, bindx, bpntrb, bpntre, routes, return_ms=False):
    """compress_with_routes on the GPU (sable_compress_*): the same arrays,
    dtypes included, bit for bit."""
    import ctypes as C
    from . import runtime
    from ._lib import lib, check
    runtime.init()
    val = np.asarray(val)
    v64 = np.ascontiguousarray(val, dtype=np.float64)
    ix = np.ascontiguousarray(indx, dtype=np.int64)
    bx = np.ascontiguousarray(bindx, dtype=np.int64)
    rp = np.ascontiguousarray(rpntr, dtype=np.int64)
    cp = np.ascontiguousarray(cpntr, dtype=np.int64)
    bb = np.ascontiguousarray(bpntrb, dtype=np.int64)
    be = np.ascontiguousarray(bpntre, dtype=np.int64)
    rt = np.ascontiguousarray(routes, dtype=np.uint8)
    nb = len(ix) - 1
    h, nv2, nub, nnz, ms = C.c_void_p(), C.c_int64(), C.c_int64(), C.c_int64(), C.c_float()
    check(lib.sable_compress_create(v64.ctypes.data, len(v64), ix.ctypes.data, bx.ctypes.data, nb, rp.ctypes.data,
                                    len(rp) - 1, cp.ctypes.data, len(cp) - 1, bb.ctypes.data, be.ctypes.data,
                                    rt.ctypes.data, C.byref(h), C.byref(nv2), C.byref(nub), C.byref(nnz),
                                    C.byref(ms)))
    out = _fetch_compressed(h, nb, nv2.value, nub.value, nnz.value, int(rp[-1]), int(cp[-1]), val.dtype)
    return (out, ms.value) if return_ms else out


def _fetch_compressed(h, nb, nv2, nub, nnz, nrows, ncols, val_dtype):
    """The arrays of a compress handle (released here), in compress_with_routes'
    order and dtypes."""
    from ._lib import lib, check
    try:
        val2 = np.empty(nv2, dtype=np.float64)
        indx2 = np.empty(nb + 1 - nub, dtype=np.int64)
        ub = np.empty(nub, dtype=np.int64)
        if nnz >= 0:
            ip = np.empty(nrows + 1, dtype=np.int64)
            ind = np.empty(nnz, dtype=np.int64)
            cv = np.empty(nnz, dtype=np.float64)
            check(lib.sable_compress_fetch(h, val2.ctypes.data, indx2.ctypes.data, ub.ctypes.data, ip.ctypes.data,
                                           ind.ctypes.data, cv.ctypes.data))
        else:
            check(lib.sable_compress_fetch(h, val2.ctypes.data, indx2.ctypes.data, ub.ctypes.data, None, None,
                                           None))
    finally:
        lib.sable_compress_destroy(h)
    val2 = val2.astype(val_dtype)
    if nnz < 0:
        return val2, indx2, ub, np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0, np.float64)
    # the index dtypes scipy gives the host construction (COO of int64 row /
    # column arrays -> CSR), and the float64 the padding entry brings in
    probe = scipy.sparse.coo_array((np.zeros(1), (np.array([nrows - 1], dtype=np.int64),
                                                  np.array([ncols - 1], dtype=np.int64)))).tocsr()
    padded = cv[-1] == 0.0
    return (val2, indx2, ub, ip.astype(probe.indptr.dtype), ind.astype(probe.indices.dtype),
            cv if padded else cv.astype(val_dtype))


def sparse_to_vbrc_device(A, rpntr, cpntr, density=15, return_ms=False):
    """convert_sparse_to_vbr then convert_vbr_to_compressed with the fixed
    threshold (utils/convert_real_to_vbr.py:98-199, :18-96), the VBR arrays
    staying on the GPU between the two (sable_inspect_create, then
    sable_inspect_compress): returns (val2, indx2, bindx, bpntrb, bpntre,
    ublocks, indptr, indices, csr_val) as convert_vbr_to_compressed does, bit
    for bit.  A is a CSC matrix with summed duplicates; empty block rows get
    -1/-1.  Routes: dense iff the block's nonzero share in percent is
    strictly above density."""
    import ctypes as C
    from . import runtime
    from ._lib import lib, check
    runtime.init()
    rp = np.ascontiguousarray(rpntr, dtype=np.int64)
    cp = np.ascontiguousarray(cpntr, dtype=np.int64)
    ip = np.ascontiguousarray(A.indptr, dtype=np.int64)
    ix = np.ascontiguousarray(A.indices, dtype=np.int32)
    v = np.ascontiguousarray(A.data, dtype=np.float64)
    h, nb, nv, ms = C.c_void_p(), C.c_int64(), C.c_int64(), C.c_float()
    check(lib.sable_inspect_create(ip.ctypes.data, ix.ctypes.data, v.ctypes.data, A.shape[0], A.shape[1],
                                   rp.ctypes.data, len(rp) - 1, cp.ctypes.data, len(cp) - 1, C.byref(h),
                                   C.byref(nb), C.byref(nv), C.byref(ms)))
    try:
        indx = np.empty(nb.value + 1, dtype=np.int64)
        bindx = np.empty(nb.value, dtype=np.int64)
        bb = np.empty(len(rp) - 1, dtype=np.int64)
        be = np.empty(len(rp) - 1, dtype=np.int64)
        bnnz = np.empty(nb.value, dtype=np.int32)
        check(lib.sable_inspect_fetch(h, None, indx.ctypes.data, bindx.ctypes.data, bb.ctypes.data,
                                      be.ctypes.data, bnnz.ctypes.data))
        routes = np.where(block_density_percent(bnnz, np.diff(indx)) > density, ROUTE_DENSE,
                          ROUTE_SPARSE).astype(np.uint8)
        hc, nv2, nub, nnz, ms2 = C.c_void_p(), C.c_int64(), C.c_int64(), C.c_int64(), C.c_float()
        check(lib.sable_inspect_compress(h, routes.ctypes.data if len(routes) else None, C.byref(hc),
                                         C.byref(nv2), C.byref(nub), C.byref(nnz), C.byref(ms2)))
    finally:
        lib.sable_inspect_destroy(h)
    val2, indx2, ub, indptr, indices, csr_val = _fetch_compressed(hc, nb.value, nv2.value, nub.value, nnz.value,
                                                                  int(rp[-1]), int(cp[-1]), v.dtype)
    out = (val2, indx2, bindx, bb, be, ub, indptr, indices, csr_val)
    return (out, ms.value + ms2.value) if return_ms else out


def convert_vbr_to_compressed(val, rpntr, cpntr, indx, bindx, bpntrb, bpntre,
                              fname, dst_dir, density=None, *, router=None,
                              routes=None, device=False):
    """utils/convert_real_to_vbr.py:18-96 (vectorised).

    ``density``: fixed threshold T (percent); a block is dense iff its
    nonzero percentage is strictly greater than T.  With ``density=None`` the
    learned ``router`` decides (falling back to ``DEFAULT_DENSITY`` when no
    router is given).  ``routes`` overrides both with explicit per-block
    routes.  The .vbrc file is written when ``fname`` and ``dst_dir`` are not
    None, as in the reference.  ``device=True`` assembles the arrays on the
    GPU (the same arrays, bit for bit).
    """
    validate_vbr(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre)
    if routes is None:
        if density is not None:
            routes = threshold_routes(val, indx, density)
        elif router is not None:
            routes = router.route_blocks(val, indx, bindx, rpntr, cpntr,
                                         bpntrb, bpntre, n_rhs=1)
        else:
            routes = threshold_routes(val, indx, DEFAULT_DENSITY)
    compress = compress_with_routes_device if device else compress_with_routes
    val2, indx2, ublocks, indptr, indices, csr_val = compress(
        val, rpntr, cpntr, indx, bindx, bpntrb, bpntre, routes)
    if fname is not None and dst_dir is not None:
        os.makedirs(dst_dir, exist_ok=True)
        write_vbrc(os.path.join(dst_dir, f"{fname}.vbrc"), val2, csr_val,
                   indptr, indices, indx2, bindx, rpntr, cpntr, bpntrb,
                   bpntre, ublocks)
    return val2, indx2, bindx, bpntrb, bpntre, ublocks, indptr, indices, csr_val


def decompress_vbrc(val2, indx2, bindx, rpntr, cpntr, bpntrb, bpntre, ublocks,
                    indptr, indices, csr_val, dtype=np.float64):
    """Rebuild the full VBR ``val``/``indx`` and the per-block route from the
    compressed (.vbrc) arrays: dense blocks are copied from val2, sparse
    blocks are materialised from the CSR (zeros elsewhere).  Inverse of
    ``compress_with_routes``; entries of the CSR that fall outside every
    sparse-routed block can only be the explicit zero padding entry of
    utils/convert_real_to_vbr.py:70-74 and are dropped."""
    brow, bcol, r0, h, c0, w = block_table(rpntr, cpntr, bindx, bpntrb, bpntre)
    nb = len(brow)
    sizes = h * w
    indx = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    route = np.zeros(nb, dtype=np.uint8)
    ub = np.asarray(ublocks, dtype=np.int64)
    route[ub] = ROUTE_SPARSE
    val = np.zeros(int(indx[-1]), dtype=dtype)
    dense_b = route == ROUTE_DENSE
    val[np.repeat(dense_b, sizes)] = np.asarray(val2, dtype=dtype)
    if len(ub) and len(csr_val):
        ip = np.asarray(indptr, dtype=np.int64)
        rows = np.repeat(np.arange(len(ip) - 1), np.diff(ip))
        cols = np.asarray(indices, dtype=np.int64)
        vals = np.asarray(csr_val, dtype=dtype)
        a = np.searchsorted(np.asarray(rpntr), rows, side="right") - 1
        b = np.searchsorted(np.asarray(cpntr), cols, side="right") - 1
        # ordinal of (a, b) among sparse blocks
        key_blocks = brow[ub] * (len(cpntr)) + bcol[ub]
        order = np.argsort(key_blocks)
        key_sorted = key_blocks[order]
        key_e = a * (len(cpntr)) + b
        pos = np.searchsorted(key_sorted, key_e)
        pos_c = np.minimum(pos, len(key_sorted) - 1)
        hit = (pos < len(key_sorted)) & (key_sorted[pos_c] == key_e)
        if np.any(~hit & (vals != 0)):
            raise ValueError("CSR entry outside every sparse-routed block")
        k = ub[order[pos_c[hit]]]
        off = indx[k] + (cols[hit] - c0[k]) * h[k] + (rows[hit] - r0[k])
        val[off] = vals[hit]
    return val, indx, route


def sparse_to_vbr_device(A, rpntr, cpntr):
    """convert_sparse_to_vbr's arrays built on the GPU (sable_inspect_*):
    returns (val (FP64), indx, bindx, bpntrb, bpntre with -1 for a block row
    without blocks, block_nnz, device_ms).  A is a CSC matrix with summed
    duplicates."""
    import ctypes as C
    from . import runtime
    from ._lib import lib, check
    runtime.init()
    rp = np.ascontiguousarray(rpntr, dtype=np.int64)
    cp = np.ascontiguousarray(cpntr, dtype=np.int64)
    ip = np.ascontiguousarray(A.indptr, dtype=np.int64)
    ix = np.ascontiguousarray(A.indices, dtype=np.int32)
    v = np.ascontiguousarray(A.data, dtype=np.float64)
    h, nb, nv, ms = C.c_void_p(), C.c_int64(), C.c_int64(), C.c_float()
    check(lib.sable_inspect_create(ip.ctypes.data, ix.ctypes.data, v.ctypes.data, A.shape[0], A.shape[1],
                                   rp.ctypes.data, len(rp) - 1, cp.ctypes.data, len(cp) - 1, C.byref(h),
                                   C.byref(nb), C.byref(nv), C.byref(ms)))
    try:
        val = np.empty(nv.value, dtype=np.float64)
        indx = np.empty(nb.value + 1, dtype=np.int64)
        bindx = np.empty(nb.value, dtype=np.int64)
        bb = np.empty(len(rp) - 1, dtype=np.int64)
        be = np.empty(len(rp) - 1, dtype=np.int64)
        bnnz = np.empty(nb.value, dtype=np.int32)
        check(lib.sable_inspect_fetch(h, val.ctypes.data, indx.ctypes.data, bindx.ctypes.data, bb.ctypes.data,
                                      be.ctypes.data, bnnz.ctypes.data))
    finally:
        lib.sable_inspect_destroy(h)
    return val, indx, bindx, bb, be, bnnz, ms.value


def convert_sparse_to_vbr(mat, rpntr, cpntr, fname, dst_dir, *,
                          allow_empty_rows=False, device=False):
    """utils/convert_real_to_vbr.py:98-199 (vectorised).

    A block is kept when it holds at least one structurally stored entry
    (explicit zeros count, as scipy's ``block.nnz`` does); kept blocks are
    materialised densely in column-major order.  The reference asserts that
    every block row keeps a block (:179-184); ``allow_empty_rows=True``
    writes -1/-1 for such rows instead.  Writes ``{fname}.vbr`` when
    ``fname``/``dst_dir`` are given.  ``device=True`` builds the arrays on
    the GPU (the same arrays, bit for bit).
    """
    A = scipy.sparse.csc_matrix(mat)
    A.sum_duplicates()
    rp = np.asarray(rpntr, dtype=np.int64)
    cp = np.asarray(cpntr, dtype=np.int64)
    nbr, nbc = len(rp) - 1, len(cp) - 1
    logger.info("mat shape: %s, num block rows: %d, num block cols : %d, nnz: %d",
                A.shape, len(rp), len(cp), A.nnz)
    if device:
        if np.iscomplexobj(A.data):
            raise ValueError("the device path takes real values")
        val64, indx, bindx, bpntrb, bpntre, _, _ = sparse_to_vbr_device(A, rp, cp)
        val = val64.astype(A.dtype)
        if np.any(bpntrb == -1) and not allow_empty_rows:
            raise AssertionError("convert_sparse_to_vbr: a block row holds no block")
        return _write_vbr_file(fname, dst_dir, val, indx, bindx, rpntr, cpntr, bpntrb, bpntre)
    coo = A.tocoo()
    i = coo.row.astype(np.int64)
    j = coo.col.astype(np.int64)
    a = np.searchsorted(rp, i, side="right") - 1
    b = np.searchsorted(cp, j, side="right") - 1
    key = a * nbc + b
    kept = np.unique(key)                      # block-row-major order
    kr, kc = kept // nbc, kept % nbc
    h = rp[kr + 1] - rp[kr]
    w = cp[kc + 1] - cp[kc]
    sizes = h * w
    indx = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    val = np.zeros(int(indx[-1]), dtype=A.dtype)
    k = np.searchsorted(kept, key)
    off = indx[k] + (j - cp[kc[k]]) * h[k] + (i - rp[kr[k]])
    val[off] = coo.data
    bindx = kc.astype(np.int64)
    counts = np.bincount(kr, minlength=nbr)
    if np.any(counts == 0) and not allow_empty_rows:
        # the reference's assert(len(bpntrb) == len(rpntr)-1), :184
        raise AssertionError("convert_sparse_to_vbr: a block row holds no block")
    ends = np.cumsum(counts)
    bpntrb = np.where(counts > 0, ends - counts, -1).astype(np.int64)
    bpntre = np.where(counts > 0, ends, -1).astype(np.int64)
    return _write_vbr_file(fname, dst_dir, val, indx, bindx, rpntr, cpntr, bpntrb, bpntre)


def _write_vbr_file(fname, dst_dir, val, indx, bindx, rpntr, cpntr, bpntrb, bpntre):
    if fname is not None and dst_dir is not None:
        with open(os.path.join(dst_dir, f"{fname}.vbr"), "w") as f:
            for name, arr in (("val", val), ("indx", indx), ("bindx", bindx),
                              ("rpntr", rpntr), ("cpntr", cpntr),
                              ("bpntrb", bpntrb), ("bpntre", bpntre)):
                _write_line(f, name, arr)
    return val, indx, bindx, bpntrb, bpntre


def vbr_to_dense(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre):
    """utils/mtx_matrices_gen.py:29-49 without the file write."""
    brow, bcol, r0, h, c0, w = block_table(rpntr, cpntr, bindx, bpntrb, bpntre)
    M = np.zeros((int(rpntr[-1]), int(cpntr[-1])), dtype=np.float64)
    ix = np.asarray(indx, dtype=np.int64)
    v = np.asarray(val, dtype=np.float64)
    for k in range(len(brow)):
        blk = v[ix[k]:ix[k + 1]].reshape(int(w[k]), int(h[k])).T
        M[r0[k]:r0[k] + h[k], c0[k]:c0[k] + w[k]] = blk
    return M


def vbr_to_csc(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre):
    """The VBR matrix as a scipy CSC matrix of its nonzero values, without a
    dense expansion: stored value p of block k sits at row r0 + (p mod h),
    column c0 + (p div h) (column-major blocks, src/vbr.py:5-14)."""
    _, _, r0, h, c0, _ = block_table(rpntr, cpntr, bindx, bpntrb, bpntre)
    val = np.asarray(val)
    ix = np.asarray(indx, dtype=np.int64)
    nz = np.flatnonzero(val)
    k = np.searchsorted(ix, nz, side="right") - 1
    off = nz - ix[k]
    rows = np.asarray(r0, dtype=np.int64)[k] + off % np.asarray(h, dtype=np.int64)[k]
    cols = np.asarray(c0, dtype=np.int64)[k] + off // np.asarray(h, dtype=np.int64)[k]
    shape = (int(np.asarray(rpntr)[-1]), int(np.asarray(cpntr)[-1]))
    return scipy.sparse.csc_matrix((val[nz].astype(np.float64), (rows, cols)), shape=shape)


def vbr_to_mtx(filename, dir_name, vbr_dir):
    """utils/mtx_matrices_gen.py:29-49: read ``{vbr_dir}/{filename}`` (.vbr),
    write ``{dir_name}/{stem}.mtx`` and return the dense matrix."""
    from .fileio import read_vbr, write_mm_file
    assert filename.endswith(".vbr")
    os.makedirs(dir_name, exist_ok=True)
    val, indx, bindx, rpntr, cpntr, bpntrb, bpntre = read_vbr(os.path.join(vbr_dir, filename))
    M = vbr_to_dense(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre)
    write_mm_file(os.path.join(dir_name, filename[:-len(".vbr")] + ".mtx"), M)
    return M
