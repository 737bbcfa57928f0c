"""Pure-Python restatement of the reference inspector and executor — TEST ONLY.

Every function here is a deliberately plain loop (no vectorisation, no scipy
conversion routines) so that it is an independent statement of the
reference's algorithm against which the vectorised product code and the GPU
kernels are checked.  Use it on small inputs only (the 11x11 example, the
1000x1000 config-1 matrix, a few hundred blocks).

Pin status (what fixes each function besides this restatement):

=============================  ===============================================
function                       pinned by
=============================  ===============================================
convert_sparse_to_vbr          test.py:122-142 KAT arrays; docstring KAT
                               utils/convert_real_to_vbr.py:105-131;
                               tests/example.vbr; round trip to dense
convert_vbr_to_compressed      test.py:64-77 (threshold 0), test.py:80-105
                               (threshold 80: val2/indx2/ublocks/CSR);
                               invariant sum nnz(route) == nnz(A)
vbr_to_dense                   test.py:35-52 dense canon + example-canon.mtx
spmv_generated_c_f64           tests/output_spmv_canon.txt (x = 1),
                               output_spmv_canon_sparse.txt (routed plan);
                               closed form A @ x in FP64 for random x
spmm_generated_c               tests/output_spmm_canon.txt (X = 1, N = 512);
                               closed form A @ X for random X
similarity2 / cut_indices2     brute force on tiny 0/1 patterns where every
                               dot product is an exact integer (closed form)
=============================  ===============================================
"""
from __future__ import annotations

import math


# ---------------------------------------------------------------------------
# helpers

def _blocks_in_ordinal_order(rpntr, cpntr, bindx, bpntrb, bpntre):
    """Yield (a, b) in the reference's block-ordinal order.

    Follows utils/convert_real_to_vbr.py:28-33 and src/codegen.py:421-427:
    block rows ascending, rows with bpntrb == -1 skipped, and inside a row
    every block column b (ascending) that appears in bindx[bpntrb:bpntre].
    """
    for a in range(len(rpntr) - 1):
        if bpntrb[a] == -1:
            continue
        valid = list(bindx[bpntrb[a]:bpntre[a]])
        for b in range(len(cpntr) - 1):
            if b in valid:
                yield a, b


# ---------------------------------------------------------------------------
# inspector

def convert_sparse_to_vbr(entries, shape, rpntr, cpntr, allow_empty_rows=False):
    """Restates utils/convert_real_to_vbr.py:98-199.

    ``entries`` maps (i, j) -> value for every STRUCTURALLY stored entry of
    the matrix (explicit zeros included, as scipy's ``block.nnz`` counts
    them, :150-157).  Blocks are enumerated block-row-major (:146-159) and
    kept when they hold at least one stored entry; each kept block is written
    densely in column-major (Fortran) order (:162-163); ``indx`` is the prefix
    sum of block sizes (:167); ``bindx`` the block-column of every kept block
    (:173).  The reference asserts that no block row is empty (:179-184); with
    ``allow_empty_rows`` such a row is written as -1/-1 instead, which is the
    synthetic generator's convention (synthesize_matrices/vbr_matrices_gen.py
    :80-101).
    """
    nrows, ncols = shape
    assert rpntr[-1] == nrows and cpntr[-1] == ncols
    val, indx, bindx = [], [0], []
    row_of_block = []
    for ri in range(len(rpntr) - 1):
        for ci in range(len(cpntr) - 1):
            r0, r1, c0, c1 = rpntr[ri], rpntr[ri + 1], cpntr[ci], cpntr[ci + 1]
            stored = 0
            for (i, j) in entries:
                if r0 <= i < r1 and c0 <= j < c1:
                    stored += 1
            if stored == 0:
                continue
            for j in range(c0, c1):
                for i in range(r0, r1):
                    val.append(float(entries.get((i, j), 0.0)))
            indx.append(indx[-1] + (r1 - r0) * (c1 - c0))
            bindx.append(ci)
            row_of_block.append(ri)
    bpntrb, bpntre = [], []
    pos = 0
    for ri in range(len(rpntr) - 1):
        cnt = sum(1 for r in row_of_block if r == ri)
        if cnt == 0:
            if not allow_empty_rows:
                raise AssertionError("empty block row %d" % ri)
            bpntrb.append(-1)
            bpntre.append(-1)
            continue
        bpntrb.append(pos)
        pos += cnt
        bpntre.append(pos)
    return val, indx, bindx, bpntrb, bpntre


def block_density_percent(dense_count, sparse_count):
    """The routing statistic of utils/convert_real_to_vbr.py:50:
    (dense/(dense+sparse))*100 evaluated in IEEE double, Python semantics."""
    return (dense_count / (dense_count + sparse_count)) * 100


def convert_vbr_to_compressed(val, rpntr, cpntr, indx, bindx, bpntrb, bpntre,
                              density):
    """Restates utils/convert_real_to_vbr.py:18-96 for a fixed threshold.

    Per block (ordinal order), values are walked column-major (:39-48); a
    value counts as nonzero iff ``!= 0.0``; the block is dense iff
    ``calc_density > density`` (strict, :52-54).  Dense blocks are copied to
    val2 and indx2 is extended (:61-63); sparse blocks contribute their
    nonzeros to a COO list and their ordinal to ublocks (:64-68).  If any COO
    entry exists and the last row or the last column index is missing, an
    explicit (nrows-1, ncols-1, 0.0) entry is appended (:70-74); the COO is
    then converted to CSR with rows ascending and, within a row, columns
    ascending (scipy's canonical order, :75-78).
    """
    nrows, ncols = rpntr[-1], cpntr[-1]
    val2, indx2, ublocks = [], [0], []
    coo = []  # (i, j, v) in walk order
    count = 0
    for a, b in _blocks_in_ordinal_order(rpntr, cpntr, bindx, bpntrb, bpntre):
        off = indx[count]
        k = 0
        nz, zeros, cur = 0, 0, []
        for j in range(cpntr[b], cpntr[b + 1]):
            for i in range(rpntr[a], rpntr[a + 1]):
                v = val[off + k]
                if v == 0.0:
                    zeros += 1
                else:
                    nz += 1
                    cur.append((i, j, v))
                k += 1
        if block_density_percent(nz, zeros) > density:
            val2.extend(val[indx[count]:indx[count + 1]])
            indx2.append(len(val2))
        else:
            coo.extend(cur)
            ublocks.append(count)
        count += 1
    if coo:
        rows = [e[0] for e in coo]
        cols = [e[1] for e in coo]
        if (nrows - 1) not in rows or (ncols - 1) not in cols:
            coo.append((nrows - 1, ncols - 1, 0.0))
        coo_sorted = sorted(coo, key=lambda e: (e[0], e[1]))
        indptr = [0] * (nrows + 1)
        for (i, _, _) in coo_sorted:
            indptr[i + 1] += 1
        for i in range(nrows):
            indptr[i + 1] += indptr[i]
        indices = [e[1] for e in coo_sorted]
        csr_val = [e[2] for e in coo_sorted]
    else:
        indptr, indices, csr_val = [], [], []
    return val2, indx2, ublocks, indptr, indices, csr_val


def vbr_to_dense(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre):
    """Restates utils/mtx_matrices_gen.py:29-49 (VBR -> dense, column-major
    walk of every stored block, block-ordinal order)."""
    nrows, ncols = rpntr[-1], cpntr[-1]
    M = [[0.0] * ncols for _ in range(nrows)]
    count = 0
    for a, b in _blocks_in_ordinal_order(rpntr, cpntr, bindx, bpntrb, bpntre):
        k = 0
        for j in range(cpntr[b], cpntr[b + 1]):
            for i in range(rpntr[a], rpntr[a + 1]):
                M[i][j] = float(val[indx[count] + k])
                k += 1
        count += 1
    return M


# ---------------------------------------------------------------------------
# executor semantics

def spmv_generated_c_f64(val2, indx2, bindx, rpntr, cpntr, bpntrb, bpntre,
                         ublocks, indptr, indices, csr_val, x):
    """Restates the single-threaded generated SpMV program, src/codegen.py
    :293-452 with the kernel shapes of :263-291, in FP64.

    y is zeroed (:416); if there are sparse-routed blocks the CSR part is
    applied first as y = A_csr x (MKL with beta = 0, :419-420; row sums as in
    src/csr-spmv.cpp:250-262); then every dense-routed block in ordinal order
    accumulates into y: height 1 -> row dot (:274-281), width 1 -> column
    axpy (:283-291), otherwise column-outer / row-inner (:263-272), with the
    value offset indx2[count] of the dense-only numbering.
    """
    nrows = rpntr[-1]
    y = [0.0] * nrows
    ub = set(ublocks)
    # sparse-routed blocks that hold only zeros produce no CSR at all
    # (utils/convert_real_to_vbr.py:69,80-83); the CSR part then adds nothing
    if len(ublocks) > 0 and len(indptr) > 0:
        for i in range(nrows):
            s = 0.0
            for p in range(indptr[i], indptr[i + 1]):
                s += float(csr_val[p]) * float(x[indices[p]])
            y[i] = s
    ordinal = 0
    count = 0
    for a, b in _blocks_in_ordinal_order(rpntr, cpntr, bindx, bpntrb, bpntre):
        if ordinal not in ub:
            i0, i1, j0, j1 = rpntr[a], rpntr[a + 1], cpntr[b], cpntr[b + 1]
            off = indx2[count]
            h = i1 - i0
            if h == 1:
                for j in range(j0, j1):
                    y[i0] += float(val2[off + j - j0]) * float(x[j])
            elif j1 - j0 == 1:
                xj = float(x[j0])
                for i in range(i0, i1):
                    y[i] += float(val2[off + i - i0]) * xj
            else:
                for j in range(j0, j1):
                    for i in range(i0, i1):
                        y[i] += float(val2[off + (j - j0) * h + (i - i0)]) * float(x[j])
            count += 1
        ordinal += 1
    return y


def spmm_generated_c(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, density,
                     X, n_rhs):
    """Restates gen_single_threaded_spmm, src/codegen.py:827-937, with the RHS
    width as a parameter (the reference hard-codes 512, :854-855).

    Blocks are routed inline with the same strict-density rule (:910-918;
    ``density == 0`` sends every block to the dense kernel).  Dense blocks run
    the i/j/k loop nest of spmm_kernel (:851-859); sparse blocks run one
    statement per stored nonzero, looped over the RHS columns (:919-922 via
    the DSL spmm, :150-163).  X and Y are row-major [rows x n_rhs].
    """
    nrows = rpntr[-1]
    Y = [0.0] * (nrows * n_rhs)
    count = 0
    for a, b in _blocks_in_ordinal_order(rpntr, cpntr, bindx, bpntrb, bpntre):
        i0, i1, j0, j1 = rpntr[a], rpntr[a + 1], cpntr[b], cpntr[b + 1]
        h = i1 - i0
        off = indx[count]
        dense = True
        if density > 0:
            nz = 0
            zeros = 0
            for k in range(h * (j1 - j0)):
                if val[off + k] == 0.0:
                    zeros += 1
                else:
                    nz += 1
            dense = block_density_percent(nz, zeros) > density
        for i in range(i0, i1):
            for j in range(j0, j1):
                v = float(val[off + (j - j0) * h + (i - i0)])
                if not dense and v == 0.0:
                    continue
                for k in range(n_rhs):
                    Y[i * n_rhs + k] += v * float(X[j * n_rhs + k])
        count += 1
    return Y


def spmv_f64_blocks(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, x):
    """FP64 y = A x straight from the VBR arrays (closed form), with
    math.fsum per row so the result is correctly rounded; the validation
    reference for random x."""
    nrows = rpntr[-1]
    terms = [[] for _ in range(nrows)]
    count = 0
    for a, b in _blocks_in_ordinal_order(rpntr, cpntr, bindx, bpntrb, bpntre):
        h = rpntr[a + 1] - rpntr[a]
        k = 0
        for j in range(cpntr[b], cpntr[b + 1]):
            for i in range(rpntr[a], rpntr[a + 1]):
                terms[i].append(float(val[indx[count] + k]) * float(x[j]))
                k += 1
        count += 1
    return [math.fsum(t) for t in terms]


# ---------------------------------------------------------------------------
# partitioner

def similarity2(a, b):
    """src/autopartition.py:73-74: max of the aligned and the two one-step
    shifted dot products, divided by the larger nonzero count."""
    n = len(a)
    d0 = sum(a[k] * b[k] for k in range(n))
    d1 = sum(a[k + 1] * b[k] for k in range(n - 1))
    d2 = sum(a[k] * b[k + 1] for k in range(n - 1))
    na = sum(1 for v in a if v != 0)
    nb = sum(1 for v in b if v != 0)
    den = max(na, nb)
    if den == 0:
        return float("nan")
    return max(d0, d1, d2) / den


def cut_indices2(dense, structural, cut_threshold):
    """src/autopartition.py:34-68 on a small dense matrix.

    ``structural[i][j]`` marks stored entries (the ``.nnz`` test of :42/:58);
    ``dense`` holds the values seen by the similarity.  Columns first, then
    rows; a run of empty lines is skipped and cut after its end; otherwise a
    cut goes at i+1 iff similarity < threshold (NaN never cuts).  Returns
    (col_indices, row_indices) like the reference.
    """
    nr = len(dense)
    nc = len(dense[0]) if nr else 0

    def col(j):
        return [dense[i][j] for i in range(nr)]

    def col_nnz(j):
        return sum(1 for i in range(nr) if structural[i][j])

    def row(i):
        return list(dense[i])

    def row_nnz(i):
        return sum(1 for j in range(nc) if structural[i][j])

    def scan(n, line, nnz):
        cuts = [0]
        i = 0
        while i < n - 1:
            if nnz(i) == 0:
                while i < n - 1 and nnz(i + 1) == 0:
                    i += 1
                cuts.append(i + 1)
            else:
                s = similarity2(line(i), line(i + 1))
                if s < cut_threshold:
                    cuts.append(i + 1)
            i += 1
        if cuts[-1] != n:
            cuts.append(n)
        return cuts

    return scan(nc, col, col_nnz), scan(nr, row, row_nnz)


def cut_indices2_dense_lines(A, cut_threshold):
    """src/autopartition.py:34-68 with similarity2 (:73-74), done the way the
    reference does it: every line is densified (``toarray().ravel()``) and
    the three dot products are numpy dots.  Usable at real-matrix scale
    (e.g. the reference's tests/Franz8.mtx); A is a scipy CSC matrix."""
    import numpy as np
    A = A.tocsc()
    R = A.tocsr()

    def sim(a, b):
        num = max(a.dot(b), a[1:].dot(b[:-1]), a[:-1].dot(b[1:]))
        den = max(np.count_nonzero(a), np.count_nonzero(b))
        with np.errstate(invalid="ignore", divide="ignore"):
            return np.float64(num) / den

    def scan(n, line, nnz):
        cuts = [0]
        i = 0
        while i < n - 1:
            if nnz(i) == 0:
                while i < n - 1 and nnz(i + 1) == 0:
                    i += 1
                cuts.append(i + 1)
            elif sim(line(i), line(i + 1)) < cut_threshold:
                cuts.append(i + 1)
            i += 1
        if cuts[-1] != n:
            cuts.append(n)
        return cuts

    cols = scan(A.shape[1], lambda i: A[:, i].toarray().ravel().astype(np.float64),
                lambda i: A.indptr[i + 1] - A.indptr[i])
    rows = scan(A.shape[0], lambda i: R[i, :].toarray().ravel().astype(np.float64),
                lambda i: R.indptr[i + 1] - R.indptr[i])
    return cols, rows
