"""The vectorised partitioner against the oracle: brute force on tiny 0/1
patterns (exact integer dot products) and the reference's own test matrix
tests/Franz8.mtx (integer values, so every comparison is exact)."""
import gzip
import os

import numpy as np
import pytest
import scipy.io
import scipy.sparse

from oracle import vbr_oracle as O
from sable_amd import autopartition as AP
from sable_amd.convert import convert_sparse_to_vbr, vbr_to_dense


@pytest.mark.parametrize("seed", range(12))
def test_cut_indices2_matches_loop_oracle(seed):
    rng = np.random.default_rng(seed)
    n, m = int(rng.integers(2, 12)), int(rng.integers(2, 12))
    P = (rng.random((n, m)) < rng.uniform(0.1, 0.7)).astype(float)
    P[:, rng.random(m) < 0.2] = 0      # empty columns / runs
    P[rng.random(n) < 0.2, :] = 0      # empty rows
    A = scipy.sparse.csc_matrix(P)
    for thr in (0.1, 0.2, 0.5):
        ref = O.cut_indices2(P.tolist(), (P != 0).tolist(), thr)
        assert AP.cut_indices2(A, thr, AP.similarity2) == tuple(ref) or list(AP.cut_indices2(A, thr, AP.similarity2)) == list(ref)
        assert AP.cut_indices2_fast(A, thr, AP.similarity2) == AP.cut_indices2(A, thr, AP.similarity2)


def test_similarity_functions_closed_forms():
    assert AP.similarity2([1, 1, 0], [1, 1, 0]) == 1.0
    assert AP.similarity2([1, 0, 0, 0], [0, 1, 0, 0]) == 1.0
    assert np.isnan(AP.similarity2([0, 0], [0, 0]))
    assert AP.similarity2_numba([0, 0], [1, 0]) == 0.0
    assert AP.similarity2_numba([1, 1, 1, 1], [1, 0, 0, 0]) == 0.25


def test_custom_similarity_callable_path():
    rng = np.random.default_rng(3)
    P = (rng.random((9, 8)) < 0.4).astype(float)
    A = scipy.sparse.csc_matrix(P)
    fast = AP.cut_indices2(A, 0.3, AP.similarity2)
    slow = AP.cut_indices2(A, 0.3, lambda a, b: AP.similarity2(a, b))
    assert fast == slow


def test_franz8_partition_matches_reference_algorithm(golden):
    with gzip.open(os.path.join(golden, "Franz8.mtx.gz"), "rt") as f:
        A = scipy.sparse.csc_matrix(scipy.io.mmread(f))
    assert A.shape == (16728, 7176) and A.nnz == 100368
    cols, rows = AP.cut_indices2(A, 0.2, AP.similarity2)
    rcols, rrows = O.cut_indices2_dense_lines(A, 0.2)
    assert cols == rcols and rows == rrows
    # the partition converts (the reference's test_partition_vals_real,
    # test.py:325-350, minus its missing canon file): nnz is preserved
    val, indx, bindx, bb, be = convert_sparse_to_vbr(A, rows, cols, None, None, allow_empty_rows=True)
    assert np.count_nonzero(val) == A.count_nonzero()
