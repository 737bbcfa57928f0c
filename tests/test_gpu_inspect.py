"""GPU block materialisation (sable_inspect_*: convert_sparse_to_vbr on the
device) against the reference's known-answer arrays, the oracle and the host
restatement (marked gpu).  Integer work and value copies: bit-identical."""
import numpy as np
import pytest
import scipy.sparse

from oracle import vbr_oracle as O
from sable_amd import convert as C
from sable_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def rt():
    from sable_amd import runtime
    runtime.init()
    return runtime


def _eq(a, b):
    return np.array_equal(np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64))


def test_reference_kats(kat, tmp_path):
    for dense_key, ex_key in (("dense_docstring", "example_vbr"),
                              ("dense_test_compression", "partition_test_compression")):
        print(f"case: KAT {dense_key}", flush=True)
        d = np.array(kat[dense_key], dtype=float)
        val, indx, bindx, bb, be = C.convert_sparse_to_vbr(scipy.sparse.csc_matrix(d), kat["rpntr"], kat["cpntr"],
                                                           "ex", str(tmp_path), device=True)
        ex = kat[ex_key]
        assert _eq(val, ex["val"]) and _eq(indx, ex["indx"]) and _eq(bindx, ex["bindx"])
        assert _eq(bb, ex["bpntrb"]) and _eq(be, ex["bpntre"])


def test_random_against_oracle_and_host():
    rng = np.random.default_rng(5)
    for t in range(6):
        n, m = 23, 19
        d = rng.uniform(-1, 1, (n, m))
        d[rng.random((n, m)) < 0.8] = 0
        A = scipy.sparse.csc_matrix(d).tolil()
        A[0, 0] = 0.0
        A = scipy.sparse.csc_matrix(A)
        A[0, 0] = 0.0                       # an explicit stored zero keeps its block
        rp, cp = [0, 5, 6, 13, 23], [0, 2, 9, 10, 19]
        coo = A.tocoo()
        entries = {(int(i), int(j)): float(v) for i, j, v in zip(coo.row, coo.col, coo.data)}
        print(f"case: random 23x19 #{t}", flush=True)
        r = O.convert_sparse_to_vbr(entries, (n, m), rp, cp, allow_empty_rows=True)
        out = C.convert_sparse_to_vbr(A, rp, cp, None, None, allow_empty_rows=True, device=True)
        for a, b in zip(out, r):
            assert _eq(a, b)
        if np.any(np.asarray(r[3]) == -1):
            with pytest.raises(AssertionError):
                C.convert_sparse_to_vbr(A, rp, cp, None, None, device=True)


def test_larger_matrices_bit_identical_to_host():
    for seed, dtype in ((1, np.float64), (2, np.float32)):
        vbr, _, _ = synth.mixed_vbr(20000, 200, 20, 200, 0.02, 0.5, (0.8, 1.0), (0.01, 0.2), seed)
        A = C.vbr_to_csc(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb, vbr.bpntre).astype(dtype)
        print(f"case: 20k mixed, {np.dtype(dtype).name}", flush=True)
        host = C.convert_sparse_to_vbr(A, vbr.rpntr, vbr.cpntr, None, None, allow_empty_rows=True)
        dev = C.convert_sparse_to_vbr(A, vbr.rpntr, vbr.cpntr, None, None, allow_empty_rows=True, device=True)
        for a, b in zip(host, dev):
            assert a.dtype == b.dtype and np.array_equal(a, b)
        # the routing count per block
        _, indx, _, _, _, bnnz, _ = C.sparse_to_vbr_device(A, vbr.rpntr, vbr.cpntr)
        assert np.array_equal(bnnz, C.block_nonzero_counts(host[0], indx)[0])


def test_c2_pattern_matches_host():
    vbr, _, _ = synth.workload_c2()
    A = C.vbr_to_csc(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb, vbr.bpntre)
    rp, cp = vbr.rpntr, vbr.cpntr
    del vbr
    print("case: 100k matrix", flush=True)
    dev = C.convert_sparse_to_vbr(A, rp, cp, None, None, allow_empty_rows=True, device=True)
    host = C.convert_sparse_to_vbr(A, rp, cp, None, None, allow_empty_rows=True)
    for a, b in zip(host, dev):
        assert np.array_equal(a, b)


def test_bad_inputs_rejected():
    from sable_amd._lib import SableError
    A = scipy.sparse.csc_matrix(np.eye(6))
    for rp, cp in (([0, 3, 5], [0, 6]), ([0, 3, 3, 6], [0, 6]), ([1, 6], [0, 6])):
        with pytest.raises(SableError):
            C.sparse_to_vbr_device(A, rp, cp)
