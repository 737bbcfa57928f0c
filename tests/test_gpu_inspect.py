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


def test_compressed_kat_threshold_80(kat, tmp_path):
    from sable_amd import fileio
    d = np.array(kat["dense_test_compression"], dtype=float)
    rp, cp = kat["rpntr"], kat["cpntr"]
    val, indx, bindx, bb, be = C.convert_sparse_to_vbr(scipy.sparse.csc_matrix(d), rp, cp, None, None, device=True)
    print("case: compressed KAT T=80", flush=True)
    out = C.convert_vbr_to_compressed(val, rp, cp, indx, bindx, bb, be, "example2", str(tmp_path), 80, device=True)
    val2, indx2, bindx2, bb2, be2, ub, indptr, indices, csr_val = out
    ex = kat["compressed_threshold_80"]
    assert _eq(val2, ex["val2"]) and _eq(indx2, ex["indx2"]) and _eq(ub, ex["ublocks"])
    assert _eq(indptr, ex["indptr"]) and _eq(indices, ex["indices"]) and _eq(csr_val, ex["csr_val"])
    r = fileio.read_vbrc(str(tmp_path / "example2.vbrc"))
    assert _eq(r[8], ex["indptr"]) and _eq(r[9], ex["indices"]) and _eq(r[10], ex["csr_val"])


@pytest.mark.parametrize("seed", range(6))
def test_compress_matches_host(seed):
    """Random routes on random VBRs (empty block rows, all-zero sparse
    blocks, the last row / column with and without entries): the device
    arrays equal the host's, dtypes included."""
    rng = np.random.default_rng(seed)
    vbr, _ = synth.vbr_matrix_gen_seeded(300, 260, 30, 25, 200, seed=seed + 40,
                                         perc_zeros=int(rng.integers(0, 95)))
    v = np.asarray(vbr.val, dtype=np.float64 if seed % 2 else np.float32)
    v[rng.random(len(v)) < rng.uniform(0, 0.95)] = 0.0
    nb = len(vbr.bindx)
    for routes in (rng.integers(0, 2, nb).astype(np.uint8), np.ones(nb, np.uint8), np.zeros(nb, np.uint8)):
        print(f"case: compress seed {seed} sparse blocks {int(routes.sum())}", flush=True)
        a = (v, vbr.rpntr, vbr.cpntr, vbr.indx, vbr.bindx, vbr.bpntrb, vbr.bpntre, routes)
        host = C.compress_with_routes(*a)
        dev = C.compress_with_routes_device(*a)
        for h, d in zip(host, dev):
            assert h.dtype == d.dtype and np.array_equal(h, d)


def test_compress_rejects_bad_vbr():
    from sable_amd._lib import SableError
    v = np.ones(8)
    with pytest.raises(SableError):       # indx step does not match the block size
        C.compress_with_routes_device(v, [0, 2], [0, 2, 4], [0, 3, 8], [0, 1], [0], [2], [1, 1])


def test_sparse_to_vbrc_on_device_kat_and_host():
    """CSC -> .vbrc arrays with the VBR arrays kept on the device between the
    two steps (sable_inspect_compress): the reference's T = 80 known answer,
    and on random matrices (empty block rows, explicit zeros, several
    thresholds) the host pipeline's arrays, dtypes included."""
    import json, os
    kat = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kat_example.json")))
    d = np.array(kat["dense_test_compression"], dtype=float)
    print("case: sparse_to_vbrc KAT T=80", flush=True)
    val2, indx2, bindx, bb, be, ub, indptr, indices, csr_val = C.sparse_to_vbrc_device(
        scipy.sparse.csc_matrix(d), kat["rpntr"], kat["cpntr"], density=80)
    ex = kat["compressed_threshold_80"]
    assert _eq(val2, ex["val2"]) and _eq(indx2, ex["indx2"]) and _eq(ub, ex["ublocks"])
    assert _eq(indptr, ex["indptr"]) and _eq(indices, ex["indices"]) and _eq(csr_val, ex["csr_val"])
    rng = np.random.default_rng(31)
    for t in range(4):
        n, m = int(rng.integers(200, 600)), int(rng.integers(200, 600))
        d = rng.standard_normal((n, m))
        d[rng.random((n, m)) < rng.uniform(0.5, 0.99)] = 0.0
        d[:, :7] = 0.0                            # empty block rows / columns somewhere
        A = scipy.sparse.csc_matrix(d)
        rp = np.unique(np.concatenate([[0, n], rng.integers(1, n, 20)]))
        cp = np.unique(np.concatenate([[0, m], rng.integers(1, m, 18)]))
        for T in (0, 15, 80):
            print(f"case: sparse_to_vbrc random #{t} T={T}", flush=True)
            dev = C.sparse_to_vbrc_device(A, rp, cp, density=T)
            val, indx, bindx, bb, be = C.convert_sparse_to_vbr(A, rp, cp, None, None, allow_empty_rows=True)
            routes = C.threshold_routes(val, indx, T)
            hv2, hix2, hub, hip, hind, hcv = C.compress_with_routes(val, rp, cp, indx, bindx, bb, be, routes)
            host = (hv2, hix2, bindx, bb, be, hub, hip, hind, hcv)
            for h, g in zip(host, dev):
                assert np.asarray(h).dtype == np.asarray(g).dtype and np.array_equal(h, g)
