"""The product inspector (vectorised) against the reference KATs and the
loop oracle: bit-exact VBR conversion, routing, CSR assembly, text formats."""
import os

import numpy as np
import pytest
import scipy.sparse

from oracle import vbr_oracle as O
from sable_amd import convert as C
from sable_amd import fileio, synth
from sable_amd.vbr import VBR, validate_vbr


def _eq(a, b):
    return np.array_equal(np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64))


def test_convert_sparse_to_vbr_kats(kat, tmp_path):
    for dense_key, ex_key in (("dense_docstring", "example_vbr"),
                              ("dense_test_compression", "partition_test_compression")):
        d = np.array(kat[dense_key], dtype=float)
        val, indx, bindx, bb, be = C.convert_sparse_to_vbr(scipy.sparse.csc_matrix(d), kat["rpntr"],
                                                           kat["cpntr"], "ex", str(tmp_path))
        ex = kat[ex_key]
        assert _eq(val, ex["val"]) and _eq(indx, ex["indx"]) and _eq(bindx, ex["bindx"])
        assert _eq(bb, ex["bpntrb"]) and _eq(be, ex["bpntre"])


def test_vbr_text_written_like_reference(kat, golden, tmp_path):
    """The .vbr text of the docstring matrix is byte-identical to the
    reference's tests/example.vbr."""
    d = np.array(kat["dense_docstring"], dtype=float)
    C.convert_sparse_to_vbr(scipy.sparse.csc_matrix(d), kat["rpntr"], kat["cpntr"], "example", str(tmp_path))
    assert open(tmp_path / "example.vbr").read() == open(os.path.join(golden, "example.vbr")).read()


def test_vbrc_text_written_like_reference(golden, tmp_path):
    """Threshold 0 on example.vbr writes the reference's tests/example.vbrc."""
    val, indx, bindx, rp, cp, bb, be = fileio.read_vbr(os.path.join(golden, "example.vbr"))
    C.convert_vbr_to_compressed(val, rp, cp, indx, bindx, bb, be, "example", str(tmp_path), 0)
    assert open(tmp_path / "example.vbrc").read() == open(os.path.join(golden, "example.vbrc")).read()


def test_compressed_kat_threshold_80(kat, tmp_path):
    d = np.array(kat["dense_test_compression"], dtype=float)
    rp, cp = kat["rpntr"], kat["cpntr"]
    val, indx, bindx, bb, be = C.convert_sparse_to_vbr(scipy.sparse.csc_matrix(d), rp, cp, None, None)
    out = C.convert_vbr_to_compressed(val, rp, cp, indx, bindx, bb, be, "example2", str(tmp_path), 80)
    val2, indx2, bindx2, bb2, be2, ub, indptr, indices, csr_val = out
    ex = kat["compressed_threshold_80"]
    assert _eq(val2, ex["val2"]) and _eq(indx2, ex["indx2"]) and _eq(ub, ex["ublocks"])
    assert _eq(indptr, ex["indptr"]) and _eq(indices, ex["indices"]) and _eq(csr_val, ex["csr_val"])
    # test.py:107-120: read back the .vbrc
    r = fileio.read_vbrc(str(tmp_path / "example2.vbrc"))
    assert _eq(r[0], ex["val2"]) and _eq(r[1], ex["indx2"]) and _eq(r[7], ex["ublocks"])
    assert _eq(r[8], ex["indptr"]) and _eq(r[9], ex["indices"]) and _eq(r[10], ex["csr_val"])


def _rand_vbr(seed):
    rng = np.random.default_rng(seed)
    vbr, x = synth.vbr_matrix_gen_seeded(60, 55, 9, 8, 30, seed=seed,
                                         perc_zeros=int(rng.integers(0, 95)))
    v = np.asarray(vbr.val, dtype=np.float64)
    v[rng.random(len(v)) < rng.uniform(0, 0.9)] = 0.0
    return VBR(v, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb, vbr.bpntre)


@pytest.mark.parametrize("seed", range(8))
def test_compressed_matches_loop_oracle(seed):
    vbr = _rand_vbr(seed)
    a = [list(np.asarray(z).tolist()) for z in (vbr.val, vbr.rpntr, vbr.cpntr, vbr.indx, vbr.bindx,
                                                vbr.bpntrb, vbr.bpntre)]
    for T in (0, 10, 25, 50, 75, 99, 100):
        val2, indx2, _, _, _, ub, ip, ind, cv = C.convert_vbr_to_compressed(
            vbr.val, vbr.rpntr, vbr.cpntr, vbr.indx, vbr.bindx, vbr.bpntrb, vbr.bpntre, None, None, T)
        r = O.convert_vbr_to_compressed(*a, T)
        assert _eq(val2, r[0]) and _eq(indx2, r[1]) and _eq(ub, r[2])
        assert _eq(ip, r[3]) and _eq(ind, r[4]) and _eq(cv, r[5])


def test_density_boundary_is_strict():
    """A block whose nonzero percentage equals T exactly stays sparse
    (strict '>' of utils/convert_real_to_vbr.py:52-54), including
    percentages that are not exact in binary (1/3 -> 33.33...)."""
    val = np.array([1, 0, 0, 0, 1, 1, 0, 0, 1, 0, 0], dtype=float)[:9]  # 3x3, 3 nonzeros
    val = np.array([1, 0, 0, 1, 0, 0, 1, 0, 0], dtype=float)
    pct = C.block_density_percent(np.array([3]), np.array([9]))[0]
    assert pct == (3 / 9) * 100
    r = C.threshold_routes(val, [0, 9], pct)
    assert r[0] == C.ROUTE_SPARSE
    r = C.threshold_routes(val, [0, 9], np.nextafter(pct, 0))
    assert r[0] == C.ROUTE_DENSE


def test_nnz_invariant_and_decompress_roundtrip():
    vbr, x, _ = synth.workload_like_c2(3000, seed=1, frac_blocks=0.05)
    for T in (0, 15, 50, 100):
        val2, indx2, bindx, bb, be, ub, ip, ind, cv = C.convert_vbr_to_compressed(
            vbr.val, vbr.rpntr, vbr.cpntr, vbr.indx, vbr.bindx, vbr.bpntrb, vbr.bpntre, None, None, T)
        # sum over routes of nonzeros equals nnz(A) (+ possibly the zero pad entry)
        nz = np.count_nonzero(val2) + np.count_nonzero(cv)
        assert nz == np.count_nonzero(vbr.val)
        val, indx, route = C.decompress_vbrc(val2, indx2, bindx, vbr.rpntr, vbr.cpntr, bb, be, ub, ip, ind, cv)
        assert np.array_equal(indx, vbr.indx)
        dense_rows = np.repeat(route == C.ROUTE_DENSE, np.diff(indx))
        assert np.array_equal(val[dense_rows], np.asarray(vbr.val, dtype=np.float64)[dense_rows])
        assert np.array_equal(val[~dense_rows], np.asarray(vbr.val, dtype=np.float64)[~dense_rows])


def test_convert_sparse_to_vbr_random_against_oracle():
    rng = np.random.default_rng(5)
    for _ in range(5):
        n, m = 23, 19
        d = rng.uniform(-1, 1, (n, m))
        d[rng.random((n, m)) < 0.8] = 0
        A = scipy.sparse.csc_matrix(d)
        # an explicit stored zero counts for block retention
        A = A.tolil()
        A[0, 0] = 0.0
        A = scipy.sparse.csc_matrix(A)
        A[0, 0] = 0.0
        rp = [0, 5, 6, 13, 23]
        cp = [0, 2, 9, 10, 19]
        coo = A.tocoo()
        entries = {(int(i), int(j)): float(v) for i, j, v in zip(coo.row, coo.col, coo.data)}
        try:
            r = O.convert_sparse_to_vbr(entries, (n, m), rp, cp)
        except AssertionError:
            with pytest.raises(AssertionError):
                C.convert_sparse_to_vbr(A, rp, cp, None, None)
            r = O.convert_sparse_to_vbr(entries, (n, m), rp, cp, allow_empty_rows=True)
            out = C.convert_sparse_to_vbr(A, rp, cp, None, None, allow_empty_rows=True)
        else:
            out = C.convert_sparse_to_vbr(A, rp, cp, None, None)
        for a, b in zip(out, r):
            assert _eq(a, b)


def test_vbr_to_dense_matches_oracle_and_mtx(kat, golden, tmp_path):
    ex = kat["example_vbr"]
    M = C.vbr_to_dense(ex["val"], ex["indx"], ex["bindx"], kat["rpntr"], kat["cpntr"], ex["bpntrb"], ex["bpntre"])
    assert np.array_equal(M, np.array(kat["dense_docstring"], dtype=float))
    import shutil
    shutil.copy(os.path.join(golden, "example.vbr"), tmp_path / "example.vbr")
    C.vbr_to_mtx("example.vbr", str(tmp_path), str(tmp_path))
    assert open(tmp_path / "example.mtx").read() == open(os.path.join(golden, "example-canon.mtx")).read()


def test_vbr_to_csc_matches_oracle_dense(kat):
    """The sparse expansion equals the oracle's dense expansion (the
    reference docstring KAT and random matrices with empty block rows and
    explicit zeros inside blocks)."""
    ex = kat["example_vbr"]
    args = (ex["val"], ex["indx"], ex["bindx"], kat["rpntr"], kat["cpntr"], ex["bpntrb"], ex["bpntre"])
    assert np.array_equal(C.vbr_to_csc(*args).toarray(), np.array(kat["dense_docstring"], dtype=float))
    for seed in range(4):
        vbr, _, _ = synth.mixed_vbr(300, 12, 5, 40, 0.2, 0.5, (0.8, 1.0), (0.01, 0.3), seed)
        a = (vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb, vbr.bpntre)
        D = np.asarray(O.vbr_to_dense(*a), dtype=np.float64)
        S = C.vbr_to_csc(*a)
        assert np.array_equal(S.toarray(), D) and S.nnz == np.count_nonzero(D)


def test_validate_rejects_malformed():
    good = dict(val=[1.0] * 4, indx=[0, 4], bindx=[0], rpntr=[0, 2], cpntr=[0, 2], bpntrb=[0], bpntre=[1])
    validate_vbr(**good)
    for bad in (dict(good, indx=[0, 3]), dict(good, rpntr=[0, 2, 2]), dict(good, bpntrb=[-1]),
                dict(good, bindx=[3]), dict(good, val=[1.0] * 3), dict(good, cpntr=[1, 2])):
        with pytest.raises(ValueError):
            validate_vbr(**bad)
    with pytest.raises(ValueError):   # unsorted bindx
        validate_vbr([1.0] * 8, [0, 4, 8], [1, 0], [0, 2], [0, 2, 4], [0], [2])


def test_read_vbrc_golden(golden):
    r = fileio.read_vbrc(os.path.join(golden, "example.vbrc"))
    v = fileio.read_vbr(os.path.join(golden, "example.vbr"))
    assert _eq(r[0], v[0]) and _eq(r[1], v[1]) and len(r[7]) == 0 and len(r[8]) == 0


def test_npz_roundtrip(tmp_path):
    vbr, x = synth.workload_c1()
    fileio.save_vbr_npz(tmp_path / "c1.npz", vbr)
    v2 = fileio.load_vbr_npz(tmp_path / "c1.npz")
    for k in ("val", "indx", "bindx", "rpntr", "cpntr", "bpntrb", "bpntre"):
        assert np.array_equal(getattr(vbr, k), getattr(v2, k))


def test_workload_generators_are_seeded_and_bounded():
    v1, x1 = synth.workload_c1()
    v2, x2 = synth.workload_c1()
    assert np.array_equal(v1.val, v2.val) and np.array_equal(x1, x2)
    assert len(v1.bindx) == 40 and np.all(np.asarray(v1.val) != 0)
    validate_vbr(v1.val, v1.indx, v1.bindx, v1.rpntr, v1.cpntr, v1.bpntrb, v1.bpntre)
    vbr, x, cls = synth.mixed_vbr(20000, 200, 20, 200, 0.01, 0.5, (0.8, 1.0), (0.01, 0.2), seed=2)
    sizes = np.diff(vbr.rpntr)
    assert sizes.min() >= 20 and sizes.max() <= 200 and sizes.sum() == 20000
    counts, tot = C.block_nonzero_counts(vbr.val, vbr.indx)
    dens = counts / tot
    assert np.all(dens[cls] >= 0.79) and np.all(dens[~cls] <= 0.21)
