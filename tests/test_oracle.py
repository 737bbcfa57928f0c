"""Pins the CPU oracle (oracle/vbr_oracle.py) before anything is checked
against it: reference KATs (test.py), the canon outputs shipped with the
reference tests, closed forms, and planted mistakes that the pins catch."""
import gzip
import os

import numpy as np
import pytest

from oracle import vbr_oracle as O


def _entries(dense):
    return {(i, j): float(v) for (i, j), v in np.ndenumerate(np.asarray(dense)) if v != 0}


def _canon(path):
    with open(path) as f:
        return [float(l) for l in f.read().split()]


def test_convert_sparse_to_vbr_docstring_kat(kat):
    """utils/convert_real_to_vbr.py:105-131 docstring and tests/example.vbr."""
    d = np.array(kat["dense_docstring"])
    val, indx, bindx, bb, be = O.convert_sparse_to_vbr(_entries(d), d.shape, kat["rpntr"], kat["cpntr"])
    ex = kat["example_vbr"]
    assert val == ex["val"] and indx == ex["indx"] and bindx == ex["bindx"]
    assert bb == ex["bpntrb"] and be == ex["bpntre"]


def test_convert_sparse_to_vbr_partition_kat(kat):
    """test.py:122-142."""
    d = np.array(kat["dense_test_compression"])
    val, indx, bindx, bb, be = O.convert_sparse_to_vbr(_entries(d), d.shape, kat["rpntr"], kat["cpntr"])
    ex = kat["partition_test_compression"]
    assert val == [float(v) for v in ex["val"]]
    assert indx == ex["indx"] and bindx == ex["bindx"] and bb == ex["bpntrb"] and be == ex["bpntre"]


def test_convert_sparse_to_vbr_empty_row_asserts():
    d = np.zeros((4, 4))
    d[0, 0] = 1.0
    with pytest.raises(AssertionError):
        O.convert_sparse_to_vbr(_entries(d), d.shape, [0, 2, 4], [0, 2, 4])
    val, indx, bindx, bb, be = O.convert_sparse_to_vbr(_entries(d), d.shape, [0, 2, 4], [0, 2, 4],
                                                       allow_empty_rows=True)
    assert bb == [0, -1] and be == [1, -1]


def test_compressed_threshold_80_kat(kat):
    """test.py:80-105: val2, indx2, ublocks and the CSR of the sparse part."""
    d = np.array(kat["dense_test_compression"])
    rp, cp = kat["rpntr"], kat["cpntr"]
    val, indx, bindx, bb, be = O.convert_sparse_to_vbr(_entries(d), d.shape, rp, cp)
    val2, indx2, ub, indptr, indices, csr_val = O.convert_vbr_to_compressed(val, rp, cp, indx, bindx, bb, be, 80)
    ex = kat["compressed_threshold_80"]
    assert val2 == ex["val2"] and indx2 == ex["indx2"] and ub == ex["ublocks"]
    assert indptr == ex["indptr"] and indices == ex["indices"] and csr_val == ex["csr_val"]


def test_compressed_threshold_0_is_identity(kat):
    """test.py:64-77: threshold 0 keeps every block dense, empty CSR."""
    ex = kat["example_vbr"]
    val2, indx2, ub, indptr, indices, csr_val = O.convert_vbr_to_compressed(
        ex["val"], kat["rpntr"], kat["cpntr"], ex["indx"], ex["bindx"], ex["bpntrb"], ex["bpntre"], 0)
    assert val2 == ex["val"] and indx2 == ex["indx"]
    assert ub == [] and indptr == [] and indices == [] and csr_val == []


def test_compressed_padding_entry():
    """utils/convert_real_to_vbr.py:70-74: explicit (n-1, n-1, 0.0) entry when
    the last row or column is absent from the sparse part."""
    d = np.zeros((4, 4))
    d[0, 0] = 1.0
    d[3, 3] = 2.0
    d[2, 2] = 5.0
    d[3, 2] = 5.0
    d[2, 3] = 5.0
    rp = cp = [0, 2, 4]
    val, indx, bindx, bb, be = O.convert_sparse_to_vbr(_entries(d), d.shape, rp, cp)
    # block (0,0) density 25% -> sparse at T=50; block (1,1) 100% -> dense
    val2, indx2, ub, indptr, indices, csr_val = O.convert_vbr_to_compressed(val, rp, cp, indx, bindx, bb, be, 50)
    assert ub == [0]
    assert indptr == [0, 1, 1, 1, 2] and indices == [0, 3] and csr_val == [1.0, 0.0]


def test_vbr_to_dense_kat(kat, golden):
    """test.py:35-52 and tests/example-canon.mtx."""
    ex = kat["example_vbr"]
    M = O.vbr_to_dense(ex["val"], ex["indx"], ex["bindx"], kat["rpntr"], kat["cpntr"], ex["bpntrb"], ex["bpntre"])
    assert np.array_equal(np.array(M), np.array(kat["dense_docstring"], dtype=float))
    lines = open(os.path.join(golden, "example-canon.mtx")).read().split("\n")
    assert lines[1].split() == ["11", "11", "47"]
    for l in lines[2:]:
        if l.strip():
            i, j, v = l.split()
            assert M[int(i) - 1][int(j) - 1] == float(v)


def test_spmv_program_canon(kat, golden):
    """tests/output_spmv_canon.txt: the generated program with x = 1."""
    ex = kat["example_vbr"]
    rp, cp = kat["rpntr"], kat["cpntr"]
    val2, indx2, ub, ip, ind, cv = O.convert_vbr_to_compressed(ex["val"], rp, cp, ex["indx"], ex["bindx"],
                                                               ex["bpntrb"], ex["bpntre"], 0)
    y = O.spmv_generated_c_f64(val2, indx2, ex["bindx"], rp, cp, ex["bpntrb"], ex["bpntre"], ub, ip, ind, cv,
                               [1.0] * 11)
    assert [round(v, 6) for v in y] == _canon(os.path.join(golden, "output_spmv_canon.txt"))


def test_spmv_program_canon_sparse(kat, golden):
    """tests/output_spmv_canon_sparse.txt: the routed program (threshold 80,
    four blocks through the CSR part) with x = 1."""
    d = np.array(kat["dense_test_compression"])
    rp, cp = kat["rpntr"], kat["cpntr"]
    val, indx, bindx, bb, be = O.convert_sparse_to_vbr(_entries(d), d.shape, rp, cp)
    val2, indx2, ub, ip, ind, cv = O.convert_vbr_to_compressed(val, rp, cp, indx, bindx, bb, be, 80)
    y = O.spmv_generated_c_f64(val2, indx2, bindx, rp, cp, bb, be, ub, ip, ind, cv, [1.0] * 11)
    assert [round(v, 6) for v in y] == _canon(os.path.join(golden, "output_spmv_canon_sparse.txt"))


def test_spmm_program_canon(kat, golden):
    """tests/output_spmm_canon.txt: Y = A 1 with 512 right-hand sides."""
    ex = kat["example_vbr"]
    with gzip.open(os.path.join(golden, "output_spmm_canon.txt.gz"), "rt") as f:
        canon = [float(v) for v in f.read().split()]
    for density in (0, 50):
        Y = O.spmm_generated_c(ex["val"], ex["indx"], ex["bindx"], kat["rpntr"], kat["cpntr"], ex["bpntrb"],
                               ex["bpntre"], density, [1.0] * (11 * 512), 512)
        assert [round(v, 6) for v in Y] == canon


def _random_vbr(rng, nbr=6, nbc=5, p=0.45, maxsz=5, zero_frac=0.3):
    rs = rng.integers(1, maxsz + 1, nbr)
    cs = rng.integers(1, maxsz + 1, nbc)
    rp = [0] + list(np.cumsum(rs))
    cp = [0] + list(np.cumsum(cs))
    val, indx, bindx, bb, be = [], [0], [], [], []
    for a in range(nbr):
        cols = [b for b in range(nbc) if rng.random() < p]
        if not cols:
            bb.append(-1)
            be.append(-1)
            continue
        bb.append(len(bindx))
        for b in cols:
            n = int(rs[a] * cs[b])
            v = rng.uniform(-1, 1, n)
            v[rng.random(n) < zero_frac] = 0.0
            val.extend(v.tolist())
            indx.append(indx[-1] + n)
            bindx.append(b)
        be.append(len(bindx))
    return val, indx, bindx, [int(v) for v in rp], [int(v) for v in cp], bb, be


@pytest.mark.parametrize("seed", range(6))
def test_spmv_program_matches_closed_form(seed):
    """Random x (the reference fixtures only pin x = 1): the routed program
    equals the FP64 closed form A x for every threshold."""
    rng = np.random.default_rng(seed)
    val, indx, bindx, rp, cp, bb, be = _random_vbr(rng)
    x = rng.uniform(-1, 1, cp[-1]).tolist()
    ref = O.spmv_f64_blocks(val, indx, bindx, rp, cp, bb, be, x)
    for T in (0, 30, 60, 100):
        val2, indx2, ub, ip, ind, cv = O.convert_vbr_to_compressed(val, rp, cp, indx, bindx, bb, be, T)
        y = O.spmv_generated_c_f64(val2, indx2, bindx, rp, cp, bb, be, ub, ip, ind, cv, x)
        assert np.allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_planted_x_index_mistake_is_caught(kat, golden):
    """A plausible mistake (indexing x by the row instead of the column)
    passes the x = 1 canon but fails the random-x closed form."""
    def wrong_spmv(val, indx, bindx, rp, cp, bb, be, x):
        nrows = rp[-1]
        y = [0.0] * nrows
        count = 0
        for a, b in O._blocks_in_ordinal_order(rp, cp, bindx, bb, be):
            k = 0
            for j in range(cp[b], cp[b + 1]):
                for i in range(rp[a], rp[a + 1]):
                    y[i] += val[indx[count] + k] * x[i % len(x)]
                    k += 1
            count += 1
        return y
    ex = kat["example_vbr"]
    args = (ex["val"], ex["indx"], ex["bindx"], kat["rpntr"], kat["cpntr"], ex["bpntrb"], ex["bpntre"])
    assert [round(v, 6) for v in wrong_spmv(*args, [1.0] * 11)] == _canon(os.path.join(golden, "output_spmv_canon.txt"))
    x = np.random.default_rng(0).uniform(-1, 1, 11).tolist()
    assert not np.allclose(wrong_spmv(*args, x), O.spmv_f64_blocks(*args, x))


def test_planted_row_major_block_mistake_is_caught(kat):
    """Reading blocks row-major instead of column-major fails the KAT."""
    ex = kat["example_vbr"]
    rp, cp = kat["rpntr"], kat["cpntr"]
    M = np.zeros((11, 11))
    count = 0
    for a, b in O._blocks_in_ordinal_order(rp, cp, ex["bindx"], ex["bpntrb"], ex["bpntre"]):
        h, w = rp[a + 1] - rp[a], cp[b + 1] - cp[b]
        blk = np.array(ex["val"][ex["indx"][count]:ex["indx"][count + 1]]).reshape(h, w)
        M[rp[a]:rp[a + 1], cp[b]:cp[b + 1]] = blk
        count += 1
    assert not np.array_equal(M, np.array(kat["dense_docstring"], dtype=float))


def test_similarity2_closed_forms():
    a = [1, 1, 0, 0, 1]
    assert O.similarity2(a, a) == 1.0                       # identical lines
    assert O.similarity2([1, 0, 0, 0], [0, 1, 0, 0]) == 1.0  # one-step shift
    assert O.similarity2([1, 0, 0, 0], [0, 0, 1, 0]) == 0.0  # two-step shift
    assert O.similarity2([1, 1, 1, 1], [1, 0, 0, 0]) == 0.25
    assert np.isnan(O.similarity2([0, 0], [0, 0]))


def test_cut_indices2_hand_example():
    """A 5x6 0/1 pattern whose cut points follow by hand from the rule.
    Columns: 0,1 identical (no cut); 2 = 1 shifted down one row (similarity
    1, no cut); 3 overlaps 2 under no shift (cut at 3); 4 empty so 3|4 has
    similarity 0 (cut at 4); 4,5 empty run, cut after it at 6 = end.
    Rows: r0/r1 = 2/3, r1/r2 = 1/3 (no cuts); r3 empty so r2|r3 cuts at 3;
    the empty run r3 ends before r4 (cut at 4); then the end 5 is added."""
    P = np.array([[1, 1, 0, 0, 0, 0],
                  [1, 1, 1, 0, 0, 0],
                  [0, 0, 1, 0, 0, 0],
                  [0, 0, 0, 0, 0, 0],
                  [0, 0, 0, 1, 0, 0]], dtype=float)
    cols, rows = O.cut_indices2(P.tolist(), (P != 0).tolist(), 0.2)
    assert cols == [0, 3, 4, 6]
    assert rows == [0, 3, 4, 5]
