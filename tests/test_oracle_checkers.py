"""Pins the FP64 checkers that the GPU parity tests and the bench call
(oracle/c_oracle.py: spmv_f64_dense_blocks, spmm_f64_dense_blocks,
scaled_error, the C restatements CProgram and spmm) against the pinned
oracle (oracle/vbr_oracle.py: vbr_to_dense is fixed by the reference's
tests/example-canon.mtx and test.py:35-52; convert_vbr_to_compressed by the
test.py:64-105 KATs; spmv_generated_c_f64 / spmm_generated_c by the canon
outputs).  Random x / X (the reference fixtures only hold x = 1), partitions
with rpntr != cpntr, unit-height and unit-width blocks, and block rows marked
-1.  Planted mistakes (transposed X, row-major blocks, x indexed through the
row partition) must fail the same comparisons."""
import os

import numpy as np
import pytest

from oracle import c_oracle as CO
from oracle import vbr_oracle as O


def _random_vbr(seed, nbr=7, nbc=6, p=0.45, sizes=(1, 2, 3, 5, 7), zero_frac=0.3, empty_row=True):
    """VBR arrays with independent row and column partitions (so rpntr !=
    cpntr), sizes including 1 (unit blocks), at least one -1 block row."""
    rng = np.random.default_rng(seed)
    rs = rng.choice(sizes, nbr)
    cs = rng.choice(sizes, nbc)
    rs[0], cs[-1] = 1, 1                      # a height-1 row and a width-1 column always exist
    rp = [0] + np.cumsum(rs).tolist()
    cp = [0] + np.cumsum(cs).tolist()
    val, indx, bindx, bb, be = [], [0], [], [], []
    for a in range(nbr):
        cols = [b for b in range(nbc) if rng.random() < p]
        if a == 0:
            cols = sorted(set(cols) | {nbc - 1})      # height-1 x width-1 block
        if (empty_row and a == nbr // 2) or not cols:
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


def _dense(args):
    return np.array(O.vbr_to_dense(*args), dtype=np.float64)


def test_dense_expansion_is_pinned(kat, golden):
    """The comparison base (vbr_oracle.vbr_to_dense) reproduces the
    reference's tests/example-canon.mtx entry for entry."""
    ex = kat["example_vbr"]
    M = _dense((ex["val"], ex["indx"], ex["bindx"], kat["rpntr"], kat["cpntr"], ex["bpntrb"], ex["bpntre"]))
    n = 0
    with open(os.path.join(golden, "example-canon.mtx")) as f:
        lines = f.read().split("\n")
    for l in lines[2:]:
        if l.strip():
            i, j, v = l.split()
            assert M[int(i) - 1, int(j) - 1] == float(v)
            n += 1
    assert n == int(lines[1].split()[2]) == np.count_nonzero(M)


@pytest.mark.parametrize("seed", range(8))
def test_spmv_checker_matches_pinned_dense(seed):
    args = _random_vbr(seed)
    A = _dense(args)
    x = np.random.default_rng(100 + seed).uniform(-1, 1, A.shape[1])
    y, ay = CO.spmv_f64_dense_blocks(*args, x)
    assert y.shape == (A.shape[0],)
    np.testing.assert_allclose(y, A @ x, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(ay, np.abs(A) @ np.abs(x), rtol=1e-13, atol=1e-13)
    # and the correctly rounded per-row sums of the oracle
    np.testing.assert_allclose(y, O.spmv_f64_blocks(*args, x.tolist()), rtol=1e-13, atol=1e-13)
    # rows of the -1 block row are exactly zero
    a = len(args[3]) // 2 - 1
    if args[5][a] == -1:
        assert not np.any(y[args[3][a]:args[3][a + 1]])


@pytest.mark.parametrize("n_rhs", [1, 16, 128, 512])
def test_spmm_checker_matches_pinned_dense(n_rhs):
    for seed in range(3):
        args = _random_vbr(10 + seed)
        A = _dense(args)
        X = np.random.default_rng(200 + seed).uniform(-1, 1, (A.shape[1], n_rhs))
        Y, AY = CO.spmm_f64_dense_blocks(*args, X)
        assert Y.shape == (A.shape[0], n_rhs)
        np.testing.assert_allclose(Y, A @ X, rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(AY, np.abs(A) @ np.abs(X), rtol=1e-13, atol=1e-13)


def test_spmm_checker_on_the_example_with_ones(kat, golden):
    """X = 1 with 512 columns reproduces tests/output_spmm_canon.txt."""
    import gzip
    ex = kat["example_vbr"]
    args = (ex["val"], ex["indx"], ex["bindx"], kat["rpntr"], kat["cpntr"], ex["bpntrb"], ex["bpntre"])
    Y, _ = CO.spmm_f64_dense_blocks(*args, np.ones((11, 512)))
    with gzip.open(os.path.join(golden, "output_spmm_canon.txt.gz"), "rt") as f:
        canon = np.array([float(v) for v in f.read().split()])
    assert np.array_equal(np.round(Y.ravel(), 6), canon)


@pytest.mark.parametrize("n_rhs", [1, 3, 16])
def test_spmm_generated_c_random_x_closed_form(n_rhs):
    """vbr_oracle.spmm_generated_c (the restated reference SpMM program) equals
    A @ X for random X, routed (density 50) and unrouted (density 0)."""
    for seed in range(3):
        args = _random_vbr(20 + seed)
        A = _dense(args)
        X = np.random.default_rng(300 + seed).uniform(-1, 1, (A.shape[1], n_rhs))
        for density in (0, 50):
            Y = np.array(O.spmm_generated_c(*args, density, X.ravel().tolist(), n_rhs)).reshape(-1, n_rhs)
            np.testing.assert_allclose(Y, A @ X, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("seed", range(6))
def test_c_program_matches_pinned_oracle(seed):
    """CProgram (the C restatement of the generated SpMV program) against the
    pure-Python restatement on the compressed arrays the KAT-pinned
    convert_vbr_to_compressed produces, at several thresholds, and against
    A x; single- and multi-threaded."""
    args = _random_vbr(30 + seed)
    val, indx, bindx, rp, cp, bb, be = args
    A = _dense(args)
    x = np.random.default_rng(400 + seed).uniform(-1, 1, A.shape[1])
    for T in (0, 30, 60, 100):
        val2, indx2, ub, ip, ind, cv = O.convert_vbr_to_compressed(val, rp, cp, indx, bindx, bb, be, T)
        ref = O.spmv_generated_c_f64(val2, indx2, bindx, rp, cp, bb, be, ub, ip, ind, cv, x.tolist())
        prog = CO.CProgram(val2, indx2, bindx, rp, cp, bb, be, ub, ip, ind, cv)
        y = prog.run(x)
        np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(y, A @ x, rtol=1e-12, atol=1e-12)
        prog.set_threads(3)
        np.testing.assert_allclose(prog.run(x), ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("n_rhs", [1, 4, 16, 128])
def test_c_spmm_matches_pinned_oracle(n_rhs):
    """oracle_spmm_f32 (FP32 restatement of gen_single_threaded_spmm) against
    the pure-Python restatement and A @ X, with density routing."""
    for seed in range(3):
        args = _random_vbr(40 + seed)
        val, indx, bindx, rp, cp, bb, be = args
        A = _dense(args)
        X = np.random.default_rng(500 + seed).uniform(-1, 1, (A.shape[1], n_rhs)).astype(np.float32)
        for density in (0, 50):
            from sable_amd.convert import threshold_routes
            route = (threshold_routes(np.asarray(val, np.float32), indx, density) if density
                     else np.zeros(len(bindx), np.uint8))
            Y = CO.spmm(np.asarray(val, np.float32), indx, bindx, rp, cp, bb, be, route, X, n_rhs)
            ref = np.array(O.spmm_generated_c(list(np.asarray(val, np.float32).astype(float)), indx, bindx, rp, cp,
                                              bb, be, density, X.astype(float).ravel().tolist(), n_rhs))
            scale = np.abs(A) @ np.abs(X.astype(np.float64))
            assert np.all(np.abs(Y.ravel() - ref) <= 1e-6 * scale.ravel() + 1e-30)
            assert np.all(np.abs(Y - A @ X) <= 1e-6 * scale + 1e-30)


def test_scaled_error_closed_forms():
    y64 = np.array([1.0, -2.0, 0.0, 4.0])
    ay = np.array([2.0, 4.0, 0.0, 8.0])
    y = y64.copy()
    y[1] += 1e-3
    e = CO.scaled_error(y, y64, ay)
    assert e["componentwise"] == pytest.approx(1e-3 / 4.0, rel=1e-12)
    assert e["normwise"] == pytest.approx(1e-3 / 4.0, rel=1e-12)
    assert e["elementwise"] == pytest.approx(1e-3 / 2.0, rel=1e-12)
    assert CO.scaled_error(y64, y64, ay)["componentwise"] == 0.0
    # a nonzero where |A||x| is zero cannot be scaled: infinite error
    y = y64.copy()
    y[2] = 1e-30
    assert CO.scaled_error(y, y64, ay)["componentwise"] == float("inf")


# ---------------------------------------------------------------------------
# planted mistakes: each plausible slip in a dense-blocks checker fails the
# comparisons above


def _planted_spmm(args, X, mistake):
    val, indx, bindx, rp, cp, bb, be = args
    v = np.asarray(val, dtype=np.float64)
    Y = np.zeros((rp[-1], X.shape[1]))
    Xm = X
    if mistake == "transposed_x":
        Xm = np.ravel(X).reshape(X.shape[1], X.shape[0]).T      # row-major buffer read column-major
    k = 0
    for a in range(len(rp) - 1):
        if bb[a] == -1:
            continue
        h = rp[a + 1] - rp[a]
        for k in range(bb[a], be[a]):
            b = bindx[k]
            w = cp[b + 1] - cp[b]
            if mistake == "row_major_block":
                blk = v[indx[k]:indx[k + 1]].reshape(h, w)
            else:
                blk = v[indx[k]:indx[k + 1]].reshape(w, h).T
            if mistake == "x_by_row_partition":
                c0 = rp[b] if b < len(rp) - 1 else cp[b]
                xs = Xm[c0:c0 + w] if c0 + w <= Xm.shape[0] else Xm[cp[b]:cp[b + 1]]
            else:
                xs = Xm[cp[b]:cp[b + 1]]
            Y[rp[a]:rp[a + 1]] += blk @ xs
    return Y


@pytest.mark.parametrize("mistake", ["transposed_x", "row_major_block", "x_by_row_partition"])
def test_planted_checker_mistakes_fail(mistake):
    caught = False
    for seed in range(4):
        args = _random_vbr(50 + seed)
        A = _dense(args)
        n = 16 if mistake == "transposed_x" else 1
        X =np.random.default_rng(600 + seed).uniform(-1, 1, (A.shape[1], n))
        good = _planted_spmm(args, X, None)
        np.testing.assert_allclose(good, A @ X, rtol=1e-13, atol=1e-13)   # the harness itself is right
        bad = _planted_spmm(args, X, mistake)
        caught |= not np.allclose(bad, A @ X, rtol=1e-9, atol=1e-9)
    assert caught
