"""Pins the oracle's CSR part against the reference's own compiled CSR
program (src/csr-spmv.cpp, built from the reference sources into
oracle/_ref/ by `make -C oracle ref`).  Skipped where that build is absent
(the reference sources are not on every host).

The program reads a MatrixMarket file and an x vector ("v0,v1,...") and
prints y = A x in FP64, 6 significant digits (std::cout default).  The
oracle's generated-program restatement with every block sparse-routed
(threshold 100: the whole matrix goes through the CSR part,
src/codegen.py:419-420) must agree to that printed precision on random x.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import vbr_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "csr-spmv")

pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="reference csr-spmv not built here")


def _write_mtx(path, nrows, ncols, rows, cols, vals):
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{nrows} {ncols} {len(vals)}\n")
        for r, c, v in zip(rows, cols, vals):
            f.write(f"{r + 1} {c + 1} {v!r}\n")


def _run_ref(tmp_path, nrows, ncols, rows, cols, vals, x):
    mtx = tmp_path / "a.mtx"
    vec = tmp_path / "x.vector"
    _write_mtx(mtx, nrows, ncols, rows, cols, vals)
    vec.write_text(",".join(repr(float(v)) for v in x) + ",")
    out = subprocess.run([REF, str(mtx), "1", "3", str(vec)], capture_output=True, text=True, check=True,
                         timeout=60).stdout.split("\n")
    assert out[0].startswith("Time:")
    return np.array([float(v) for v in out[1:] if v.strip()])


def _random_vbr(rng, nbr=12, nbc=10, p=0.4, maxsz=9, zero_frac=0.5):
    rs = rng.integers(1, maxsz + 1, nbr)
    cs = rng.integers(1, maxsz + 1, nbc)
    rp = [0] + [int(v) for v in np.cumsum(rs)]
    cp = [0] + [int(v) for v in np.cumsum(cs)]
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
    return val, indx, bindx, rp, cp, bb, be


@pytest.mark.parametrize("seed", range(4))
def test_oracle_csr_part_matches_reference_program(tmp_path, seed):
    rng = np.random.default_rng(100 + seed)
    val, indx, bindx, rp, cp, bb, be = _random_vbr(rng)
    x = rng.uniform(-1, 1, cp[-1])
    val2, indx2, ub, ip, ind, cv = O.convert_vbr_to_compressed(val, rp, cp, indx, bindx, bb, be, 100)
    assert len(val2) == 0 and len(ub) == len(bindx)        # everything in the CSR part
    y_oracle = np.array(O.spmv_generated_c_f64(val2, indx2, bindx, rp, cp, bb, be, ub, ip, ind, cv, x.tolist()))
    rows = np.repeat(np.arange(rp[-1]), np.diff(ip))
    y_ref = _run_ref(tmp_path, rp[-1], cp[-1], rows, ind, cv, x)
    assert y_ref.shape == y_oracle.shape
    # 6 significant digits printed
    assert np.allclose(y_ref, y_oracle, rtol=1e-5, atol=1e-12)
    # and both are the closed form A x
    ref = O.spmv_f64_blocks(val, indx, bindx, rp, cp, bb, be, x.tolist())
    assert np.allclose(y_oracle, ref, rtol=1e-12, atol=1e-12)


def test_reference_program_on_canon_example(tmp_path, golden):
    """tests/example-canon.mtx with x = 1 gives tests/output_spmv_canon.txt
    (the reference's own check, test.py:249-256)."""
    lines = open(os.path.join(golden, "example-canon.mtx")).read().split("\n")
    body = [l for l in lines if l and not l.startswith("%")]
    nr, nc, nz = map(int, body[0].split())
    ent = [l.split() for l in body[1:1 + nz]]
    rows = [int(e[0]) - 1 for e in ent]
    cols = [int(e[1]) - 1 for e in ent]
    vals = [float(e[2]) for e in ent]
    y = _run_ref(tmp_path, nr, nc, rows, cols, vals, np.ones(nc))
    canon = np.array([float(v) for v in open(os.path.join(golden, "output_spmv_canon.txt")).read().split()])
    assert np.allclose(y, canon, rtol=1e-6, atol=1e-6)
