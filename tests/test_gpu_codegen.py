"""The drop-in generate -> compile -> run flow on the GPU (marked gpu):
vbr_spmv_codegen / vbr_spmm_codegen (src/codegen.py:1438-1458), then the
run step and the program's stdout (src/codegen.py:440-447, :927-934),
checked line for line against the reference's canon outputs (x = 1, as the
reference tests write it, test.py:35-52) and, for random x, against the FP64
checkers.  The CLI (python -m sable_amd.codegen run) is called in-process.

Bars: canon lines equal after the reference's float comparison (test.py:19-33);
line 0 parses with extract_mul_nums into `bench` timings; random inputs
within 1e-5 componentwise of FP64."""
import gzip
import os
import shutil
import sys

import numpy as np
import pytest

from oracle import c_oracle
from sable_amd import codegen, convert, synth
from sable_amd.fileio import write_dense_matrix, write_dense_vector

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def rt():
    from sable_amd import runtime
    runtime.init()
    return runtime


def _stdout(capsys, path):
    """The program's stdout through the CLI entry (earlier output dropped)."""
    capsys.readouterr()
    assert codegen.main(["run", path]) == 0
    return capsys.readouterr().out.split("\n")


def _canon_floats(path, gz=False):
    op = gzip.open if gz else open
    with op(path, "rt") as f:
        return [float(v) for v in f.read().split()]


def test_spmv_flow_example_canon(rt, golden, tmp_path, capsys):
    """test.py run_spmv: vbr_spmv_codegen on tests/example.vbrc, x = 1,
    stdout = output_spmv_canon.txt; at 1 and 4 threads; bench = 5 timings."""
    shutil.copy(os.path.join(golden, "example.vbrc"), tmp_path / "example.vbrc")
    write_dense_vector(1.0, 11)
    canon = _canon_floats(os.path.join(golden, "output_spmv_canon.txt"))
    for threads in (1, 4):
        print(f"case: codegen spmv example threads={threads}", file=sys.stderr, flush=True)
        d = str(tmp_path / f"t{threads}")
        codegen.vbr_spmv_codegen("example", d, str(tmp_path), threads=threads, bench=5)
        lines = _stdout(capsys, os.path.join(d, "example"))
        assert lines[0].startswith("example = ")
        ts = codegen.extract_mul_nums(lines[0])
        assert len(ts) == 5 and all(int(t) > 0 for t in ts)
        body = [l for l in lines[1:] if l.strip()]
        assert [float(v) for v in body] == canon


def test_spmv_flow_routed_canon(rt, kat, golden, tmp_path, capsys):
    """test.py run_spmv_unroll: the threshold-80 .vbrc (CSR part, the MKL
    boundary), x = 1, stdout = output_spmv_canon_sparse.txt."""
    import scipy.sparse as sp
    d = np.array(kat["dense_test_compression"], dtype=np.float64)
    rp, cp = kat["rpntr"], kat["cpntr"]
    val, indx, bindx, bb, be = convert.convert_sparse_to_vbr(sp.csc_matrix(d), rp, cp, "example2", str(tmp_path))
    convert.convert_vbr_to_compressed(val, rp, cp, indx, bindx, bb, be, "example2", str(tmp_path), density=80)
    write_dense_vector(1.0, 11)
    print("case: codegen spmv example2 threshold 80", file=sys.stderr, flush=True)
    codegen.vbr_spmv_codegen("example2", str(tmp_path), str(tmp_path), threads=1)
    lines = _stdout(capsys, str(tmp_path / "example2"))
    body = [float(l) for l in lines[1:] if l.strip()]
    assert body == _canon_floats(os.path.join(golden, "output_spmv_canon_sparse.txt"))


@pytest.mark.parametrize("precision,threads,n_rhs", [("fp32", 1, 512), ("fp32", 4, 512), ("bf16x6", 1, 256),
                                                     ("bf16x6", 4, 256)])
def test_spmm_flow_example_canon(rt, golden, tmp_path, capsys, precision, threads, n_rhs):
    """test.py run_spmm: vbr_spmm_codegen (density 0) on tests/example.vbr
    with X = 1: stdout = output_spmm_canon.txt (512 right-hand sides; the
    MFMA route takes at most 256, whose Y is the canon's first 256 columns)."""
    shutil.copy(os.path.join(golden, "example.vbr"), tmp_path / "example.vbr")
    write_dense_matrix(1.0, 11, n_rhs)
    print(f"case: codegen spmm example {precision} threads={threads} N={n_rhs}", file=sys.stderr, flush=True)
    codegen.vbr_spmm_codegen("example", 0, str(tmp_path), str(tmp_path), threads=threads, n_rhs=n_rhs,
                             precision=precision)
    lines = _stdout(capsys, str(tmp_path / "example"))
    assert lines[0].startswith("example = ") and int(lines[0].split("=")[1]) > 0
    body = [float(l) for l in lines[1:] if l.strip()]
    canon = np.array(_canon_floats(os.path.join(golden, "output_spmm_canon.txt.gz"), gz=True)).reshape(11, 512)
    assert body == canon[:, :n_rhs].ravel().tolist()


def test_spmv_flow_random_x_c1(rt, tmp_path):
    """The flow on the 1000 x 1000 workload (40 dense blocks) at two
    thresholds (0: every block dense; 100: every block in the CSR part) with a
    random x passed to run_program."""
    vbr, x = synth.workload_c1()
    y64, ay = c_oracle.spmv_f64_dense_blocks(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb,
                                             vbr.bpntre, x)
    for T in (0, 100):
        print(f"case: codegen spmv c1 T={T}", file=sys.stderr, flush=True)
        convert.convert_vbr_to_compressed(vbr.val, vbr.rpntr, vbr.cpntr, vbr.indx, vbr.bindx, vbr.bpntrb,
                                          vbr.bpntre, "c1", str(tmp_path), density=T)
        codegen.vbr_spmv_codegen("c1", str(tmp_path), str(tmp_path), threads=1, bench=3)
        times, y = codegen.run_program(str(tmp_path), "c1", x=x)
        assert len(times) == 3
        assert c_oracle.scaled_error(y, y64, ay)["componentwise"] <= TOL


@pytest.mark.parametrize("n_rhs", [16, 128])
def test_spmm_flow_random_x(rt, tmp_path, n_rhs):
    """vbr_spmm_codegen at density 15 (dense blocks on MFMA, 3-term split) on
    a C3-like matrix, random X: within 1e-5 of FP64."""
    from sable_amd.fileio import write_vbr_matrix
    from sable_amd.vbr import VBR
    vbr, _, _ = synth.mixed_vbr(3000, 30, 32, 192, 0.05, 0.7, (0.9, 1.0), (0.02, 0.1), seed=7)
    write_vbr_matrix("m", VBR(np.asarray(vbr.val, np.float64), vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr,
                              vbr.bpntrb, vbr.bpntre), str(tmp_path))
    X = np.random.default_rng(9).uniform(-1, 1, (3000, n_rhs)).astype(np.float32)
    Y64, AY = c_oracle.spmm_f64_dense_blocks(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb,
                                             vbr.bpntre, X)
    print(f"case: codegen spmm mixed 3000 N={n_rhs}", file=sys.stderr, flush=True)
    codegen.vbr_spmm_codegen("m", 15, str(tmp_path), str(tmp_path), threads=1, n_rhs=n_rhs, precision="bf16x6")
    _, Y = codegen.run_program(str(tmp_path), "m", x=X)
    assert c_oracle.scaled_error(Y.ravel(), Y64.ravel(), AY.ravel())["componentwise"] <= TOL
