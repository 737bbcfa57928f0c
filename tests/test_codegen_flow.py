"""The per-matrix generate step of the drop-in flow (reference:
src/codegen.py:165-186 split_chunks, :293-452 / :827-1075 the generators,
:1438-1458 vbr_spmv_codegen / vbr_spmm_codegen; utils/utils.py:30-33
extract_mul_nums) on the CPU: artefacts, hipRTC precompile (no GPU needed),
the stdout contract's formatting, and the precision guard.  The run step is
in tests/test_gpu_codegen.py."""
import io
import json
import os
import shutil

import numpy as np
import pytest

from oracle import c_oracle
from sable_amd import codegen
from sable_amd.fileio import read_vbr, read_vbrc


def test_split_chunks_known_answers():
    """Hand-derived from the rule: zero entries dropped, largest first (index
    order among equals), each into the first chunk with the smallest sum."""
    assert codegen.split_chunks([5, 0, 3, 3, 8, 1], 3) == [[4], [0, 5], [2, 3]]
    assert codegen.split_chunks([0, 2, 0, 7], 4) == [[3], [1]]
    assert codegen.split_chunks([1, 1, 1, 1, 1], 2) == [[0, 2, 4], [1, 3]]
    assert codegen.split_chunks([0, 0], 3) == []
    # the reference's SpMM work list has one entry per rpntr element (the
    # last always 0) and zero for rows without dense work: those are in no chunk
    assert codegen.split_chunks([4, 0, 9, 4, 0], 2) == [[2], [0, 3]]


def test_split_chunks_agrees_with_the_oracle_lpt():
    rng = np.random.default_rng(1)
    for _ in range(50):
        w = rng.integers(0, 20, rng.integers(1, 60)).tolist()
        k = int(rng.integers(1, 9))
        assert codegen.split_chunks(w, k) == c_oracle.lpt_chunks(w, k)


def test_extract_mul_nums():
    assert codegen.extract_mul_nums("example = 10,20,30,40,50,") == [" 10", "20", "30", "40", "50"]
    assert len(codegen.extract_mul_nums("example = 1,2,3,4,5,\n".strip())) == 5


def _cmp_lines(lines, canon_path):
    """tests' cmp_file (test.py:19-33): numeric lines compared as floats."""
    with open(canon_path) as f:
        canon = [l.strip() for l in f.read().split("\n")]
    for a, b in zip(lines, canon):
        a, b = a.strip(), b.strip()
        try:
            if float(a) != float(b):
                return False
        except ValueError:
            if a != b:
                return False
    return True


def test_print_program_output_spmv_contract(golden):
    """Line 0 'name = t0,...,' (src/codegen.py:440-444), then '%lf' per row
    (:445-447): y = A 1 of tests/example.vbr reproduces the canon file."""
    canon = [float(v) for v in open(os.path.join(golden, "output_spmv_canon.txt")).read().split()]
    buf = io.StringIO()
    codegen.print_program_output("example", [101, 102, 103, 104, 105], np.array(canon, np.float32), "spmv", out=buf)
    lines = buf.getvalue().split("\n")
    assert lines[0] == "example = 101,102,103,104,105,"
    assert len(codegen.extract_mul_nums(lines[0])) == 5
    assert all(l == "%f" % v for l, v in zip(lines[1:], canon))
    assert _cmp_lines(lines[1:], os.path.join(golden, "output_spmv_canon.txt"))


def test_print_program_output_spmm_contract():
    buf = io.StringIO()
    Y = np.array([[1.5, -2.0], [0.0, 3.25]], np.float32)
    codegen.print_program_output("m", [777], Y, "spmm", out=buf)
    assert buf.getvalue() == "m = 777\n1.500000\n-2.000000\n0.000000\n3.250000\n"


def test_vbr_spmv_codegen_artefacts(golden, tmp_path):
    """vbr_spmv_codegen on tests/example.vbrc: the program artefact holds the
    matrix (threshold 0: every block dense, the .vbr's values) and the
    kernel source, and the kernel is compiled into the code-object cache."""
    shutil.copy(os.path.join(golden, "example.vbrc"), tmp_path / "example.vbrc")
    ms = codegen.vbr_spmv_codegen("example", str(tmp_path / "out"), str(tmp_path), threads=1, bench=5)
    assert ms >= 0
    meta, a = codegen.load_program(str(tmp_path / "out"), "example")
    assert meta["op"] == "spmv" and meta["bench"] == 5 and meta["filename"] == "example"
    v = read_vbr(os.path.join(golden, "example.vbr"))
    for got, want in zip((a["val"], a["indx"], a["bindx"], a["rpntr"], a["cpntr"], a["bpntrb"], a["bpntre"]), v):
        assert np.array_equal(np.asarray(got, np.float64), np.asarray(want, np.float64))
    assert not np.any(a["routes"])                                      # ublocks empty
    src = open(tmp_path / "out" / "example.hip").read()
    assert "sable_program" in src and "#define SB_WG" in src
    # the threaded generator emits the same program
    codegen.vbr_spmv_codegen("example", str(tmp_path / "mt"), str(tmp_path), threads=4)
    meta2, a2 = codegen.load_program(str(tmp_path / "mt"), "example")
    assert all(np.array_equal(a[k], a2[k]) for k in a)
    assert open(tmp_path / "mt" / "example.hip").read() == src


def test_vbr_spmv_codegen_routed(kat, tmp_path):
    """The threshold-80 .vbrc (four blocks in the CSR part, test.py:80-105):
    the program's routes mark exactly ublocks and its values re-expand to the
    original blocks."""
    from sable_amd import convert
    d = np.array(kat["dense_test_compression"], dtype=np.float64)
    import scipy.sparse as sp
    A = sp.csc_matrix(d)
    rp, cp = kat["rpntr"], kat["cpntr"]
    val, indx, bindx, bb, be = convert.convert_sparse_to_vbr(A, rp, cp, "example2", str(tmp_path))
    convert.convert_vbr_to_compressed(val, rp, cp, indx, bindx, bb, be, "example2", str(tmp_path), density=80)
    codegen.vbr_spmv_codegen("example2", str(tmp_path), str(tmp_path), threads=1)
    meta, a = codegen.load_program(str(tmp_path), "example2")
    assert np.flatnonzero(a["routes"]).tolist() == kat["compressed_threshold_80"]["ublocks"]
    assert np.array_equal(np.asarray(a["val"], np.float64), np.asarray(val, np.float64))


def test_vbr_spmm_codegen_precision_guard(golden, tmp_path):
    shutil.copy(os.path.join(golden, "example.vbr"), tmp_path / "example.vbr")
    with pytest.raises(ValueError, match="bf16x3"):
        codegen.vbr_spmm_codegen("example", 0, str(tmp_path), str(tmp_path), threads=1, n_rhs=32,
                                 precision="bf16x3")
    with pytest.raises(ValueError):
        codegen.vbr_spmm_codegen("example", 0, str(tmp_path), str(tmp_path), threads=1, n_rhs=32,
                                 precision="fp16")
    assert codegen.MFMA_SPLIT_BOUND["bf16x6"] < 1e-6 < 1e-5 < codegen.MFMA_SPLIT_BOUND["bf16x3"]
    # accepted when the caller's tolerance covers the bound
    codegen.vbr_spmm_codegen("example", 0, str(tmp_path), str(tmp_path), threads=1, n_rhs=32,
                             precision="bf16x3", tol=5e-5)
    meta, a = codegen.load_program(str(tmp_path), "example")
    assert meta["mfma_split"] == 2 and np.all(a["routes"] == 2)


def test_vbr_spmm_codegen_threads_same_program(golden, tmp_path):
    shutil.copy(os.path.join(golden, "example.vbr"), tmp_path / "example.vbr")
    for threads, sub in ((1, "st"), (4, "mt")):
        codegen.vbr_spmm_codegen("example", 80, str(tmp_path / sub), str(tmp_path), threads=threads, n_rhs=32,
                                 precision="bf16x6")
    m1, a1 = codegen.load_program(str(tmp_path / "st"), "example")
    m2, a2 = codegen.load_program(str(tmp_path / "mt"), "example")
    assert m1 == m2 and all(np.array_equal(a1[k], a2[k]) for k in a1)
    # density 80 routes the example's blocks of density 2/3 and 3/4 (ordinals
    # 2, 4, 9) to the sparse path, the rest to MFMA
    assert np.flatnonzero(a1["routes"] == 1).tolist() == [2, 4, 9] and set(a1["routes"].tolist()) == {1, 2}
    with pytest.raises(ValueError):
        codegen.gen_multi_threaded_spmm(0, *read_vbr(str(tmp_path / "example.vbr")), 0, str(tmp_path), "x",
                                        str(tmp_path))


def test_print_program_output_follows_redirected_stdout(capsys):
    """The default stream is the current sys.stdout, not the one at import."""
    codegen.print_program_output("m", [5], np.array([1.0], np.float32), "spmv")
    assert capsys.readouterr().out == "m = 5,\n1.000000\n"
