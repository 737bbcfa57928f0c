"""The multi-threaded CPU baseline (reference: gen_multi_threaded_spmv and
split_chunks, src/codegen.py:165-186,454-643), restated in oracle/.

Block rows are split into LPT chunks, one thread per chunk, each block row
computed by one thread in the single-threaded order, so the result is the
single-threaded program's bit for bit at every thread count.  The
reference's own multi-threaded emitter passes wrong value offsets to two of
its kernels (indx[count] for indx[indx_start + count]); that defect is not
reproduced: tests/example.vbr, whose block row 2 has height 1 and would take
the wrong offset, matches the canonical output at 2 to 8 threads."""
import os

import numpy as np
import pytest

from oracle import c_oracle
from oracle import vbr_oracle as O
from sable_amd import synth
from sable_amd.convert import compress_with_routes, threshold_routes


def test_lpt_chunks_known_answer():
    """split_chunks([5, 0, 3, 3, 8, 1], 3): zero-work rows dropped, rows taken
    largest first (stable among equals) into the first lightest chunk."""
    assert c_oracle.lpt_chunks([5, 0, 3, 3, 8, 1], 3) == [[4], [0, 5], [2, 3]]
    # fewer non-empty rows than chunks: one chunk per row
    assert c_oracle.lpt_chunks([0, 2, 0, 7], 4) == [[3], [1]]
    # equal work: round robin in row order
    assert c_oracle.lpt_chunks([1, 1, 1, 1, 1], 2) == [[0, 2, 4], [1, 3]]
    # every row lands in exactly one chunk, sums within the largest item
    w = list(np.random.default_rng(0).integers(1, 1000, 300))
    ch = c_oracle.lpt_chunks(w, 7)
    assert sorted(i for c in ch for i in c) == list(range(300))
    sums = [sum(w[i] for i in c) for c in ch]
    assert max(sums) - min(sums) <= max(w)


def _canon(path):
    return [float(v) for v in open(path).read().split()]


def test_example_matches_canon_at_every_thread_count(kat, golden):
    ex = kat["example_vbr"]
    rp, cp = kat["rpntr"], kat["cpntr"]
    val2, indx2, ub, ip, ind, cv = O.convert_vbr_to_compressed(ex["val"], rp, cp, ex["indx"], ex["bindx"],
                                                               ex["bpntrb"], ex["bpntre"], 0)
    prog = c_oracle.CProgram(val2, indx2, ex["bindx"], rp, cp, ex["bpntrb"], ex["bpntre"], ub, ip, ind, cv)
    canon = _canon(os.path.join(golden, "output_spmv_canon.txt"))
    for t in (1, 2, 3, 4, 8):
        prog.set_threads(t)
        y = prog.run(np.ones(11))
        assert [round(float(v), 6) for v in y] == canon, t


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_threads_bit_identical_to_single_thread(dtype):
    vbr, x, _ = synth.mixed_vbr(3000, 40, 20, 140, 0.3, 0.5, (0.8, 1.0), (0.02, 0.3), 11)
    routes = threshold_routes(vbr.val, vbr.indx, 15)
    val2, indx2, ub, ip, ind, cv = compress_with_routes(vbr.val, vbr.rpntr, vbr.cpntr, vbr.indx, vbr.bindx,
                                                        vbr.bpntrb, vbr.bpntre, routes)
    prog = c_oracle.CProgram(val2, indx2, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb, vbr.bpntre, ub, ip, ind,
                             cv, dtype=dtype)
    y1 = prog.run(x)
    for t in (2, 5, 8):
        prog.set_threads(t)
        assert np.array_equal(prog.run(x).view(np.uint8), y1.view(np.uint8)), t


def test_remove_outliers_deciles_known_answers():
    """eval.py:14-21: fewer than 10 values kept as they are; else values in
    [P10, P90] (numpy linear percentiles: 1..20 -> P10 = 2.9, P90 = 18.1)."""
    from oracle.cpu_protocol import remove_outliers_deciles, parse_line0
    assert remove_outliers_deciles([5, 1, 9]) == [5, 1, 9]
    assert remove_outliers_deciles(list(range(1, 21))) == list(range(3, 19))
    assert remove_outliers_deciles([7.0] * 12) == [7.0] * 12
    assert parse_line0("prog = 10,20,30,") == [10.0, 20.0, 30.0]


def test_process_protocol_runs_children(tmp_path):
    """Three child processes of a small program (SpMV and threaded SpMM):
    line 0 parsed, medians kept (fewer than 10 runs: no filter), mean > 0."""
    from oracle import cpu_protocol as P
    vbr, x = synth.workload_c1()
    routes = threshold_routes(vbr.val, vbr.indx, 15)
    val2, indx2, ub, ip, ind, cv = compress_with_routes(vbr.val, vbr.rpntr, vbr.cpntr, vbr.indx, vbr.bindx,
                                                        vbr.bpntrb, vbr.bpntre, routes)
    path = str(tmp_path / "c1.npz")
    P.save_spmv_program(path, val2, indx2, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb, vbr.bpntre, ub, ip, ind,
                        cv, x)
    r = P.time_processes(path, 1, 5, procs=3)
    assert r["runs"] == 3 and r["kept_runs"] == 3 and r["mean_s"] > 0 and len(r["run_medians_s"]) == 3
    X = np.random.default_rng(0).uniform(-1, 1, (1000, 8)).astype(np.float32)
    path = str(tmp_path / "c1m.npz")
    P.save_spmm_program(path, vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb, vbr.bpntre, routes, X)
    r = P.time_processes(path, 3, 2, procs=2)
    assert r["mean_s"] > 0


def test_threaded_spmm_bit_identical():
    vbr, _, _ = synth.mixed_vbr(3000, 30, 32, 192, 0.05, 0.7, (0.9, 1.0), (0.02, 0.1), seed=7)
    X = np.random.default_rng(1).uniform(-1, 1, (3000, 16)).astype(np.float32)
    routes = threshold_routes(vbr.val, vbr.indx, 15)
    args = (vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb, vbr.bpntre, routes, X, 16)
    Y1 = c_oracle.spmm(*args)
    for t in (2, 5, 8):
        assert np.array_equal(c_oracle.spmm(*args, threads=t), Y1)
