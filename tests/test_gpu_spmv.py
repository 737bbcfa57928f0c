"""GPU parity of the SpMV path against the oracle (marked gpu).

Bar: FP32 GPU result vs FP64 oracle, componentwise-scaled error
max_i |y_i - y64_i| / (|A||x|)_i <= 1e-5; the reference's canon outputs
(x = 1, printed with 6 decimals) reproduced exactly after rounding; bit-exact
repeatability across launches, scheduling modes and kernel builds.
"""
import os

import numpy as np
import pytest

from oracle import c_oracle
from oracle import vbr_oracle as O
from sable_amd import synth
from sable_amd.convert import (convert_vbr_to_compressed, threshold_routes, ROUTE_DENSE,
                               ROUTE_SPARSE)
from sable_amd.fileio import read_vbr

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def rt():
    from sable_amd import runtime
    runtime.init()
    return runtime


def _err(vbr, x, y):
    y64, ay = c_oracle.spmv_f64_dense_blocks(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr,
                                             vbr.bpntrb, vbr.bpntre, x)
    return c_oracle.scaled_error(y, y64, ay)


def _canon(path):
    return [float(v) for v in open(path).read().split()]


def test_example_canon_all_dense(rt, golden):
    val, indx, bindx, rpntr, cpntr, bb, be = read_vbr(os.path.join(golden, "example.vbr"))
    print("case: example.vbr all dense", flush=True)
    plan = rt.Plan(val, indx, bindx, rpntr, cpntr, bb, be)
    y = plan.spmv(np.ones(11, np.float32))
    assert [round(float(v), 6) for v in y] == _canon(os.path.join(golden, "output_spmv_canon.txt"))


def test_example_canon_routed(rt, kat, golden):
    import scipy.sparse
    from sable_amd.convert import convert_sparse_to_vbr
    d = np.array(kat["dense_test_compression"], dtype=float)
    rp, cp = kat["rpntr"], kat["cpntr"]
    val, indx, bindx, bb, be = convert_sparse_to_vbr(scipy.sparse.csc_matrix(d), rp, cp, None, None)
    routes = threshold_routes(val, indx, 80)
    assert list(np.nonzero(routes)[0]) == [2, 4, 9, 12]
    print("case: example2 threshold 80", flush=True)
    plan = rt.Plan(val, indx, bindx, rp, cp, bb, be, routes)
    y = plan.spmv(np.ones(11, np.float32))
    assert [round(float(v), 6) for v in y] == _canon(os.path.join(golden, "output_spmv_canon_sparse.txt"))


def test_c1_workload_matches_oracles(rt):
    vbr, x = synth.workload_c1()
    print("case: c1 1000x1000 all dense", flush=True)
    plan = rt.Plan.from_vbr(vbr)
    y = plan.spmv(x)
    assert _err(vbr, x, y)["componentwise"] <= TOL
    # against the restated generated program (FP64, reference semantics)
    val2, indx2, bindx, bb, be, ub, ip, ind, cv = convert_vbr_to_compressed(
        vbr.val, vbr.rpntr, vbr.cpntr, vbr.indx, vbr.bindx, vbr.bpntrb, vbr.bpntre, None, None, density=0)
    prog = c_oracle.CProgram(val2, indx2, bindx, vbr.rpntr, vbr.cpntr, bb, be, ub, ip, ind, cv)
    yc = prog.run(x.astype(np.float64))
    _, ay = c_oracle.spmv_f64_dense_blocks(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr,
                                           vbr.bpntrb, vbr.bpntre, x)
    assert c_oracle.scaled_error(y, yc, ay)["componentwise"] <= TOL


@pytest.mark.parametrize("T", [0, 15, 50, 100])
def test_mixed_matches_oracle_for_thresholds(rt, T):
    vbr, x, _ = synth.workload_like_c2(8000, seed=21, frac_blocks=0.03)
    routes = threshold_routes(vbr.val, vbr.indx, T)
    print(f"case: mixed n=8000 T={T}", flush=True)
    plan = rt.Plan.from_vbr(vbr, routes)
    y = plan.spmv(x)
    assert _err(vbr, x, y)["componentwise"] <= TOL


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_modes_bit_identical(rt, mode):
    vbr, x, _ = synth.workload_like_c2(6000, seed=4, frac_blocks=0.03)
    routes = threshold_routes(vbr.val, vbr.indx, 15)
    print(f"case: modes mode={mode}", flush=True)
    ref = rt.Plan.from_vbr(vbr, routes, mode=0).spmv(x)
    plan = rt.Plan.from_vbr(vbr, routes, mode=mode, grid=37)
    xd = rt.DeviceArray.from_numpy(x)
    outs = [plan.spmv(xd).numpy() for _ in range(3)]
    for o in outs:
        assert np.array_equal(o, ref)


def test_rtc_and_aot_agree(rt):
    vbr, x, _ = synth.workload_like_c2(5000, seed=8, frac_blocks=0.03)
    routes = threshold_routes(vbr.val, vbr.indx, 15)
    print("case: rtc vs aot", flush=True)
    a = rt.Plan.from_vbr(vbr, routes, use_rtc=False, wg=256, unroll=4)   # the generic AOT kernel's shape
    b = rt.Plan.from_vbr(vbr, routes, use_rtc=True, wg=256, unroll=4)
    assert b.stats["rtc_used"] == 1 and a.stats["rtc_used"] == 0
    assert np.array_equal(a.spmv(x), b.spmv(x))


def test_edge_shapes(rt):
    """Unit partitions (height-1 and width-1 blocks), empty block rows, a
    block wider than the staging capacity, a block row taller than a
    workgroup, and all-zero sparse blocks."""
    vbr, x = synth.vbr_matrix_gen_seeded(300, 280, 60, 60, 200, seed=9)
    print("case: unit partitions", flush=True)
    assert np.any(np.asarray(vbr.bpntrb) == -1)
    y = rt.Plan.from_vbr(vbr).spmv(x)
    assert _err(vbr, x, y)["componentwise"] <= TOL
    rng = np.random.default_rng(1)
    h, w = 700, 5000
    val = rng.standard_normal(h * w).astype(np.float32)
    val[rng.random(h * w) < 0.7] = 0
    blk = dict(val=val, indx=[0, h * w], bindx=[0], rpntr=[0, h], cpntr=[0, w], bpntrb=[0], bpntre=[1])
    xw = rng.uniform(-1, 1, w).astype(np.float32)
    from sable_amd.vbr import VBR
    V = VBR(**blk)
    for route, xcap in ((ROUTE_DENSE, 4095), (ROUTE_SPARSE, 4095), (ROUTE_DENSE, 100), (ROUTE_SPARSE, 64)):
        print(f"case: 700x5000 route={route} xcap={xcap}", flush=True)
        y = rt.Plan(**blk, routes=[route], xcap=xcap).spmv(xw)
        assert _err(V, xw, y)["componentwise"] <= TOL
    zero = dict(blk, val=np.zeros(h * w, np.float32))
    y = rt.Plan(**zero, routes=[ROUTE_SPARSE]).spmv(xw)
    assert np.all(y == 0)


@pytest.mark.parametrize("wg", [64, 256, 512])
def test_short_block_row_spans(rt, wg):
    """Runs of short block rows planned as one tile each (every block on the
    sparse path, dense-routed blocks with their zeros; min_tiles 1, as for a
    group member): against FP64 on the
    threshold, all-dense and all-sparse routes; a span with empty block rows
    and block rows taller than the span limit in between."""
    vbr, x, _ = synth.mixed_vbr(3000, 600, 1, 12, 0.02, 0.5, (0.8, 1.0), (0.05, 0.3), seed=41)
    tall, xt, _ = synth.mixed_vbr(2000, 40, 20, 120, 0.1, 0.5, (0.8, 1.0), (0.05, 0.3), seed=42)
    for name, V, xv in (("short rows", vbr, x), ("short and tall rows", tall, xt)):
        thr = threshold_routes(V.val, V.indx, 15)
        for rname, routes in (("threshold", thr), ("dense", np.full_like(thr, ROUTE_DENSE)),
                              ("sparse", np.full_like(thr, ROUTE_SPARSE))):
            print(f"case: spans {name} wg={wg} routes={rname}", flush=True)
            plan = rt.Plan.from_vbr(V, routes, wg=wg, min_tiles=1)
            assert plan.stats["nspans"] > 0
            if name == "short rows":
                assert plan.stats["ntiles"] < (len(V.rpntr) - 1) // 4
            assert _err(V, xv, plan.spmv(xv))["componentwise"] <= TOL


@pytest.mark.parametrize("wg,seg", [(64, 64), (128, 3), (256, 256), (256, 8), (256, 1), (512, 512)])
def test_sparse_wave_slices(rt, wg, seg):
    """Sparse rows in wave slices: tile heights from 1 row to the workgroup
    size give every lanes-per-row choice from 64 down to 1, with ragged rows
    (per-row lengths vary with the block density) and 1-4 wavefronts."""
    vbr, x, _ = synth.workload_like_c2(4000, seed=13, frac_blocks=0.03)
    routes = np.full(len(vbr.bindx), ROUTE_SPARSE, np.uint8)
    print(f"case: sparse slices wg={wg} max_seg_rows={seg}", flush=True)
    plan = rt.Plan.from_vbr(vbr, routes, wg=wg, max_seg_rows=seg)
    glogs = {int(t["sp_glog"]) for t in plan.export("tiles") if t["sp_map"] >= 0}
    y = plan.spmv(x)
    assert _err(vbr, x, y)["componentwise"] <= TOL
    assert glogs  # sparse tiles exist


def test_sparse_slices_repeatable_and_mixed(rt):
    """Dense and sparse pieces in one tile, several sparse groups per tile
    (small staging capacity): results are bit-identical across launches."""
    vbr, x, _ = synth.workload_like_c2(3000, seed=17, frac_blocks=0.05)
    routes = threshold_routes(vbr.val, vbr.indx, 15)
    print("case: mixed slices xcap=64", flush=True)
    plan = rt.Plan.from_vbr(vbr, routes, xcap=64)
    y1 = plan.spmv(x)
    y2 = plan.spmv(x)
    assert np.array_equal(y1, y2)
    assert _err(vbr, x, y1)["componentwise"] <= TOL


def test_device_tensors_and_timing(rt):
    vbr, x, _ = synth.workload_like_c2(10000, seed=3)
    plan = rt.Plan.from_vbr(vbr, threshold_routes(vbr.val, vbr.indx, 15))
    xd = rt.DeviceArray.from_numpy(x)
    yd = rt.DeviceArray(plan.nrows)
    print("case: timing", flush=True)
    t = plan.time(xd, yd, warmup=2, reps=5, flush_bytes=64 << 20)
    assert np.all(t > 0)
    assert _err(vbr, x, yd.numpy())["componentwise"] <= TOL


@pytest.mark.skip(reason="per-tile clocks were removed from the library after their first GPU run went "
                         "silent and was stopped; the cause was not shown, so the path is not run again")
def test_tile_clock(rt):
    vbr, x, _ = synth.workload_like_c2(5000, seed=5)
    plan = rt.Plan.from_vbr(vbr, threshold_routes(vbr.val, vbr.indx, 15))
    plan.enable_tile_clock(True)
    print("case: tile clock", flush=True)
    y = plan.spmv(x)
    ns = plan.tile_times()
    assert len(ns) == plan.stats["ntiles"] and np.all(ns >= 0) and np.any(ns > 0)
    assert _err(vbr, x, y)["componentwise"] <= TOL


def test_planted_error_would_be_caught(rt):
    """The parity check above is strong enough to see a one-value error: the
    GPU result compared against an oracle whose matrix has one value changed
    fails the bar."""
    vbr, x = synth.workload_c1()
    y = rt.Plan.from_vbr(vbr).spmv(x)
    val = np.array(vbr.val, copy=True)
    val[12345] += 0.01
    from sable_amd.vbr import VBR
    bad = VBR(val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb, vbr.bpntre)
    assert _err(bad, x, y)["componentwise"] > TOL


def test_mixed_routes_256_threads_unroll_2(rt):
    """256-thread workgroups at unroll 2 (the 256-matrix batch's SpMV
    members): dense and sparse pieces in the same tiles, split roles."""
    vbr, x, _ = synth.workload_like_c2(6000, seed=17)
    from sable_amd.convert import threshold_routes
    routes = threshold_routes(vbr.val, vbr.indx, 15)
    print("case: mixed wg=256 unroll=2", flush=True)
    plan = rt.Plan.from_vbr(vbr, routes, wg=256, unroll=2, min_tiles=1)
    assert plan.stats["wg"] == 256
    assert any(int(t["sp_waves"]) > 0 for t in plan.export("tiles"))
    assert _err(vbr, x, plan.spmv(x))["componentwise"] <= TOL
