"""GPU parity of the SpMM paths (FP32 vector path and the bf16-split MFMA
path) against the FP64 oracle and the restated reference SpMM program
(marked gpu).  Bar: componentwise-scaled error <= 1e-5 for the FP32 path and
the exact 3-term split (6 passes); the 2-term hi+lo split (3 passes) is held
to its derived worst case 3 * 2^-16 per product."""
import gzip
import os

import numpy as np
import pytest

from oracle import c_oracle
from sable_amd import synth
from sable_amd.convert import threshold_routes, ROUTE_DENSE, ROUTE_SPARSE, ROUTE_DENSE_MFMA
from sable_amd.fileio import read_vbr

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def rt():
    from sable_amd import runtime
    runtime.init()
    return runtime


def _err(vbr, X, Y):
    Y64, AY = c_oracle.spmm_f64_dense_blocks(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr,
                                             vbr.bpntrb, vbr.bpntre, X)
    return c_oracle.scaled_error(np.asarray(Y).ravel(), Y64.ravel(), AY.ravel())


def _c3_like(n=3000, seed=3):
    vbr, _, cls = synth.mixed_vbr(n, n // 100, 32, 192, 0.05, 0.7, (0.9, 1.0), (0.02, 0.1), seed)
    return vbr


def test_example_canon_n512(rt, golden):
    """tests/output_spmm_canon.txt: Y = A 1 with the reference's 512 columns."""
    val, indx, bindx, rp, cp, bb, be = read_vbr(os.path.join(golden, "example.vbr"))
    with gzip.open(os.path.join(golden, "output_spmm_canon.txt.gz"), "rt") as f:
        canon = [float(v) for v in f.read().split()]
    print("case: example n=512 fp32", flush=True)
    plan = rt.Plan(val, indx, bindx, rp, cp, bb, be, op="spmm", n_rhs=512)
    Y = plan.spmm(np.ones((11, 512), np.float32))
    assert [round(float(v), 6) for v in Y.ravel()] == canon


@pytest.mark.parametrize("route", [ROUTE_DENSE, ROUTE_DENSE_MFMA])
@pytest.mark.parametrize("n_rhs", [16, 128])
def test_dense_routes_match_oracle(rt, route, n_rhs):
    vbr = _c3_like()
    X = np.random.default_rng(1).uniform(-1, 1, (vbr.cpntr[-1], n_rhs)).astype(np.float32)
    routes = np.where(threshold_routes(vbr.val, vbr.indx, 15) == ROUTE_DENSE, route, ROUTE_SPARSE)
    print(f"case: c3-like route={route} n={n_rhs}", flush=True)
    plan = rt.Plan.from_vbr(vbr, routes, op="spmm", n_rhs=n_rhs)
    Y = plan.spmm(X)
    e = _err(vbr, X, Y)
    print(f"  componentwise {e['componentwise']:.3e}", flush=True)
    assert e["componentwise"] <= TOL


@pytest.mark.parametrize("n_rhs", [4, 8, 12, 64, 256])
def test_fp32_widths(rt, n_rhs):
    vbr = _c3_like(2000, seed=5)
    X = np.random.default_rng(2).uniform(-1, 1, (vbr.cpntr[-1], n_rhs)).astype(np.float32)
    routes = threshold_routes(vbr.val, vbr.indx, 15)
    print(f"case: fp32 n={n_rhs}", flush=True)
    Y = rt.Plan.from_vbr(vbr, routes, op="spmm", n_rhs=n_rhs).spmm(X)
    assert _err(vbr, X, Y)["componentwise"] <= TOL


@pytest.mark.parametrize("n_rhs", [32, 48, 64, 80, 256])
def test_mfma_widths(rt, n_rhs):
    vbr = _c3_like(2000, seed=6)
    X = np.random.default_rng(3).uniform(-1, 1, (vbr.cpntr[-1], n_rhs)).astype(np.float32)
    routes = np.where(threshold_routes(vbr.val, vbr.indx, 15) == ROUTE_DENSE, ROUTE_DENSE_MFMA, ROUTE_SPARSE)
    print(f"case: mfma n={n_rhs}", flush=True)
    Y = rt.Plan.from_vbr(vbr, routes, op="spmm", n_rhs=n_rhs).spmm(X)
    assert _err(vbr, X, Y)["componentwise"] <= TOL


def test_three_routes_in_one_block_row(rt):
    """MFMA, FP32-dense and sparse blocks share block rows: the MFMA tiles
    write Y in the first launch and the vector tiles add to it in the
    second."""
    vbr = _c3_like(2500, seed=9)
    rng = np.random.default_rng(4)
    routes = rng.choice([ROUTE_DENSE, ROUTE_SPARSE, ROUTE_DENSE_MFMA], len(vbr.bindx)).astype(np.uint8)
    X = rng.uniform(-1, 1, (vbr.cpntr[-1], 64)).astype(np.float32)
    print("case: three routes", flush=True)
    Y = rt.Plan.from_vbr(vbr, routes, op="spmm", n_rhs=64).spmm(X)
    assert _err(vbr, X, Y)["componentwise"] <= TOL


def test_mfma_short_and_unit_blocks(rt):
    """Block rows shorter than an MFMA tile (heights 1..15) and width-1
    blocks: padding rows/columns of the MFMA layout stay out of Y."""
    vbr, _ = synth.vbr_matrix_gen_seeded(200, 180, 40, 30, 150, seed=12)
    X = np.random.default_rng(5).uniform(-1, 1, (180, 32)).astype(np.float32)
    routes = np.full(len(vbr.bindx), ROUTE_DENSE_MFMA, np.uint8)
    print("case: mfma unit blocks", flush=True)
    Y = rt.Plan.from_vbr(vbr, routes, op="spmm", n_rhs=32).spmm(X)
    assert _err(vbr, X, Y)["componentwise"] <= TOL


@pytest.mark.parametrize("n_rhs", [16, 48, 128])
def test_spmm_edge_shapes(rt, n_rhs):
    """Unit partitions and empty block rows on every route; one 700 x 5000
    block (wider than the FP32 path's LDS staging and the MFMA staging
    capacity, taller than an MFMA tile) on every route; an all-zero sparse
    block gives Y = 0."""
    from sable_amd.vbr import VBR
    vbr, _ = synth.vbr_matrix_gen_seeded(300, 280, 60, 60, 200, seed=19)
    assert np.any(np.asarray(vbr.bpntrb) == -1)
    X = np.random.default_rng(8).uniform(-1, 1, (280, n_rhs)).astype(np.float32)
    for route in (ROUTE_DENSE, ROUTE_SPARSE, ROUTE_DENSE_MFMA):
        print(f"case: spmm unit partitions n={n_rhs} route={route}", flush=True)
        Y = rt.Plan.from_vbr(vbr, np.full(len(vbr.bindx), route, np.uint8), op="spmm", n_rhs=n_rhs).spmm(X)
        assert _err(vbr, X, Y)["componentwise"] <= TOL
    rng = np.random.default_rng(2)
    h, w = 700, 5000
    val = rng.standard_normal(h * w).astype(np.float32)
    val[rng.random(h * w) < 0.7] = 0
    blk = dict(val=val, indx=[0, h * w], bindx=[0], rpntr=[0, h], cpntr=[0, w], bpntrb=[0], bpntre=[1])
    Xw = rng.uniform(-1, 1, (w, n_rhs)).astype(np.float32)
    V = VBR(**blk)
    for route in (ROUTE_DENSE, ROUTE_SPARSE, ROUTE_DENSE_MFMA):
        print(f"case: spmm 700x5000 n={n_rhs} route={route}", flush=True)
        Y = rt.Plan(**blk, routes=[route], op="spmm", n_rhs=n_rhs).spmm(Xw)
        assert _err(V, Xw, Y)["componentwise"] <= TOL
    zero = dict(blk, val=np.zeros(h * w, np.float32))
    Y = rt.Plan(**zero, routes=[ROUTE_SPARSE], op="spmm", n_rhs=n_rhs).spmm(Xw)
    assert np.all(Y == 0)


def test_matches_reference_spmm_program(rt):
    """Against the restated reference SpMM program (src/codegen.py:827-937,
    N as a parameter, density-routed inline)."""
    vbr = _c3_like(1500, seed=7)
    X = np.random.default_rng(6).uniform(-1, 1, (vbr.cpntr[-1], 16)).astype(np.float32)
    route = threshold_routes(vbr.val, vbr.indx, 15)
    Yc = c_oracle.spmm(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb, vbr.bpntre, route, X, 16)
    _, AY = c_oracle.spmm_f64_dense_blocks(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr,
                                           vbr.bpntrb, vbr.bpntre, X)
    for r in (ROUTE_DENSE, ROUTE_DENSE_MFMA):
        routes = np.where(route == ROUTE_DENSE, r, ROUTE_SPARSE)
        print(f"case: vs reference program route={r}", flush=True)
        Y = rt.Plan.from_vbr(vbr, routes, op="spmm", n_rhs=16).spmm(X)
        assert c_oracle.scaled_error(Y.ravel(), Yc.astype(np.float64).ravel(), AY.ravel())["componentwise"] <= 2 * TOL


@pytest.mark.parametrize("n_rhs", [16, 128])
def test_fp32_generic_kernel(rt, n_rhs):
    """The ahead-of-time FP32 kernel (no hipRTC, N read at run time) takes
    the thread shape of the specialised one (wide at N = 128, narrow at
    N = 16) on the same tiles; both within the tolerance of FP64, with
    dense and sparse pieces in the same tiles and sparse-only tiles after
    MFMA tiles."""
    vbr = _c3_like(2000, seed=31)
    rng = np.random.default_rng(32)
    X = rng.uniform(-1, 1, (vbr.cpntr[-1], n_rhs)).astype(np.float32)
    thr = threshold_routes(vbr.val, vbr.indx, 15)
    mixed = np.where(thr == ROUTE_DENSE, ROUTE_DENSE_MFMA, ROUTE_SPARSE).astype(np.uint8)
    for name, routes in (("fp32", thr), ("mfma+sparse", mixed)):
        for use_rtc in (True, False):
            print(f"case: generic kernel n={n_rhs} {name} rtc={use_rtc}", flush=True)
            plan = rt.Plan.from_vbr(vbr, routes, op="spmm", n_rhs=n_rhs, wg=256, use_rtc=use_rtc)
            assert plan.stats["rtc_used"] == int(use_rtc)
            assert _err(vbr, X, plan.spmm(X))["componentwise"] <= TOL


def test_spmm_device_arrays_and_modes(rt):
    vbr = _c3_like(2000, seed=8)
    X = np.random.default_rng(7).uniform(-1, 1, (vbr.cpntr[-1], 128)).astype(np.float32)
    routes = np.where(threshold_routes(vbr.val, vbr.indx, 15) == ROUTE_DENSE, ROUTE_DENSE_MFMA, ROUTE_SPARSE)
    ref = None
    for mode in (0, 1, 2):
        print(f"case: spmm mode={mode}", flush=True)
        plan = rt.Plan.from_vbr(vbr, routes, op="spmm", n_rhs=128, mode=mode)
        Xd = rt.DeviceArray.from_numpy(X)
        Yd = rt.DeviceArray((plan.nrows, 128))
        plan.spmm(Xd, Yd)
        Y = Yd.numpy()
        if ref is None:
            ref = Y
        assert np.array_equal(Y, ref)
        t = plan.time(Xd, Yd, warmup=1, reps=3)
        assert np.all(t > 0)
    assert _err(vbr, X, ref)["componentwise"] <= TOL


def test_split_accuracy_bounds(rt):
    """Exact 3-term split (default): FP32-level, within 1e-5 and no worse
    than a few times the FP32 vector path; hi+lo split: within its worst-case
    bound 3 * 2^-16 and markedly less accurate than the exact split."""
    vbr = _c3_like(3000, seed=17)
    rng = np.random.default_rng(9)
    X = rng.uniform(-1, 1, (vbr.cpntr[-1], 64)).astype(np.float32)
    routes = rng.choice([ROUTE_SPARSE, ROUTE_DENSE_MFMA], len(vbr.bindx)).astype(np.uint8)
    e = {}
    for split in (2, 3):
        print(f"case: split={split}", flush=True)
        Y = rt.Plan.from_vbr(vbr, routes, op="spmm", n_rhs=64, mfma_split=split).spmm(X)
        e[split] = _err(vbr, X, Y)["componentwise"]
    fp32 = np.where(routes == ROUTE_DENSE_MFMA, ROUTE_DENSE, routes)
    e["fp32"] = _err(vbr, X, rt.Plan.from_vbr(vbr, fp32, op="spmm", n_rhs=64).spmm(X))["componentwise"]
    print(f"  errors {e}", flush=True)
    assert e[3] <= TOL and e[3] <= 4 * e["fp32"] + 1e-7
    assert e[2] <= 3 * 2.0 ** -16


@pytest.mark.parametrize("n_rhs", [16, 32])
def test_small_workgroups_all_routes(rt, n_rhs):
    """128-thread workgroups (the tuned choice for N = 16): MFMA, FP32-dense
    and sparse blocks in the same block rows, against the FP64 oracle."""
    vbr = _c3_like(2500, seed=23)
    rng = np.random.default_rng(24)
    routes = rng.choice([ROUTE_DENSE, ROUTE_SPARSE, ROUTE_DENSE_MFMA], len(vbr.bindx)).astype(np.uint8)
    X = rng.uniform(-1, 1, (vbr.cpntr[-1], n_rhs)).astype(np.float32)
    print(f"case: wg 128 n={n_rhs}", flush=True)
    plan = rt.Plan.from_vbr(vbr, routes, op="spmm", n_rhs=n_rhs, wg=128)
    assert plan.stats["wg"] == 128
    assert _err(vbr, X, plan.spmm(X))["componentwise"] <= TOL


@pytest.mark.parametrize("kchunk", [1, 4, -1])
def test_kchunk_split_tiles(rt, kchunk):
    """MFMA row tiles split along k (mfma_kchunk k-steps per chunk, partial
    sums added by the fixup launch in chunk order): within 1e-5 of FP64 and
    the same bits on every launch; -1 splits nothing."""
    vbr = _c3_like(6000, seed=17)
    X = np.random.default_rng(18).uniform(-1, 1, (6000, 128)).astype(np.float32)
    routes = np.full(len(vbr.bindx), ROUTE_DENSE_MFMA, np.uint8)
    print(f"case: mfma k-split chunk {kchunk}", flush=True)
    p = rt.Plan.from_vbr(vbr, routes, op="spmm", n_rhs=128, mfma_kchunk=kchunk)
    st = p.stats
    assert (st["ntiles_split"] > 0) == (kchunk > 0)
    assert (st["scratch_bytes"] > 0) == (kchunk > 0)
    Y1 = p.spmm(X)
    Y2 = p.spmm(X)
    assert np.array_equal(Y1, Y2)
    assert _err(vbr, X, Y1)["componentwise"] <= TOL
