"""GPU parity of batched execution (groups) against per-plan launches and the
FP64 oracle (marked gpu).

A group runs the tiles of many plans in persistent launches captured once in
a hipGraph.  Every tile computes exactly what it computes in its own plan, so
a group's outputs must be bit-identical to the members' own launches, on
every replay, with eager work between replays.
"""
import numpy as np
import pytest

from oracle import c_oracle
from sable_amd import synth
from sable_amd.convert import threshold_routes, ROUTE_DENSE, ROUTE_DENSE_MFMA

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def rt():
    from sable_amd import runtime
    runtime.init()
    return runtime


def _mats(count, seed=5):
    return list(synth.workload_c5(count=count, seed=seed, n_range=(2000, 9000), p_range=(50, 300)))


def _spmv_err(vbr, x, y):
    y64, ay = c_oracle.spmv_f64_dense_blocks(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr,
                                             vbr.bpntrb, vbr.bpntre, x)
    return c_oracle.scaled_error(y, y64, ay)["componentwise"]


def _spmm_err(vbr, X, Y):
    Y64, AY = c_oracle.spmm_f64_dense_blocks(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr,
                                             vbr.bpntrb, vbr.bpntre, X)
    return c_oracle.scaled_error(np.asarray(Y).ravel(), Y64.ravel(), AY.ravel())["componentwise"]


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_spmv_group_matches_members(rt, mode):
    mats = _mats(6)
    plans, xs, ys, own = [], [], [], []
    for vbr, x, _ in mats:
        p = rt.Plan.from_vbr(vbr, threshold_routes(vbr.val, vbr.indx, 15))
        plans.append(p)
        xs.append(rt.DeviceArray.from_numpy(x))
        ys.append(rt.DeviceArray(p.nrows))
        own.append(p.spmv(x))
    print(f"case: spmv group of 6, mode {mode}", flush=True)
    g = rt.Group(plans, xs, ys, mode=mode)
    g.launch()
    for (vbr, x, _), y, o in zip(mats, ys, own):
        assert np.array_equal(y.numpy(), o)
        assert _spmv_err(vbr, x, o) <= TOL
    g.close()


def test_mixed_group_spmv_and_spmm_replays(rt):
    """SpMV plans and SpMM (N = 16, MFMA dense blocks) plans in one group
    (default mode): two launches in one graph.  Replayed 3 times with eager
    launches and host synchronisation in between; every replay matches."""
    mats = _mats(5, seed=7)
    plans, xs, ys, own = [], [], [], []
    for vbr, x, X in mats:
        r = threshold_routes(vbr.val, vbr.indx, 15)
        pv = rt.Plan.from_vbr(vbr, r)
        rm = np.where(r == ROUTE_DENSE, ROUTE_DENSE_MFMA, r).astype(np.uint8)
        pm = rt.Plan.from_vbr(vbr, rm, op="spmm", n_rhs=16)
        for p, inp, shape in ((pv, x, (pv.nrows,)), (pm, X, (pm.nrows, 16))):
            plans.append(p)
            xs.append(rt.DeviceArray.from_numpy(inp))
            ys.append(rt.DeviceArray(shape))
        own.append(pv.spmv(x))
        own.append(pm.spmm(X))
    print("case: mixed group of 10 (spmv + spmm n=16)", flush=True)
    g = rt.Group(plans, xs, ys)
    for rep in range(3):
        for y in ys:
            y.copy_from(np.full(y.shape, np.nan, np.float32))
        g.launch()
        for y, o in zip(ys, own):
            assert np.array_equal(y.numpy(), o), f"replay {rep}"
        # eager work between replays
        plans[0].spmv(mats[0][1])
    for i, (vbr, x, X) in enumerate(mats):
        assert _spmv_err(vbr, x, own[2 * i]) <= TOL
        assert _spmm_err(vbr, X, own[2 * i + 1]) <= TOL
    t = g.time(warmup=1, reps=3, flush_bytes=64 << 20)
    assert np.all(t > 0)
    g.close()


def test_unbatched_timing_runs_members(rt):
    mats = _mats(3, seed=9)
    plans = [rt.Plan.from_vbr(v, threshold_routes(v.val, v.indx, 15)) for v, _, _ in mats]
    xs = [rt.DeviceArray.from_numpy(x) for _, x, _ in mats]
    ys = [rt.DeviceArray(p.nrows) for p in plans]
    print("case: time_plans over 3 plans", flush=True)
    t = rt.time_plans(plans, xs, ys, warmup=1, reps=3, flush_bytes=64 << 20)
    assert np.all(t > 0)
    for p, (v, x, _), y in zip(plans, mats, ys):
        assert np.array_equal(y.numpy(), p.spmv(x))


def test_group_rejects_host_only_member(rt):
    vbr, x, _ = _mats(1)[0]
    p = rt.Plan.from_vbr(vbr, None, host_only=True)
    with pytest.raises(Exception):
        rt.Group([p], [rt.DeviceArray.from_numpy(x)], [rt.DeviceArray(p.nrows)])


def test_group_keeps_each_members_split(rt):
    """SpMM members built with the 3-term split and with hi+lo in one group:
    they land in different launches (the part key holds the split), and each
    output is bit-identical to that plan's own launch."""
    mats = _mats(4, seed=13)
    plans, xs, ys, own = [], [], [], []
    for j, (vbr, _, X) in enumerate(mats):
        r = threshold_routes(vbr.val, vbr.indx, 15)
        rm = np.where(r == ROUTE_DENSE, ROUTE_DENSE_MFMA, r).astype(np.uint8)
        p = rt.Plan.from_vbr(vbr, rm, op="spmm", n_rhs=16, mfma_split=2 if j % 2 else 3)
        plans.append(p)
        xs.append(rt.DeviceArray.from_numpy(X))
        ys.append(rt.DeviceArray((p.nrows, 16)))
        own.append(p.spmm(X))
    # the two splits give different results on these inputs
    assert not all(np.array_equal(a, b) for a, b in zip(own, [q.spmm(X) for q, (_, _, X) in zip(
        [rt.Plan.from_vbr(v, np.where(threshold_routes(v.val, v.indx, 15) == ROUTE_DENSE, ROUTE_DENSE_MFMA,
                                      threshold_routes(v.val, v.indx, 15)).astype(np.uint8), op="spmm",
                          n_rhs=16, mfma_split=3) for v, _, _ in mats], mats)]))
    print("case: group of split-3 and split-2 spmm members", flush=True)
    g = rt.Group(plans, xs, ys)
    g.launch()
    for y, o in zip(ys, own):
        assert np.array_equal(y.numpy(), o)
    g.close()


def test_group_of_k_split_members(rt):
    """SpMM members whose MFMA row tiles are split along k (4 k-steps per
    chunk): the group's fixup launch adds each member's partial sums, and
    every output is bit-identical to that plan's own launch over two
    replays."""
    mats = _mats(3, seed=19)
    plans, xs, ys, own = [], [], [], []
    for vbr, _, X in mats:
        rm = np.full(len(vbr.bindx), ROUTE_DENSE_MFMA, np.uint8)
        p = rt.Plan.from_vbr(vbr, rm, op="spmm", n_rhs=16, mfma_kchunk=4, min_tiles=1)
        plans.append(p)
        xs.append(rt.DeviceArray.from_numpy(X))
        ys.append(rt.DeviceArray((p.nrows, 16)))
        own.append(p.spmm(X))
    assert any(p.stats["ntiles_split"] > 0 for p in plans)
    print("case: group of k-split spmm members", flush=True)
    g = rt.Group(plans, xs, ys)
    for _ in range(2):
        g.launch()
        for y, o in zip(ys, own):
            assert np.array_equal(y.numpy(), o)
    g.close()
