"""sable_spmm_noalloc, the variant of sable_spmm that allocates nothing (its
split MFMA tiles' partial sums and counters live in the library's
file-scope device scratch array), through the C-ABI (marked gpu).

Bar: bit-identical to sable_spmm on the same plan (same kernels, same
partial-sum order), on two calls in a row, the second with a smaller
input; within 1e-5 of FP64; a plan needing more scratch than the array
holds and host pointers are refused with an error."""
import numpy as np
import pytest

from oracle import c_oracle
from sable_amd import synth
from sable_amd.convert import ROUTE_DENSE_MFMA

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    from sable_amd import runtime
    runtime.init()
    return runtime


def _matrix(n, seed):
    vbr, _, _ = synth.mixed_vbr(n, n // 100, 32, 192, 0.02, 0.7, (0.9, 1.0), (0.02, 0.1), seed=seed)
    X = np.random.default_rng(seed + 1).uniform(-1, 1, (n, 128)).astype(np.float32)
    return vbr, X


def test_noalloc_matches_spmm_on_two_calls(rt):
    from sable_amd._lib import lib, check
    for n, seed in ((20000, 41), (6000, 43)):
        vbr, X = _matrix(n, seed)
        routes = np.full(len(vbr.bindx), ROUTE_DENSE_MFMA, np.uint8)
        plan = rt.Plan.from_vbr(vbr, routes, op="spmm", n_rhs=128, mfma_kchunk=8)
        assert plan.stats["ntiles_split"] > 0
        Xd = rt.DeviceArray.from_numpy(X)
        Y1 = rt.DeviceArray.zeros((n, 128))
        Y2 = rt.DeviceArray.zeros((n, 128))
        print(f"case: noalloc spmm n={n} (split tiles {plan.stats['ntiles_split']})", flush=True)
        check(lib.sable_spmm(plan._h, Xd.ptr, Y1.ptr, 0))
        check(lib.sable_spmm_noalloc(plan._h, Xd.ptr, Y2.ptr, 0))
        a, b = Y1.numpy(), Y2.numpy()
        assert np.array_equal(a, b)
        Y64, AY = c_oracle.spmm_f64_dense_blocks(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb,
                                                 vbr.bpntre, X)
        assert c_oracle.scaled_error(b.ravel(), Y64.ravel(), AY.ravel())["componentwise"] <= 1e-5
        # again on the same array (its counters were left at zero)
        check(lib.sable_spmm_noalloc(plan._h, Xd.ptr, Y2.ptr, 0))
        assert np.array_equal(a, Y2.numpy())
        plan.close()


def test_noalloc_refuses_oversized_scratch_and_host_pointers(rt):
    from sable_amd._lib import lib, SABLE_ERR_UNSUPPORTED, SABLE_ERR_INVALID
    vbr, X = _matrix(20000, 45)
    routes = np.full(len(vbr.bindx), ROUTE_DENSE_MFMA, np.uint8)
    plan = rt.Plan.from_vbr(vbr, routes, op="spmm", n_rhs=128, mfma_kchunk=1)
    assert plan.stats["scratch_bytes"] > 64 << 20
    Xd = rt.DeviceArray.from_numpy(X)
    Yd = rt.DeviceArray((20000, 128))
    print("case: noalloc refusals", flush=True)
    assert lib.sable_spmm_noalloc(plan._h, Xd.ptr, Yd.ptr, 0) == SABLE_ERR_UNSUPPORTED
    assert b"scratch" in lib.sable_last_error()
    small = rt.Plan.from_vbr(*_matrix(6000, 47)[:1], routes=None, op="spmm", n_rhs=128)
    Xh = np.zeros((6000, 128), np.float32)
    Yh = np.zeros((6000, 128), np.float32)
    assert lib.sable_spmm_noalloc(small._h, Xh.ctypes.data, Yh.ctypes.data, 1) == SABLE_ERR_INVALID
