"""Host-side routing rules (no GPU): which routes a block may take.  MFMA
needs N a multiple of 16 up to 256; the fixed MFMA policy also needs both
block edges to reach 16 (MFMA tiles are 16 x 32)."""
import numpy as np

from sable_amd.convert import ROUTE_DENSE, ROUTE_DENSE_MFMA, ROUTE_SPARSE
from sable_amd.router.router import allowed_mask, fixed_routes, mfma_edges_ok
from sable_amd.vbr import VBR


def test_allowed_mask_by_width():
    for n in (16, 128, 256):
        assert allowed_mask(n, 4)[:, 2].all()
    for n in (1, 8, 24, 272):
        assert not allowed_mask(n, 4)[:, 2].any()
    assert allowed_mask(1, 4)[:, :2].all()


def test_edge_rule():
    assert mfma_edges_ok([8, 16, 32, 15], [32, 16, 8, 200]).tolist() == [False, True, False, False]


def _vbr_two_blocks():
    # block row 0: 8 rows x (16 + 32) columns; block row 1: 32 rows x 16 columns
    rpntr = np.array([0, 8, 40], np.int32)
    cpntr = np.array([0, 16, 48], np.int32)
    bindx = np.array([0, 1, 0], np.int32)
    bpntrb = np.array([0, 2], np.int32)
    bpntre = np.array([2, 3], np.int32)
    sizes = [8 * 16, 8 * 32, 32 * 16]
    indx = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    val = np.ones(indx[-1], np.float64)
    return VBR(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre)


def test_fixed_mfma_policy_keeps_short_blocks_on_fp32():
    v = _vbr_two_blocks()
    assert fixed_routes(v, 1).tolist() == [ROUTE_DENSE] * 3
    assert fixed_routes(v, 128, mfma=True).tolist() == [ROUTE_DENSE, ROUTE_DENSE, ROUTE_DENSE_MFMA]
    assert fixed_routes(v, 24, mfma=True).tolist() == [ROUTE_DENSE] * 3
    v.val[:] = 0
    v.val[0] = 1.0                                   # density 1/128 < 15 %: sparse
    assert fixed_routes(v, 128, mfma=True)[0] == ROUTE_SPARSE
