"""Per-block features of the router (vectorised over all blocks).

rows, cols, nnz (values != 0.0, the reference's routing count), density
(nnz / (rows*cols)), the variance of the per-row nonzero counts inside the
block, and the RHS width N.  The network sees them log-scaled and
standardised (``encode``).
"""
from __future__ import annotations

import numpy as np

from ..vbr import block_table

FEATURE_NAMES = ("rows", "cols", "nnz", "density", "row_len_var", "n_rhs")


def block_features(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, n_rhs):
    """Returns a float64 array [nblocks, 6] in FEATURE_NAMES order."""
    brow, bcol, r0, h, c0, w = block_table(rpntr, cpntr, bindx, bpntrb, bpntre)
    val = np.asarray(val)
    ix = np.asarray(indx, dtype=np.int64)
    nb = len(h)
    sizes = h * w
    nzmask = val != 0
    csum = np.concatenate([[0], np.cumsum(nzmask, dtype=np.int64)])
    nnz = csum[ix[1:]] - csum[ix[:-1]]
    # per-row nonzero counts: element e of block k lies in local row e % h[k]
    blk = np.repeat(np.arange(nb), sizes)
    local = np.arange(len(val), dtype=np.int64) - ix[blk]
    row_id = (np.concatenate([[0], np.cumsum(h)])[:-1])[blk] + local % h[blk]
    counts = np.bincount(row_id[nzmask], minlength=int(h.sum())).astype(np.float64)
    rstart = np.concatenate([[0], np.cumsum(h)])
    s1 = np.add.reduceat(counts, rstart[:-1]) if nb else np.zeros(0)
    s2 = np.add.reduceat(counts * counts, rstart[:-1]) if nb else np.zeros(0)
    mean = s1 / h
    var = np.maximum(s2 / h - mean * mean, 0.0)
    F = np.empty((nb, 6), dtype=np.float64)
    F[:, 0] = h
    F[:, 1] = w
    F[:, 2] = nnz
    F[:, 3] = nnz / sizes
    F[:, 4] = var
    F[:, 5] = n_rhs
    return F


def encode(F):
    """Log-scale the size-like features (density kept linear)."""
    E = np.empty_like(F, dtype=np.float64)
    E[:, 0] = np.log2(F[:, 0])
    E[:, 1] = np.log2(F[:, 1])
    E[:, 2] = np.log2(F[:, 2] + 1.0)
    E[:, 3] = F[:, 3]
    E[:, 4] = np.log2(F[:, 4] + 1.0)
    E[:, 5] = np.log2(F[:, 5])
    return E
