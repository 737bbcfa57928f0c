"""The Variable Block Row container (reference: src/vbr.py:5-14).

Same seven attributes and constructor order as the reference's ``VBR``; the
additions are ``validate()`` (the reference checks nothing) and a few derived
views used by the plan builder.

Layout conventions (shared with every producer in the reference):

* ``rpntr``/``cpntr``: partition boundaries, length P+1, start at 0, strictly
  increasing, end at the matrix size.
* ``bindx[bpntrb[a]:bpntre[a]]``: block columns of the stored blocks of block
  row ``a``, ascending; ``bpntrb[a] == bpntre[a] == -1`` marks an empty block
  row (synthesize_matrices/vbr_matrices_gen.py:80-101).
* ``indx``: prefix offsets into ``val``, one entry per stored block plus one.
* ``val``: every stored block densely, column-major.

The reference enumerates blocks by block row, then by ascending block column
(utils/convert_real_to_vbr.py:28-33, src/codegen.py:421-427) while reading
values at ``indx[count]``; that equals storage order only when rows are laid
out consecutively with ascending ``bindx``.  ``validate()`` enforces exactly
that, so ordinal ``k`` and storage position ``k`` coincide.
"""
from __future__ import annotations

import numpy as np


class VBR:
    """VBR matrix; constructor signature of src/vbr.py:7."""

    def __init__(self, val, indx, bindx, rpntr, cpntr, bpntrb, bpntre):
        self.val = val
        self.indx = indx
        self.bindx = bindx
        self.rpntr = rpntr
        self.cpntr = cpntr
        self.bpntrb = bpntrb
        self.bpntre = bpntre

    # -- derived -----------------------------------------------------------
    @property
    def shape(self):
        return int(self.rpntr[-1]), int(self.cpntr[-1])

    @property
    def nblocks(self):
        return len(self.bindx)

    def arrays(self):
        """The seven arrays as numpy (val keeps its dtype; indices int64)."""
        return (np.asarray(self.val), np.asarray(self.indx, dtype=np.int64),
                np.asarray(self.bindx, dtype=np.int64),
                np.asarray(self.rpntr, dtype=np.int64),
                np.asarray(self.cpntr, dtype=np.int64),
                np.asarray(self.bpntrb, dtype=np.int64),
                np.asarray(self.bpntre, dtype=np.int64))

    def validate(self):
        validate_vbr(self.val, self.indx, self.bindx, self.rpntr, self.cpntr,
                     self.bpntrb, self.bpntre)
        return self

    def block_table(self):
        return block_table(self.rpntr, self.cpntr, self.bindx, self.bpntrb,
                           self.bpntre)


def validate_vbr(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre):
    """Reject malformed VBR arrays with ValueError.

    Checks: partitions start at 0 and strictly increase; ``bpntrb``/``bpntre``
    have one entry per block row, use -1/-1 only together, and lay the
    non-empty rows out consecutively from 0; ``bindx`` is strictly ascending
    inside each row and within range; ``indx`` starts at 0, has one entry per
    block plus one, and each step equals the block's height x width;
    ``len(val) == indx[-1]``.
    """
    rp = np.asarray(rpntr, dtype=np.int64)
    cp = np.asarray(cpntr, dtype=np.int64)
    bx = np.asarray(bindx, dtype=np.int64)
    bb = np.asarray(bpntrb, dtype=np.int64)
    be = np.asarray(bpntre, dtype=np.int64)
    ix = np.asarray(indx, dtype=np.int64)
    if rp.ndim != 1 or cp.ndim != 1 or len(rp) < 2 or len(cp) < 2:
        raise ValueError("rpntr/cpntr must be 1-D with at least two entries")
    if rp[0] != 0 or cp[0] != 0:
        raise ValueError("rpntr/cpntr must start at 0")
    if np.any(np.diff(rp) <= 0) or np.any(np.diff(cp) <= 0):
        raise ValueError("rpntr/cpntr must be strictly increasing")
    nbr, nbc = len(rp) - 1, len(cp) - 1
    if len(bb) != nbr or len(be) != nbr:
        raise ValueError("bpntrb/bpntre need one entry per block row "
                         f"({len(bb)}, {len(be)} vs {nbr})")
    empty = bb == -1
    if np.any((be == -1) != empty):
        raise ValueError("bpntrb/bpntre must mark empty block rows with -1 together")
    counts = np.where(empty, 0, be - bb)
    if np.any(counts < 0):
        raise ValueError("bpntre < bpntrb")
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    if np.any(bb[~empty] != starts[~empty]):
        raise ValueError("block rows must be stored consecutively in row order")
    nb = int(counts.sum())
    if len(bx) != nb:
        raise ValueError(f"bindx has {len(bx)} entries, rows describe {nb}")
    if nb:
        if bx.min() < 0 or bx.max() >= nbc:
            raise ValueError("bindx out of range")
        row_of = np.repeat(np.arange(nbr), counts)
        same_row = row_of[1:] == row_of[:-1]
        if np.any(same_row & (bx[1:] <= bx[:-1])):
            raise ValueError("bindx must be strictly ascending within a block row")
    if len(ix) != nb + 1 or ix[0] != 0:
        raise ValueError("indx must have nblocks+1 entries starting at 0")
    if nb:
        row_of = np.repeat(np.arange(nbr), counts)
        sizes = (rp[row_of + 1] - rp[row_of]) * (cp[bx + 1] - cp[bx])
        if np.any(np.diff(ix) != sizes):
            raise ValueError("indx steps must equal block height x width")
    if len(val) != ix[-1]:
        raise ValueError(f"len(val)={len(val)} but indx[-1]={int(ix[-1])}")


def block_table(rpntr, cpntr, bindx, bpntrb, bpntre):
    """Per stored block (ordinal order): block row, block col, r0, h, c0, w.

    Assumes a validated VBR (ordinal == storage order).
    """
    rp = np.asarray(rpntr, dtype=np.int64)
    cp = np.asarray(cpntr, dtype=np.int64)
    bx = np.asarray(bindx, dtype=np.int64)
    bb = np.asarray(bpntrb, dtype=np.int64)
    be = np.asarray(bpntre, dtype=np.int64)
    counts = np.where(bb == -1, 0, be - bb)
    brow = np.repeat(np.arange(len(rp) - 1), counts)
    r0 = rp[brow]
    h = rp[brow + 1] - r0
    c0 = cp[bx]
    w = cp[bx + 1] - c0
    return brow, bx, r0, h, c0, w
