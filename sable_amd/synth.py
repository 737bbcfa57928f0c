"""Seeded synthetic VBR workloads.

Two generators:

* ``vbr_matrix_gen_seeded`` keeps the law of the reference's generator
  (synthesize_matrices/vbr_matrices_gen.py:24-113): partition cut points drawn
  uniformly without replacement from [1, n) (``random_splits``, :24-34),
  ``k`` of the R*C blocks chosen without replacement, blocks stored in sorted
  order with -1/-1 for empty block rows (:80-101).  It uses numpy's PCG64
  instead of ``random.sample`` (the reference's stream is not reproduced).

* ``mixed_vbr`` builds the mixed dense/sparse matrices of the scale
  workloads: bounded partition sizes, a fraction of blocks stored, each
  stored block either "dense" (density drawn from one range) or "sparse"
  (another range) with exactly round(d*r*c) (at least 1) nonzeros at
  uniformly drawn positions, values N(0,1) or U[-1,1] in float32.

Every stream comes from ``numpy.random.SeedSequence(seed).spawn(..)`` so each
property (partitions, block choice, densities, positions, values, x) has its
own independent, reproducible sub-stream.

Bounded partitions: sizes are drawn uniformly in [lo, hi], then the surplus
over lo is rescaled (and clipped at hi) to make the sizes sum to at most n,
and the integer remainder is handed out one unit at a time to randomly
chosen parts that are still below hi.  This keeps the spread of sizes while hitting n exactly.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .vbr import VBR

_STREAMS = ("rows", "cols", "blocks", "density", "positions", "values", "x")


def _rngs(seed):
    ss = np.random.SeedSequence(seed)
    return {name: np.random.Generator(np.random.PCG64(child))
            for name, child in zip(_STREAMS, ss.spawn(len(_STREAMS)))}


def random_splits(n, parts, rng):
    """Partition sizes from ``parts-1`` distinct cut points in [1, n), the law
    of synthesize_matrices/vbr_matrices_gen.py:24-34 (sizes may be 1)."""
    if n <= 0 or parts <= 0 or parts > n:
        raise ValueError("need 0 < parts <= n")
    cuts = np.sort(rng.choice(np.arange(1, n), parts - 1, replace=False))
    return np.diff(np.concatenate([[0], cuts, [n]])).astype(np.int64)


def bounded_splits(n, parts, lo, hi, rng):
    """``parts`` sizes in [lo, hi] summing exactly to n (procedure in the
    module docstring)."""
    if not (parts * lo <= n <= parts * hi):
        raise ValueError(f"cannot split {n} into {parts} parts within [{lo},{hi}]")
    s = rng.integers(lo, hi + 1, size=parts).astype(np.int64)
    surplus_target = n - parts * lo
    surplus = int(s.sum()) - parts * lo
    if surplus > 0:
        s = lo + np.minimum(((s - lo) * surplus_target) // surplus, hi - lo)
    else:
        s = np.full(parts, lo, dtype=np.int64)
    rem = n - int(s.sum())
    while rem > 0:
        room = np.nonzero(s < hi)[0]
        take = rng.choice(room, size=min(rem, len(room)), replace=False)
        s[take] += 1
        rem -= len(take)
    assert int(s.sum()) == n and s.min() >= lo and s.max() <= hi
    return s


def cumsum_list(sizes):
    """[a, b, c] -> [0, a, a+b, a+b+c] (vbr_matrices_gen.py:13-22)."""
    return np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)


def _assemble(rpntr, cpntr, chosen, fill_block, val_dtype=np.float32):
    """Lay out the chosen blocks (sorted flat ids r*C+c) as VBR arrays;
    ``fill_block(k, h, w)`` returns the column-major values of block k."""
    nbr, nbc = len(rpntr) - 1, len(cpntr) - 1
    chosen = np.sort(np.asarray(chosen, dtype=np.int64))
    br, bc = chosen // nbc, chosen % nbc
    h = rpntr[br + 1] - rpntr[br]
    w = cpntr[bc + 1] - cpntr[bc]
    sizes = h * w
    indx = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    val = np.empty(int(indx[-1]), dtype=val_dtype)
    for k in range(len(chosen)):
        val[indx[k]:indx[k + 1]] = fill_block(k, int(h[k]), int(w[k]))
    counts = np.bincount(br, minlength=nbr)
    ends = np.cumsum(counts)
    bpntrb = np.where(counts > 0, ends - counts, -1).astype(np.int64)
    bpntre = np.where(counts > 0, ends, -1).astype(np.int64)
    return VBR(val, indx, bc.astype(np.int64), rpntr, cpntr, bpntrb, bpntre)


def _nonzero_normal(rng, n, dtype=np.float32):
    v = rng.standard_normal(n, dtype=np.float32).astype(dtype, copy=False)
    bad = v == 0
    while np.any(bad):
        v[bad] = rng.standard_normal(int(bad.sum()), dtype=np.float32)
        bad = v == 0
    return v


def _nonzero_uniform(rng, n, dtype=np.float32):
    v = rng.uniform(-1.0, 1.0, n).astype(dtype)
    bad = v == 0
    while np.any(bad):
        v[bad] = rng.uniform(-1.0, 1.0, int(bad.sum())).astype(dtype)
        bad = v == 0
    return v


def vbr_matrix_gen_seeded(m, n, row_split, col_split, num_blocks, seed,
                          partitioning="nonuniform", perc_zeros=0):
    """Reference-law generator (vbr_matrices_gen.py:36-113) with seeded
    numpy streams and U[-1,1] nonzero values instead of 1.0.  Returns
    (VBR, x)."""
    g = _rngs(seed)
    if partitioning == "nonuniform":
        rpntr = cumsum_list(random_splits(m, row_split, g["rows"]))
        cpntr = cumsum_list(random_splits(n, col_split, g["cols"]))
    elif partitioning == "uniform":
        rpntr = np.arange(0, m + 1, m // row_split, dtype=np.int64)
        cpntr = np.arange(0, n + 1, n // col_split, dtype=np.int64)
    else:
        raise ValueError(partitioning)
    chosen = g["blocks"].choice(row_split * col_split, num_blocks, replace=False)

    def fill(k, h, w):
        size = h * w
        v = _nonzero_uniform(g["values"], size)
        nzero = (size * perc_zeros) // 100
        if nzero:
            v[g["positions"].choice(size, nzero, replace=False)] = 0
        return v

    vbr = _assemble(rpntr, cpntr, chosen, fill)
    x = _nonzero_uniform(g["x"], n)
    return vbr, x


def mixed_vbr(n, parts, lo, hi, frac_blocks, frac_dense, dense_range,
              sparse_range, seed, value_law="normal", ncols=None, col_parts=None):
    """Mixed dense/sparse VBR (see module docstring).  Returns (VBR, x,
    is_dense_block) where is_dense_block marks the "dense" class (the
    generator's label, not a routing decision)."""
    g = _rngs(seed)
    ncols = ncols or n
    col_parts = col_parts or parts
    rpntr = cumsum_list(bounded_splits(n, parts, lo, hi, g["rows"]))
    cpntr = cumsum_list(bounded_splits(ncols, col_parts, lo, hi, g["cols"]))
    total = parts * col_parts
    k = int(round(frac_blocks * total))
    chosen = np.sort(g["blocks"].choice(total, k, replace=False))
    ndense = int(round(frac_dense * k))
    cls = np.zeros(k, dtype=bool)
    cls[g["blocks"].choice(k, ndense, replace=False)] = True
    dens = np.where(cls, g["density"].uniform(*dense_range, size=k),
                    g["density"].uniform(*sparse_range, size=k))
    vals_fn = _nonzero_normal if value_law == "normal" else _nonzero_uniform

    def fill(kk, h, w):
        size = h * w
        nnz = max(1, int(round(dens[kk] * size)))
        out = np.zeros(size, dtype=np.float32)
        if nnz >= size:
            out[:] = vals_fn(g["values"], size)
        else:
            pos = g["positions"].choice(size, nnz, replace=False)
            out[pos] = vals_fn(g["values"], nnz)
        return out

    vbr = _assemble(rpntr, cpntr, chosen, fill)
    x = _nonzero_uniform(g["x"], ncols)
    return vbr, x, cls


@dataclass(frozen=True)
class Workload:
    name: str
    seed: int
    description: str


WORKLOADS = {
    "c1": Workload("c1", 1, "1000x1000 VBR, 20x20 uniform-random partitions, "
                   "40 of 400 blocks, fully dense, U[-1,1]"),
    "c2": Workload("c2", 2, "100000x100000 VBR, 1000x1000 partitions sized "
                   "20-200, 1% blocks, half density 0.8-1.0 / half 0.01-0.2, N(0,1)"),
    "c3": Workload("c3", 3, "50000x50000 VBR, 500x500 partitions sized 32-192, "
                   "2% blocks, 70% density 0.9-1.0 / 30% 0.02-0.1, N(0,1), N=128"),
    "c5": Workload("c5", 5, "256 VBR matrices, n log-uniform 2000-64000, P uniform 50-800 "
                   "partitions sized [ceil(n/5P), floor(2n/P)], 1% blocks, density mix of c2, "
                   "SpMV and SpMM with N=16"),
}


def workload_c1(seed=1):
    """1000x1000, 19 uniform distinct cut points per axis, 40 of 400 blocks,
    fully dense, values and x U[-1,1]."""
    return vbr_matrix_gen_seeded(1000, 1000, 20, 20, 40, seed)


def workload_c2(seed=2, n=100_000, parts=1000):
    """100k x 100k, 1000x1000 partitions in [20,200], 1% blocks, half dense
    (0.8-1.0), half sparse (0.01-0.2), values N(0,1), x U[-1,1]."""
    return mixed_vbr(n, parts, 20, 200, 0.01, 0.5, (0.8, 1.0), (0.01, 0.2), seed)


def workload_c3(seed=3, n=50_000, parts=500, n_rhs=128):
    """50k x 50k, 500x500 partitions in [32,192], 2% blocks, 70% dense
    (0.9-1.0), 30% sparse (0.02-0.1), values N(0,1); X [n x n_rhs] U[-1,1]."""
    vbr, _, cls = mixed_vbr(n, parts, 32, 192, 0.02, 0.7, (0.9, 1.0), (0.02, 0.1), seed)
    g = _rngs(seed + 1000)
    X = g["x"].uniform(-1.0, 1.0, (n, n_rhs)).astype(np.float32)
    return vbr, X, cls


def c5_shapes(count=256, seed=5, n_range=(2000, 64000), p_range=(50, 800)):
    """(n, P) of the batched workload's matrices: n log-uniform integer in
    n_range, P uniform integer in p_range (at most n)."""
    g = np.random.Generator(np.random.PCG64(np.random.SeedSequence([seed, 0])))
    lo, hi = np.log(n_range[0]), np.log(n_range[1] + 1)
    ns = np.floor(np.exp(g.uniform(lo, hi, count))).astype(np.int64)
    ns = np.clip(ns, n_range[0], n_range[1])
    ps = g.integers(p_range[0], p_range[1] + 1, count)
    return [(int(n), int(min(p, n))) for n, p in zip(ns, ps)]


def workload_c5_matrix(i, n, parts, seed=5, n_rhs=16):
    """Matrix i of the batched workload: partition sizes in [ceil(n/5P),
    floor(2n/P)], 1% of blocks, the 100k workload's density mix; x U[-1,1]
    and an n x n_rhs row-major X U[-1,1]."""
    lo = max(1, -(-n // (5 * parts)))
    hi = max(lo, (2 * n) // parts)
    vbr, x, _ = mixed_vbr(n, parts, lo, hi, 0.01, 0.5, (0.8, 1.0), (0.01, 0.2), seed * 100_003 + i)
    g = np.random.Generator(np.random.PCG64(np.random.SeedSequence([seed, 1, i])))
    X = g.uniform(-1.0, 1.0, (n, n_rhs)).astype(np.float32)
    return vbr, x, X


def workload_c5(count=256, seed=5, n_rhs=16, **kw):
    """Yields (vbr, x, X) for the batched workload's matrices, one at a time."""
    for i, (n, p) in enumerate(c5_shapes(count, seed, **kw)):
        yield workload_c5_matrix(i, n, p, seed, n_rhs)


def workload_like_c2(n, seed, frac_blocks=0.01):
    """A matrix of size n generated like the 100k workload: partition sizes in
    [20, 200] with n/100 parts, 1% blocks, the same density mix."""
    parts = max(1, n // 100)
    return mixed_vbr(n, parts, 20, 200, frac_blocks, 0.5, (0.8, 1.0), (0.01, 0.2), seed)
