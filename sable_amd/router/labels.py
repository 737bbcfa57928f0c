"""Router training data: seeded synthetic blocks timed on the GPU kernels.

Blocks follow the router workload law: rows and cols log-uniform integers in
[4, 512] (independent, so rectangular), density log-uniform in [0.005, 1],
exactly max(1, round(d*r*c)) nonzeros at uniform positions, N uniform in
{1, 8, 64, 256}.

Labels are per-block costs measured by the kernels themselves.  Blocks are
grouped by (log2 rows, log2 cols, density bin, N); for every group, every
tile shape (the launch's workgroup size, SHAPES) and every route that
applies (dense FP32, sparse, and the MFMA route when N is a multiple of 16)
one block-diagonal matrix holding the group's blocks
(replicated until the launch streams at least ``min_bytes``) is timed with
HIP events, starting cold (a 300 MB buffer larger than the last-level cache
is overwritten before each launch).  The per-block cost is (median launch
time - empty-launch time) / number of blocks; every block of the group gets
its group's costs as labels, and the network learns the cost surface from
the blocks' exact features.
"""
from __future__ import annotations

import numpy as np

from .. import convert
from ..vbr import VBR
from .features import block_features, encode

N_CHOICES = (1, 8, 64, 256)
DENSITY_EDGES = np.array([0.005, 0.0125, 0.03, 0.08, 0.2, 0.5, 1.0000001])
ROUTES = (convert.ROUTE_DENSE, convert.ROUTE_SPARSE, convert.ROUTE_DENSE_MFMA)
# tile shapes the router chooses between: the workgroup size of the launch
# (narrow, wide), per operation.  Narrow workgroups suit short tiles (small
# blocks, small matrices); wide ones long streams.
SHAPES = {"spmv": (256, 512), "spmm": (128, 256)}


def shape_wg(op, s):
    return SHAPES["spmv" if op == "spmv" else "spmm"][s]


def sample_blocks(n_blocks, seed):
    """(r, c, density, N) of the training blocks."""
    rng = np.random.default_rng(np.random.SeedSequence(seed))
    r = np.floor(np.exp(rng.uniform(np.log(4), np.log(513), n_blocks))).astype(np.int64)
    c = np.floor(np.exp(rng.uniform(np.log(4), np.log(513), n_blocks))).astype(np.int64)
    d = np.exp(rng.uniform(np.log(0.005), np.log(1.0), n_blocks))
    N = rng.choice(N_CHOICES, n_blocks)
    return np.clip(r, 4, 512), np.clip(c, 4, 512), d, N


def group_keys(r, c, d, N):
    db = np.searchsorted(DENSITY_EDGES, d, side="right") - 1
    return np.stack([np.floor(np.log2(r)).astype(np.int64), np.floor(np.log2(c)).astype(np.int64),
                     db, np.asarray(N, dtype=np.int64)], axis=1)


def block_diagonal(rs, cs, ds, rng):
    """A VBR with one block per block row: block i is rs[i] x cs[i] with
    max(1, round(d*r*c)) N(0,1) nonzeros, placed at block column i."""
    rs = np.asarray(rs, dtype=np.int64)
    cs = np.asarray(cs, dtype=np.int64)
    rpntr = np.concatenate([[0], np.cumsum(rs)])
    cpntr = np.concatenate([[0], np.cumsum(cs)])
    sizes = rs * cs
    indx = np.concatenate([[0], np.cumsum(sizes)])
    val = np.zeros(int(indx[-1]), dtype=np.float32)
    for k in range(len(rs)):
        n = int(sizes[k])
        nnz = max(1, int(round(ds[k] * n)))
        if nnz >= n:
            v = rng.standard_normal(n).astype(np.float32)
            v[v == 0] = 1.0
            val[indx[k]:indx[k + 1]] = v
        else:
            pos = rng.choice(n, nnz, replace=False)
            v = rng.standard_normal(nnz).astype(np.float32)
            v[v == 0] = 1.0
            val[indx[k] + pos] = v
    nb = len(rs)
    ar = np.arange(nb, dtype=np.int64)
    return VBR(val, indx, ar, rpntr, cpntr, ar, ar + 1)


def measure_groups(r, c, d, N, seed=0, min_bytes=6 << 20, max_blocks=4096, reps=5, progress=None):
    """Times every (group, shape, route).  Returns (keys, costs_ns [G, 6]
    (column shape * 3 + route) with NaN where a route does not apply, group
    index of every block)."""
    from ..runtime import Plan, DeviceArray
    keys = group_keys(r, c, d, N)
    uniq, inv = np.unique(keys, axis=0, return_inverse=True)
    inv = inv.ravel()
    rng = np.random.default_rng(seed)
    ns = len(SHAPES["spmv"])
    costs = np.full((len(uniq), ns * len(ROUTES)), np.nan)
    # empty-launch overhead per op and shape: one tiny block
    base = {}
    tiny = block_diagonal([4], [4], [1.0], rng)
    for n in N_CHOICES:
        op = "spmv" if n == 1 else "spmm"
        xd = DeviceArray.from_numpy(np.ones((4,) if n == 1 else (4, n), np.float32))
        yd = DeviceArray((4,) if n == 1 else (4, n))
        for s in range(ns):
            p = Plan.from_vbr(tiny, [0], op=op, n_rhs=n, wg=shape_wg(op, s))
            base[n, s] = float(np.median(p.time(xd, yd, warmup=3, reps=reps, flush_bytes=300 << 20)))
            p.close()
    for g in range(len(uniq)):
        members = np.nonzero(inv == g)[0]
        n = int(uniq[g, 3])
        rs, cs, ds = r[members], c[members], d[members]
        per = rs * cs * 4 + rs * n * 4
        reps_needed = int(np.clip(np.ceil(min_bytes / max(per.sum(), 1)), 1, 10 ** 6))
        idx = np.tile(np.arange(len(members)), reps_needed)[:max(len(members), min(max_blocks, len(members) * reps_needed))]
        vbr = block_diagonal(rs[idx], cs[idx], ds[idx], rng)
        count = len(idx)
        op = "spmv" if n == 1 else "spmm"
        ncols, nrows = int(vbr.cpntr[-1]), int(vbr.rpntr[-1])
        X = rng.uniform(-1, 1, (ncols,) if n == 1 else (ncols, n)).astype(np.float32)
        xd = DeviceArray.from_numpy(X)
        yd = DeviceArray((nrows,) if n == 1 else (nrows, n))
        for s in range(ns):
            for ri, route in enumerate(ROUTES):
                if route == convert.ROUTE_DENSE_MFMA and (n == 1 or n % 16):
                    continue
                plan = Plan.from_vbr(vbr, np.full(count, route, np.uint8), op=op, n_rhs=n, wg=shape_wg(op, s))
                t = float(np.median(plan.time(xd, yd, warmup=2, reps=reps, flush_bytes=300 << 20)))
                plan.close()
                costs[g, s * len(ROUTES) + ri] = max(t - base[n, s], 1e-6) * 1e6 / count     # ns per block
        xd.free()
        yd.free()
        if progress and (g % 50 == 0 or g == len(uniq) - 1):
            progress(f"labels: group {g + 1}/{len(uniq)}")
    return uniq, costs, inv


def training_set(n_blocks=200_000, seed=4, **kw):
    """Features (encoded), log2-cost targets and masks for n_blocks blocks."""
    r, c, d, N = sample_blocks(n_blocks, seed)
    keys, costs, inv = measure_groups(r, c, d, N, seed=seed, **kw)
    nnz = np.maximum(1, np.round(d * r * c))
    F = np.stack([r, c, nnz, nnz / (r * c), d * (1 - d) * c, N], axis=1).astype(np.float64)
    # row-length variance of a uniform-position block: binomial c*d*(1-d)
    E = encode(F)
    C = costs[inv]
    M = np.isfinite(C).astype(np.float32)
    T = np.where(M > 0, np.log2(np.where(np.isfinite(C), C, 1.0)), 0.0)
    return E, T.astype(np.float32), M, dict(keys=keys, costs=costs)
