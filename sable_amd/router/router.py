"""Route selection: the learned MLP router and the keep-if-faster rule.

``Router.route_and_shape`` predicts the per-block cost of every applicable route
(dense FP32, sparse, and MFMA when the RHS width allows it) under every
tile shape, and picks the cheapest.  ``choose_plan`` builds the fixed-threshold plan (the fallback,
threshold 15 as in the reference's mixed flows) and the learned plans, checks
them against the same FP64 tolerance, times them cold, and keeps the fastest
learned plan only if it beats the fixed one.

The training labels are the costs of blocks run in isolation.  Inside a tile a
block shares the x staging and the tile's one dense stream, so its marginal
cost differs from the isolated one, and differently per route.  A route bias
(log2 ns subtracted from that route's predicted cost; 0.5 makes the route win
while its prediction is below 2^0.5 times the cheapest other) corrects for
that; ``choose_plan`` tries the small grid ``BIAS_GRID`` and keeps the
fastest.  On the 100k x 100k mixed SpMV a sparse bias of 0.5 gave 49.4 us
against 51.2 us at 0; on the 50k SpMM with N = 128 an MFMA bias of 1.0 (every
block on MFMA) gave 0.171 ms against 0.181 ms.
"""
from __future__ import annotations

import os

import numpy as np

from .. import convert
from ..vbr import block_table
from .features import block_features, encode

ROUTES = (convert.ROUTE_DENSE, convert.ROUTE_SPARSE, convert.ROUTE_DENSE_MFMA)
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights.npz")

# (sparse_bias, mfma_bias) candidates in log2 ns, per operation
BIAS_GRID = {"spmv": ((0.0, 0.0), (0.5, 0.0), (1.0, 0.0)),
             "spmm": ((0.0, 0.0), (0.0, 0.5), (0.0, 1.0))}


# MFMA tiles are 16 rows by 32 columns.  The fixed MFMA policy puts a dense
# block on MFMA only when both of its edges reach 16; the router may put a
# smaller block there too when its predicted cost (the padding included, as
# the labels were timed on the padded tiles) is the lowest: on the
# 256-matrix batch, whose smaller matrices have blocks of a few rows, the
# edge rule in the router made the SpMM launches 1.5x slower.
MFMA_MIN_EDGE = 16


def mfma_edges_ok(rows, cols):
    return (np.asarray(rows) >= MFMA_MIN_EDGE) & (np.asarray(cols) >= MFMA_MIN_EDGE)


def allowed_mask(n_rhs, nblocks):
    """[nblocks, 3] routes each block may take: MFMA needs N a multiple of 16
    up to 256."""
    m = np.ones((nblocks, 3), dtype=bool)
    if n_rhs == 1 or n_rhs % 16 or n_rhs > 256:
        m[:, 2] = False
    return m


class Router:
    """A trained MLP plus the feature standardisation it was trained with.
    The network predicts, per block, the log2 ns cost of every route under
    every tile shape (output shape * 3 + route; labels.SHAPES)."""

    def __init__(self, params, mean, std, meta=""):
        from .mlp import DeviceMLP, N_ROUTES
        self.mlp = DeviceMLP(params)
        self.mean = np.asarray(mean, dtype=np.float64)
        self.std = np.asarray(std, dtype=np.float64)
        self.meta = meta
        self.n_shapes = max(1, self.mlp.K // N_ROUTES)

    @classmethod
    def load(cls, path=DEFAULT_WEIGHTS):
        from .mlp import load
        params, mean, std, meta = load(path)
        return cls(params, mean, std, meta)

    @classmethod
    def available(cls, path=DEFAULT_WEIGHTS):
        return os.path.exists(path)

    def predict_costs(self, F):
        """log2 ns per block, [n, shapes * 3] (column shape * 3 + route)."""
        E = (encode(F) - self.mean) / self.std
        return self.mlp.predict(E.astype(np.float32))

    def _costs(self, F, n_rhs, sparse_bias, mfma_bias):
        pred = self.predict_costs(F).astype(np.float64).reshape(len(F), self.n_shapes, 3)
        pred[:, :, 1] -= float(sparse_bias)
        pred[:, :, 2] -= float(mfma_bias)
        pred[np.broadcast_to(~allowed_mask(n_rhs, len(F))[:, None, :], pred.shape)] = np.inf
        return pred

    def route_and_shape(self, val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, n_rhs=1, sparse_bias=0.0,
                        mfma_bias=0.0, shape=None):
        """The tile shape and the route of every block: per shape, each block
        takes its cheapest allowed route (``sparse_bias`` / ``mfma_bias``,
        log2 ns, are subtracted from the predicted costs of the sparse and
        MFMA routes first); the plan takes the shape whose blocks' predicted
        costs (2^log2) add up to the least, unless ``shape`` fixes it.
        Returns (routes, shape index, workgroup size)."""
        from .labels import shape_wg
        F = block_features(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, n_rhs)
        op = "spmv" if n_rhs == 1 else "spmm"
        if len(F) == 0:
            s = 0 if shape is None else int(shape)
            return np.zeros(0, dtype=np.uint8), s, shape_wg(op, s)
        pred = self._costs(F, n_rhs, sparse_bias, mfma_bias)
        best = np.argmin(pred, axis=2)                            # [n, shapes]
        if shape is None:
            tot = np.sum(np.exp2(np.take_along_axis(pred, best[:, :, None], axis=2)[:, :, 0]), axis=0)
            s = int(np.argmin(tot))
        else:
            s = int(shape)
        routes = np.asarray(ROUTES, dtype=np.uint8)[best[:, s]]
        return routes, s, shape_wg(op, s)

    def routes_per_shape(self, val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, n_rhs=1, sparse_bias=0.0,
                         mfma_bias=0.0):
        """The routes under every tile shape and the shape the router picks
        (the one route_and_shape returns): ([(routes, workgroup size) per
        shape], picked shape index)."""
        from .labels import shape_wg
        F = block_features(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, n_rhs)
        op = "spmv" if n_rhs == 1 else "spmm"
        if len(F) == 0:
            return [(np.zeros(0, dtype=np.uint8), shape_wg(op, s)) for s in range(self.n_shapes)], 0
        pred = self._costs(F, n_rhs, sparse_bias, mfma_bias)
        best = np.argmin(pred, axis=2)
        tot = np.sum(np.exp2(np.take_along_axis(pred, best[:, :, None], axis=2)[:, :, 0]), axis=0)
        table = np.asarray(ROUTES, dtype=np.uint8)
        return [(table[best[:, s]], shape_wg(op, s)) for s in range(self.n_shapes)], int(np.argmin(tot))

    def route_blocks(self, val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, n_rhs=1, sparse_bias=0.0,
                     mfma_bias=0.0, shape=None):
        """The cheapest allowed route per block (under the chosen shape, see
        route_and_shape)."""
        return self.route_and_shape(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, n_rhs, sparse_bias,
                                    mfma_bias, shape)[0]


def fixed_routes(vbr, n_rhs=1, threshold=convert.DEFAULT_DENSITY, mfma=False):
    r = convert.threshold_routes(vbr.val, vbr.indx, threshold)
    if mfma and n_rhs > 1 and n_rhs % 16 == 0 and n_rhs <= 256:
        _, _, _, h, _, w = block_table(vbr.rpntr, vbr.cpntr, vbr.bindx, vbr.bpntrb, vbr.bpntre)
        r = np.where((r == convert.ROUTE_DENSE) & mfma_edges_ok(h, w), convert.ROUTE_DENSE_MFMA, r).astype(np.uint8)
    return r


def choose_plan(vbr, x, router=None, n_rhs=1, threshold=convert.DEFAULT_DENSITY, tol=1e-5, reps=20,
                flush_bytes=1 << 30, biases=None, **plan_kw):
    """Builds the fixed-threshold plan (for SpMM with the dense blocks on
    MFMA when N allows it) and, when a router is available, the learned
    plans: per (sparse_bias, mfma_bias) in ``biases`` (default
    ``BIAS_GRID[op]``), the router's routes under every tile shape, plus
    the routes of the shape the router picks in the plan's default
    workgroup size when no shape has it (candidates giving the same routes
    and workgroup size are built once); returns (plan, report).  The
    router's per-block costs do not see a plan's tile count, so its shape
    pick is timed beside the others, not trusted (``router_pick`` marks it
    in the report).  The fastest learned plan is kept only if its error
    against the FP64 oracle is within ``tol`` like the fixed plan and its
    cold median launch time is lower."""
    from ..runtime import Plan, DeviceArray
    from ..autotune import tuned_options
    op = "spmv" if n_rhs == 1 else "spmm"
    xd = DeviceArray.from_numpy(np.asarray(x, dtype=np.float32))
    shape_y = (int(vbr.rpntr[-1]),) if n_rhs == 1 else (int(vbr.rpntr[-1]), n_rhs)
    yd = DeviceArray(shape_y)

    def run(routes, wg=None):
        kw = dict(plan_kw)
        if wg is not None:
            kw["wg"] = wg
        p = Plan.from_vbr(vbr, routes, op=op, n_rhs=n_rhs, **kw)
        t = float(np.median(p.time(xd, yd, warmup=3, reps=reps, flush_bytes=flush_bytes)))
        y = yd.numpy()
        return p, t, y

    # the best fixed policy: threshold routes, dense blocks on MFMA where N allows
    fixed = fixed_routes(vbr, n_rhs, threshold, mfma=True)
    p_fixed, t_fixed, y_fixed = run(fixed)
    report = {"fixed_ms": t_fixed, "kept": "fixed"}
    if router is None:
        return p_fixed, report
    fixed_ok = _within(vbr, x, y_fixed, n_rhs, tol)
    default_wg = plan_kw.get("wg") or tuned_options(op, n_rhs).get("wg") or 256
    best = None
    seen = []
    report["candidates"] = []
    for sb, mb in (BIAS_GRID[op] if biases is None else biases):
        shapes, pick = router.routes_per_shape(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb,
                                               vbr.bpntre, n_rhs, sparse_bias=sb, mfma_bias=mb)
        cands = [(r, w, s == pick) for s, (r, w) in enumerate(shapes)]
        if all(w != default_wg for _, w in shapes):
            cands.append((shapes[pick][0], default_wg, False))
        for learned, wg, is_pick in cands:
            if any(np.array_equal(learned, r) and wg == w for r, w in seen):
                continue
            seen.append((learned, wg))
            p, t, y = run(learned, wg)
            ok = fixed_ok and _within(vbr, x, y, n_rhs, tol)
            report["candidates"].append({"sparse_bias": sb, "mfma_bias": mb, "wg": wg, "router_pick": bool(is_pick),
                                         "ms": t, "within_tol": bool(ok),
                                         "routes": {int(k): int(v) for k, v in
                                                    zip(*np.unique(learned, return_counts=True))}})
            if ok and (best is None or t < best[1]):
                if best is not None:
                    best[0].close()
                best = (p, t, sb, mb, learned, wg, is_pick)
            else:
                p.close()
    report["learned_within_tol"] = best is not None
    if best is not None:
        report["learned_ms"] = best[1]
        report["sparse_bias"], report["mfma_bias"], report["wg"] = best[2], best[3], best[5]
        report["router_pick"] = bool(best[6])
        report["learned_routes"] = {int(k): int(v) for k, v in zip(*np.unique(best[4], return_counts=True))}
        if best[1] < t_fixed:
            report["kept"] = "learned"
            p_fixed.close()
            return best[0], report
        best[0].close()
    return p_fixed, report


def _within(vbr, x, y, n_rhs, tol):
    """Componentwise-scaled error against FP64, computed block by block."""
    rp = np.asarray(vbr.rpntr, dtype=np.int64)
    cp = np.asarray(vbr.cpntr, dtype=np.int64)
    bx = np.asarray(vbr.bindx, dtype=np.int64)
    bb = np.asarray(vbr.bpntrb, dtype=np.int64)
    be = np.asarray(vbr.bpntre, dtype=np.int64)
    ix = np.asarray(vbr.indx, dtype=np.int64)
    X = np.asarray(x, dtype=np.float64).reshape(int(cp[-1]), -1)
    Y = np.zeros((int(rp[-1]), X.shape[1]))
    AY = np.zeros_like(Y)
    v = np.asarray(vbr.val)
    for a in range(len(rp) - 1):
        if bb[a] == -1:
            continue
        r0, r1 = rp[a], rp[a + 1]
        for k in range(bb[a], be[a]):
            c0, c1 = cp[bx[k]], cp[bx[k] + 1]
            blk = v[ix[k]:ix[k + 1]].astype(np.float64).reshape(c1 - c0, r1 - r0).T
            Y[r0:r1] += blk @ X[c0:c1]
            AY[r0:r1] += np.abs(blk) @ np.abs(X[c0:c1])
    d = np.abs(np.asarray(y, dtype=np.float64).reshape(Y.shape) - Y)
    m = AY > 0
    if np.any(~m & (d > 0)):
        return False
    return bool(np.all(d[m] <= tol * AY[m]))
