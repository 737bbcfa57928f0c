"""Route selection: the learned MLP router and the keep-if-faster rule.

``Router.route_blocks`` predicts the per-block cost of every applicable route
(dense FP32, sparse, and MFMA when the RHS width allows it) and picks the
cheapest.  ``choose_plan`` builds the fixed-threshold plan (the fallback,
threshold 15 as in the reference's mixed flows) and the learned plan, checks
both against the same FP64 tolerance, times both cold, and keeps the learned
plan only if it is faster.
"""
from __future__ import annotations

import os

import numpy as np

from .. import convert
from .features import block_features, encode

ROUTES = (convert.ROUTE_DENSE, convert.ROUTE_SPARSE, convert.ROUTE_DENSE_MFMA)
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights.npz")


def allowed_mask(n_rhs, nblocks):
    m = np.ones((nblocks, 3), dtype=bool)
    if n_rhs == 1 or n_rhs % 16 or n_rhs > 256:
        m[:, 2] = False
    return m


class Router:
    """A trained MLP plus the feature standardisation it was trained with."""

    def __init__(self, params, mean, std, meta=""):
        from .mlp import DeviceMLP
        self.mlp = DeviceMLP(params)
        self.mean = np.asarray(mean, dtype=np.float64)
        self.std = np.asarray(std, dtype=np.float64)
        self.meta = meta

    @classmethod
    def load(cls, path=DEFAULT_WEIGHTS):
        from .mlp import load
        params, mean, std, meta = load(path)
        return cls(params, mean, std, meta)

    @classmethod
    def available(cls, path=DEFAULT_WEIGHTS):
        return os.path.exists(path)

    def predict_costs(self, F):
        """log2 ns per block for each route, [n, 3]."""
        E = (encode(F) - self.mean) / self.std
        return self.mlp.predict(E.astype(np.float32))

    def route_blocks(self, val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, n_rhs=1):
        F = block_features(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, n_rhs)
        if len(F) == 0:
            return np.zeros(0, dtype=np.uint8)
        pred = self.predict_costs(F).astype(np.float64)
        pred[~allowed_mask(n_rhs, len(F))] = np.inf
        return np.asarray(ROUTES, dtype=np.uint8)[np.argmin(pred, axis=1)]


def fixed_routes(vbr, n_rhs=1, threshold=convert.DEFAULT_DENSITY, mfma=False):
    r = convert.threshold_routes(vbr.val, vbr.indx, threshold)
    if mfma and n_rhs > 1 and n_rhs % 16 == 0 and n_rhs <= 256:
        r = np.where(r == convert.ROUTE_DENSE, convert.ROUTE_DENSE_MFMA, r).astype(np.uint8)
    return r


def choose_plan(vbr, x, router=None, n_rhs=1, threshold=convert.DEFAULT_DENSITY, tol=1e-5, reps=20,
                flush_bytes=1 << 30, **plan_kw):
    """Builds the fixed-threshold plan (for SpMM with the dense blocks on
    MFMA when N allows it) and, when a router is available, the learned
    plan; returns (plan, report).  The learned plan is kept only if its
    error against the FP64 oracle is within ``tol`` like the fixed plan and
    its cold median launch time is lower."""
    from ..runtime import Plan, DeviceArray
    op = "spmv" if n_rhs == 1 else "spmm"
    xd = DeviceArray.from_numpy(np.asarray(x, dtype=np.float32))
    shape_y = (int(vbr.rpntr[-1]),) if n_rhs == 1 else (int(vbr.rpntr[-1]), n_rhs)
    yd = DeviceArray(shape_y)

    def run(routes):
        p = Plan.from_vbr(vbr, routes, op=op, n_rhs=n_rhs, **plan_kw)
        t = float(np.median(p.time(xd, yd, warmup=3, reps=reps, flush_bytes=flush_bytes)))
        y = yd.numpy()
        return p, t, y

    # the best fixed policy: threshold routes, dense blocks on MFMA where N allows
    fixed = fixed_routes(vbr, n_rhs, threshold, mfma=True)
    p_fixed, t_fixed, y_fixed = run(fixed)
    report = {"fixed_ms": t_fixed, "kept": "fixed"}
    if router is None:
        return p_fixed, report
    learned = router.route_blocks(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb, vbr.bpntre,
                                  n_rhs)
    p_learned, t_learned, y_learned = run(learned)
    report["learned_ms"] = t_learned
    report["learned_routes"] = {int(k): int(v) for k, v in zip(*np.unique(learned, return_counts=True))}
    ok = _within(vbr, x, y_learned, n_rhs, tol) and _within(vbr, x, y_fixed, n_rhs, tol)
    report["learned_within_tol"] = bool(ok)
    if ok and t_learned < t_fixed:
        report["kept"] = "learned"
        p_fixed.close()
        return p_learned, report
    p_learned.close()
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
