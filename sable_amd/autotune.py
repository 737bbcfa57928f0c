"""Tile-shape autotuner: measures candidate plan options on the GPU and keeps
the fastest.

The options that shape the work are plan-level: the row-segment target
(``tile_bytes``: how much of a block row one tile streams), the workgroup
size, the dense-stream unroll, the scheduling mode (one workgroup per tile
vs a persistent work queue), and for SpMM the MFMA split.  ``tune`` builds a
plan per candidate, checks its result against the first candidate's (all
candidates compute the same sums in the same per-row order except for the
tile boundaries, so results agree to rounding), times it cold, and returns
the table.  ``tuning.json`` next to this module holds the winners measured
on MI355X for the workloads of this repository; ``tuned_options`` looks them
up.
"""
from __future__ import annotations

import itertools
import json
import os
import sys

import numpy as np

TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")

SPMV_SPACE = {
    "unroll": (2, 4),
    "tile_bytes": (128 << 10, 256 << 10, 512 << 10, 1 << 20),
    "mode": (0,),
    "wg": (256, 512),
}
SPMM_SPACE = {
    "tile_bytes": (256 << 10, 512 << 10, 1 << 20, 2 << 20),
    "mode": (0,),
    "wg": (256, 512),
}


def load_table(path=TABLE_PATH):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def tuned_options(op, n_rhs=1, path=TABLE_PATH):
    """Best measured options for an operation (empty dict when untuned)."""
    key = "spmv" if op == "spmv" else f"spmm_n{n_rhs}"
    return dict(load_table(path).get(key, {}).get("best", {}))


def tune(vbr, routes, x, op="spmv", n_rhs=1, space=None, reps=10, flush_bytes=1 << 30, extra=None,
         progress=None):
    """Measures every combination in ``space``; returns (best_options, rows)."""
    from .runtime import Plan, DeviceArray
    space = space or (SPMV_SPACE if op == "spmv" else SPMM_SPACE)
    keys = sorted(space)
    xd = DeviceArray.from_numpy(np.asarray(x, dtype=np.float32))
    shape_y = (int(vbr.rpntr[-1]),) if op == "spmv" else (int(vbr.rpntr[-1]), n_rhs)
    yd = DeviceArray(shape_y)
    rows, ref = [], None
    for combo in itertools.product(*[space[k] for k in keys]):
        opts = dict(zip(keys, combo))
        opts.update(extra or {})
        plan = Plan.from_vbr(vbr, routes, op=op, n_rhs=n_rhs, **opts)
        t = plan.time(xd, yd, warmup=3, reps=reps, flush_bytes=flush_bytes)
        y = yd.numpy()
        if ref is None:
            ref = y
            agree = 0.0
        else:
            scale = np.maximum(np.abs(ref), 1e-30)
            agree = float(np.max(np.abs(y - ref) / np.maximum(scale, np.max(np.abs(ref)) * 1e-6)))
        row = dict(opts, ms=float(np.median(t)), mean_ms=float(np.mean(t)), ntiles=plan.stats["ntiles"],
                   grid=plan.stats["grid"], bytes_alg=plan.stats["bytes_alg"], max_rel_diff=agree)
        rows.append(row)
        if progress:
            progress(f"tune {op}: {opts} -> {row['ms']:.4f} ms")
        plan.close()
    best = min(rows, key=lambda r: r["ms"])
    return {k: best[k] for k in keys}, rows


def save(key, best, rows, path=TABLE_PATH):
    table = load_table(path)
    table[key] = {"best": best, "rows": rows}
    with open(path, "w") as f:
        json.dump(table, f, indent=1)


def main(argv=None):
    """python -m sable_amd.autotune [spmv|spmm] [--out path]: tunes on the
    repository's workloads (100k SpMV, 50k SpMM with N=128)."""
    import argparse
    from . import synth
    from .convert import threshold_routes, ROUTE_DENSE, ROUTE_DENSE_MFMA
    from .runtime import init
    ap = argparse.ArgumentParser()
    ap.add_argument("op", choices=["spmv", "spmm"])
    ap.add_argument("--out", default=TABLE_PATH)
    ap.add_argument("--reps", type=int, default=10)
    args = ap.parse_args(argv)
    init()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    if args.op == "spmv":
        vbr, x, _ = synth.workload_c2()
        routes = threshold_routes(vbr.val, vbr.indx, 15)
        best, rows = tune(vbr, routes, x, "spmv", reps=args.reps, progress=log)
        save("spmv", best, rows, args.out)
    else:
        vbr, X, _ = synth.workload_c3()
        r = threshold_routes(vbr.val, vbr.indx, 15)
        routes = np.where(r == ROUTE_DENSE, ROUTE_DENSE_MFMA, r).astype(np.uint8)
        best, rows = tune(vbr, routes, X, "spmm", n_rhs=X.shape[1], reps=args.reps, progress=log)
        save("spmm_n128", best, rows, args.out)
    print(json.dumps(best))


if __name__ == "__main__":
    main()
