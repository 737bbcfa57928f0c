"""Train the block router on GPU-measured labels and evaluate it.

    python -m sable_amd.router.train [--blocks 200000] [--steps 3000]
                                     [--out weights.npz] [--report r.json]
                                     [--eval 20]

1. samples the training blocks (seeded, the router workload law) and
   measures per-block route costs on the GPU (labels.py);
2. trains the MLP with the library's forward/backward/Adam kernels;
3. evaluates on held-out matrices of 20k-80k rows generated like the 100k
   mixed workload: cold launch time of the fixed-threshold plan (T = 15),
   the all-dense plan, the all-sparse plan and the learned plan, each
   checked against the FP64 oracle at 1e-5.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from .. import convert, synth
from . import labels, mlp
from .router import Router, _within


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def train(n_blocks=200_000, steps=3000, seed=4, progress=log):
    t0 = time.time()
    E, T, M, info = labels.training_set(n_blocks, seed=seed, progress=progress)
    t_labels = time.time() - t0
    progress(f"labels: {len(info['keys'])} groups measured in {t_labels:.1f}s")
    mean = E.mean(axis=0)
    std = E.std(axis=0) + 1e-6
    Z = ((E - mean) / std).astype(np.float32)
    # hold out 5% of the blocks to report generalisation
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(Z))
    nval = max(1, len(Z) // 20)
    va, tr = perm[:nval], perm[nval:]
    net = mlp.DeviceMLP(mlp.init_params(seed))
    t1 = time.time()
    hist = net.train(Z[tr], T[tr], M[tr], steps=steps, batch=4096, lr=3e-3, seed=seed,
                     log_every=max(1, steps // 10))
    t_train = time.time() - t1
    for s, l in hist:
        progress(f"train step {s}: loss {l:.4f}")
    pred = net.predict(Z[va])
    mse = float(np.sum(M[va] * (pred - T[va]) ** 2) / max(M[va].sum(), 1))
    # how often the predicted cheapest route is the measured cheapest
    allowed = M[va] > 0
    p2 = np.where(allowed, pred, np.inf)
    t2 = np.where(allowed, T[va], np.inf)
    agree = float(np.mean(np.argmin(p2, 1) == np.argmin(t2, 1)))
    params = net.params()
    report = {"blocks": n_blocks, "groups": int(len(info["keys"])), "label_seconds": t_labels,
              "train_seconds": t_train, "steps": steps, "val_mse_log2": mse, "val_route_agreement": agree,
              "loss_history": hist}
    return params, mean, std, report


def evaluate(router, n_matrices=20, seed=4000, reps=15, progress=log):
    from ..runtime import Plan, DeviceArray
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n_matrices):
        n = int(rng.integers(20_000, 80_001))
        n -= n % 100
        vbr, x, _ = synth.workload_like_c2(n, seed=seed + i)
        xd = DeviceArray.from_numpy(x)
        yd = DeviceArray(int(vbr.rpntr[-1]))
        learned, shape, wg = router.route_and_shape(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr,
                                                    vbr.bpntrb, vbr.bpntre, 1)
        plans = {
            "fixed": convert.threshold_routes(vbr.val, vbr.indx, convert.DEFAULT_DENSITY),
            "all_dense": np.zeros(len(vbr.bindx), np.uint8),
            "all_sparse": np.ones(len(vbr.bindx), np.uint8),
            "learned": learned,
        }
        row = {"n": n, "blocks": int(len(vbr.bindx)), "nnz": int(np.count_nonzero(vbr.val)), "learned_wg": wg}
        for name, routes in plans.items():
            # the learned plan runs in the tile shape the router chose; the
            # others in the tuned default shape
            p = Plan.from_vbr(vbr, routes, **({"wg": wg} if name == "learned" else {}))
            t = float(np.median(p.time(xd, yd, warmup=3, reps=reps, flush_bytes=1 << 30)))
            row[name + "_ms"] = t
            row[name + "_ok"] = _within(vbr, x, yd.numpy(), 1, 1e-5)
            p.close()
        row["learned_dense_blocks"] = int(np.sum(plans["learned"] == 0))
        row["fixed_dense_blocks"] = int(np.sum(plans["fixed"] == 0))
        rows.append(row)
        progress(f"eval {i + 1}/{n_matrices}: n={n} fixed {row['fixed_ms']:.4f} learned {row['learned_ms']:.4f} "
                 f"dense {row['all_dense_ms']:.4f} sparse {row['all_sparse_ms']:.4f} ms")
    g = lambda k: float(np.exp(np.mean(np.log([r["fixed_ms"] / r[k] for r in rows]))))
    summary = {"geomean_speedup_learned_vs_fixed": g("learned_ms"),
               "geomean_speedup_all_dense_vs_fixed": g("all_dense_ms"),
               "geomean_speedup_all_sparse_vs_fixed": g("all_sparse_ms"),
               "all_within_tol": bool(all(r[k] for r in rows for k in r if k.endswith("_ok")))}
    return rows, summary


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--blocks", type=int, default=200_000)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--out", default="weights.npz")
    ap.add_argument("--report", default="router_report.json")
    ap.add_argument("--eval", type=int, default=20)
    ap.add_argument("--seed", type=int, default=4)
    args = ap.parse_args(argv)
    params, mean, std, report = train(args.blocks, args.steps, args.seed)
    mlp.save(args.out, params, mean, std, {"blocks": args.blocks, "steps": args.steps, "seed": args.seed})
    log(f"saved {args.out}")
    if args.eval:
        router = Router(params, mean, std)
        rows, summary = evaluate(router, args.eval)
        report["eval"] = rows
        report["summary"] = summary
        log(json.dumps(summary))
    with open(args.report, "w") as f:
        json.dump(report, f, indent=1)


if __name__ == "__main__":
    main()
