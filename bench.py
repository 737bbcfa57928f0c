"""Benchmark of the VBR SpMV/SpMM hot path on one MI355X.

Default workload (the headline): FP32 SpMV on the 100,000 x 100,000 VBR
matrix with 1,000 x 1,000 partitions sized 20-200, 1% of blocks stored, half
of them with density 0.8-1.0 and half 0.01-0.2, values N(0,1), x U[-1,1]
(seed 2).  A "step" is one SpMV launch over the whole matrix with x and the
matrix resident in HBM.  By default every step starts cold: a 1 GiB buffer is
overwritten between launches (outside the timed events), so the 256 MB
last-level cache holds none of the matrix.

Plans measured: the plain one (fixed density threshold T = 15) and, when
router weights are present, the router's plans at sparse-route biases 0, 0.5
and 1 (log2 ns); the fastest of those within 1e-5 of the FP64 oracle is the
learned plan, kept only if it is faster than the plain one.  Prints ONE JSON line: value = GFLOP/s (2 * nnz / time) of the kept
plan, the roofline of its kernel (algorithmic bytes of the executed format
per launch / measured kernel time vs 8 TB/s HBM), both plans' figures, and
the reference CPU path (the restated generated-C program, FP64 like the
reference) timed on this host's cores on a bounded sample.

    python bench.py [--steps K] [--warmup W] [--workload c2|c1|c3|c5|partition|inspect|io] [--warm]
                    [--routing plain|learned|both] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0
FLUSH_BYTES = 1 << 30


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _native_oracle():
    """The oracle's C program built with the reference's CFLAGS for this host
    (src/consts.py:7), else the portable build."""
    from oracle import c_oracle
    try:
        d = tempfile.mkdtemp(prefix="sable_oracle_")
        return c_oracle.build(native_dir=d), "reference CFLAGS (src/consts.py:7, -march=native)"
    except Exception as e:  # gcc missing or flags unsupported
        log(f"native oracle build failed ({e}); using the portable build")
        return c_oracle.build(), "portable -O3 -mavx2 -mfma -ffast-math"


def cpu_baseline_spmv(vbr, routes, x, procs=10):
    """The reference CPU path (the restated generated program: FP64, CSR part
    then the dense block calls, src/codegen.py:293-452) timed with the
    reference's protocol (eval.py:14-45): `procs` processes, each the median
    of its timed loop, the runs outside the 10th-90th percentile dropped, the
    mean of the rest; one thread pinned to one core, and all allowed cores
    (LPT block-row chunks, src/codegen.py:165-186).  Runs before the GPU is
    initialised (its children are separate processes)."""
    from oracle import c_oracle, cpu_protocol
    from sable_amd.convert import compress_with_routes
    lib_path, flags = _native_oracle()
    val2, indx2, ub, ip, ind, cv = compress_with_routes(vbr.val, vbr.rpntr, vbr.cpntr, vbr.indx, vbr.bindx,
                                                        vbr.bpntrb, vbr.bpntre, routes)
    d = tempfile.mkdtemp(prefix="sable_cpu_")
    path = os.path.join(d, "spmv.npz")
    cpu_protocol.save_spmv_program(path, val2, indx2, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb, vbr.bpntre,
                                   ub, ip, ind, cv, x)
    cores = len(os.sched_getaffinity(0))
    out = {}
    for label, nthreads, bench in (("1thread", 1, 50), ("allcores", cores, 100)):
        t0 = time.time()
        r = cpu_protocol.time_processes(path, nthreads, bench, procs=procs, lib_path=lib_path)
        r.update({"threads": nthreads, "bench": bench, "wall_s": time.time() - t0})
        out[label] = r
        log(f"cpu spmv {label}: {r['mean_s'] * 1e3:.3f} ms ({r['kept_runs']}/{procs} runs kept)")
    # the program's output, checked against FP64 by the caller
    prog = c_oracle.CProgram(val2, indx2, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb, vbr.bpntre, ub, ip, ind,
                             cv, dtype=np.float64, lib_path=lib_path)
    prog.set_threads(cores)
    out["y"] = prog.run(np.asarray(x, np.float64))
    shutil.rmtree(d, ignore_errors=True)
    return out, flags, cores


def cpu_baseline_spmm(vbr, routes, X, procs=10):
    """The reference SpMM program (FP32, dense blocks as the i/j/k loop nest,
    sparse-routed blocks per nonzero, src/codegen.py:827-937 with N = 128)
    with the same process protocol: one thread pinned, and all cores (the
    threaded program, block rows in LPT chunks, :939-1075)."""
    from oracle import c_oracle, cpu_protocol
    lib_path, flags = _native_oracle()
    d = tempfile.mkdtemp(prefix="sable_cpu_")
    path = os.path.join(d, "spmm.npz")
    cpu_protocol.save_spmm_program(path, vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb,
                                   vbr.bpntre, routes, X)
    cores = len(os.sched_getaffinity(0))
    out = {}
    for label, nthreads, bench in (("1thread", 1, 3), ("allcores", cores, 10)):
        t0 = time.time()
        r = cpu_protocol.time_processes(path, nthreads, bench, procs=procs, lib_path=lib_path)
        r.update({"threads": nthreads, "bench": bench, "wall_s": time.time() - t0})
        out[label] = r
        log(f"cpu spmm {label}: {r['mean_s'] * 1e3:.1f} ms ({r['kept_runs']}/{procs} runs kept)")
    out["Y"] = c_oracle.spmm(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb, vbr.bpntre, routes,
                             X, X.shape[1], lib_path=lib_path, threads=cores)
    shutil.rmtree(d, ignore_errors=True)
    return out, flags, cores


def traffic_for(workload, key, stats, **match):
    """HBM bytes per launch of the kept plan from the committed rocprofv3
    FETCH_SIZE + WRITE_SIZE summary (profiles/traffic_<workload>.json), or
    None unless the profiled entry is the same plan: same key (the plan's
    label, which names its route bias and workgroup size), the same
    algorithmic bytes and the same workgroup size."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        e = json.load(f).get(key)
    if not e or e.get("bytes_alg") != stats["bytes_alg"]:
        return None
    for k, v in match.items():
        if k in e and e[k] != v:
            return None
    return e.get("hbm_bytes_per_launch")


def retime(plan, x, y, args, flush):
    """The kept plan timed again with fresh launches (its selection used
    other samples), so the reported time carries no selection bias."""
    return plan.time(x, y, warmup=args.warmup, reps=args.steps, flush_bytes=flush)


SPINUP_S = 0.15


def retime_spun_up(plan, x, y, args, flush):
    """The same measurement after SPINUP_S seconds of the plan's own untimed
    launches.  After the host-side work between variants the GPU idles, and
    the first ~70 timed launches of the 50k SpMM then read 3-15 % slower
    than the steady ones; this figure is reported beside the standard one,
    never in its place."""
    t0 = time.time()
    while time.time() - t0 < SPINUP_S:
        plan.time(x, y, warmup=0, reps=10, flush_bytes=flush)
    return plan.time(x, y, warmup=args.warmup, reps=args.steps, flush_bytes=flush)


def spmm_c3(args, runtime, vbr, X, base, cpu=None):
    """SpMM on the 50,000 x 50,000 VBR (500 x 500 partitions sized 32-192, 2 %
    blocks, 70 % density 0.9-1.0 and 30 % 0.02-0.1, N(0,1); X 50,000 x 128
    U[-1,1]; seed 3).  Plain = fixed threshold T = 15 with dense blocks on
    the FP32 vector path; MFMA = the same routes with dense blocks on the
    bf16-split MFMA path (3-term split, 6 passes; and hi+lo, 3 passes);
    learned = the router's plans at MFMA-route biases 0, 0.5 and 1 (3-term
    split), each under every tile shape of the router and in the default
    workgroup size (its own shape pick marked); every-block-on-MFMA is timed
    beside them as a fixed policy.  The fastest learned
    plan within 1e-5 of FP64 is kept if faster than MFMA x6 on the threshold
    routes; the kept plan is then timed again with fresh launches.  Returns
    the result dict (roofline of the kept plan, every variant's time and
    error, the CPU baseline when given)."""
    from sable_amd.convert import ROUTE_DENSE, ROUTE_DENSE_MFMA
    from oracle import c_oracle
    n_rhs = X.shape[1]
    Y64, AY = c_oracle.spmm_f64_dense_blocks(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb,
                                             vbr.bpntre, X)
    Xd = runtime.DeviceArray.from_numpy(X)
    Yd = runtime.DeviceArray((int(vbr.rpntr[-1]), n_rhs))
    flush = 0 if args.warm else FLUSH_BYTES
    res = {}
    mf = np.where(base == ROUTE_DENSE, ROUTE_DENSE_MFMA, base)
    allm = np.full(len(base), ROUTE_DENSE_MFMA, np.uint8)
    variants = [("fp32_plain", base, 3, {}), ("mfma_x6", mf, 3, {}), ("mfma_x3", mf, 2, {}),
                ("mfma_x6_all", allm, 3, {})]
    from sable_amd.router import Router
    from sable_amd.router.router import BIAS_GRID
    if args.routing in ("learned", "both") and Router.available():
        from sable_amd.autotune import tuned_options
        router = Router.load()
        default_wg = tuned_options("spmm", n_rhs).get("wg") or 256
        seen = []
        for sb, mb in BIAS_GRID["spmm"]:
            # the router's routes under each tile shape (workgroup size), its
            # own pick marked, and the pick's routes in the default workgroup
            # size when no shape has it
            shapes, pick = router.routes_per_shape(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr,
                                                   vbr.bpntrb, vbr.bpntre, n_rhs, sparse_bias=sb, mfma_bias=mb)
            cands = [(r, w, s == pick) for s, (r, w) in enumerate(shapes)]
            if all(w != default_wg for _, w in shapes):
                cands.append((shapes[pick][0], default_wg, False))
            for lr, wg, is_pick in cands:
                if any(np.array_equal(lr, r) and wg == w for r, w in seen):
                    continue
                seen.append((lr, wg))
                variants.append((f"learned_x6 mfma_bias={mb:g} wg={wg}", lr, 3,
                                 {"wg": wg, "mfma_bias": mb, "router_pick": is_pick}))
    if args.spmm_variants:
        keep = [v for v in args.spmm_variants.split(",") if v]
        variants = [v for v in variants if any(v[0] == k or v[0].startswith(k + " ") for k in keep)]
    plans = {}
    for name, routes, split, kw in variants:
        kw = dict(kw)
        meta = {"mfma_bias": kw.pop("mfma_bias", None), "router_pick": kw.pop("router_pick", None),
                "split": split}
        t1 = time.time()
        plan = runtime.Plan.from_vbr(vbr, routes.astype(np.uint8), op="spmm", n_rhs=n_rhs, mfma_split=split, **kw)
        insp = time.time() - t1
        times = plan.time(Xd, Yd, warmup=args.warmup, reps=args.steps, flush_bytes=flush)
        err = c_oracle.scaled_error(Yd.numpy().ravel(), Y64.ravel(), AY.ravel())
        st = plan.stats
        res[name] = {"ms": float(np.mean(times)), "median_ms": float(np.median(times)), "stats": st,
                     "err": err, "inspector_s": insp, **meta}
        log(f"{name}: {res[name]['ms']:.4f} ms, err {err['componentwise']:.2e}, rtc={st['rtc_used']}")
        plans[name] = plan
    cands = [k for k in res if k.startswith("learned_x6 ")]
    good = [k for k in cands if res[k]["err"]["componentwise"] <= 1e-5]
    kept = "mfma_x6" if "mfma_x6" in res else next(iter(res))
    if good:
        best = min(good, key=lambda k: res[k]["ms"])
        if "mfma_x6" not in res or res[best]["ms"] < res["mfma_x6"]["ms"]:
            kept = best
    # fresh launches of the kept plan and of the FP32 plain plan (the speedup's denominator)
    for name in dict.fromkeys(k for k in (kept, "fp32_plain") if k in res):
        t = retime(plans[name], Xd, Yd, args, flush)
        res[name]["retimed_ms"] = float(np.mean(t))
        res[name]["retimed_median_ms"] = float(np.median(t))
        res[name]["spun_up_ms"] = float(np.mean(retime_spun_up(plans[name], Xd, Yd, args, flush)))
        if name == kept:
            assert c_oracle.scaled_error(Yd.numpy().ravel(), Y64.ravel(), AY.ravel())["componentwise"] <= 1e-5
    for p in plans.values():
        p.close()
    best = res[kept]
    st = best["stats"]
    ms = best["retimed_ms"]
    gflops = st["flops_alg"] / (ms * 1e-3) / 1e9
    mfma_flops = 2 * st["mfma_values"] * n_rhs * 6
    t_mfma = mfma_flops / (BF16_PEAK_TFLOPS * 1e12)
    t_hbm = st["bytes_alg"] / (HBM_PEAK_GBS * 1e9)
    if t_hbm >= t_mfma:
        roof = {"bound": "hbm", "achieved": st["bytes_alg"] / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                "unit": "GB/s"}
    else:
        roof = {"bound": "mfma", "achieved": mfma_flops / (ms * 1e-3) / 1e12, "peak": BF16_PEAK_TFLOPS,
                "unit": "TFLOP/s (bf16, 6 passes of the 3-term split)"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = traffic_for("spmm_c3", kept, st, wg=st["wg"])
    plain = res.get("fp32_plain")
    fp32_ms = plain["retimed_ms"] if plain else None
    fp32_roof = None
    if plain:
        # the published FP32 peak counts packed FMAs; the kernels are built
        # without packed FP32 instructions, whose FMA peak is half
        fp32_roof = {"bound": "fp32 vector", "achieved_tflops": plain["stats"]["flops_alg"] / (fp32_ms * 1e-3) / 1e12,
                     "peak_tflops": FP32_PEAK_TFLOPS, "peak_tflops_unpacked": FP32_PEAK_TFLOPS / 2}
        fp32_roof["frac"] = fp32_roof["achieved_tflops"] / FP32_PEAK_TFLOPS
        fp32_roof["frac_of_unpacked"] = fp32_roof["achieved_tflops"] / (FP32_PEAK_TFLOPS / 2)
    out = {
        "workload": "SpMM, 50000x50000 VBR, 500x500 partitions sized 32-192, 2% blocks, 70% density 0.9-1.0 / "
                    "30% 0.02-0.1, N(0,1), X 50000x128 U[-1,1], seed 3",
        "kept": kept, "ms": ms, "gflops": gflops, "error_vs_fp64": best["err"], "roofline": roof,
        "roofline_times_us": {"mfma_6pass": t_mfma * 1e6, "hbm": t_hbm * 1e6},
        "speedup_vs_fp32_plain": fp32_ms / ms if plain else None, "fp32_plain_ms": fp32_ms,
        "spun_up": {"note": f"the same timing after {SPINUP_S:g} s of the plan's own untimed launches "
                            "(GPU clocks out of their idle state); reported beside ms, never in its place",
                    "ms": best.get("spun_up_ms"), "fp32_plain_ms": plain.get("spun_up_ms") if plain else None},
        "fp32_plain_roofline": fp32_roof,
        "speedup_x3_vs_fp32_plain": plain["ms"] / res["mfma_x3"]["ms"] if plain and "mfma_x3" in res else None,
        "variants": {k: {"ms": v["ms"], "median_ms": v["median_ms"], "err": v["err"]["componentwise"],
                         "gflops": v["stats"]["flops_alg"] / (v["ms"] * 1e-3) / 1e9,
                         "inspector_s": v["inspector_s"], "bytes_alg": v["stats"]["bytes_alg"],
                         "ntiles": v["stats"]["ntiles"], "ntiles_mfma": v["stats"]["ntiles_mfma"],
                         "wg": v["stats"]["wg"], "split": v["split"], "mfma_bias": v["mfma_bias"],
                         "router_pick": v["router_pick"]}
                     for k, v in res.items()},
        "bytes_alg": st["bytes_alg"], "flops_alg": st["flops_alg"], "cpu_baseline": None,
    }
    if cpu is not None:
        res_cpu, flags, cores = cpu
        e = c_oracle.scaled_error(res_cpu["Y"].ravel(), Y64.ravel(), AY.ravel())["componentwise"]
        fl = st["flops_alg"]
        out["cpu_baseline"] = {
            "value": fl / res_cpu["allcores"]["mean_s"] / 1e9, "unit": "GFLOP/s", "cores": cores, "kind": "port",
            "sample": (f"whole matrix, restated reference SpMM program (FP32, N=128, threshold routes), "
                       f"eval.py protocol: {res_cpu['allcores']['runs']} processes x "
                       f"{res_cpu['allcores']['bench']} passes on {cores} threads (LPT block-row chunks), "
                       f"per-process median, 10-90% decile filter, mean; built with {flags}"),
            "single_thread_gflops": fl / res_cpu["1thread"]["mean_s"] / 1e9,
            "single_thread_sample": (f"{res_cpu['1thread']['runs']} processes x {res_cpu['1thread']['bench']} "
                                     f"passes pinned to one core"),
            "error_vs_fp64_oracle": e, "gpu_speedup_vs_allcores": res_cpu["allcores"]["mean_s"] / (ms * 1e-3)}
    return out


def bench_spmm(args, runtime, info, cpu=None, vbr=None, X=None):
    """--workload c3: the SpMM result as its own JSON line."""
    from sable_amd import synth
    from sable_amd.convert import threshold_routes
    if vbr is None:
        vbr, X, _ = synth.workload_c3()
    base = threshold_routes(vbr.val, vbr.indx, args.threshold)
    r = spmm_c3(args, runtime, vbr, X, base, cpu)
    line = {
        "metric": "VBR SpMV/SpMM GFLOP/s + % HBM/MFMA peak, learned-routed vs plain, 1 GPU",
        "value": r["gflops"], "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": r["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (bf16x6 split on MFMA, f32 accumulate)", "data": "synthetic",
        "config": {"workload": r["workload"],
                   "routing": f"{r['kept']} (dense blocks on MFMA, 3-term bf16 split, FP32 accumulate)",
                   "cache": "warm" if args.warm else "cold (1 GiB flush between launches)"},
        "roofline": r["roofline"], "cpu_baseline": r["cpu_baseline"], "spmm": r, "device": info,
    }
    print(json.dumps(line), flush=True)


def bench_batch(args, runtime, info):
    """Batched workload: 256 VBR matrices (n log-uniform 2,000-64,000, P
    uniform 50-800 partitions sized [ceil(n/5P), floor(2n/P)], 1% blocks, the
    100k workload's density mix, seed 5), each with an SpMV and an SpMM with
    N = 16.  Batched = learned routes (MFMA allowed for the SpMM; router
    biases (0, 0), (0, 1) and (1, 1), the fastest group kept), autotuned
    tiles sized for the group (no per-plan minimum tile count), every plan in
    one group (one launch per operation over the concatenated tiles, both
    captured in one hipGraph).  Baseline = plain plans (fixed threshold
    T = 15, FP32 dense path, default tiling) launched one after another."""
    from sable_amd import synth
    from sable_amd.convert import threshold_routes
    from sable_amd.router import Router
    from oracle import c_oracle
    n_rhs = 16
    router = Router.load() if Router.available() else None
    shapes = synth.c5_shapes(args.count)
    # router biases (sparse bias of the SpMV members, MFMA bias of the SpMM
    # members, log2 ns): the group is timed at each, the faster is kept
    biases = ((0.0, 0.0), (0.0, 1.0), (1.0, 1.0)) if router else ((0.0, 0.0),)
    plain, xs, ys_p, ys_l, samples = [], [], [], [], []
    cand = [[] for _ in biases]
    cand_bytes = [0 for _ in biases]
    flops = bytes_plain = values = 0
    t0 = time.time()
    for i, (n, parts) in enumerate(shapes):
        vbr, x, X = synth.workload_c5_matrix(i, n, parts, n_rhs=n_rhs)
        values += len(vbr.val)
        fixed = threshold_routes(vbr.val, vbr.indx, args.threshold)
        for op, inp in (("spmv", x), ("spmm", X)):
            k = 1 if op == "spmv" else n_rhs
            shape = (n,) if op == "spmv" else (n, n_rhs)
            pp = runtime.Plan.from_vbr(vbr, fixed, op=op, n_rhs=k)
            for j, (sb, mb) in enumerate(biases):
                # built for the group: no per-plan minimum tile count (the group
                # fills the GPU; small matrices keep block-row-sized tiles), all
                # members of an operation in the router's narrow tile shape
                # (one launch per operation; SpMV 256 threads, SpMM N = 16 128
                # threads: the group measured 3.10 ms so, 3.46 with SpMV at 512,
                # 3.65 with SpMM at 256 too), routed for that shape
                lr, _, wgl = (router.route_and_shape(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb,
                                                     vbr.bpntre, k, sparse_bias=sb if k == 1 else 0.0,
                                                     mfma_bias=mb if k > 1 else 0.0, shape=0)
                              if router else (fixed, 0, 128 if op == "spmm" else 256))
                pl = runtime.Plan.from_vbr(vbr, lr, op=op, n_rhs=k, min_tiles=1, wg=wgl)
                cand[j].append(pl)
                cand_bytes[j] += pl.stats["bytes_alg"]
            plain.append(pp)
            xs.append(runtime.DeviceArray.from_numpy(inp))
            ys_p.append(runtime.DeviceArray(shape))
            ys_l.append(runtime.DeviceArray(shape))
            flops += pp.stats["flops_alg"]
            bytes_plain += pp.stats["bytes_alg"]
        if i % 32 == 0:
            samples.append((i, vbr, x, X))
        del vbr
        if (i + 1) % 16 == 0:
            log(f"c5: {i + 1}/{len(shapes)} matrices planned, {values / 1e9:.2f}e9 stored values, "
                f"{time.time() - t0:.0f}s")
    flush = 0 if args.warm else 4 << 30
    log("c5: timing the plain plans launched one by one")
    t_plain = runtime.time_plans(plain, xs, ys_p, warmup=2, reps=args.steps, flush_bytes=flush)
    def sample_errors(ys):
        errs = []
        for i, vbr, x, X in samples:
            y64, ay = c_oracle.spmv_f64_dense_blocks(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr,
                                                     vbr.bpntrb, vbr.bpntre, x)
            errs.append(c_oracle.scaled_error(ys[2 * i].numpy(), y64, ay)["componentwise"])
            Y64, AY = c_oracle.spmm_f64_dense_blocks(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr,
                                                     vbr.bpntrb, vbr.bpntre, X)
            errs.append(c_oracle.scaled_error(ys[2 * i + 1].numpy().ravel(), Y64.ravel(),
                                              AY.ravel())["componentwise"])
        return errs

    t_cands, e_cands = [], []
    for j, b in enumerate(biases):
        log(f"c5: timing the batched group at router biases {b}")
        g = runtime.Group(cand[j], xs, ys_l)
        t_cands.append(g.time(warmup=2, reps=args.steps, flush_bytes=flush))
        g.launch()
        e_cands.append(max(sample_errors(ys_l)))
        g.close()
    # the fastest group whose sampled products are within 1e-5 of FP64
    ok = [j for j in range(len(biases)) if e_cands[j] <= 1e-5] or list(range(len(biases)))
    kb = min(ok, key=lambda j: np.mean(t_cands[j]))
    learned, bytes_learned, t_batch = cand[kb], cand_bytes[kb], t_cands[kb]
    for j in range(len(biases)):
        if j != kb:
            for p in cand[j]:
                p.close()
    g = runtime.Group(learned, xs, ys_l)
    log("c5: timing the plain plans as a group, the learned plans one by one")
    g_plain = runtime.Group(plain, xs, ys_p)
    t_batch_plain = g_plain.time(warmup=2, reps=args.steps, flush_bytes=flush)
    g_plain.close()
    t_learned_unbatched = runtime.time_plans(learned, xs, ys_l, warmup=1, reps=max(3, args.steps // 2),
                                             flush_bytes=flush)
    # the kept group again after SPINUP_S seconds of its own untimed replays
    # (reported beside ms_per_step, never in its place)
    t0 = time.time()
    while time.time() - t0 < SPINUP_S:
        g.time(warmup=0, reps=3, flush_bytes=flush)
    t_batch_spun = g.time(warmup=2, reps=args.steps, flush_bytes=flush)
    g.launch()
    errs = sample_errors(ys_l)
    g.close()
    # the batched SpMV and SpMM launches on their own: groups of only the
    # SpMV members, then only the SpMM members (plans alternate SpMV, SpMM)
    per_op = {}
    for name, sl in (("spmv", slice(0, None, 2)), ("spmm", slice(1, None, 2))):
        g_op = runtime.Group(learned[sl], xs[sl], ys_l[sl])
        t_op = float(np.mean(g_op.time(warmup=2, reps=args.steps, flush_bytes=flush)))
        g_op.close()
        b_op = int(sum(p.stats["bytes_alg"] for p in learned[sl]))
        per_op[name] = {"ms": t_op, "bytes_alg": b_op, "achieved_gbs": b_op / (t_op * 1e-3) / 1e9,
                        "frac_of_hbm_peak": b_op / (t_op * 1e-3) / 1e9 / HBM_PEAK_GBS}
        log(f"c5: batched {name} alone {t_op:.3f} ms, {per_op[name]['achieved_gbs']:.0f} GB/s algorithmic")
    ms = float(np.mean(t_batch))
    ms_plain = float(np.mean(t_plain))
    achieved = bytes_learned / (ms * 1e-3) / 1e9
    line = {
        "metric": "VBR SpMV/SpMM GFLOP/s + % HBM/MFMA peak, learned-routed vs plain, 1 GPU",
        "value": flops / (ms * 1e-3) / 1e9, "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": 2,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (SpMM dense blocks on bf16x6 MFMA where routed)", "data": "synthetic",
        "config": {"workload": (f"{len(shapes)} VBR matrices, n log-uniform 2000-64000, P uniform 50-800, "
                                f"1% blocks, c2 density mix, seed 5; SpMV + SpMM N=16 each"),
                   "routing": (f"learned at router biases {biases[kb]} (sparse for SpMV, MFMA for SpMM; the "
                               f"faster group of {list(biases)})" if router else "fixed T=15 (no router weights)"),
                   "cache": "warm" if args.warm else "cold (4 GiB flush before each repetition)",
                   "stored_values": values, "plans": len(learned)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "cpu_baseline": None,
        "unbatched_plain_ms": ms_plain, "batched_plain_ms": float(np.mean(t_batch_plain)),
        "unbatched_learned_ms": float(np.mean(t_learned_unbatched)),
        "batched_learned_ms_by_bias": {str(bb): float(np.mean(t)) for bb, t in zip(biases, t_cands)},
        "speedup_vs_unbatched_plain": ms_plain / ms,
        "spun_up_ms": float(np.mean(t_batch_spun)),
        "bytes_alg_learned": bytes_learned, "bytes_alg_plain": bytes_plain, "flops_alg": flops,
        "sample_errors_vs_fp64": {"max": max(errs), "checked_products": len(errs)},
        "sample_errors_by_bias": {str(bb): e for bb, e in zip(biases, e_cands)},
        "batched_by_operation": per_op,
        "device": info,
    }
    print(json.dumps(line), flush=True)


def bench_partition(args, runtime, info):
    """Block partitioner workload (src/autopartition.py:34-68, cut_indices2
    with similarity2 at threshold 0.2, both axes) on the 100,000 x 100,000
    SpMV matrix's nonzero pattern (values 1.0, FP64).  The
    GPU time is the two kernels of sable_axis_cuts per axis (inputs copied
    to the device before the events); the cuts are checked against the host
    partitioner bit for bit.  CPU baseline: the reference algorithm (the
    oracle's restatement: every line densified, numpy dot products) on a
    square corner, scaled by line length to lines per second."""
    import ctypes as C
    from sable_amd import synth, autopartition as AP
    from sable_amd._lib import lib, check
    from sable_amd.convert import vbr_to_csc
    vbr, _, _ = synth.workload_c2()
    A = vbr_to_csc(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb, vbr.bpntre)
    del vbr
    A.data[:] = 1.0          # its pattern: columns of a dense block are alike, the cuts find the blocks
    A.sum_duplicates()
    A.sort_indices()
    AT = A.T.tocsc()
    AT.sort_indices()
    log(f"partition: {A.shape[0]}x{A.shape[1]} matrix, {A.nnz} nonzeros")
    axes = []
    bytes_alg = 0
    for M in (A, AT):
        n = M.shape[1]
        arrs = (np.ascontiguousarray(M.indptr, dtype=np.int64), np.ascontiguousarray(M.indices, dtype=np.int32),
                np.ascontiguousarray(M.data, dtype=np.float64))
        axes.append((M, n, arrs, np.zeros(n + 1, np.uint8)))
        # read indptr, indices and values once; write the similarities and cut flags
        bytes_alg += 8 * (n + 1) + 12 * M.nnz + 8 * (n - 1) + (n + 1)
    times = []
    for rep in range(args.warmup + args.steps):
        total = 0.0
        for M, n, (ip, ix, v), cut in axes:
            ms = C.c_float(0.0)
            check(lib.sable_axis_cuts(ip.ctypes.data, ix.ctypes.data, v.ctypes.data, n, M.shape[0], 0, 0.2,
                                      cut.ctypes.data, None, C.byref(ms)))
            total += ms.value
        if rep >= args.warmup:
            times.append(total)
    gpu_cuts = [np.flatnonzero(cut).tolist() for _, _, _, cut in axes]
    t0 = time.time()
    host_cuts = [AP._axis_cuts(M, 0.2, AP.similarity2) for M, _, _, _ in axes]
    host_s = time.time() - t0
    assert gpu_cuts == host_cuts, "GPU and host partitions differ"
    ms = float(np.mean(times))
    lines = A.shape[0] + A.shape[1]
    out = {
        "metric": "block partitioner (cut_indices2, similarity2): adjacent-line similarity scan, both axes",
        "value": bytes_alg / (ms * 1e-3) / 1e9, "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 (values), int32 (indices)", "data": "synthetic",
        "config": {"workload": "cut_indices2(A, 0.2, similarity2) on the pattern of the 100000x100000 SpMV "
                               f"matrix ({A.nnz} nonzeros), both axes"},
        "roofline": {"bound": "hbm", "achieved": bytes_alg / (ms * 1e-3) / 1e9, "peak": 8000.0, "unit": "GB/s",
                     "frac": bytes_alg / (ms * 1e-3) / 1e9 / 8000.0, "traffic": None},
        "lines_per_s": lines / (ms * 1e-3), "cuts": [len(c) for c in gpu_cuts], "bytes_alg": bytes_alg,
        "host_partitioner_s": host_s, "matches_host": True, "cpu_baseline": None, "device": info,
    }
    if not args.no_cpu:
        # the reference algorithm densifies every line (cost linear in the line
        # length): time it on a square corner and scale by line elements
        from oracle import vbr_oracle
        m = 10000
        S = A[:m, :m].tocsc()
        t0 = time.time()
        vbr_oracle.cut_indices2_dense_lines(S, 0.2)
        dt = time.time() - t0
        rate = 2 * m * m / dt
        full_s = 2.0 * A.shape[0] * A.shape[1] / rate
        out["cpu_baseline"] = {"value": lines / full_s, "unit": "lines/s", "cores": 1, "kind": "port",
                               "sample": f"the reference algorithm (oracle restatement: every line densified, "
                                         f"numpy dots) on the {m}x{m} corner of the matrix ({2 * m} lines of "
                                         f"{m}, {dt:.1f} s), scaled by line length to the full matrix",
                               "est_full_matrix_s": full_s}
    print(json.dumps(out), flush=True)


def bench_io(args, runtime, info):
    """Text and binary I/O of the 100,000 x 100,000 SpMV matrix (the .vbr
    text of utils/fileio.py:13-64, about 1 GB): the native writer, the native
    parallel reader (arrays checked bit for bit against what was written),
    the binary container (save, load by memory map, load + upload of the
    arrays to HBM).  The file sits in the page cache after the write, so the
    reads measure parsing, not the disk.  CPU baseline: the reference's parse
    (Python float() / int() per token) on the first 2 M values of the val
    line, scaled to the file."""
    import shutil
    from sable_amd import synth, fileio
    from sable_amd.vbr import VBR
    from oracle.fileio_oracle import parse_line_reference
    vbr, _, _ = synth.workload_c2()
    v = VBR(np.asarray(vbr.val, dtype=np.float64), *(np.asarray(getattr(vbr, n), dtype=np.int64) for n in
                                                    ("indx", "bindx", "rpntr", "cpntr", "bpntrb", "bpntre")))
    d = tempfile.mkdtemp(prefix="sable_io_")
    try:
        cores = len(os.sched_getaffinity(0))
        t0 = time.time()
        fileio.write_vbr_matrix("c2", v, d)
        write_s = time.time() - t0
        path = os.path.join(d, "c2.vbr")
        nbytes = os.path.getsize(path)
        reads = []
        for rep in range(max(1, args.steps)):
            t0 = time.time()
            arrs = fileio.read_vbr(path)
            reads.append(time.time() - t0)
        names = ("val", "indx", "bindx", "rpntr", "cpntr", "bpntrb", "bpntre")
        assert all(np.array_equal(a.view(np.uint64) if a.dtype == np.float64 else a,
                                  np.asarray(getattr(v, n)).view(np.uint64) if n == "val" else np.asarray(getattr(v, n)))
                   for a, n in zip(arrs, names)), "text round trip differs"
        read_s = min(reads)
        log(f"io: {nbytes / 1e9:.3f} GB of text; write {write_s:.2f} s, read {read_s:.3f} s on {cores} threads")
        bpath = os.path.join(d, "c2.bin")
        t0 = time.time()
        fileio.save_vbr_bin(bpath, v)
        bsave_s = time.time() - t0
        bbytes = os.path.getsize(bpath)
        t0 = time.time()
        w = fileio.load_vbr_bin(bpath, mmap=False)
        bload_s = time.time() - t0
        assert all(np.array_equal(np.asarray(getattr(w, n)), np.asarray(getattr(v, n))) for n in names)
        t0 = time.time()
        w = fileio.load_vbr_bin(bpath, mmap=True)
        dv = runtime.DeviceArray.from_numpy(np.asarray(w.val, dtype=np.float32))
        bhbm_s = time.time() - t0
        del dv
        # the reference's parse on a bounded sample of the val line
        with open(path) as f:
            line = f.readline()
        cut = 0
        for _ in range(2_000_000):
            cut = line.index(",", cut) + 1
        sample = line[:cut - 1] + "]\n"
        t0 = time.time()
        ref_vals = parse_line_reference(sample, float)
        ref_s = time.time() - t0
        assert np.array_equal(np.asarray(ref_vals).view(np.uint64), arrs[0][:len(ref_vals)].view(np.uint64))
        ref_rate = len(sample) / ref_s
        # and its writer (str() per value, utils/fileio.py:28) on the same values
        t0 = time.time()
        ref_text = ",".join(map(str, ref_vals))
        ref_write_s = time.time() - t0
        assert ref_text == sample[sample.index("[") + 1:-2]
        cpu = {"value": ref_rate / 1e9, "unit": "GB/s of text", "cores": 1, "kind": "port",
               "sample": f"the reference's read_vbr parse (Python float() per token, utils/fileio.py:55-64) on the "
                         f"first {len(ref_vals)} values of the val line ({len(sample) / 1e6:.1f} MB, {ref_s:.1f} s)",
               "est_full_file_s": nbytes / ref_rate,
               "write_sample_s": ref_write_s, "est_full_write_s": ref_write_s * nbytes / len(sample)}
    finally:
        shutil.rmtree(d, ignore_errors=True)
    line = {
        "metric": "VBR text I/O: .vbr parse throughput of the 100k matrix (native parallel reader)",
        "value": nbytes / read_s / 1e9, "unit": "GB/s of text", "n_gpus": 1, "steps": max(1, args.steps),
        "warmup": 0, "ms_per_step": read_s * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 text", "data": "synthetic",
        "config": {"workload": "100,000 x 100,000 VBR, 1000 x 1000 partitions, 51 M stored values, .vbr text "
                               f"({nbytes / 1e9:.2f} GB) in the page cache"},
        "text_bytes": nbytes, "threads": cores, "write_s": write_s, "read_s": read_s,
        "binary_bytes": bbytes, "binary_save_s": bsave_s, "binary_load_s": bload_s,
        "binary_load_gbs": bbytes / bload_s / 1e9,
        "binary_load_and_upload_val_s": bhbm_s,
        "cpu_baseline": cpu, "speedup_vs_reference_parse": (nbytes / ref_rate) / read_s,
        "device": info,
    }
    print(json.dumps(line))
    return line


def bench_inspect(args, runtime, info):
    """Inspector workload (utils/convert_real_to_vbr.py:18-199,
    convert_sparse_to_vbr then convert_vbr_to_compressed at T = 15): the
    100,000 x 100,000 SpMV matrix as CSC (its 51 M stored nonzeros, FP64)
    and its 1000 x 1000 partitions to the VBR arrays, the routes (from the
    device's per-block nonzero counts) and the compressed arrays on the GPU.
    Time: the two device pipelines after their input uploads; every array is
    checked against the host restatement bit for bit.  CPU baseline: the oracle's per-entry
    restatement on the first 30 x 30 blocks, as entries per second."""
    from sable_amd import synth
    from sable_amd import convert as CV
    vbr, _, _ = synth.workload_c2()
    A = CV.vbr_to_csc(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb, vbr.bpntre)
    rp, cp = np.asarray(vbr.rpntr, np.int64), np.asarray(vbr.cpntr, np.int64)
    del vbr
    log(f"inspect: {A.shape[0]}x{A.shape[1]} matrix, {A.nnz} nonzeros, {len(rp) - 1}x{len(cp) - 1} partitions")
    times = []
    for rep in range(args.warmup + args.steps):
        out = CV.sparse_to_vbr_device(A, rp, cp)
        if rep >= args.warmup:
            times.append(out[-1])
    val, indx, bindx, bb, be, bnnz, _ = out
    # routing (T = 15, the reference's strict > on the density percentage)
    # from the device's per-block counts, then the compressed arrays
    sizes = np.diff(indx)
    routes = np.where(CV.block_density_percent(bnnz, sizes) > args.threshold, CV.ROUTE_DENSE,
                      CV.ROUTE_SPARSE).astype(np.uint8)
    ctimes = []
    for rep in range(args.warmup + args.steps):
        comp, cms = CV.compress_with_routes_device(val, rp, cp, indx, bindx, bb, be, routes, return_ms=True)
        if rep >= args.warmup:
            ctimes.append(cms)
    t0 = time.time()
    host = CV.convert_sparse_to_vbr(A, rp, cp, None, None, allow_empty_rows=True)
    host_routes = CV.threshold_routes(host[0], host[1], args.threshold)
    host_comp = CV.compress_with_routes(host[0], rp, cp, host[1], host[2], host[3], host[4], host_routes)
    host_s = time.time() - t0
    assert all(np.array_equal(a, b) for a, b in zip(host, (val, indx, bindx, bb, be))), "device != host"
    assert np.array_equal(routes, host_routes), "routes differ"
    assert all(np.array_equal(a, b) for a, b in zip(host_comp, comp)), "compressed arrays differ"
    # end to end (uploads and downloads included): two calls, val going to
    # the host and back in between, against one call keeping it on the device
    t0 = time.time()
    v_, ix_, bx_, bb_, be_, _, _ = CV.sparse_to_vbr_device(A, rp, cp)
    CV.compress_with_routes_device(v_, rp, cp, ix_, bx_, bb_, be_, routes)
    two_call_s = time.time() - t0
    t0 = time.time()
    one = CV.sparse_to_vbrc_device(A, rp, cp, density=args.threshold)
    one_call_s = time.time() - t0
    assert all(np.array_equal(a, b) for a, b in zip(one, (comp[0], comp[1], bindx, bb, be, comp[2], comp[3],
                                                        comp[4], comp[5]))), "one-call arrays differ"
    ms = float(np.mean(times)) + float(np.mean(ctimes))
    # materialisation: entries read, val written; compression: val read,
    # val2 and the CSR written
    bytes_alg = (12 * A.nnz + 8 * (A.shape[1] + 1) + 8 * len(val) + 16 * len(bindx) + 16 * (len(rp) - 1)
                 + 8 * len(val) + 8 * len(comp[0]) + 16 * len(comp[4]) + 8 * len(comp[3]))
    res = {
        "metric": "inspector (convert_sparse_to_vbr + routing + convert_vbr_to_compressed): CSC to .vbrc arrays",
        "value": bytes_alg / (ms * 1e-3) / 1e9, "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 (values), int64/int32 (indices)", "data": "synthetic",
        "config": {"workload": f"convert_sparse_to_vbr of the 100000x100000 SpMV matrix ({A.nnz} nonzeros, "
                               f"{len(rp) - 1}x{len(cp) - 1} partitions, {len(bindx)} blocks kept)"},
        "roofline": {"bound": "hbm", "achieved": bytes_alg / (ms * 1e-3) / 1e9, "peak": 8000.0, "unit": "GB/s",
                     "frac": bytes_alg / (ms * 1e-3) / 1e9 / 8000.0, "traffic": None},
        "bytes_alg": bytes_alg, "entries_per_s": A.nnz / (ms * 1e-3), "host_restatement_s": host_s,
        "end_to_end_two_calls_s": two_call_s, "end_to_end_one_call_s": one_call_s,
        "materialise_ms": float(np.mean(times)), "compress_ms": float(np.mean(ctimes)),
        "sparse_blocks": int(np.sum(routes == CV.ROUTE_SPARSE)), "csr_nnz": int(len(comp[4])),
        "matches_host": True, "cpu_baseline": None, "device": info,
    }
    if not args.no_cpu:
        from oracle import vbr_oracle
        kr = kc = 30
        S = A[:int(rp[kr]), :int(cp[kc])].tocoo()
        entries = {(int(i), int(j)): float(v) for i, j, v in zip(S.row, S.col, S.data)}
        t0 = time.time()
        vbr_oracle.convert_sparse_to_vbr(entries, S.shape, rp[:kr + 1].tolist(), cp[:kc + 1].tolist(),
                                         allow_empty_rows=True)
        dt = time.time() - t0
        res["cpu_baseline"] = {"value": len(entries) / dt, "unit": "entries/s", "cores": 1, "kind": "port",
                               "sample": f"the oracle's convert_sparse_to_vbr (per-entry restatement) on the first "
                                         f"{kr}x{kc} blocks ({len(entries)} entries, {dt:.1f} s); its cost "
                                         f"grows faster than the entry count, so the full-matrix figure is "
                                         f"a lower bound",
                               "est_full_matrix_s_lower_bound": A.nnz / (len(entries) / dt)}
    print(json.dumps(res), flush=True)


def spmv_options(args):
    """Plan options: the autotuned table's winners unless given on the command line."""
    from sable_amd.autotune import tuned_options
    opts = {k: v for k, v in tuned_options("spmv").items() if k in ("mode", "tile_bytes", "unroll", "wg")}
    for k in ("mode", "tile_bytes", "unroll", "wg"):
        if getattr(args, k) is not None:
            opts[k] = getattr(args, k)
    return opts


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c5", "partition", "inspect", "io"])
    ap.add_argument("--count", type=int, default=256, help="c5: number of matrices")
    ap.add_argument("--warm", action="store_true", help="no cache flush between launches")
    ap.add_argument("--routing", default="both", choices=["plain", "learned", "both"],
                    help="both: measure the learned plan beside the plain one, keep it if faster")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-spmm", action="store_true", help="c2: leave out the 50k SpMM object")
    ap.add_argument("--spmm-variants", default="", help="c3: comma list of the variants to run (profiling)")
    ap.add_argument("--spmv-plans", default="", help="c2: comma list of the plan labels to run (profiling)")
    ap.add_argument("--cpu-procs", type=int, default=10, help="processes of the CPU timing protocol")
    ap.add_argument("--mode", type=int, default=None, help="default: the autotuned value")
    ap.add_argument("--tile-bytes", type=int, default=None, help="default: the autotuned value")
    ap.add_argument("--wg", type=int, default=None, help="SpMV workgroup size; default: the autotuned value")
    ap.add_argument("--unroll", type=int, default=None, help="default: the autotuned value")
    ap.add_argument("--threshold", type=float, default=15)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from sable_amd import synth
    from sable_amd.convert import threshold_routes
    from oracle import c_oracle

    # workloads and the CPU baselines first: the baseline's timing processes
    # start before this process touches the GPU
    t0 = time.time()
    cpu = cpu3 = None
    vbr3 = X3 = None
    if args.workload == "c2":
        vbr, x, _ = synth.workload_c2()
        desc = ("SpMV FP32, 100000x100000 VBR, 1000x1000 partitions sized 20-200, 1% blocks, "
                "half density 0.8-1.0 / half 0.01-0.2, N(0,1), seed 2")
    elif args.workload == "c1":
        vbr, x = synth.workload_c1()
        desc = "SpMV FP32, 1000x1000 VBR, 20x20 random partitions, 40 dense blocks, seed 1"
    if args.workload in ("c1", "c2"):
        log(f"workload generated in {time.time() - t0:.1f}s: {len(vbr.val)} stored values, {len(vbr.bindx)} blocks")
        routes = threshold_routes(vbr.val, vbr.indx, args.threshold)
        if not args.no_cpu:
            cpu = cpu_baseline_spmv(vbr, routes, x, procs=args.cpu_procs)
    if (args.workload == "c2" and not args.no_spmm) or args.workload == "c3":
        vbr3, X3, _ = synth.workload_c3()
        if not args.no_cpu:
            cpu3 = cpu_baseline_spmm(vbr3, threshold_routes(vbr3.val, vbr3.indx, args.threshold), X3,
                                     procs=args.cpu_procs)

    from sable_amd import runtime
    runtime.init(0)
    info = runtime.device_info()
    info.update(runtime.bandwidth_probe(256 << 20, reps=3))
    log(f"device: {info}")
    if args.workload == "c5":
        return bench_batch(args, runtime, info)
    elif args.workload == "partition":
        return bench_partition(args, runtime, info)
    elif args.workload == "inspect":
        return bench_inspect(args, runtime, info)
    elif args.workload == "io":
        return bench_io(args, runtime, info)
    elif args.workload == "c3":
        return bench_spmm(args, runtime, info, cpu3, vbr3, X3)
    opts = spmv_options(args)
    xd = runtime.DeviceArray.from_numpy(x)
    yd = runtime.DeviceArray(int(vbr.rpntr[-1]))
    flush = 0 if args.warm else FLUSH_BYTES
    y64, ay = c_oracle.spmv_f64_dense_blocks(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr,
                                             vbr.bpntrb, vbr.bpntre, x)
    plans = {}

    def measure(routes, label, wg=None):
        t1 = time.time()
        plan = runtime.Plan.from_vbr(vbr, routes, **dict(opts, **({"wg": wg} if wg else {})))
        inspector_s = time.time() - t1
        st = plan.stats
        log(f"{label} plan: {st['ntiles']} tiles, rtc={st['rtc_used']}, grid={st['grid']}, "
            f"bytes_alg={st['bytes_alg']}, built in {inspector_s:.2f}s")
        times = plan.time(xd, yd, warmup=args.warmup, reps=args.steps, flush_bytes=flush)
        errs = c_oracle.scaled_error(yd.numpy(), y64, ay)
        ms = float(np.mean(times))
        r = {"ms": ms, "median_ms": float(np.median(times)), "error_vs_fp64": errs, "bytes_alg": st["bytes_alg"],
             "flops_alg": st["flops_alg"], "inspector_s": inspector_s, "stats": st,
             "plan": {k: st[k] for k in ("ntiles", "npieces", "dense_values", "sparse_nnz", "sparse_padded",
                                         "rowptr_entries", "xmap_entries", "nblocks_dense", "nblocks_sparse",
                                         "rtc_used", "grid", "mode", "wg")}}
        log(f"{label}: {ms * 1e3:.1f} us, {st['bytes_alg'] / (ms * 1e-3) / 1e9:.0f} GB/s, "
            f"error {errs['componentwise']:.2e}")
        plans[label] = plan
        return r

    only = [v for v in args.spmv_plans.split(",") if v]      # profiling: a subset of the plans
    res = {"plain": measure(routes, "plain")} if not only or "plain" in only else {}
    kept = "plain"
    if args.routing in ("learned", "both"):
        from sable_amd.router import Router
        if Router.available():
            from sable_amd.router.router import BIAS_GRID
            router = Router.load()
            # one learned plan per sparse-route bias; the fastest within 1e-5 is "learned"
            cands = []
            for sb, mb in BIAS_GRID["spmv"]:
                # the router's routes under each tile shape (workgroup
                # size); its own shape pick is marked, not trusted: the
                # per-block costs do not see the plan's tile count
                shapes, pick = router.routes_per_shape(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr,
                                                       vbr.bpntrb, vbr.bpntre, 1, sparse_bias=sb, mfma_bias=mb)
                for s_i, (lr, wgl) in enumerate(shapes):
                    lab = f"learned sparse_bias={sb:g} wg={wgl}"
                    if only and lab not in only:
                        continue
                    r = measure(lr, lab, wg=wgl)
                    r["sparse_bias"] = sb
                    r["wg"] = wgl
                    r["router_pick"] = s_i == pick
                    r["routes_changed"] = int(np.sum(lr != routes))
                    r["label"] = lab
                    cands.append(r)
            good = [r for r in cands if r["error_vs_fp64"]["componentwise"] <= 1e-5]
            if cands:
                res["learned"] = min(good or cands, key=lambda r: r["ms"])
                res["learned_candidates"] = [{k: r[k] for k in ("sparse_bias", "wg", "router_pick", "ms",
                                                                "median_ms", "bytes_alg")} for r in cands]
                ok = res["learned"]["error_vs_fp64"]["componentwise"] <= 1e-5
                # keep-if-faster at identical tolerance
                if (args.routing == "learned" or "plain" not in res
                        or (ok and res["learned"]["ms"] < res["plain"]["ms"])):
                    kept = "learned"
        else:
            log("no router weights; learned routing skipped")
    head = res[kept]
    label = head.get("label", "plain")
    # the kept plan's reported time: fresh launches (the selection used the samples above)
    times = retime(plans[label], xd, yd, args, flush)
    errs = c_oracle.scaled_error(yd.numpy(), y64, ay)
    assert errs["componentwise"] <= 1e-5, errs
    warm = plans[label].time(xd, yd, warmup=3, reps=max(5, args.steps // 2), flush_bytes=0)
    spun = retime_spun_up(plans[label], xd, yd, args, flush)
    for p in plans.values():
        p.close()
    ms = float(np.mean(times))
    st = head["stats"]
    gflops = st["flops_alg"] / (ms * 1e-3) / 1e9
    achieved = st["bytes_alg"] / (ms * 1e-3) / 1e9
    traffic = traffic_for("spmv_c2" if args.workload == "c2" else "spmv_c1", label, st, wg=st["wg"])

    cpu_line = None
    if cpu is not None:
        cres, kind_flags, cores = cpu
        mc = cres["allcores"]["mean_s"]
        m1 = cres["1thread"]["mean_s"]
        fl = st["flops_alg"]
        cpu_line = {"value": fl / mc / 1e9, "unit": "GFLOP/s", "cores": cores, "kind": "port",
                    "sample": (f"whole matrix, FP64 restated generated program (CSR part + dense block calls, "
                               f"threshold routes), eval.py protocol: {cres['allcores']['runs']} processes x "
                               f"{cres['allcores']['bench']} timed passes on {cores} threads (LPT block-row "
                               f"chunks), per-process median, 10-90% decile filter, mean; built with "
                               f"{kind_flags}"),
                    "single_thread_gflops": fl / m1 / 1e9,
                    "single_thread_sample": (f"{cres['1thread']['runs']} processes x {cres['1thread']['bench']} "
                                             f"passes pinned to one core"),
                    "runs_kept": {k: cres[k]["kept_runs"] for k in ("1thread", "allcores")},
                    "error_vs_fp64_oracle": c_oracle.scaled_error(cres["y"], y64, ay)["componentwise"],
                    "gpu_speedup_vs_allcores": mc / (ms * 1e-3)}

    for r in res.values():
        if isinstance(r, dict):
            r.pop("stats", None)
    line = {
        "metric": "VBR SpMV/SpMM GFLOP/s + % HBM/MFMA peak, learned-routed vs plain, 1 GPU",
        "value": gflops, "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": desc,
                   "routing": (f"{kept} (learned = fastest of the router's plans at sparse biases "
                               f"{sorted(set(c['sparse_bias'] for c in res['learned_candidates']))} under its "
                               f"tile shapes (workgroup sizes "
                               f"{sorted(set(c['wg'] for c in res['learned_candidates']))}), kept only if faster "
                               f"at 1e-5, then timed again with fresh launches; plain = fixed threshold "
                               f"T={args.threshold:g})" if "learned" in res
                               else f"plain fixed threshold T={args.threshold:g}"),
                   "cache": "warm" if args.warm else "cold (1 GiB flush between launches)",
                   "nrows": int(vbr.rpntr[-1]), "ncols": int(vbr.cpntr[-1]), "nnz": head["flops_alg"] // 2,
                   "stored_values": int(len(vbr.val)), "blocks": int(len(vbr.bindx))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu_line,
        "kept": kept, "kept_plan": label,
        "median_ms": float(np.median(times)), "p10_ms": float(np.percentile(times, 10)),
        "p90_ms": float(np.percentile(times, 90)), "warm_ms": float(np.mean(warm)),
        "spun_up_ms": float(np.mean(spun)),
        "error_vs_fp64": errs, "bytes_alg": head["bytes_alg"],
        "routings": res,
        "learned_speedup_vs_plain": (res["plain"]["ms"] / res["learned"]["ms"]) if "learned" in res and "plain" in res
        else None,
        "options": opts,
        "device": info,
    }
    if vbr3 is not None:
        del vbr
        line["spmm_c3"] = spmm_c3(args, runtime, vbr3, X3, threshold_routes(vbr3.val, vbr3.indx, args.threshold),
                                  cpu3)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
