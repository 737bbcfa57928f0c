"""The reference's CPU timing protocol (eval.py:14-45) — TEST / CPU-BASELINE ONLY.

The reference times a generated program by running it as 100 separate
processes pinned to core 0 (``taskset -a -c 0``, eval.py:30), taking the
median of each run's in-program timings (the ``bench`` iterations printed on
stdout line 0, parsed with utils/utils.py:30-33), dropping the runs outside
the 10th-90th percentile (``remove_outliers_deciles``, eval.py:14-21) and
reporting the mean of the rest (eval.py:37-39).

Here each run is a child process (``python -m oracle.cpu_protocol``) that
loads the program's arrays, runs the restated program's timed loop
(oracle/spmv_ref.c) and prints the same stdout line 0 (``name = t0,t1,...,``,
src/codegen.py:440-444); the parent parses it, filters and averages the same
way.  Single-threaded runs are pinned to one core (the first this process may
use); threaded runs (LPT chunks of block rows, src/codegen.py:165-186) are
not pinned, unlike eval.py, which pins every run to core 0 whatever its
thread count.  Children are started before anything touches the GPU.
"""
from __future__ import annotations

import os
import statistics
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def remove_outliers_deciles(data):
    """eval.py:14-21: keep values within [P10, P90] (numpy's linear
    percentile); fewer than 10 values are kept as they are."""
    if len(data) < 10:
        return list(data)
    d1 = np.percentile(data, 10)
    d9 = np.percentile(data, 90)
    return [x for x in data if d1 <= x <= d9]


def parse_line0(line):
    """Timings of stdout line 0, ``name = t0,t1,...,`` (utils/utils.py:30-33)."""
    return [float(f) for f in line.split("=")[1].split(",") if f.strip() != ""]


def save_spmv_program(path, val2, indx2, bindx, rpntr, cpntr, bpntrb, bpntre, ublocks, indptr, indices, csr_val,
                      x):
    np.savez(path, kind=np.array("spmv"), val2=np.asarray(val2, np.float64), indx2=np.asarray(indx2, np.int64),
             bindx=bindx, rpntr=rpntr, cpntr=cpntr, bpntrb=bpntrb, bpntre=bpntre,
             ublocks=np.asarray(ublocks, np.int64), indptr=np.asarray(indptr, np.int64),
             indices=np.asarray(indices, np.int64), csr_val=np.asarray(csr_val, np.float64),
             x=np.asarray(x, np.float64))


def save_spmm_program(path, val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, route, X):
    np.savez(path, kind=np.array("spmm"), val=np.asarray(val, np.float32), indx=indx, bindx=bindx, rpntr=rpntr,
             cpntr=cpntr, bpntrb=bpntrb, bpntre=bpntre, route=np.asarray(route, np.uint8),
             X=np.asarray(X, np.float32))


def run_child(path, threads, bench, lib_path, name="prog"):
    """One process's work: returns the stdout line 0 of the program."""
    from oracle import c_oracle
    with np.load(path, allow_pickle=False) as z:
        a = {k: z[k] for k in z.files}
    if str(a["kind"]) == "spmv":
        prog = c_oracle.CProgram(a["val2"], a["indx2"], a["bindx"], a["rpntr"], a["cpntr"], a["bpntrb"],
                                 a["bpntre"], a["ublocks"], a["indptr"], a["indices"], a["csr_val"],
                                 dtype=np.float64, lib_path=lib_path)
        if threads > 1:
            prog.set_threads(threads)
        ns, _ = prog.time(a["x"], bench=bench)
    else:
        X = a["X"]
        _, ns = c_oracle.spmm(a["val"], a["indx"], a["bindx"], a["rpntr"], a["cpntr"], a["bpntrb"], a["bpntre"],
                              a["route"], X, X.shape[1], lib_path=lib_path, reps=bench, threads=threads)
    return f"{name} = " + "".join(f"{int(t)}," for t in ns)


def time_processes(path, threads, bench, procs=10, lib_path=None, timeout_s=600):
    """Runs ``procs`` child processes of the program at ``path`` and returns
    {"mean_s", "kept_runs", "runs", "run_medians_s"}: the mean over the runs
    kept by remove_outliers_deciles of each run's median iteration time."""
    pin = threads <= 1
    cmd = [sys.executable, "-m", "oracle.cpu_protocol", path, str(threads), str(bench), lib_path or "",
           "1" if pin else "0"]
    env = dict(os.environ)
    env["PYTHONPATH"] = ROOT + os.pathsep + env.get("PYTHONPATH", "")
    env["OMP_NUM_THREADS"] = str(max(1, threads))
    medians = []
    for _ in range(procs):
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout_s, env=env, cwd=ROOT)
        if out.returncode != 0:
            raise RuntimeError(f"CPU baseline process failed: {out.stderr[-2000:]}")
        line0 = out.stdout.split("\n")[0]
        medians.append(statistics.median(parse_line0(line0)) * 1e-9)
    kept = remove_outliers_deciles(medians)
    return {"mean_s": float(statistics.mean(kept)), "kept_runs": len(kept), "runs": procs,
            "run_medians_s": medians}


def main(argv):
    path, threads, bench, lib_path, pin = argv[0], int(argv[1]), int(argv[2]), argv[3] or None, argv[4] == "1"
    if pin:
        os.sched_setaffinity(0, {min(os.sched_getaffinity(0))})
    print(run_child(path, threads, bench, lib_path), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
