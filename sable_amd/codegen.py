"""Per-matrix generate / compile / run flow (reference: src/codegen.py).

The reference emits one C program per matrix, compiles it with gcc and runs
it, parsing stdout.  Here the same three steps are:

generate  ``gen_single_threaded_spmv`` / ``gen_single_threaded_spmm`` (same
          names and arguments as src/codegen.py:293 / :827) write the
          program artefact ``{dir_name}/{filename}.sable.npz`` (the routed
          matrix and run options) and ``{dir_name}/{filename}.hip`` (the
          kernel specialisation chosen for this matrix);
compile   the specialisation is compiled by hipRTC for gfx950 into the
          code-object cache during generation (no GPU needed), replacing
          ``gcc {f}.c`` (bench_launcher.py:85);
run       ``run_program(dir_name, filename)`` (or ``python -m
          sable_amd.codegen run {dir_name}/{filename}``) executes it on the
          GPU and reproduces the binary's stdout contract: line 0
          ``{filename} = t0,t1,...,`` (ns per timed iteration, the first of
          bench+1 dropped, src/codegen.py:414-444), then one ``%lf`` value of
          y per line (:445-447); SpMM prints ``{filename} = t`` then every
          Y value with ``%f`` row-major (:927-934).

Deviations (documented): SpMV runs in FP32 on the GPU (the reference's C is
FP64); the SpMM RHS width is a parameter (``n_rhs``, default 512 as the
reference); the threaded generators produce the same GPU program as the
single-threaded ones (the thread count only shapes the host CPU program in
the reference).  Two defects of the reference's threaded emitters are not
reproduced: its SpMV passes ``indx[count]`` instead of
``indx[indx_start + count]`` to two kernel shapes (src/codegen.py:615-624),
and its SpMM gives block rows whose blocks are all sparse-routed zero work,
so ``split_chunks`` assigns them to no thread and their rows stay zero
(src/codegen.py:961-986, :987).
"""
from __future__ import annotations

import json
import os
import pathlib
import sys
import time

import numpy as np

from . import convert
from .fileio import read_vbr, read_vbrc, read_dense_values

FILEPATH = pathlib.Path(__file__).resolve().parent
BASE_PATH = os.path.join(FILEPATH, "..")
DENSE_DIR = os.path.join(BASE_PATH, "Generated_dense_tensors")

KIND_SPMV, KIND_SPMM_F32, KIND_SPMM_MFMA, KIND_SPMM_MFMA2 = 0, 1, 2, 3


def kernel_spec(plan_host):
    """The hipRTC specialisation the runtime selects for a plan:
    (kind, wg, unroll, xcap, nrhs, rtmax)."""
    return plan_host.kernel_spec()


def _kernel_source(spec):
    kind, wg, unroll, xcap, nrhs, rtmax = spec
    split = 2 if kind == KIND_SPMM_MFMA2 else 3
    body = {KIND_SPMV: "sb_spmv_body<SB_WG, SB_UNROLL, SB_XCAP>(a);",
            KIND_SPMM_F32: "sb_spmm_f32_body<SB_WG, SB_RTMAX, SB_NRHS, SB_XCAP>(a);",
            KIND_SPMM_MFMA: "sb_spmm_mfma_body<SB_WG, SB_NRHS, SB_XCAP, SB_SPLIT>(a);",
            KIND_SPMM_MFMA2: "sb_spmm_mfma_body<SB_WG, SB_NRHS, SB_XCAP, SB_SPLIT>(a);"}[kind]
    args = "SbSpmvArgs" if kind == KIND_SPMV else "SbSpmmArgs"
    bounds = {KIND_SPMV: "SB_WG, SB_SPMV_WAVES(SB_UNROLL)",
              KIND_SPMM_F32: "SB_WG, SB_F32_WAVES(SB_NRHS)"}.get(kind, "SB_WG, SB_MFMA_WAVES")
    return (f"// kernel specialisation for this matrix (compiled by hipRTC for gfx950)\n"
            f"#define SB_WG {wg}\n#define SB_UNROLL {unroll}\n#define SB_XCAP {xcap}\n"
            f"#define SB_NRHS {nrhs}\n#define SB_RTMAX {rtmax}\n#define SB_SPLIT {split}\n"
            f"#include \"sable_kernels.h\"\n#include \"sable_mfma.h\"\n"
            f"extern \"C\" __global__ void __launch_bounds__({bounds}) sable_program({args} a)\n"
            f"{{ {body} }}\n")


def _compile(spec):
    from ._lib import lib, check
    check(lib.sable_rtc_precompile(*spec, b"gfx950"))


def _emit(dir_name, filename, arrays, meta, op, n_rhs, routes):
    split = meta.get("mfma_split", 3)
    from .runtime import Plan
    os.makedirs(dir_name, exist_ok=True)
    np.savez(os.path.join(dir_name, filename + ".sable.npz"), routes=routes, meta=np.array(json.dumps(meta)),
             **arrays)
    host = Plan(arrays["val"], arrays["indx"], arrays["bindx"], arrays["rpntr"], arrays["cpntr"],
                arrays["bpntrb"], arrays["bpntre"], routes, op=op, n_rhs=n_rhs, host_only=True, mfma_split=split)
    spec = kernel_spec(host)
    with open(os.path.join(dir_name, filename + ".hip"), "w") as f:
        f.write(_kernel_source(spec))
    _compile(spec)
    return spec


def gen_single_threaded_spmv(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, ublocks, indptr, indices,
                             csr_val, dir_name, filename, vbr_dir, bench: int = 5) -> int:
    """src/codegen.py:293-452: takes the compressed (.vbrc) arrays, writes the
    program and compiles its kernel.  Returns elapsed milliseconds."""
    t1 = time.time_ns() // 1_000_000
    v, ix, routes = convert.decompress_vbrc(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, ublocks, indptr,
                                            indices, csr_val)
    arrays = dict(val=v, indx=ix, bindx=np.asarray(bindx), rpntr=np.asarray(rpntr), cpntr=np.asarray(cpntr),
                  bpntrb=np.asarray(bpntrb), bpntre=np.asarray(bpntre))
    meta = {"op": "spmv", "n_rhs": 1, "bench": int(bench), "filename": filename,
            "vector": os.path.join(DENSE_DIR, f"generated_vector_{int(cpntr[-1])}.vector")}
    _emit(dir_name, filename, arrays, meta, "spmv", 1, routes)
    return time.time_ns() // 1_000_000 - t1


def gen_multi_threaded_spmv(threads, val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, ublocks, indptr, indices,
                            csr_val, dir_name: str, filename: str, vbr_dir: str, bench: int = 5) -> int:
    """src/codegen.py:454-643.  The GPU program does not depend on the host
    thread count, so this emits the same program as the single-threaded
    generator."""
    return gen_single_threaded_spmv(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, ublocks, indptr, indices,
                                    csr_val, dir_name, filename, vbr_dir, bench)


def vbr_spmv_codegen(filename: str, dir_name: str, vbr_dir: str, threads: int, bench: int = 5) -> int:
    """src/codegen.py:1438-1447: reads ``{vbr_dir}/{filename}.vbrc`` and
    generates the program.  Returns the generation time in ms (the read is
    outside the timed region, as in the reference)."""
    a = read_vbrc(os.path.join(vbr_dir, filename + ".vbrc"))
    val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, ublocks, indptr, indices, csr_val = a
    t1 = time.time_ns() // 1_000_000
    if threads == 1:
        gen_single_threaded_spmv(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, ublocks, indptr, indices,
                                 csr_val, dir_name, filename, vbr_dir, bench)
    else:
        gen_multi_threaded_spmv(threads, val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, ublocks, indptr,
                                indices, csr_val, dir_name, filename, vbr_dir, bench)
    return time.time_ns() // 1_000_000 - t1


# Worst-case componentwise error of one product term a*x on the MFMA route,
# relative to |a||x|, from the bf16 terms each split keeps (u = 2^-8, the
# unit roundoff of bf16 round to nearest): the 3-term split drops a2x3,
# a3x2 and a3x3 (a1 + a2 + a3 == a exactly), about 2u^3; hi+lo drops a2x2
# and the two products with the residual beyond two terms, about 3u^2.
# FP32 accumulation adds its own rounding on top of either.
MFMA_SPLIT_BOUND = {"bf16x6": 2.0 * 2.0 ** -24 + 2.0 ** -32, "bf16x3": 3.0 * 2.0 ** -16 + 2.0 ** -24}


def split_chunks(values, num_chunks):
    """src/codegen.py:165-186: greedy longest-processing-time split of the
    indices of the nonzero entries of ``values`` into at most ``num_chunks``
    lists (fewer when fewer entries are nonzero).  Entries are taken by value,
    largest first (equal values keep index order), each into the first chunk
    with the smallest running sum; zero entries are in no chunk."""
    nonzero = [(i, v) for i, v in enumerate(values) if v != 0]
    num_chunks = min(int(num_chunks), len(nonzero))
    chunks = [[] for _ in range(num_chunks)]
    sums = [0] * num_chunks
    for i, v in sorted(nonzero, key=lambda e: e[1], reverse=True):
        j = sums.index(min(sums))
        chunks[j].append(i)
        sums[j] += v
    return chunks


def extract_mul_nums(output) -> list:
    """utils/utils.py:30-33: the timing fields of a program's first stdout
    line ``name = t0,t1,...,`` (strings, empty fields dropped)."""
    return [f for f in output.split("=")[1].split(",") if f != ""]


def _spmm_routes(val, indx, nblocks, density, precision, tol, n_rhs):
    if precision not in ("fp32", "bf16x6", "bf16x3"):
        raise ValueError(f"unknown precision {precision!r}")
    if precision != "fp32" and (n_rhs % 16 or n_rhs > 256):
        raise ValueError(f"precision {precision!r} (the MFMA route) needs n_rhs a multiple of 16 and at most "
                         f"256, got {n_rhs}")
    if precision != "fp32" and MFMA_SPLIT_BOUND[precision] > tol:
        raise ValueError(f"precision {precision!r} is bounded by {MFMA_SPLIT_BOUND[precision]:.2e} relative "
                         f"to |A||X| per row, above the tolerance {tol:g}; use 'bf16x6' or pass a larger tol")
    if density > 0:
        routes = convert.threshold_routes(val, indx, density)
    else:
        routes = np.zeros(nblocks, dtype=np.uint8)
    if precision != "fp32":
        routes = np.where(routes == convert.ROUTE_DENSE, convert.ROUTE_DENSE_MFMA, routes).astype(np.uint8)
    return routes


def gen_single_threaded_spmm(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, density: int, dir_name: str,
                             filename: str, vbr_dir: str, n_rhs: int = 512, precision: str = "fp32",
                             tol: float = 1e-5):
    """src/codegen.py:827-937 with the RHS width as a parameter.  Blocks are
    routed inline like the reference (:906-922): with density > 0 a block is
    dense iff its nonzero percentage > density, otherwise every block is
    dense.  precision: "fp32" (vector FP32 path), "bf16x6" (dense blocks on
    MFMA with the exact 3-term bf16 split, 6 passes) or "bf16x3" (hi+lo
    split, 3 passes).  A precision whose worst-case error bound
    (MFMA_SPLIT_BOUND) exceeds ``tol`` is refused with ValueError, so
    "bf16x3" needs tol >= about 4.6e-5.  Returns elapsed microseconds."""
    t1 = time.time_ns() // 1_000
    val = np.asarray(val)
    routes = _spmm_routes(val, indx, len(bindx), density, precision, tol, int(n_rhs))
    arrays = dict(val=val, indx=np.asarray(indx), bindx=np.asarray(bindx), rpntr=np.asarray(rpntr),
                  cpntr=np.asarray(cpntr), bpntrb=np.asarray(bpntrb), bpntre=np.asarray(bpntre))
    meta = {"op": "spmm", "n_rhs": int(n_rhs), "bench": 1, "filename": filename,
            "mfma_split": 2 if precision == "bf16x3" else 3,
            "matrix": os.path.join(DENSE_DIR, f"generated_matrix_{int(cpntr[-1])}x{int(n_rhs)}.matrix")}
    _emit(dir_name, filename, arrays, meta, "spmm", int(n_rhs), routes)
    return time.time_ns() // 1_000 - t1


def gen_multi_threaded_spmm(threads, val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, density: int,
                            dir_name: str, filename: str, vbr_dir: str, n_rhs: int = 512,
                            precision: str = "fp32", tol: float = 1e-5):
    """src/codegen.py:939-1075.  The reference splits the block rows over
    ``threads`` pthreads by split_chunks of their dense work; the GPU program
    covers every block row in one launch whatever the thread count, so this
    emits the single-threaded generator's program.  Block rows whose blocks
    are all sparse-routed (zero dense work, which the reference's chunks
    leave out, :986-987) are computed.  Returns elapsed microseconds."""
    if int(threads) < 1:
        raise ValueError("threads must be at least 1")
    return gen_single_threaded_spmm(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, density, dir_name, filename,
                                    vbr_dir, n_rhs, precision, tol)


def vbr_spmm_codegen(filename: str, density: int, dir_name: str, vbr_dir: str, threads: int,
                     n_rhs: int = 512, precision: str = "fp32", tol: float = 1e-5):
    """src/codegen.py:1449-1458: reads ``{vbr_dir}/{filename}.vbr`` and
    dispatches on ``threads`` (1: single-threaded generator, else the
    multi-threaded one).  Returns the generation time in microseconds."""
    val, indx, bindx, rpntr, cpntr, bpntrb, bpntre = read_vbr(os.path.join(vbr_dir, filename + ".vbr"))
    t1 = time.time_ns() // 1_000
    if threads == 1:
        gen_single_threaded_spmm(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, density, dir_name, filename,
                                 vbr_dir, n_rhs, precision, tol)
    else:
        gen_multi_threaded_spmm(threads, val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, density, dir_name,
                                filename, vbr_dir, n_rhs, precision, tol)
    return time.time_ns() // 1_000 - t1


def load_program(dir_name, filename):
    with np.load(os.path.join(dir_name, filename + ".sable.npz"), allow_pickle=False) as z:
        meta = json.loads(str(z["meta"]))
        arrays = {k: z[k] for k in ("val", "indx", "bindx", "rpntr", "cpntr", "bpntrb", "bpntre", "routes")}
    return meta, arrays


def run_program(dir_name, filename, x=None):
    """Executes a generated program on the GPU.  Returns (times_ns, y).  x
    defaults to the dense tensor file the reference reads
    (src/codegen.py:298, :830)."""
    from .runtime import Plan, DeviceArray
    meta, a = load_program(dir_name, filename)
    op, n = meta["op"], meta["n_rhs"]
    plan = Plan(a["val"], a["indx"], a["bindx"], a["rpntr"], a["cpntr"], a["bpntrb"], a["bpntre"], a["routes"],
                op=op, n_rhs=n, mfma_split=meta.get("mfma_split", 3))
    ncols = int(a["cpntr"][-1])
    if x is None:
        path = meta["vector"] if op == "spmv" else meta["matrix"]
        if not os.path.exists(path):
            raise FileNotFoundError(f"Error opening file2: {path}")
        x = read_dense_values(path)[: ncols * n]
    x = np.ascontiguousarray(x, dtype=np.float32).reshape((ncols,) if op == "spmv" else (ncols, n))
    xd = DeviceArray.from_numpy(x)
    yd = DeviceArray((plan.nrows,) if op == "spmv" else (plan.nrows, n))
    bench = meta["bench"]
    ms = plan.time(xd, yd, warmup=1, reps=bench)
    times = [int(round(t * 1e6)) for t in ms]
    return times, yd.numpy()


def print_program_output(filename, times, y, op="spmv", out=None):
    """The generated binary's stdout (src/codegen.py:440-447, :927-934),
    written to ``out`` (default: the current sys.stdout)."""
    out = sys.stdout if out is None else out
    if op == "spmv":
        out.write(f"{filename} = " + "".join(f"{t}," for t in times) + "\n")
        for v in np.asarray(y, dtype=np.float64).ravel():
            out.write("%lf\n".replace("l", "") % v)
    else:
        out.write(f"{filename} = {times[0]}\n")
        for v in np.asarray(y, dtype=np.float64).ravel():
            out.write("%f\n" % v)


def main(argv=None):
    argv = argv or sys.argv[1:]
    if len(argv) != 2 or argv[0] != "run":
        print("usage: python -m sable_amd.codegen run <dir>/<filename>", file=sys.stderr)
        return 2
    d, f = os.path.split(argv[1])
    meta, _ = load_program(d or ".", f)
    times, y = run_program(d or ".", f)
    print_program_output(f, times, y, meta["op"])
    return 0


if __name__ == "__main__":
    sys.exit(main())
