"""ctypes wrapper of oracle/spmv_ref.c — TEST / CPU-BASELINE ONLY.

``CProgram`` mirrors one generated reference program (src/codegen.py:293-452):
it is built once from the compressed (.vbrc) arrays, then ``run`` executes
the program body and ``time`` the timed loop of :414-439.  ``spmm`` restates
gen_single_threaded_spmm (:827-937) with the RHS width as a parameter.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PORTABLE = os.path.join(HERE, "_build", "liboracle.so")

CALL_DTYPE = np.dtype([("kind", "<i4"), ("i0", "<i4"), ("i1", "<i4"), ("j0", "<i4"),
                       ("j1", "<i4"), ("brow", "<i4"), ("off", "<i8")])


def build(native_dir=None):
    """Build the portable library (or, with native_dir, a -march=native one
    with the reference's CFLAGS into native_dir).  Returns its path."""
    if native_dir:
        subprocess.run(["make", "-s", "-C", HERE, "native", f"OUT={native_dir}"], check=True,
                       capture_output=True)
        return os.path.join(native_dir, "liboracle_native.so")
    subprocess.run(["make", "-s", "-C", HERE], check=True, capture_output=True)
    return PORTABLE


_libs = {}


def load(path=None):
    path = path or PORTABLE
    if path not in _libs:
        if not os.path.exists(path):
            build()
        lib = C.CDLL(path)
        P, I32, I64 = C.c_void_p, C.c_int32, C.c_int64
        lib.oracle_build_calls.restype = I64
        lib.oracle_build_calls.argtypes = [I32, I32, P, P, P, P, P, P, P, P, I64]
        for suf in ("f64", "f32"):
            getattr(lib, f"oracle_spmv_{suf}").argtypes = [P, I64, P, I32, I32, P, P, P, P, P]
            getattr(lib, f"oracle_spmv_mt_{suf}").argtypes = [P, P, P, P, P, I32, P, I32, I32, P, P, P, P, P]
            getattr(lib, f"oracle_time_spmv_{suf}").argtypes = [P, I64, P, P, P, P, I32, P, I32, I32,
                                                               P, P, P, P, P, I32, P]
        lib.oracle_spmm_f32.argtypes = [I32, I32, P, P, P, P, P, P, P, P, P, I32, P]
        lib.oracle_time_spmm_f32.argtypes = [I32, I32, P, P, P, P, P, P, P, P, P, I32, P, I32, P]
        lib.oracle_spmm_mt_f32.argtypes = [I32, I32, P, P, P, P, P, P, P, P, P, I32, P, P, P, I32]
        lib.oracle_time_spmm_mt_f32.argtypes = [I32, I32, P, P, P, P, P, P, P, P, P, I32, P, P, P, I32, I32, P]
        lib.oracle_max_threads.restype = I32
        _libs[path] = lib
    return _libs[path]


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def lpt_chunks(work, nchunks):
    """Greedy LPT over block rows (src/codegen.py:165-186): rows with zero
    work are dropped, the rest assigned largest first to the lightest chunk."""
    idx = [i for i, w in enumerate(work) if w != 0]
    nchunks = max(1, min(nchunks, len(idx)))
    order = sorted(idx, key=lambda i: work[i], reverse=True)
    chunks = [[] for _ in range(nchunks)]
    sums = [0] * nchunks
    for i in order:
        j = sums.index(min(sums))
        chunks[j].append(i)
        sums[j] += work[i]
    return chunks


class CProgram:
    """One reference program (SpMV) over compressed arrays."""

    def __init__(self, val2, indx2, bindx, rpntr, cpntr, bpntrb, bpntre, ublocks, indptr,
                 indices, csr_val, dtype=np.float64, lib_path=None):
        self.lib = load(lib_path)
        self.dtype = np.dtype(dtype)
        self.suf = "f64" if self.dtype == np.float64 else "f32"
        self.rpntr, self.cpntr = _i32(rpntr), _i32(cpntr)
        self.bindx, self.bpntrb, self.bpntre = _i32(bindx), _i32(bpntrb), _i32(bpntre)
        self.nrows, self.ncols = int(self.rpntr[-1]), int(self.cpntr[-1])
        nbr, nbc = len(self.rpntr) - 1, len(self.cpntr) - 1
        nblocks = len(self.bindx)
        sparse = np.zeros(max(nblocks, 1), dtype=np.uint8)
        ub = np.asarray(ublocks, dtype=np.int64)
        if len(ub):
            sparse[ub] = 1
        self.val2 = np.ascontiguousarray(val2, dtype=self.dtype)
        if len(self.val2) == 0:
            self.val2 = np.zeros(1, dtype=self.dtype)
        dense_off = _i64(indx2)
        calls = np.zeros(max(nblocks, 1), dtype=CALL_DTYPE)
        n = self.lib.oracle_build_calls(nbr, nbc, self.rpntr.ctypes.data, self.cpntr.ctypes.data,
                                        self.bindx.ctypes.data, self.bpntrb.ctypes.data,
                                        self.bpntre.ctypes.data, sparse.ctypes.data,
                                        dense_off.ctypes.data, calls.ctypes.data, len(calls))
        if n < 0:
            raise RuntimeError("oracle_build_calls failed")
        self.calls = calls[:n]
        self.has_csr = int(len(ub) > 0 and len(indptr) > 0)
        self.indptr = _i32(indptr if self.has_csr else [0])
        self.indices = _i32(indices if self.has_csr else [0])
        self.csr_val = np.ascontiguousarray(csr_val if self.has_csr else [0], dtype=self.dtype)
        # per block row call ranges (for the threaded body)
        counts = np.bincount(self.calls["brow"], minlength=nbr) if n else np.zeros(nbr, dtype=np.int64)
        self.call_ptr = _i64(np.concatenate([[0], np.cumsum(counts)]))
        work = np.zeros(nbr, dtype=np.int64)
        if n:
            np.add.at(work, self.calls["brow"],
                      (self.calls["i1"] - self.calls["i0"]).astype(np.int64) *
                      (self.calls["j1"] - self.calls["j0"]))
        if self.has_csr:
            row_nnz = np.diff(self.indptr.astype(np.int64))
            work += np.add.reduceat(row_nnz, self.rpntr[:-1]) if len(row_nnz) else 0
        self.work = work
        self.chunk_ptr = _i32([0])
        self.rows_of = _i32([0])
        self.nchunks = 1

    def set_threads(self, nthreads):
        """Split block rows into nthreads LPT chunks (all rows included so y
        is fully written)."""
        work = [int(w) if w else 1 for w in self.work]
        chunks = lpt_chunks(work, nthreads)
        self.nchunks = len(chunks)
        self.chunk_ptr = _i32(np.concatenate([[0], np.cumsum([len(c) for c in chunks])]))
        self.rows_of = _i32([r for c in chunks for r in c])

    def _args(self, x, y):
        return (self.calls.ctypes.data, len(self.calls), self.val2.ctypes.data, self.nrows,
                self.has_csr, self.indptr.ctypes.data, self.indices.ctypes.data,
                self.csr_val.ctypes.data, x.ctypes.data, y.ctypes.data)

    def run(self, x):
        x = np.ascontiguousarray(x, dtype=self.dtype)
        y = np.zeros(self.nrows, dtype=self.dtype)
        if self.nchunks <= 1:
            getattr(self.lib, f"oracle_spmv_{self.suf}")(*self._args(x, y))
        else:
            getattr(self.lib, f"oracle_spmv_mt_{self.suf}")(
                self.calls.ctypes.data, self.call_ptr.ctypes.data, self.rpntr.ctypes.data,
                self.chunk_ptr.ctypes.data, self.rows_of.ctypes.data, self.nchunks,
                self.val2.ctypes.data, self.nrows, self.has_csr, self.indptr.ctypes.data,
                self.indices.ctypes.data, self.csr_val.ctypes.data, x.ctypes.data, y.ctypes.data)
        return y

    def time(self, x, bench=100):
        """Per-iteration ns of the timed loop (first iteration dropped)."""
        x = np.ascontiguousarray(x, dtype=self.dtype)
        y = np.zeros(self.nrows, dtype=self.dtype)
        out = np.zeros(bench, dtype=np.int64)
        getattr(self.lib, f"oracle_time_spmv_{self.suf}")(
            self.calls.ctypes.data, len(self.calls), self.call_ptr.ctypes.data,
            self.rpntr.ctypes.data, self.chunk_ptr.ctypes.data, self.rows_of.ctypes.data,
            self.nchunks, self.val2.ctypes.data, self.nrows, self.has_csr,
            self.indptr.ctypes.data, self.indices.ctypes.data, self.csr_val.ctypes.data,
            x.ctypes.data, y.ctypes.data, bench, out.ctypes.data)
        return out, y


def spmm(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, route, X, n_rhs, lib_path=None,
         reps=0, threads=1):
    """Reference SpMM program body (FP32).  route[k] != 0 -> sparse block.
    With reps > 0 also returns per-run ns.  threads > 1 runs the threaded
    program (src/codegen.py:939-1075): block rows in LPT chunks of their
    stored size (every row in a chunk), bit-identical to one thread."""
    lib = load(lib_path)
    rp, cp = _i32(rpntr), _i32(cpntr)
    bx, bb, be = _i32(bindx), _i32(bpntrb), _i32(bpntre)
    ix = _i64(indx)
    v = np.ascontiguousarray(val, dtype=np.float32)
    r = np.ascontiguousarray(route, dtype=np.uint8)
    if len(r) == 0:
        r = np.zeros(1, dtype=np.uint8)
    Xc = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.zeros((int(rp[-1]), n_rhs), dtype=np.float32)
    args = (len(rp) - 1, len(cp) - 1, rp.ctypes.data, cp.ctypes.data, bx.ctypes.data,
            bb.ctypes.data, be.ctypes.data, ix.ctypes.data, v.ctypes.data, r.ctypes.data,
            Xc.ctypes.data, n_rhs, Y.ctypes.data)
    if threads > 1:
        bb_, be_ = np.asarray(bpntrb, np.int64), np.asarray(bpntre, np.int64)
        work = [int(ix[e] - ix[b]) if b != -1 else 0 for b, e in zip(bb_, be_)]
        chunks = lpt_chunks([w if w else 1 for w in work], threads)
        cptr = _i32(np.concatenate([[0], np.cumsum([len(c) for c in chunks])]))
        rows = _i32([r for c in chunks for r in c])
        targs = args + (cptr.ctypes.data, rows.ctypes.data, len(chunks))
        if reps > 0:
            t = np.zeros(reps, dtype=np.int64)
            lib.oracle_time_spmm_mt_f32(*targs, reps, t.ctypes.data)
            return Y, t
        lib.oracle_spmm_mt_f32(*targs)
        return Y
    if reps > 0:
        t = np.zeros(reps, dtype=np.int64)
        lib.oracle_time_spmm_f32(*args, reps, t.ctypes.data)
        return Y, t
    lib.oracle_spmm_f32(*args)
    return Y


def spmv_f64_dense_blocks(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, x):
    """FP64 y = A x from the VBR arrays, block by block in numpy (closed form;
    used at sizes where the pure-Python oracle is too slow)."""
    rp = np.asarray(rpntr, dtype=np.int64)
    cp = np.asarray(cpntr, dtype=np.int64)
    bx = np.asarray(bindx, dtype=np.int64)
    bb = np.asarray(bpntrb, dtype=np.int64)
    be = np.asarray(bpntre, dtype=np.int64)
    ix = np.asarray(indx, dtype=np.int64)
    v = np.asarray(val)
    x64 = np.asarray(x, dtype=np.float64)
    y = np.zeros(int(rp[-1]), dtype=np.float64)
    ay = np.zeros(int(rp[-1]), dtype=np.float64)   # (|A||x|)
    for a in range(len(rp) - 1):
        if bb[a] == -1:
            continue
        r0, r1 = rp[a], rp[a + 1]
        h = r1 - r0
        for k in range(bb[a], be[a]):
            c0, c1 = cp[bx[k]], cp[bx[k] + 1]
            blk = v[ix[k]:ix[k + 1]].astype(np.float64).reshape(c1 - c0, h).T
            y[r0:r1] += blk @ x64[c0:c1]
            ay[r0:r1] += np.abs(blk) @ np.abs(x64[c0:c1])
    return y, ay


def spmm_f64_dense_blocks(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, X):
    """FP64 Y = A X and |A||X| from the VBR arrays (closed form)."""
    rp = np.asarray(rpntr, dtype=np.int64)
    cp = np.asarray(cpntr, dtype=np.int64)
    bx = np.asarray(bindx, dtype=np.int64)
    bb = np.asarray(bpntrb, dtype=np.int64)
    be = np.asarray(bpntre, dtype=np.int64)
    ix = np.asarray(indx, dtype=np.int64)
    v = np.asarray(val)
    X64 = np.asarray(X, dtype=np.float64)
    Y = np.zeros((int(rp[-1]), X64.shape[1]), dtype=np.float64)
    AY = np.zeros_like(Y)
    for a in range(len(rp) - 1):
        if bb[a] == -1:
            continue
        r0, r1 = rp[a], rp[a + 1]
        h = r1 - r0
        for k in range(bb[a], be[a]):
            c0, c1 = cp[bx[k]], cp[bx[k] + 1]
            blk = v[ix[k]:ix[k + 1]].astype(np.float64).reshape(c1 - c0, h).T
            Y[r0:r1] += blk @ X64[c0:c1]
            AY[r0:r1] += np.abs(blk) @ np.abs(X64[c0:c1])
    return Y, AY


def scaled_error(y, y64, ay):
    """max_i |y_i - y64_i| / (|A||x|)_i over rows with (|A||x|)_i > 0, plus
    the normwise ||y - y64||_inf / ||y64||_inf and the elementwise max
    relative error (rows where y64 != 0)."""
    y = np.asarray(y, dtype=np.float64)
    d = np.abs(y - y64)
    m = ay > 0
    comp = float(np.max(d[m] / ay[m])) if np.any(m) else 0.0
    nrm = float(np.max(d) / max(np.max(np.abs(y64)), 1e-300)) if d.size else 0.0
    nz = y64 != 0
    elem = float(np.max(d[nz] / np.abs(y64[nz]))) if np.any(nz) else 0.0
    if np.any(~m & (d > 0)):
        comp = float("inf")
    return {"componentwise": comp, "normwise": nrm, "elementwise": elem}
