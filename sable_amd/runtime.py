"""Python face of the native executor: device init, plans, launches, timing.

A ``Plan`` is the in-process replacement of one generated-and-compiled
SABLE program (src/codegen.py:293-452): it holds the matrix laid out in HBM
for one operation (SpMV, or SpMM with a given RHS width) and the kernel
specialised for it.  x/y may be numpy arrays (copied through host memory) or
``DeviceArray``s (device memory owned by the library's allocator).

The library runs on the system ROCm runtime.  A process that also
initialises torch's bundled ROCm runtime would hold two HIP/HSA runtimes, so
the GPU path here never goes through torch.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib

_initialised = False


def init(device: int = 0):
    """Select the GPU (idempotent).  Raises SableError without a GPU."""
    global _initialised
    if not _initialised:
        check(lib.sable_init(device))
        _initialised = True


def device_info():
    init()
    n = _lib.I32()
    hbm = _lib.I64()
    arch = C.create_string_buffer(64)
    check(lib.sable_device_info(C.byref(n), C.byref(hbm), arch, 64))
    return {"n_cu": n.value, "hbm_bytes": hbm.value, "arch": arch.value.decode()}


def bandwidth_probe(nbytes=2 << 30, reps=5):
    """Achievable streaming read and copy rates of the device, GB/s."""
    init()
    r = C.c_float()
    c = C.c_float()
    check(lib.sable_bandwidth_probe(int(nbytes), reps, C.byref(r), C.byref(c)))
    return {"read_gbs": r.value, "copy_gbs": c.value}


def _vbr_struct(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre):
    keep = {
        "rpntr": np.ascontiguousarray(rpntr, dtype=np.int32),
        "cpntr": np.ascontiguousarray(cpntr, dtype=np.int32),
        "bindx": np.ascontiguousarray(bindx, dtype=np.int32),
        "bpntrb": np.ascontiguousarray(bpntrb, dtype=np.int32),
        "bpntre": np.ascontiguousarray(bpntre, dtype=np.int32),
        "indx": np.ascontiguousarray(indx, dtype=np.int64),
        "val": np.ascontiguousarray(val, dtype=np.float32),
    }
    if len(keep["bindx"]) == 0:
        keep["bindx"] = np.zeros(1, dtype=np.int32)
    if len(keep["val"]) == 0:
        keep["val"] = np.zeros(1, dtype=np.float32)
    s = _lib.SableVbr()
    s.nbr = len(keep["rpntr"]) - 1
    s.nbc = len(keep["cpntr"]) - 1
    for k in ("rpntr", "cpntr", "bindx", "bpntrb", "bpntre", "indx", "val"):
        setattr(s, k, keep[k].ctypes.data)
    s.nblocks = len(np.asarray(bindx))
    return s, keep


class DeviceArray:
    """A device buffer from the library's allocator (sable_malloc)."""

    def __init__(self, shape, dtype=np.float32):
        init()
        self.shape = tuple(np.atleast_1d(shape).tolist()) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * self.dtype.itemsize
        p = _lib.P()
        check(lib.sable_malloc(C.byref(p), self.nbytes))
        self._p = p

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.copy_from(a)
        return d

    @classmethod
    def zeros(cls, shape, dtype=np.float32):
        return cls.from_numpy(np.zeros(shape, dtype=dtype))

    @property
    def ptr(self):
        return self._p.value

    def copy_from(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size != self.size:
            raise ValueError("size mismatch")
        check(lib.sable_memcpy_h2d(self._p, a.ctypes.data, self.nbytes))

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib.sable_memcpy_d2h(out.ctypes.data, self._p, self.nbytes))
        return out

    def free(self):
        if getattr(self, "_p", None) is not None and self._p.value:
            lib.sable_free(self._p)
            self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _ptr(t, numel=None):
    """Device pointer of a DeviceArray, or host pointer of a numpy array;
    checks dtype float32, contiguity and (when given) the element count so a
    wrong buffer is rejected on the host before any kernel sees it."""
    if isinstance(t, np.ndarray):
        if t.dtype != np.float32 or not t.flags.c_contiguous:
            raise ValueError("host arrays must be contiguous float32")
        if numel is not None and t.size != numel:
            raise ValueError(f"array has {t.size} elements, expected {numel}")
        return t.ctypes.data, True
    if not isinstance(t, DeviceArray):
        raise TypeError("x/y must be numpy arrays or DeviceArrays")
    if t.dtype != np.float32:
        raise ValueError("device arrays must be float32")
    if numel is not None and t.size != numel:
        raise ValueError(f"device array has {t.size} elements, expected {numel}")
    return t.ptr, False


class Plan:
    """A matrix laid out for one operation on the GPU.

    ``routes[k]`` (per stored block, ordinal order) is ROUTE_DENSE,
    ROUTE_SPARSE or ROUTE_DENSE_MFMA; None means all dense.

    Shape options left as None take the autotuned values of ``tuning.json``
    for the operation when ``tuned`` is true, else the library defaults.
    """

    def __init__(self, val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, routes=None, *,
                 op="spmv", n_rhs=1, host_only=False, wg=None, xcap=4095, tile_bytes=None,
                 max_seg_rows=0, min_seg_rows=0, mode=None,
                 grid=0, unroll=None, use_rtc=True, mfma_split=3, min_tiles=0, mfma_kchunk=0, tuned=True):
        if not host_only:
            init()
        given = {"wg": wg, "tile_bytes": tile_bytes, "mode": mode, "unroll": unroll}
        if tuned:
            from .autotune import tuned_options
            for k, v in tuned_options(op, n_rhs).items():
                if k in given and given[k] is None:
                    given[k] = v
        wg = given["wg"] or 256
        tile_bytes = given["tile_bytes"] or 0
        mode = _lib.MODE_GRID if given["mode"] is None else given["mode"]
        unroll = given["unroll"] or 0
        s, keep = _vbr_struct(val, indx, bindx, rpntr, cpntr, bpntrb, bpntre)
        o = _lib.SablePlanOpts()
        o.op = _lib.OP_SPMM if op == "spmm" else _lib.OP_SPMV
        o.n_rhs = n_rhs if op == "spmm" else 1
        o.wg, o.xcap, o.tile_bytes = wg, xcap, tile_bytes
        o.max_seg_rows, o.min_seg_rows = max_seg_rows, min_seg_rows
        o.mode, o.grid, o.unroll = mode, grid, unroll
        o.use_rtc, o.host_only = int(bool(use_rtc)), int(bool(host_only))
        o.mfma_split = int(mfma_split)
        o.min_tiles = int(min_tiles)
        o.mfma_kchunk = int(mfma_kchunk)
        r = None
        if routes is not None:
            r = np.ascontiguousarray(routes, dtype=np.uint8)
            if len(r) != s.nblocks:
                raise ValueError(f"routes has {len(r)} entries for {s.nblocks} blocks")
        h = _lib.P()
        check(lib.sable_plan_create(C.byref(s), r.ctypes.data if r is not None else None,
                                    C.byref(o), C.byref(h)))
        self._h = h
        self.op = op
        self.n_rhs = o.n_rhs
        self.host_only = host_only
        st = _lib.SablePlanStats()
        check(lib.sable_plan_stats(h, C.byref(st)))
        self.stats = st.as_dict()
        self.nrows, self.ncols = self.stats["nrows"], self.stats["ncols"]

    @classmethod
    def from_vbr(cls, vbr, routes=None, **kw):
        return cls(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb, vbr.bpntre,
                   routes, **kw)

    def close(self):
        if getattr(self, "_h", None):
            lib.sable_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- inspection -----------------------------------------------------
    _EXPORT = {"tiles": (0, None), "pieces": (1, None), "xranges": (2, np.int32),
               "dval": (3, np.float32), "sval": (4, np.float32), "sidx": (5, np.uint16),
               "srowptr": (6, np.int32), "xmap": (7, np.int32)}
    TILE_DTYPE = np.dtype([("row0", "<i4"), ("nrows", "<i4"), ("piece_begin", "<i4"),
                           ("piece_end", "<i4"), ("mat", "<i4"), ("dense_tpar", "<i4"),
                           ("dense_mcols", "<i4"), ("sp_glog", "<i4"),
                           ("sp_map", "<i4"), ("sp_waves", "<i4"),
                           ("y_accum", "<i4"), ("stage_map", "<i4"), ("stage_len", "<i4"),
                           ("ksplit", "<i4"), ("kpart", "<i4"), ("part_blk", "<i4")])
    PIECE_DTYPE = np.dtype([("data_off", "<i8"), ("kind", "<i4"), ("ncols", "<i4"),
                            ("xr_begin", "<i4"), ("xr_end", "<i4"), ("rowptr_off", "<i4"),
                            ("ksteps", "<i4"), ("grp_end", "<i4"), ("stage_len", "<i4"),
                            ("xmap_off", "<i4"), ("pitch", "<i4")])

    def export(self, name):
        which, dt = self._EXPORT[name]
        if name == "tiles":
            dt = self.TILE_DTYPE
        elif name == "pieces":
            dt = self.PIECE_DTYPE
        n = _lib.I64()
        check(lib.sable_plan_export(self._h, which, None, C.byref(n)))
        buf = np.zeros(n.value, dtype=np.uint8)
        if n.value:
            check(lib.sable_plan_export(self._h, which, buf.ctypes.data, C.byref(n)))
        out = buf.view(dt)
        if name == "xranges":
            out = out.reshape(-1, 2)
        return out

    # -- execution ------------------------------------------------------
    def _run(self, fn, x, y, shape_y):
        xp, xhost = _ptr(x, self.ncols * self.n_rhs)
        if y is None:
            y = np.empty(shape_y, dtype=np.float32) if xhost else DeviceArray(shape_y)
        yp, yhost = _ptr(y, int(np.prod(shape_y)))
        if xhost != yhost:
            raise ValueError("x and y must both be host arrays or both device tensors")
        check(fn(self._h, xp, yp, _lib.HOST_PTRS if xhost else 0))
        return y

    def spmv(self, x, y=None):
        """y = A x (x: ncols float32)."""
        if self.op != "spmv":
            raise ValueError("plan was built for SpMM")
        if isinstance(x, np.ndarray):
            x = np.ascontiguousarray(x, dtype=np.float32)
            if x.shape != (self.ncols,):
                raise ValueError(f"x must have shape ({self.ncols},)")
        return self._run(lib.sable_spmv, x, y, (self.nrows,))

    def spmm(self, X, Y=None):
        """Y = A X (X: ncols x n_rhs float32, row-major)."""
        if self.op != "spmm":
            raise ValueError("plan was built for SpMV")
        if isinstance(X, np.ndarray):
            X = np.ascontiguousarray(X, dtype=np.float32)
            if X.shape != (self.ncols, self.n_rhs):
                raise ValueError(f"X must have shape ({self.ncols}, {self.n_rhs})")
        return self._run(lib.sable_spmm, X, Y, (self.nrows, self.n_rhs))

    def time(self, x, y, warmup=10, reps=100, flush_bytes=0):
        """Per-launch durations (ms) from HIP events on the plan's stream."""
        out = np.zeros(reps, dtype=np.float32)
        xp, xh = _ptr(x, self.ncols * self.n_rhs)
        yp, yh = _ptr(y, self.nrows * self.n_rhs)
        if xh or yh:
            raise ValueError("timing needs device tensors")
        check(lib.sable_time(self._h, xp, yp, warmup, reps, int(flush_bytes),
                             out.ctypes.data_as(_lib.F32P)))
        return out

    def kernel_spec(self):
        """(kind, wg, unroll, xcap, nrhs, rtmax) of the chosen kernel."""
        spec = (_lib.I32 * 6)()
        check(lib.sable_plan_kernel_spec(self._h, spec))
        return tuple(int(v) for v in spec)

    def set_stream(self, stream_ptr):
        check(lib.sable_plan_set_stream(self._h, stream_ptr))


def _dev_ptrs(plans, xs, ys):
    """Device pointers of every plan's x and y (DeviceArrays only: these
    launches have no host staging)."""
    n = len(plans)
    if len(xs) != n or len(ys) != n:
        raise ValueError("one x and one y per plan")
    xp, yp = [], []
    for p, x, y in zip(plans, xs, ys):
        if not isinstance(x, DeviceArray) or not isinstance(y, DeviceArray):
            raise TypeError("batched launches take DeviceArrays")
        xp.append(_ptr(x, p.ncols * p.n_rhs)[0])
        yp.append(_ptr(y, p.nrows * p.n_rhs)[0])
    return (_lib.P * n)(*[p._h for p in plans]), (_lib.P * n)(*xp), (_lib.P * n)(*yp)


def time_plans(plans, xs, ys, warmup=3, reps=10, flush_bytes=0):
    """Launches every plan's own kernel back to back (the unbatched way to
    run many matrices) and times the whole sequence; ms per repetition."""
    init()
    n = len(plans)
    hs, xp, yp = _dev_ptrs(plans, xs, ys)
    out = np.zeros(reps, dtype=np.float32)
    check(lib.sable_time_plans(hs, n, xp, yp, warmup, reps, int(flush_bytes), out.ctypes.data_as(_lib.F32P)))
    return out


class Group:
    """Many plans executed by one launch per (op, N, wg) class over their
    concatenated tiles, replayed from a hipGraph.  GRID mode (one workgroup
    per tile) is the default: it measured faster than the persistent modes
    on batches of SpMV and SpMM plans."""

    def __init__(self, plans, xs, ys, mode=_lib.MODE_GRID, grid=0):
        init()
        n = len(plans)
        hs, xp, yp = _dev_ptrs(plans, xs, ys)
        g = _lib.P()
        check(lib.sable_group_create(hs, n, xp, yp, mode, grid, C.byref(g)))
        self._g = g
        self._keep = (plans, xs, ys)

    def launch(self, sync=True):
        check(lib.sable_group_launch(self._g, int(sync)))

    def time(self, warmup=5, reps=20, flush_bytes=0):
        out = np.zeros(reps, dtype=np.float32)
        check(lib.sable_group_time(self._g, warmup, reps, int(flush_bytes), out.ctypes.data_as(_lib.F32P)))
        return out

    def close(self):
        if getattr(self, "_g", None):
            lib.sable_group_destroy(self._g)
            self._g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
