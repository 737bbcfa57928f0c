"""ctypes binding of libsable_amd.so (declarations mirror include/sable.h).

Importing this module loads the native library and raises ``ImportError``
when it has not been built: the GPU path has no CPU fallback.  The
environment variable ``SABLE_AMD_LIBRARY`` names another build of the
library to load instead (same C-ABI).
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SABLE_AMD_LIBRARY") or os.path.join(PKG, "libsable_amd.so")
RTC_CACHE = os.path.join(PKG, "_rtc_cache")

SABLE_OK = 0
SABLE_ERR_INVALID = 1
SABLE_ERR_HIP = 2
SABLE_ERR_OOM = 3
SABLE_ERR_RTC = 4
SABLE_ERR_NO_DEVICE = 5
SABLE_ERR_UNSUPPORTED = 6

ROUTE_DENSE = 0
ROUTE_SPARSE = 1
ROUTE_DENSE_MFMA = 2

OP_SPMV = 0
OP_SPMM = 1

MODE_GRID = 0
MODE_STATIC = 1
MODE_QUEUE = 2

HOST_PTRS = 1
NO_SYNC = 2

P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
F32P = C.POINTER(C.c_float)


class SableVbr(C.Structure):
    _fields_ = [("nbr", I32), ("nbc", I32), ("rpntr", P), ("cpntr", P), ("bindx", P),
                ("bpntrb", P), ("bpntre", P), ("indx", P), ("val", P), ("nblocks", I64)]


class SablePlanOpts(C.Structure):
    _fields_ = [("op", I32), ("n_rhs", I32), ("wg", I32), ("xcap", I32), ("tile_bytes", I64),
                ("max_seg_rows", I32), ("min_seg_rows", I32),
                ("mode", I32), ("grid", I32), ("unroll", I32), ("use_rtc", I32),
                ("host_only", I32), ("mfma_split", I32),
                ("min_tiles", I32), ("mfma_kchunk", I32)]


class SablePlanStats(C.Structure):
    _fields_ = [("nrows", I32), ("ncols", I32), ("ntiles", I32), ("npieces", I32),
                ("nxranges", I32), ("ntiles_mfma", I32), ("n_rhs", I32), ("nblocks", I64), ("nblocks_dense", I64),
                ("nblocks_sparse", I64), ("nblocks_mfma", I64), ("nnz", I64),
                ("dense_values", I64), ("dense_padded", I64), ("mfma_values", I64),
                ("mfma_padded", I64), ("sparse_nnz", I64), ("sparse_padded", I64),
                ("rowptr_entries", I64), ("xmap_entries", I64),
                ("descriptor_bytes", I64), ("bytes_alg", I64), ("flops_alg", I64),
                ("device_bytes", I64), ("rtc_used", I32), ("mode", I32), ("grid", I32),
                ("wg", I32), ("ntiles_split", I32), ("scratch_bytes", I64), ("fixup_bytes", I64),
                ("nspans", I64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class SableError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sable error {code}: {msg}")
        self.code = code


if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run `python -m sable_amd.build` "
                      "(the GPU path has no CPU fallback)")

lib = C.CDLL(LIB_PATH)

_SIGS = {
    "sable_init": (I32, [I32]),
    "sable_last_error": (C.c_char_p, []),
    "sable_device_info": (I32, [C.POINTER(I32), C.POINTER(I64), C.c_char_p, I32]),
    "sable_set_cache_dir": (I32, [C.c_char_p]),
    "sable_rtc_precompile": (I32, [I32, I32, I32, I32, I32, I32, C.c_char_p]),
    "sable_plan_create": (I32, [C.POINTER(SableVbr), P, C.POINTER(SablePlanOpts), C.POINTER(P)]),
    "sable_plan_destroy": (I32, [P]),
    "sable_plan_stats": (I32, [P, C.POINTER(SablePlanStats)]),
    "sable_plan_set_stream": (I32, [P, P]),
    "sable_plan_export": (I32, [P, I32, P, C.POINTER(I64)]),
    "sable_plan_kernel_spec": (I32, [P, C.POINTER(I32)]),
    "sable_spmv": (I32, [P, P, P, I32]),
    "sable_spmm": (I32, [P, P, P, I32]),
    "sable_spmm_noalloc": (I32, [P, P, P, I32]),
    "sable_time": (I32, [P, P, P, I32, I32, I64, F32P]),
    "sable_group_create": (I32, [C.POINTER(P), I32, C.POINTER(P), C.POINTER(P), I32, I32, C.POINTER(P)]),
    "sable_group_launch": (I32, [P, I32]),
    "sable_group_time": (I32, [P, I32, I32, I64, F32P]),
    "sable_group_destroy": (I32, [P]),
    "sable_time_plans": (I32, [C.POINTER(P), I32, C.POINTER(P), C.POINTER(P), I32, I32, I64, F32P]),
    "sable_mlp_forward": (I32, [P, I32, I32, I32, P, I32, P, P]),
    "sable_mlp_backward": (I32, [P, I32, I32, I32, P, I32, P, P, P]),
    "sable_mlp_mse_grad": (I32, [P, P, P, I64, C.c_float, P]),
    "sable_adam_step": (I32, [P, P, P, P, I64, C.c_float, C.c_float, C.c_float, C.c_float, I32]),
    "sable_bandwidth_probe": (I32, [I64, I32, F32P, F32P]),
    "sable_axis_cuts": (I32, [P, P, P, I64, I64, I32, C.c_double, P, P, F32P]),
    "sable_inspect_create": (I32, [P, P, P, I64, I64, P, I32, P, I32, C.POINTER(P), C.POINTER(I64),
                                   C.POINTER(I64), F32P]),
    "sable_inspect_fetch": (I32, [P, P, P, P, P, P, P]),
    "sable_inspect_destroy": (I32, [P]),
    "sable_compress_create": (I32, [P, I64, P, P, I64, P, I32, P, I32, P, P, P, C.POINTER(P), C.POINTER(I64),
                                    C.POINTER(I64), C.POINTER(I64), F32P]),
    "sable_compress_fetch": (I32, [P, P, P, P, P, P, P]),
    "sable_compress_destroy": (I32, [P]),
    "sable_inspect_compress": (I32, [P, P, C.POINTER(P), C.POINTER(I64), C.POINTER(I64), C.POINTER(I64), F32P]),
    "sable_text_open": (I32, [C.c_char_p, I32, I32, C.POINTER(P), C.POINTER(I64)]),
    "sable_text_read_f64": (I32, [P, I32, P]),
    "sable_text_read_i64": (I32, [P, I32, P]),
    "sable_text_close": (None, [P]),
    "sable_text_write": (I32, [C.c_char_p, I32, C.POINTER(C.c_char_p), C.POINTER(I32), C.POINTER(P),
                               C.POINTER(I64), I32, I32]),
    "sable_text_last_error": (C.c_char_p, []),
    "sable_malloc": (I32, [C.POINTER(P), I64]),
    "sable_free": (I32, [P]),
    "sable_memcpy_h2d": (I32, [P, P, I64]),
    "sable_memcpy_d2h": (I32, [P, P, I64]),
    "sable_synchronize": (I32, []),
}

EXPORTED = tuple(_SIGS)

for _name, (_res, _args) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def last_error():
    msg = lib.sable_last_error()
    return msg.decode() if msg else ""


def check(rc):
    if rc != SABLE_OK:
        raise SableError(rc, last_error())
    return rc


lib.sable_set_cache_dir(RTC_CACHE.encode())
