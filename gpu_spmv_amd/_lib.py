"""ctypes binding of libaxspmv.so (include/axspmv.h).

The library is built in-tree (``make -C gpu_spmv_amd/csrc``); loading fails
loudly when it is missing -- there is no Python or CPU fallback for the GPU
path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libaxspmv.so")

u32 = C.c_uint32
u64 = C.c_uint64
f64 = C.c_double
vp = C.c_void_p
P = C.POINTER


class AxspmvError(RuntimeError):
    """A negative status returned by the library, with its message."""

    def __init__(self, status: int, where: str, msg: str):
        super().__init__(f"{where}: status {status}: {msg}")
        self.status = status


# status codes (enum axspmv_status)
OK, E_INVALID, E_HIP, E_NOMEM, E_UNSUPPORTED, E_OVERFLOW, E_IO = 0, -1, -2, -3, -4, -5, -6
# formats / kernels / value types / AXT modes / flags
FMT_CSR, FMT_K1, FMT_AXC, FMT_AXT, FMT_RCT = 0, 1, 2, 3, 4
KER_DEFAULT, KER_NAIVE, KER_STREAM = 0, 1, 2
F64, F32 = 0, 1
AXT_UNC, AXT_COM = 0, 1
BUILD_AX, BUILD_PLANES = 1, 2
PLAN_TRANSPOSE = 1
PLAN_RCT_SPLIT = 2
PLAN_AX_LAYOUT = 4
PLAN_STAGE_X = 8
PLAN_DETERMINISTIC = 16
PLAN_RCT_ATOMIC = 32
PLAN_REORDER = 64
PLAN_RCT_WIDE_KEY = 128


class Csr(C.Structure):
    _fields_ = [("nrows", u32), ("ncols", u32), ("nnz", u32), ("owner", u32),
                ("row", P(u32)), ("col", P(u32)), ("val", P(f64)), ("rl", P(u32)),
                ("rmin", u32), ("rmax", u32), ("bw", u32), ("pad_", u32),
                ("rave", f64), ("rsd", f64)]


class FmtData(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("mfp", f64), ("beta", f64), ("ct", f64)]


class K1(C.Structure):
    _fields_ = [("nrows", u32), ("nnz", u32), ("chunk", u32), ("chunkNum", u32), ("lenVC", u32),
                ("reorder", u32), ("permi", P(u32)), ("nmc", P(u32)), ("chp", P(u32)),
                ("val", P(f64)), ("col", P(u32))]


class Axc(C.Structure):
    _fields_ = [("nrows", u32), ("nnz", u32), ("hbs", u32), ("lenAX", u32), ("lenBRP", u32),
                ("lenMAPX", u32), ("ax", P(f64)), ("brp", P(u32)), ("mapx", P(u32)),
                ("vplane", P(f64))]


class Axt(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("nrows", u32), ("nnz", u32), ("mode", u32),
                ("tileHW", u32), ("tileH", u32), ("logTH", u32), ("tileN", u32),
                ("lenAX", u32), ("lenSEC", u32), ("lenCON", u32), ("log", u32), ("bs", u32),
                ("ax", P(f64)), ("sec", P(u32)), ("con", P(u32)), ("vplane", P(f64)),
                ("cplane", P(u32))]


class Dcsr(C.Structure):
    _fields_ = [("nrows", u32), ("ncols", u32), ("nnz", u32), ("pad_", u32),
                ("row", C.c_void_p), ("col", C.c_void_p), ("val", C.c_void_p)]


class Dfmt(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("format", u32), ("mode", u32), ("hw", u32), ("th", u32),
                ("bs", u32), ("log", u32), ("nrows", u32), ("nnz", u32), ("tileN", u32), ("lenAX", u32),
                ("lenSEC", u32), ("lenCON", u32), ("lenMAPX", u32), ("pad_", u32),
                ("ax", C.c_void_p), ("sec", C.c_void_p), ("con", C.c_void_p), ("plane", C.c_void_p),
                ("ms", f64)]


class Opts(C.Structure):
    _fields_ = [("format", u32), ("kernel", u32), ("mode", u32), ("thw", u32), ("th", u32),
                ("bs", u32), ("hbs", u32), ("chunk", u32), ("reorder", u32), ("vtype", u32),
                ("block", u32), ("budget", u32), ("flags", u32), ("ktile", u32), ("tile_rows", u32),
                ("slices", u32)]


class PlanInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("nrows", u32), ("ncols", u32), ("nnz", u32),
                ("format", u32), ("slots", u64), ("device_bytes", u64), ("stream_bytes", u64),
                ("alg_bytes", u64), ("beta", f64), ("launches", u32), ("nblocks", u32),
                ("staged_blocks", u32), ("pad_", u32)]


class BatchInfo(C.Structure):
    _fields_ = [("nrows", u64), ("ncols", u64), ("nnz", u64), ("slots", u64), ("alg_bytes", u64),
                ("device_bytes", u64), ("count", u32), ("nitems", u32), ("nwaves", u32), ("pad_", u32)]


class Timing(C.Structure):
    _fields_ = [("mean_ms", f64), ("median_ms", f64), ("min_ms", f64), ("max_ms", f64),
                ("iters", u32), ("mode", u32), ("main_ms", f64)]


# enum axspmv_op
OP_SPMV, OP_ADJOINT, OP_VBAR, OP_FWD_ADJ, OP_FWD_ADJ_SPLIT, OP_SPMM, OP_XREFRESH = 0, 1, 2, 3, 4, 5, 6


class Operands(C.Structure):
    _fields_ = [("x", vp), ("b", vp), ("ybar", vp), ("X", vp), ("y", vp), ("r", vp), ("xbar", vp),
                ("vbar", vp), ("Y", vp), ("k", u32), ("pad_", u32)]


# name -> (restype, argtypes); every entry point declared in include/axspmv.h
SIGNATURES = {
    "axspmv_last_error": (C.c_char_p, []),
    "axspmv_abi_version": (C.c_int, []),
    "axspmv_csr_alloc": (C.c_int, [u32, u32, u32, P(P(Csr))]),
    "axspmv_csr_free": (None, [P(Csr)]),
    "axspmv_csr_check": (C.c_int, [P(Csr)]),
    "axspmv_csr_stats": (C.c_int, [P(Csr)]),
    "axspmv_csr_write_sta": (C.c_int, [P(Csr), C.c_char_p, C.c_char_p]),
    "axspmv_csr_read": (C.c_int, [C.c_char_p, P(P(Csr))]),
    "axspmv_csr_write": (C.c_int, [P(Csr), C.c_char_p]),
    "axspmv_csr_transpose": (C.c_int, [P(Csr), P(P(Csr))]),
    "axspmv_csr_rcm": (C.c_int, [P(Csr), P(C.c_uint32)]),
    "axspmv_csr_permute": (C.c_int, [P(Csr), P(C.c_uint32), P(P(Csr))]),
    "axspmv_rng_u64": (C.c_int, [u64, u64, P(u64), C.c_size_t]),
    "axspmv_rng_uniform": (C.c_int, [u64, u64, f64, f64, P(f64), C.c_size_t]),
    "axspmv_rng_normal": (C.c_int, [u64, u64, f64, f64, P(f64), C.c_size_t]),
    "axspmv_gen_laplace2d": (C.c_int, [u32, f64, f64, P(P(Csr))]),
    "axspmv_gen_powerlaw": (C.c_int, [u32, f64, u32, f64, u64, P(P(Csr))]),
    "axspmv_gen_stencil27": (C.c_int, [u32, f64, f64, C.c_int, u64, P(P(Csr))]),
    "axspmv_gen_banded": (C.c_int, [u32, u32, f64, u64, P(P(Csr))]),
    "axspmv_gen_random": (C.c_int, [u32, u32, u32, u64, P(P(Csr))]),
    "axspmv_fill_values_normal": (C.c_int, [P(Csr), f64, f64, u64]),
    "axspmv_build_k1": (C.c_int, [P(Csr), u32, C.c_int, P(P(K1)), P(FmtData)]),
    "axspmv_k1_free": (None, [P(K1)]),
    "axspmv_build_axc": (C.c_int, [P(Csr), P(f64), u32, u32, P(P(Axc)), P(FmtData)]),
    "axspmv_axc_free": (None, [P(Axc)]),
    "axspmv_build_axt": (C.c_int, [P(Csr), P(f64), u32, u32, u32, u32, u32, P(P(Axt)), P(FmtData)]),
    "axspmv_axt_free": (None, [P(Axt)]),
    "axspmv_dev_build_axt": (C.c_int, [P(Dcsr), vp, u32, u32, u32, u32, vp, P(P(Dfmt))]),
    "axspmv_dev_build_axc": (C.c_int, [P(Dcsr), vp, u32, vp, P(P(Dfmt))]),
    "axspmv_dfmt_free": (None, [P(Dfmt)]),
    "axspmv_dev_build_axt_noalloc": (C.c_int, [P(Dcsr), vp, u32, u32, u32, u32, vp, P(Dfmt)]),
    "axspmv_opts_default": (C.c_int, [u32, P(Opts)]),
    "axspmv_plan_create": (C.c_int, [P(Csr), P(Opts), vp, P(vp)]),
    "axspmv_plan_destroy": (C.c_int, [vp]),
    "axspmv_plan_get_info": (C.c_int, [vp, P(PlanInfo)]),
    "axspmv_spmv": (C.c_int, [vp, vp, vp, vp]),
    "axspmv_adjoint": (C.c_int, [vp, vp, vp, vp]),
    "axspmv_plan_vbar_positions": (C.c_int, [vp, P(u32), u64]),
    "axspmv_vbar": (C.c_int, [vp, vp, vp, vp, vp]),
    "axspmv_fwd_adj": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, C.c_int, vp]),
    "axspmv_spmm": (C.c_int, [vp, u32, vp, vp, vp]),
    "axspmv_time_spmv": (C.c_int, [vp, vp, vp, u32, u32, u32, u32, vp, P(Timing)]),
    "axspmv_time_op": (C.c_int, [vp, u32, P(Operands), u32, u32, u32, u32, vp, P(Timing)]),
    "axspmv_batch_create": (C.c_int, [P(P(Csr)), u32, P(Opts), C.c_int, vp, P(vp)]),
    "axspmv_batch_destroy": (C.c_int, [vp]),
    "axspmv_batch_get_info": (C.c_int, [vp, P(BatchInfo)]),
    "axspmv_batch_spmv": (C.c_int, [vp, vp, vp, vp]),
    "axspmv_batch_time": (C.c_int, [vp, vp, vp, u32, u32, u32, u32, vp, P(Timing)]),
    "axspmv_plan_graph": (C.c_int, [vp, u32, P(Operands), u32, P(vp)]),
    "axspmv_batch_graph": (C.c_int, [vp, vp, vp, u32, P(vp)]),
    "axspmv_graph_launch": (C.c_int, [vp, vp]),
    "axspmv_graph_destroy": (C.c_int, [vp]),
    "axspmv_autotune_defaults": (C.c_int, [u32, P(Opts), u32, P(u32)]),
    "axspmv_autotune": (C.c_int, [P(Csr), u32, u32, P(Opts), u32, u32, vp, P(u32), P(f64)]),
    "axspmv_autotune_batch": (C.c_int, [P(P(Csr)), u32, P(Opts), u32, u32, vp, P(Opts), P(u32)]),
    "axspmv_dev_alloc": (C.c_int, [C.c_size_t, P(vp)]),
    "axspmv_dev_free": (C.c_int, [vp]),
    "axspmv_memcpy_h2d": (C.c_int, [vp, vp, C.c_size_t, vp]),
    "axspmv_memcpy_d2h": (C.c_int, [vp, vp, C.c_size_t, vp]),
    "axspmv_stream_sync": (C.c_int, [vp]),
    "axspmv_device_sync": (C.c_int, []),
    "axspmv_device_count": (C.c_int, [P(C.c_int)]),
    "axspmv_copy_f64": (C.c_int, [vp, vp, C.c_size_t, vp]),
    "axspmv_x_refresh": (C.c_int, [vp, u64, u32, u32, vp, vp, vp]),
    "axspmv_wave_selftest": (C.c_int, [u32, vp, vp, vp, vp, u64, vp]),
}

_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load libaxspmv.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built: run `make -C gpu_spmv_amd/csrc` "
                          "(or __graft_entry__.build()); there is no fallback path")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int, where: str) -> None:
    if status != OK:
        msg = load().axspmv_last_error().decode(errors="replace")
        raise AxspmvError(status, where, msg)


def call(name: str, *args):
    """Call an int-returning entry point and raise on a negative status."""
    st = getattr(load(), name)(*args)
    check(st, name)
    return st
