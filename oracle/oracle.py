"""ORACLE -- test infrastructure, not part of the library.

ctypes wrapper of liboracle.so (oracle.c), the C restatement of the
reference's CPU path.  Only tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg use it, as the checker.  Matrices are passed as any object
with numpy `row`, `col`, `val` arrays (uint32, uint32, float64).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

u32, f64 = C.c_uint32, C.c_double
P = C.POINTER
_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        lib = C.CDLL(LIB_PATH)
        sig = {
            "orc_ncsr": [C.c_int, u32, P(u32), P(u32), P(f64), P(f64), P(f64)],
            "orc_ncsr_f32val": [C.c_int, u32, P(u32), P(u32), P(f64), P(f64), P(f64)],
            "orc_spmm": [C.c_int, u32, u32, P(u32), P(u32), P(f64), P(f64), P(f64)],
            "orc_ncsr_t": [u32, u32, P(u32), P(u32), P(f64), P(f64), P(f64)],
            "orc_vbar": [u32, P(u32), P(u32), P(f64), P(f64), P(f64)],
            "orc_get_errors": [u32, P(f64), P(f64), C.c_int, P(f64), P(f64), P(u32)],
            "orc_row_lengths": [u32, P(u32), P(u32)],
            "orc_k1_len": [u32, P(u32), u32, C.c_int, P(u32), P(u32), P(u32)],
            "orc_k1_fill": [u32, P(u32), P(u32), P(f64), u32, P(u32), P(u32), P(f64), P(u32)],
            "orc_axc_len": [u32, P(u32), u32, P(u32)],
            "orc_axc_fill": [u32, P(u32), P(u32), P(f64), P(f64), u32, P(u32), P(f64), P(u32)],
            "orc_axt_len": [u32, P(u32), u32, u32, u32, P(u32), P(u32)],
            "orc_axt_fill": [u32, P(u32), P(u32), P(f64), P(f64), u32, u32, u32, u32, P(u32), P(f64), P(u32)],
        }
        for name, args in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = u32 if name in ("orc_k1_len", "orc_axc_len", "orc_axt_fill") else None
        _lib = lib
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(P(ct))


def _csr(m):
    row = np.ascontiguousarray(m.row, np.uint32)
    col = np.ascontiguousarray(m.col, np.uint32)
    val = np.ascontiguousarray(m.val, np.float64)
    return row.size - 1, row, col, val


def ncsr(m, x, nthreads: int = 1) -> np.ndarray:
    """cudaSpmv.cu:451-466"""
    n, row, col, val = _csr(m)
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty(n)
    load().orc_ncsr(nthreads, n, _p(row, u32), _p(col, u32), _p(val, f64), _p(x, f64), _p(y, f64))
    return y


def ncsr_f32val(m, x, nthreads: int = 1) -> np.ndarray:
    n, row, col, val = _csr(m)
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty(n)
    load().orc_ncsr_f32val(nthreads, n, _p(row, u32), _p(col, u32), _p(val, f64), _p(x, f64), _p(y, f64))
    return y


def spmm(m, X: np.ndarray, nthreads: int = 1) -> np.ndarray:
    """X is (ncols, k) row-major; returns Y (nrows, k)."""
    n, row, col, val = _csr(m)
    X = np.ascontiguousarray(X, np.float64)
    k = X.shape[1]
    Y = np.empty((n, k))
    load().orc_spmm(nthreads, n, k, _p(row, u32), _p(col, u32), _p(val, f64), _p(X, f64), _p(Y, f64))
    return Y


def ncsr_t(m, ybar: np.ndarray, ncols: int) -> np.ndarray:
    n, row, col, val = _csr(m)
    ybar = np.ascontiguousarray(ybar, np.float64)
    xbar = np.empty(ncols)
    load().orc_ncsr_t(n, ncols, _p(row, u32), _p(col, u32), _p(val, f64), _p(ybar, f64), _p(xbar, f64))
    return xbar


def vbar(m, ybar: np.ndarray, x: np.ndarray) -> np.ndarray:
    n, row, col, _ = _csr(m)
    ybar = np.ascontiguousarray(ybar, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty(col.size)
    load().orc_vbar(n, _p(row, u32), _p(col, u32), _p(ybar, f64), _p(x, f64), _p(out, f64))
    return out


def get_errors(ref, res, sign_blind: bool = False):
    """cudaSpmv.cu:474-495 -> (aErr, rErr, pos)."""
    ref = np.ascontiguousarray(ref, np.float64)
    res = np.ascontiguousarray(res, np.float64)
    a, r, p = f64(), f64(), u32()
    load().orc_get_errors(ref.size, _p(ref, f64), _p(res, f64), int(sign_blind), C.byref(a), C.byref(r), C.byref(p))
    return a.value, r.value, p.value


def row_lengths(m) -> np.ndarray:
    n, row, _, _ = _csr(m)
    rl = np.empty(n, np.uint32)
    load().orc_row_lengths(n, _p(row, u32), _p(rl, u32))
    return rl


def k1(m, chunk: int, reorder: bool = True) -> dict:
    """cudaSpmv.cu:716-846"""
    n, row, col, val = _csr(m)
    cn = (n + chunk - 1) // chunk
    permi = np.zeros((cn + 1) * chunk, np.uint32)
    nmc = np.zeros(max(cn, 1), np.uint32)
    chp = np.zeros(max(cn, 1), np.uint32)
    lib = load()
    lenvc = lib.orc_k1_len(n, _p(row, u32), chunk, int(reorder), _p(permi, u32), _p(nmc, u32), _p(chp, u32))
    kval = np.zeros(max(lenvc, 1))
    kcol = np.zeros(max(lenvc, 1), np.uint32)
    lib.orc_k1_fill(n, _p(row, u32), _p(col, u32), _p(val, f64), chunk, _p(permi, u32), _p(chp, u32),
                    _p(kval, f64), _p(kcol, u32))
    return dict(chunkNum=cn, lenVC=lenvc, permi=permi, nmc=nmc[:cn], chp=chp[:cn], val=kval[:lenvc], col=kcol[:lenvc])


def axc(m, x, hbs: int) -> dict:
    """cudaSpmv.cu:937-1000"""
    n, row, col, val = _csr(m)
    x = np.ascontiguousarray(x, np.float64)
    brp = np.zeros(n + 1, np.uint32)
    lib = load()
    lenax = lib.orc_axc_len(n, _p(row, u32), hbs, _p(brp, u32))
    ax = np.zeros(max(lenax, 1))
    mapx = np.zeros(lenax // 2 + 8, np.uint32)
    lib.orc_axc_fill(n, _p(row, u32), _p(col, u32), _p(val, f64), _p(x, f64), hbs, _p(brp, u32), _p(ax, f64),
                     _p(mapx, u32))
    return dict(lenAX=lenax, brp=brp, ax=ax[:lenax], mapx=mapx)


def axt(m, x, bs: int, thw: int, th: int, mode: int) -> dict:
    """cudaSpmv.cu:1158-1387 (mode 0 = UNC, 1 = COM)"""
    n, row, col, val = _csr(m)
    x = np.ascontiguousarray(x, np.float64)
    con = np.zeros(n + 1, np.uint32)
    o3 = np.zeros(3, np.uint32)
    lib = load()
    lib.orc_axt_len(n, _p(row, u32), thw, th, mode, _p(con, u32), _p(o3, u32))
    tilen, lenax, lensec = (int(v) for v in o3)
    ax = np.zeros(max(lenax, 1))
    sec = np.zeros(max(lensec, 1), np.uint32)
    log = lib.orc_axt_fill(n, _p(row, u32), _p(col, u32), _p(val, f64), _p(x, f64), bs, thw, th, mode,
                           _p(con, u32), _p(ax, f64), _p(sec, u32))
    return dict(tileN=tilen, lenAX=lenax, lenSEC=lensec, log=log, con=con, ax=ax[:lenax], sec=sec[:lensec])
