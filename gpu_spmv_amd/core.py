"""Python objects over the C ABI: host CSR matrices, generators, format
builders, device buffers and plans.  Used by the tests and the benchmark;
the computation happens in libaxspmv.so."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L

# x, b, X vectors use this RNG stream of the configuration's seed (the
# generators use stream 0 and streams 1..n for rows).
VEC_STREAM = 1 << 40


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class CsrMatrix:
    """Host CSR matrix backed by numpy arrays (row/col uint32, val float64)."""

    def __init__(self, row, col, val, ncols: int | None = None):
        self.row = np.ascontiguousarray(row, dtype=np.uint32)
        self.col = np.ascontiguousarray(col, dtype=np.uint32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        if self.row.ndim != 1 or self.row.size == 0:
            raise ValueError("CsrMatrix: row must be a 1-D array of nrows + 1 offsets (at least one)")
        if self.col.size != self.val.size:  # nnz is taken from col; the library reads that many values
            raise ValueError(f"CsrMatrix: {self.col.size} columns but {self.val.size} values")
        nrows = self.row.size - 1
        if ncols is None:
            ncols = max(nrows, int(self.col.max()) + 1 if self.col.size else 0)
        self.rl = np.zeros(max(nrows, 1), dtype=np.uint32)
        self._s = L.Csr(nrows=nrows, ncols=ncols, nnz=self.col.size, owner=0,
                        row=_ptr(self.row, L.u32), col=_ptr(self.col, L.u32),
                        val=_ptr(self.val, L.f64), rl=_ptr(self.rl, L.u32))

    @classmethod
    def _adopt(cls, p) -> "CsrMatrix":
        """Copy a library-owned CSR into numpy arrays and free it."""
        s = p.contents
        n, nnz = s.nrows, s.nnz
        row = np.ctypeslib.as_array(s.row, shape=(n + 1,)).copy()
        col = np.ctypeslib.as_array(s.col, shape=(nnz,)).copy() if nnz else np.zeros(0, np.uint32)
        val = np.ctypeslib.as_array(s.val, shape=(nnz,)).copy() if nnz else np.zeros(0)
        m = cls(row, col, val, ncols=s.ncols)
        L.load().axspmv_csr_free(p)
        m.stats()
        return m

    @property
    def nrows(self) -> int:
        return self._s.nrows

    @property
    def ncols(self) -> int:
        return self._s.ncols

    @property
    def nnz(self) -> int:
        return self._s.nnz

    @property
    def ref(self):
        return C.byref(self._s)

    def check(self) -> None:
        L.call("axspmv_csr_check", self.ref)

    def stats(self) -> dict:
        L.call("axspmv_csr_stats", self.ref)
        s = self._s
        return dict(nrows=s.nrows, nnz=s.nnz, rmin=s.rmin, rave=s.rave, rmax=s.rmax, rsd=s.rsd, bw=s.bw)

    def transpose(self) -> "CsrMatrix":
        out = C.POINTER(L.Csr)()
        L.call("axspmv_csr_transpose", self.ref, C.byref(out))
        return CsrMatrix._adopt(out)

    def rcm(self) -> np.ndarray:
        """Reverse Cuthill-McKee order: perm[k] = row/column numbered k."""
        perm = np.zeros(self.nrows, dtype=np.uint32)
        L.call("axspmv_csr_rcm", self.ref, _ptr(perm, C.c_uint32))
        return perm

    def permute(self, perm: np.ndarray) -> "CsrMatrix":
        """P A P^T: row k is row perm[k], column perm[k] becomes k."""
        perm = np.ascontiguousarray(perm, dtype=np.uint32)
        out = C.POINTER(L.Csr)()
        L.call("axspmv_csr_permute", self.ref, _ptr(perm, C.c_uint32), C.byref(out))
        return CsrMatrix._adopt(out)

    def write(self, path: str) -> None:
        L.call("axspmv_csr_write", self.ref, path.encode())

    @staticmethod
    def read(path: str) -> "CsrMatrix":
        out = C.POINTER(L.Csr)()
        L.call("axspmv_csr_read", path.encode(), C.byref(out))
        return CsrMatrix._adopt(out)

    def row_ids(self) -> np.ndarray:
        return np.repeat(np.arange(self.nrows, dtype=np.uint32), np.diff(self.row))

    def to_dense(self) -> np.ndarray:
        d = np.zeros((self.nrows, self.ncols))
        d[self.row_ids(), self.col] = self.val
        return d


# ------------------------------------------------------------------ data
def _gen(name: str, *args) -> CsrMatrix:
    out = C.POINTER(L.Csr)()
    L.call(name, *args, C.byref(out))
    return CsrMatrix._adopt(out)


def laplace2d(grid: int, diag: float = 4.0, off: float = -1.0) -> CsrMatrix:
    return _gen("axspmv_gen_laplace2d", grid, diag, off)


def powerlaw(n: int, alpha: float = 2.0, rlmax: int = 2048, mean: float = 20.0, seed: int = 1) -> CsrMatrix:
    return _gen("axspmv_gen_powerlaw", n, alpha, rlmax, mean, seed)


def stencil27(grid: int, diag: float = 26.0, off: float = -1.0, permute: bool = True, seed: int = 2) -> CsrMatrix:
    return _gen("axspmv_gen_stencil27", grid, diag, off, int(permute), seed)


def banded(n: int, hbw: int, fill: float = 0.3, seed: int = 5) -> CsrMatrix:
    return _gen("axspmv_gen_banded", n, hbw, fill, seed)


def random_sparse(n: int, rlmin: int, rlmax: int, seed: int) -> CsrMatrix:
    return _gen("axspmv_gen_random", n, rlmin, rlmax, seed)


def fill_values_normal(m: CsrMatrix, mu: float, sigma: float, seed: int) -> None:
    L.call("axspmv_fill_values_normal", m.ref, mu, sigma, seed)


def uniform(seed: int, n: int, lo: float = -1.0, hi: float = 1.0, stream: int = VEC_STREAM) -> np.ndarray:
    out = np.empty(n)
    L.call("axspmv_rng_uniform", seed, stream, lo, hi, _ptr(out, L.f64), n)
    return out


def normal(seed: int, n: int, mu: float = 0.0, sigma: float = 1.0, stream: int = VEC_STREAM) -> np.ndarray:
    out = np.empty(n)
    L.call("axspmv_rng_normal", seed, stream, mu, sigma, _ptr(out, L.f64), n)
    return out


def rng_u64(seed: int, stream: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint64)
    L.call("axspmv_rng_u64", seed, stream, _ptr(out, L.u64), n)
    return out


# -------------------------------------------------------------- builders
def _arr(p, n, dtype):
    if not p or n == 0:
        return None if not p else np.zeros(0, dtype)
    return np.ctypeslib.as_array(p, shape=(n,)).astype(dtype, copy=True)


def _fd(fd: L.FmtData) -> dict:
    return dict(name=fd.name.decode(), mfp=fd.mfp, beta=fd.beta, ct=fd.ct)


def build_k1(m: CsrMatrix, chunk: int = 64, reorder: bool = True) -> dict:
    out, fd = C.POINTER(L.K1)(), L.FmtData()
    L.call("axspmv_build_k1", m.ref, chunk, int(reorder), C.byref(out), C.byref(fd))
    s = out.contents
    r = dict(chunk=s.chunk, chunkNum=s.chunkNum, lenVC=s.lenVC, fd=_fd(fd),
             permi=_arr(s.permi, (s.chunkNum + 1) * s.chunk, np.uint32),
             nmc=_arr(s.nmc, s.chunkNum, np.uint32), chp=_arr(s.chp, s.chunkNum, np.uint32),
             val=_arr(s.val, s.lenVC, np.float64), col=_arr(s.col, s.lenVC, np.uint32))
    L.load().axspmv_k1_free(out)
    return r


def build_axc(m: CsrMatrix, x: np.ndarray | None, hbs: int, flags: int = L.BUILD_AX | L.BUILD_PLANES) -> dict:
    out, fd = C.POINTER(L.Axc)(), L.FmtData()
    xp = _ptr(np.ascontiguousarray(x, np.float64), L.f64) if x is not None else None
    L.call("axspmv_build_axc", m.ref, xp, hbs, flags, C.byref(out), C.byref(fd))
    s = out.contents
    r = dict(hbs=s.hbs, lenAX=s.lenAX, lenBRP=s.lenBRP, lenMAPX=s.lenMAPX, fd=_fd(fd),
             ax=_arr(s.ax, s.lenAX, np.float64), brp=_arr(s.brp, s.lenBRP, np.uint32),
             mapx=_arr(s.mapx, s.lenMAPX, np.uint32), vplane=_arr(s.vplane, s.lenAX // 2, np.float64))
    L.load().axspmv_axc_free(out)
    return r


def build_axt(m: CsrMatrix, x: np.ndarray | None, bs: int, thw: int, th: int, mode: int,
              flags: int = L.BUILD_AX | L.BUILD_PLANES) -> dict:
    out, fd = C.POINTER(L.Axt)(), L.FmtData()
    xp = _ptr(np.ascontiguousarray(x, np.float64), L.f64) if x is not None else None
    L.call("axspmv_build_axt", m.ref, xp, bs, thw, th, mode, flags, C.byref(out), C.byref(fd))
    s = out.contents
    r = dict(name=s.name.decode(), tileN=s.tileN, lenAX=s.lenAX, lenSEC=s.lenSEC, lenCON=s.lenCON,
             log=s.log, logTH=s.logTH, fd=_fd(fd),
             ax=_arr(s.ax, s.lenAX, np.float64), sec=_arr(s.sec, s.lenSEC, np.uint32),
             con=_arr(s.con, s.lenCON + 1, np.uint32), vplane=_arr(s.vplane, s.lenAX // 2, np.float64),
             cplane=_arr(s.cplane, s.lenAX // 2, np.uint32))
    L.load().axspmv_axt_free(out)
    return r


# ---------------------------------------------------------------- device
def device_count() -> int:
    n = C.c_int(0)
    L.call("axspmv_device_count", C.byref(n))
    return n.value


class DeviceBuffer:
    """A hipMalloc'd buffer owned by Python."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        L.call("axspmv_dev_alloc", self.nbytes, C.byref(p))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        L.call("axspmv_memcpy_h2d", self.ptr, a.ctypes.data, a.nbytes, None)

    def download(self, dtype=np.float64, count: int | None = None) -> np.ndarray:
        dtype = np.dtype(dtype)
        count = self.nbytes // dtype.itemsize if count is None else count
        out = np.empty(count, dtype)
        L.call("axspmv_memcpy_d2h", out.ctypes.data, self.ptr, out.nbytes, None)
        return out

    def free(self) -> None:
        if self.ptr:
            L.load().axspmv_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def opts(fmt: int, **kw) -> L.Opts:
    o = L.Opts()
    L.call("axspmv_opts_default", fmt, C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise TypeError(f"unknown option {k}")
        setattr(o, k, int(v))
    return o


TIME_EVENTS, TIME_GRAPH, TIME_REGION, TIME_KERNEL, TIME_COLD = 0, 1, 2, 3, 4


@dataclass
class TimingResult:
    mean_ms: float
    median_ms: float
    min_ms: float
    max_ms: float
    iters: int
    mode: int
    main_ms: float = 0.0


class Plan:
    """One matrix uploaded in one format (axspmv_plan)."""

    def __init__(self, m: CsrMatrix, fmt: int = L.FMT_AXT, **kw):
        self.opts = opts(fmt, **kw)
        p = C.c_void_p()
        L.call("axspmv_plan_create", m.ref, C.byref(self.opts), None, C.byref(p))
        self.ptr = p.value
        self.nrows, self.ncols, self.nnz = m.nrows, m.ncols, m.nnz

    def info(self) -> dict:
        i = L.PlanInfo()
        L.call("axspmv_plan_get_info", self.ptr, C.byref(i))
        return dict(name=i.name.decode(), nrows=i.nrows, ncols=i.ncols, nnz=i.nnz, format=i.format,
                    slots=i.slots, device_bytes=i.device_bytes, stream_bytes=i.stream_bytes,
                    alg_bytes=i.alg_bytes, beta=i.beta, launches=i.launches, nblocks=i.nblocks,
                    staged_blocks=i.staged_blocks)

    def spmv(self, dx, dy, stream=None) -> None:
        L.call("axspmv_spmv", self.ptr, _dptr(dx), _dptr(dy), stream)

    def time(self, dx, dy, iters: int = 100, warmup: int = 10, mode: int = TIME_EVENTS,
             batch: int = 1, stream=None) -> TimingResult:
        """Kernel time per SpMV (TIME_EVENTS: per-launch events as the
        reference; TIME_GRAPH: graph replays of `batch` SpMVs; TIME_REGION:
        one event pair around `iters` back-to-back launches)."""
        t = L.Timing()
        L.call("axspmv_time_spmv", self.ptr, _dptr(dx), _dptr(dy), iters, warmup, mode, batch,
               stream, C.byref(t))
        return TimingResult(t.mean_ms, t.median_ms, t.min_ms, t.max_ms, t.iters, t.mode, t.main_ms)

    def adjoint(self, dybar, dxbar, stream=None) -> None:
        """x_bar = A^T y_bar (plan built with flags=PLAN_TRANSPOSE)."""
        L.call("axspmv_adjoint", self.ptr, _dptr(dybar), _dptr(dxbar), stream)

    def vbar(self, dybar, dx, dvbar, stream=None) -> None:
        """v_bar[k] = y_bar[row(k)] * x[col(k)] in CSR order."""
        L.call("axspmv_vbar", self.ptr, _dptr(dybar), _dptr(dx), _dptr(dvbar), stream)

    def fwd_adj(self, dx, db, dy, dr, dxbar, dvbar, fused: bool = True, stream=None) -> None:
        """y = A x, r = y - b, x_bar = A^T r, v_bar = r[row] * x[col]."""
        L.call("axspmv_fwd_adj", self.ptr, _dptr(dx), _dptr(db), _dptr(dy), _dptr(dr), _dptr(dxbar),
               _dptr(dvbar), int(fused), stream)

    def spmm(self, k: int, dX, dY, stream=None) -> None:
        """Y = A X, X (ncols x k) and Y (nrows x k) row-major."""
        L.call("axspmv_spmm", self.ptr, k, _dptr(dX), _dptr(dY), stream)

    def vbar_positions(self) -> np.ndarray:
        """CSR position of every value-gradient entry (identity except RCT
        and reordered plans)."""
        pos = np.empty(max(self.nnz, 1), np.uint32)
        L.call("axspmv_plan_vbar_positions", self.ptr, _ptr(pos, L.u32), self.nnz)
        return pos[:self.nnz]

    def vbar_csr(self, v: np.ndarray) -> np.ndarray:
        """A value gradient in CSR order (undoes the entry order of RCT and
        reordered plans)."""
        out = np.empty_like(v)
        out[self.vbar_positions()] = v
        return out

    def time_op(self, op: int, iters: int = 100, warmup: int = 10, mode: int = TIME_REGION, batch: int = 1,
                stream=None, **operands) -> TimingResult:
        """Time one operation (L.OP_*); operands by name (x, b, ybar, X, y, r, xbar, vbar, Y, k)."""
        o = L.Operands()
        for name, v in operands.items():
            setattr(o, name, v if name == "k" else _dptr(v))
        t = L.Timing()
        L.call("axspmv_time_op", self.ptr, op, C.byref(o), iters, warmup, mode, batch, stream, C.byref(t))
        return TimingResult(t.mean_ms, t.median_ms, t.min_ms, t.max_ms, t.iters, t.mode, t.main_ms)

    def close(self) -> None:
        if self.ptr:
            L.load().axspmv_plan_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _dptr(d):
    if d is None:
        return None
    if isinstance(d, DeviceBuffer):
        return d.ptr
    if hasattr(d, "data_ptr"):
        return d.data_ptr()
    return int(d)


def _dev_build(m: CsrMatrix, x, fn: str, *args) -> dict:
    """Upload m (and x), run a device builder, download its arrays."""
    drow = DeviceBuffer.from_array(m.row)
    dcol = DeviceBuffer.from_array(m.col if m.nnz else np.zeros(1, np.uint32))
    dval = DeviceBuffer.from_array(m.val if m.nnz else np.zeros(1))
    dx = DeviceBuffer.from_array(np.ascontiguousarray(x, np.float64)) if x is not None else None
    a = L.Dcsr(m.nrows, m.ncols, m.nnz, 0, drow.ptr, dcol.ptr, dval.ptr)
    out = C.POINTER(L.Dfmt)()
    L.call(fn, C.byref(a), dx.ptr if dx is not None else None, *args, None, C.byref(out))
    f = out.contents

    def down(ptr, n, dt):
        h = np.empty(n, dt)
        if n:
            L.call("axspmv_memcpy_d2h", h.ctypes.data, ptr, h.nbytes, None)
        return h
    try:
        r = dict(name=f.name.decode(), tileN=f.tileN, lenAX=f.lenAX, lenSEC=f.lenSEC, lenCON=f.lenCON,
                 lenMAPX=f.lenMAPX, log=f.log, ms=f.ms, ax=down(f.ax, f.lenAX, np.float64),
                 con=down(f.con, f.nrows + 1, np.uint32), plane=down(f.plane, f.lenMAPX, np.uint32),
                 sec=down(f.sec, f.lenSEC, np.uint32) if f.sec else None)
        L.call("axspmv_stream_sync", None)
    finally:
        L.load().axspmv_dfmt_free(out)
    return r


def dev_build_axt(m: CsrMatrix, x: np.ndarray | None, bs: int, thw: int, th: int, mode: int) -> dict:
    """axspmv_dev_build_axt: the AXT arrays built on the device (downloaded)."""
    return _dev_build(m, x, "axspmv_dev_build_axt", bs, thw, th, mode)


def dev_build_axc(m: CsrMatrix, x: np.ndarray | None, hbs: int) -> dict:
    """axspmv_dev_build_axc: the AXC arrays built on the device (downloaded)."""
    return _dev_build(m, x, "axspmv_dev_build_axc", hbs)


def x_refresh(dmap, nslots: int, hw: int, nx: int, dx, dax, stream=None) -> None:
    """Fill the x-halves of an interleaved ax array on the device:
    ax[(q // hw) * 2hw + hw + q % hw] = x[map[q]] (0.0 where map[q] >= nx)."""
    L.call("axspmv_x_refresh", _dptr(dmap), nslots, hw, nx, _dptr(dx), _dptr(dax), stream)


def spmv_host(plan: Plan, x: np.ndarray) -> np.ndarray:
    """Upload x, run one SpMV, download y (test convenience)."""
    dx = DeviceBuffer.from_array(np.ascontiguousarray(x, np.float64))
    dy = DeviceBuffer(8 * max(plan.nrows, 1))
    plan.spmv(dx, dy)
    L.call("axspmv_device_sync")
    return dy.download(np.float64, plan.nrows)


# ----------------------------------------------------------------- batches
class Batch:
    """Many matrices, SpMV in one launch (axspmv_batch): FMT_AXT, one
    persistent launch over every matrix's AXT UNC layout (per-matrix options
    allowed); FMT_RCT, the RCT plan of the block-diagonal matrix.  x / y are
    the concatenations of the matrices' vectors in batch order."""

    def __init__(self, mats, opts_list=None, fmt: int = L.FMT_AXT, **kw):
        self.mats = list(mats)
        n = len(self.mats)
        arr = (C.POINTER(L.Csr) * n)(*[C.pointer(m._s) for m in self.mats])
        if opts_list is None:
            o = (L.Opts * 1)(opts(fmt, **kw))
            per = 0
        else:
            o = (L.Opts * n)(*opts_list)
            per = 1
        p = C.c_void_p()
        L.call("axspmv_batch_create", arr, n, o, per, None, C.byref(p))
        self.ptr = p.value
        self.xoff = np.concatenate([[0], np.cumsum([m.ncols for m in self.mats])]).astype(np.int64)
        self.yoff = np.concatenate([[0], np.cumsum([m.nrows for m in self.mats])]).astype(np.int64)

    def info(self) -> dict:
        i = L.BatchInfo()
        L.call("axspmv_batch_get_info", self.ptr, C.byref(i))
        return {f: getattr(i, f) for f, _ in L.BatchInfo._fields_ if f != "pad_"}

    def spmv(self, dx, dy, stream=None) -> None:
        L.call("axspmv_batch_spmv", self.ptr, _dptr(dx), _dptr(dy), stream)

    def time(self, dx, dy, iters=50, warmup=5, mode=TIME_REGION, batch=1, stream=None) -> TimingResult:
        t = L.Timing()
        L.call("axspmv_batch_time", self.ptr, _dptr(dx), _dptr(dy), iters, warmup, mode, batch, stream, C.byref(t))
        return TimingResult(t.mean_ms, t.median_ms, t.min_ms, t.max_ms, t.iters, t.mode, t.main_ms)

    def close(self) -> None:
        if self.ptr:
            L.load().axspmv_batch_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune_defaults(op: int = L.OP_SPMV) -> list:
    """The library's built-in candidate option sets for `op`."""
    cnt = C.c_uint32()
    L.call("axspmv_autotune_defaults", op, None, 0, C.byref(cnt))
    out = (L.Opts * cnt.value)()
    L.call("axspmv_autotune_defaults", op, out, cnt.value, C.byref(cnt))
    return list(out)


def autotune(m: CsrMatrix, op: int = L.OP_SPMV, k: int = 0, candidates=None, iters: int = 20):
    """Index of the fastest candidate (list of Opts; None = the built-in set,
    see autotune_defaults) and each candidate's time in ms (-1 = rejected)."""
    if candidates is None:
        candidates = autotune_defaults(op)
    n = len(candidates)
    cand = (L.Opts * n)(*candidates)
    times = np.full(n, -1.0)
    best = C.c_uint32()
    L.call("axspmv_autotune", m.ref, op, k, cand, n, iters, None, C.byref(best), _ptr(times, L.f64))
    return best.value, times


def autotune_batch(mats, candidates=None, iters: int = 20):
    """Per-matrix options for a Batch; returns (list of Opts, number of groups)."""
    mats = list(mats)
    n = len(mats)
    arr = (C.POINTER(L.Csr) * n)(*[C.pointer(m._s) for m in mats])
    nc = len(candidates) if candidates is not None else 0
    cand = (L.Opts * nc)(*candidates) if candidates is not None else None
    out = (L.Opts * n)()
    ng = C.c_uint32()
    L.call("axspmv_autotune_batch", arr, n, cand, nc, iters, None, out, C.byref(ng))
    return list(out), ng.value


class Graph:
    """A captured hipGraph of repeated launches (axspmv_graph)."""

    def __init__(self, ptr):
        self.ptr = ptr

    @classmethod
    def of_plan(cls, plan: Plan, op: int, repeats: int = 1, **operands) -> "Graph":
        o = L.Operands()
        for name, v in operands.items():
            setattr(o, name, v if name == "k" else _dptr(v))
        p = C.c_void_p()
        L.call("axspmv_plan_graph", plan.ptr, op, C.byref(o), repeats, C.byref(p))
        return cls(p.value)

    @classmethod
    def of_batch(cls, batch: Batch, dx, dy, repeats: int = 1) -> "Graph":
        p = C.c_void_p()
        L.call("axspmv_batch_graph", batch.ptr, _dptr(dx), _dptr(dy), repeats, C.byref(p))
        return cls(p.value)

    def launch(self, stream=None) -> None:
        L.call("axspmv_graph_launch", self.ptr, stream)

    def close(self) -> None:
        if self.ptr:
            L.load().axspmv_graph_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
