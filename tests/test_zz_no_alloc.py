"""GPU: axspmv_dev_build_axt_noalloc (the device AXT builder with its
scratch in a static device array and the caller's output arrays) against
axspmv_dev_build_axt, bit for bit, on two calls in a row -- the second on a
smaller matrix, so the shared scratch holds the first call's contents -- and
in every layout family; lengths that are not the layout's, and matrices whose
scratch exceeds the static array, are errors."""
import ctypes as C

import numpy as np
import pytest

import gpu_spmv_amd as ax
from gpu_spmv_amd import _lib as L

pytestmark = pytest.mark.gpu
SCRATCH = 48 << 20  # AXSPMV_DEV_BUILD_NOALLOC_SCRATCH


class DevCsr:
    def __init__(self, m):
        self.bufs = [ax.DeviceBuffer.from_array(a if a.size else np.zeros(1, a.dtype)) for a in (m.row, m.col, m.val)]
        self.s = L.Dcsr(m.nrows, m.ncols, m.nnz, 0, *(b.ptr for b in self.bufs))


def noalloc(m, x, bs, thw, th, mode, lengths=None):
    """The no-allocation build into fresh caller arrays of the allocating
    build's lengths (or `lengths`); returns the downloaded arrays."""
    ref = ax.dev_build_axt(m, x, bs, thw, th, mode)
    lens = lengths or dict(lenAX=ref["lenAX"], lenSEC=ref["lenSEC"], lenCON=ref["lenCON"], lenMAPX=ref["lenMAPX"])
    bufs = dict(ax=ax.DeviceBuffer(8 * max(lens["lenAX"], 1)), sec=ax.DeviceBuffer(4 * max(lens["lenSEC"], 1)),
                con=ax.DeviceBuffer(4 * (lens["lenCON"] + 1)), plane=ax.DeviceBuffer(4 * max(lens["lenMAPX"], 1)))
    # stale contents: the build must clear what it relies on being zero
    for b in bufs.values():
        b.upload(np.full(b.nbytes, 0xAB, np.uint8))
    f = L.Dfmt()
    for k, v in lens.items():
        setattr(f, k, v)
    for k, b in bufs.items():
        setattr(f, k, b.ptr)
    a = DevCsr(m)
    dx = ax.DeviceBuffer.from_array(x) if x is not None else None
    L.call("axspmv_dev_build_axt_noalloc", C.byref(a.s), dx.ptr if dx else None, bs, thw, th, mode, None, C.byref(f))
    got = dict(name=f.name.decode(), tileN=f.tileN, lenAX=f.lenAX, lenSEC=f.lenSEC, lenCON=f.lenCON,
               lenMAPX=f.lenMAPX, log=f.log, ax=bufs["ax"].download(np.float64, f.lenAX),
               con=bufs["con"].download(np.uint32, f.nrows + 1), plane=bufs["plane"].download(np.uint32, f.lenMAPX),
               sec=bufs["sec"].download(np.uint32, f.lenSEC))
    return ref, got


def same(ref, got):
    for k in ("name", "tileN", "lenAX", "lenSEC", "lenCON", "lenMAPX", "log"):
        assert ref[k] == got[k], k
    for k in ("ax", "con", "plane", "sec"):
        assert ref[k].tobytes() == got[k].tobytes(), k


LAYOUTS = [(256, 4, 1, ax.AXT_UNC), (256, 64, 4, ax.AXT_UNC), (256, 64, 1, ax.AXT_COM), (256, 64, 8, ax.AXT_COM),
           (64, 32, 1, ax.AXT_COM), (256, 2, 1, ax.AXT_UNC)]


@pytest.mark.parametrize("bs,thw,th,mode", LAYOUTS)
def test_noalloc_matches_allocating_build(bs, thw, th, mode):
    big = ax.powerlaw(1 << 17, 2.0, 2048, 20.0, seed=1)
    small = ax.random_sparse(3001, 0, 40, seed=9)
    # first call on the larger matrix, second on the smaller one (the
    # scratch still holds the first call's counts and scan)
    for m, xs in ((big, 21), (small, 22)):
        x = ax.uniform(xs, m.ncols)
        ref, got = noalloc(m, x, bs, thw, th, mode)
        same(ref, got)
    # no x: the x halves stay 0, as in the allocating build
    ref, got = noalloc(small, None, bs, thw, th, mode)
    same(ref, got)


def test_noalloc_empty_and_tiny():
    for m in (ax.CsrMatrix(np.zeros(5, np.uint32), np.zeros(0, np.uint32), np.zeros(0), ncols=3),
              ax.CsrMatrix(np.array([0, 1], np.uint32), np.array([0], np.uint32), np.array([2.5]))):
        for bs, thw, th, mode in LAYOUTS[:3]:
            ref, got = noalloc(m, np.ones(m.ncols), bs, thw, th, mode)
            same(ref, got)


def test_noalloc_rejects_wrong_lengths():
    m = ax.random_sparse(3001, 0, 40, seed=9)
    ref = ax.dev_build_axt(m, None, 256, 4, 1, ax.AXT_UNC)
    lens = dict(lenAX=ref["lenAX"] + 8, lenSEC=ref["lenSEC"] + 1, lenCON=ref["lenCON"], lenMAPX=ref["lenMAPX"] + 4)
    with pytest.raises(L.AxspmvError, match="the layout needs") as e:
        noalloc(m, None, 256, 4, 1, ax.AXT_UNC, lengths=lens)
    assert e.value.status == L.E_INVALID
    # the caller's descriptor is left as it was
    f = L.Dfmt()
    for k, v in lens.items():
        setattr(f, k, v)
    a = DevCsr(m)
    bufs = [ax.DeviceBuffer(8 * (lens["lenAX"] + 1)) for _ in range(4)]
    f.ax, f.sec, f.con, f.plane = (b.ptr for b in bufs)
    before = bytes(f)
    with pytest.raises(L.AxspmvError):
        L.call("axspmv_dev_build_axt_noalloc", C.byref(a.s), None, 256, 4, 1, ax.AXT_UNC, None, C.byref(f))
    assert bytes(f) == before


def test_noalloc_rejects_scratch_beyond_the_static_array():
    # a UNC build keeps 2 x 8 bytes per row in the scratch: 4M rows need 64 MB
    n = 4 << 20
    m = ax.CsrMatrix(np.zeros(n + 1, np.uint32), np.zeros(0, np.uint32), np.zeros(0), ncols=1)
    f = L.Dfmt()
    a = DevCsr(m)
    with pytest.raises(L.AxspmvError, match="static workspace") as e:
        L.call("axspmv_dev_build_axt_noalloc", C.byref(a.s), None, 256, 4, 1, ax.AXT_UNC, None, C.byref(f))
    assert e.value.status == L.E_UNSUPPORTED
    # COM layouts need only the flag: the same rows build
    ref, got = noalloc(m, None, 256, 64, 1, ax.AXT_COM)
    same(ref, got)
