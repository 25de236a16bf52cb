"""CPU: the format arrays are indexed with 32-bit offsets, as in the
reference (UIN lengths, cudaSpmv.cu:1158-1387, where a longer layout wraps
silently).  The builders and generators compute every length in 64 bits
first and return AXSPMV_E_OVERFLOW before allocating the arrays, so each
case below costs only its input CSR."""
import numpy as np
import pytest

import gpu_spmv_amd as ax
from gpu_spmv_amd import _lib as L


def one_per_row(n):
    """n rows of one entry each (column 0)."""
    return ax.CsrMatrix(np.arange(n + 1, dtype=np.uint32), np.zeros(n, np.uint32), np.ones(n), ncols=1)


def empty_rows(n):
    return ax.CsrMatrix(np.zeros(n + 1, np.uint32), np.zeros(0, np.uint32), np.zeros(0), ncols=1)


def overflow(fn, *args, match):
    with pytest.raises(L.AxspmvError, match=match) as e:
        fn(*args)
    assert e.value.status == L.E_OVERFLOW


def test_axt_unc_h1_length():
    # one tile per row: lenAX = 2 * 1024 * 2^21 = 2^32 slots
    overflow(ax.build_axt, one_per_row(1 << 21), None, 0, 1024, 1, L.AXT_UNC, L.BUILD_PLANES,
             match="lenAX")


def test_axt_unc_h_length():
    # one tile column per row: 2052 tiles of 1024 columns, each 2 * 1024 * th
    # slots -- 8.4M at th = 2, 4.2994e9 >= 2^32 at th = 1023
    n = (1 << 21) + (1 << 12)
    m = one_per_row(n)
    tiles = -(-n // 1024)
    r = ax.build_axt(m, None, 0, 1024, 2, L.AXT_UNC, L.BUILD_PLANES)
    assert r["tileN"] == tiles and r["lenAX"] == tiles * 2 * 1024 * 2
    assert tiles * 2 * 1024 * 1023 >= 1 << 32
    overflow(ax.build_axt, m, None, 0, 1024, 1023, L.AXT_UNC, L.BUILD_PLANES, match="lenAX")


def test_axt_com_header():
    # COM th == 1 packs row << log2(bs) into a 32-bit header: rows - 1 < 2^(32 - 10) at bs = 1024
    overflow(ax.build_axt, empty_rows((1 << 22) + 1), None, 1024, 64, 1, L.AXT_COM, L.BUILD_PLANES,
             match="header")
    r = ax.build_axt(empty_rows(1 << 22), None, 1024, 64, 1, L.AXT_COM, L.BUILD_PLANES)
    assert r["log"] == 10 and r["lenAX"] == 0
    # COM th > 1 packs row << log2(th)
    overflow(ax.build_axt, empty_rows((1 << 23) + 1), None, 0, 64, 512, L.AXT_COM, L.BUILD_PLANES,
             match="header")


def test_axc_length():
    overflow(ax.build_axc, one_per_row(1 << 21), None, 1024, L.BUILD_PLANES, match="lenAX")


def test_k1_length_and_reorder():
    # 4096 chunks of 1024 rows, each chunk's first row 1024 long: in row order
    # every chunk pads to 1024 x 1024 slots (2^32 in all); sorted by length the
    # long rows share 4 chunks and the layout holds 2^22 slots.
    chunk, nchunk = 1024, 4096
    n = chunk * nchunk
    rl = np.zeros(n, np.uint32)
    rl[::chunk] = chunk
    row = np.zeros(n + 1, np.uint32)
    np.cumsum(rl, out=row[1:])
    nnz = int(row[-1])
    col = np.tile(np.arange(chunk, dtype=np.uint32), nchunk)
    m = ax.CsrMatrix(row, col, np.ones(nnz), ncols=n)
    overflow(ax.build_k1, m, chunk, False, match="lenVC")
    r = ax.build_k1(m, chunk, True)
    assert r["lenVC"] == nnz == 1 << 22
    assert r["nmc"][:4].tolist() == [chunk] * 4 and not r["nmc"][4:].any()


def test_generators():
    overflow(ax.laplace2d, 30000, match="grid too large")  # 5 * 30000^2 > 2^32
    overflow(ax.stencil27, 600, match="grid too large")  # 27 * 600^3 > 2^32


def test_lane_group_grids():
    # CSR vector and AXC launch nrows x lanes work-items, and the dispatch
    # grid is 32 bits: 2^26 rows at 64 lanes per row is rejected before any
    # device work (this container has no GPU)
    m = empty_rows(1 << 26)
    with pytest.raises(L.AxspmvError, match="32-bit grid") as e:
        ax.Plan(m, ax.FMT_CSR, kernel=ax.KER_STREAM, hbs=64)
    assert e.value.status == L.E_OVERFLOW
    with pytest.raises(L.AxspmvError, match="32-bit grid") as e:
        ax.Plan(m, ax.FMT_AXC, hbs=64)
    assert e.value.status == L.E_OVERFLOW


def test_csr_matrix_shape_checks():
    with pytest.raises(ValueError, match="offsets"):
        ax.CsrMatrix(np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
    with pytest.raises(ValueError, match="values"):
        ax.CsrMatrix(np.array([0, 1], np.uint32), np.zeros(1, np.uint32), np.zeros(2))
    # inconsistent offsets are the library's to reject (before any kernel)
    bad = ax.CsrMatrix(np.array([0, 2], np.uint32), np.zeros(1, np.uint32), np.zeros(1))
    with pytest.raises(L.AxspmvError, match="row"):
        bad.check()
