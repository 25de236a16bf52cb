"""GPU parity: every kernel's y against the oracle's ncsr on the same inputs.

Bar (FP64): signed norm-wise relative error max|y - y_ref| / max|y_ref| <=
1e-12 (summation order and FMA contraction differ from the left-to-right
CPU sum).  FP32 value storage is compared against ncsr on the same
float-rounded values at the same 1e-12.
"""
import os

import numpy as np
import pytest

import gpu_spmv_amd as ax
from conftest import FIXTURES, FixtureMatrix, rel_err
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
TOL = 1e-12

KERNELS = [
    ("csr_naive", ax.FMT_CSR, dict(kernel=ax.KER_NAIVE)),
    ("csr_vector_auto", ax.FMT_CSR, dict(kernel=ax.KER_STREAM)),
    ("csr_vector_g1", ax.FMT_CSR, dict(kernel=ax.KER_STREAM, hbs=1)),
    ("csr_vector_g8", ax.FMT_CSR, dict(kernel=ax.KER_STREAM, hbs=8)),
    ("csr_vector_g64", ax.FMT_CSR, dict(kernel=ax.KER_STREAM, hbs=64)),
    ("unc_h1_thw4", ax.FMT_AXT, dict(thw=4, th=1)),
    ("unc_h1_thw8", ax.FMT_AXT, dict(thw=8, th=1)),
    ("unc_h1_thw16", ax.FMT_AXT, dict(thw=16, th=1)),
    ("unc_h1_thw32", ax.FMT_AXT, dict(thw=32, th=1)),
    ("unc_h1_thw64", ax.FMT_AXT, dict(thw=64, th=1)),
    ("unc_h1_thw2", ax.FMT_AXT, dict(thw=2, th=1)),
    ("unc_h4_thw64", ax.FMT_AXT, dict(thw=64, th=4)),
    ("unc_h8_thw64", ax.FMT_AXT, dict(thw=64, th=8)),
    ("unc_h6_thw32", ax.FMT_AXT, dict(thw=32, th=6)),
    ("unc_h12_thw32", ax.FMT_AXT, dict(thw=32, th=12)),
    ("unc_h5_thw64", ax.FMT_AXT, dict(thw=64, th=5)),
    ("unc_h1_thw4_small_budget", ax.FMT_AXT, dict(thw=4, th=1, budget=64)),
    ("unc_h1_thw4_s4", ax.FMT_AXT, dict(thw=4, slices=4)),
    ("unc_h1_thw2_s2", ax.FMT_AXT, dict(thw=2, slices=2)),
    ("unc_h1_thw1_s8", ax.FMT_AXT, dict(thw=1, slices=8)),
    ("unc_h1_thw8_s4_b256", ax.FMT_AXT, dict(thw=8, slices=4, budget=256)),
    ("unc_h1_thw2_s4_det", ax.FMT_AXT, dict(thw=2, slices=4, flags=ax.PLAN_DETERMINISTIC)),
    ("unc_h1_thw4_s8_det", ax.FMT_AXT, dict(thw=4, slices=8, flags=ax.PLAN_DETERMINISTIC)),
    ("k1_c64", ax.FMT_K1, dict(chunk=64)),
    ("k1_c64_natural", ax.FMT_K1, dict(chunk=64, reorder=0)),
    ("k1_c32", ax.FMT_K1, dict(chunk=32)),
    ("axc_h64", ax.FMT_AXC, dict(hbs=64)),
    ("axc_h32", ax.FMT_AXC, dict(hbs=32)),
    ("axc_h8", ax.FMT_AXC, dict(hbs=8)),
    ("axc_h4", ax.FMT_AXC, dict(hbs=4)),
    ("com_h1_thw64_bs256", ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=1, bs=256)),
    ("com_h1_thw64_bs1024", ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=1, bs=1024)),
    ("com_h1_thw32_bs64", ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=32, th=1, bs=64)),
    ("com_h1_thw4_bs2", ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=4, th=1, bs=2)),
    ("com_h4_thw64", ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=4)),
    ("com_h8_thw64", ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=8)),
    ("com_h16_thw64", ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=16)),
    ("com_h4_thw32", ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=32, th=4)),
    ("com_h32_thw64", ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=32)),
    ("rct_auto", ax.FMT_RCT, dict()),
    ("rct_r1024_s4", ax.FMT_RCT, dict(tile_rows=1024, slices=4)),
    ("rct_r4096_s8", ax.FMT_RCT, dict(tile_rows=4096, slices=8)),
    ("rct_r1024_s4_atomic", ax.FMT_RCT, dict(tile_rows=1024, slices=4, flags=ax.PLAN_RCT_ATOMIC)),
    ("rct_r2048_s2_atomic_split", ax.FMT_RCT, dict(tile_rows=2048, slices=2, flags=ax.PLAN_RCT_ATOMIC | ax.PLAN_RCT_SPLIT)),
    ("rct_r8192_s2_b256", ax.FMT_RCT, dict(tile_rows=8192, slices=2, block=256)),
    ("rct_r512_s1", ax.FMT_RCT, dict(tile_rows=512, slices=1)),
    ("rct_r2048_s4_split", ax.FMT_RCT, dict(tile_rows=2048, slices=4, flags=ax.PLAN_RCT_SPLIT)),
    ("rct_r16384_s1_split_b512", ax.FMT_RCT, dict(tile_rows=16384, slices=1, block=512, flags=ax.PLAN_RCT_SPLIT)),
]


def long_rows():
    """Rows of 1..5000 entries (several passes of one wave) between short ones."""
    rng = np.random.default_rng(3)
    n = 3000
    rl = rng.integers(0, 6, n)
    rl[rng.choice(n, 8, replace=False)] = rng.integers(600, 5000, 8)
    row = np.concatenate([[0], np.cumsum(rl)]).astype(np.uint32)
    col = np.concatenate([np.sort(rng.choice(n * 2, k, replace=False)) for k in rl]).astype(np.uint32)
    return ax.CsrMatrix(row, col, rng.standard_normal(col.size), ncols=2 * n)


def matrices():
    for name in FIXTURES:
        yield name, FixtureMatrix(name).csr()
    yield "laplace256", ax.laplace2d(256)
    yield "powerlaw64k", ax.powerlaw(1 << 16, 2.0, 2048, 20.0, seed=1)
    yield "stencil20p", ax.stencil27(20, permute=True, seed=2)
    yield "banded_empty", ax.banded(4000, 2, 0.3, seed=6)
    yield "long_rows", long_rows()
    yield "random_odd", ax.random_sparse(3001, 0, 40, seed=9)


MATS = dict(matrices())


def test_native_library_loaded():
    lib = ax._lib.load()
    assert ax.device_count() >= 1
    maps = open("/proc/self/maps").read()
    assert os.path.abspath(ax._lib.LIB_PATH) in maps
    # the library's HIP runtime (ROCm's); torch, when another test imported it
    # for CPU tensors, maps its own copy, which never touches the device here
    hip_libs = {line.split()[-1] for line in maps.splitlines() if "libamdhip64" in line}
    assert any(h.startswith("/opt/rocm") for h in hip_libs), hip_libs
    assert lib.axspmv_abi_version() == 1


@pytest.mark.parametrize("mname", list(MATS))
@pytest.mark.parametrize("kname,fmt,kw", KERNELS, ids=[k[0] for k in KERNELS])
def test_spmv_f64(mname, kname, fmt, kw):
    m = MATS[mname]
    x = ax.uniform(11, m.ncols)
    ref = orc.ncsr(m, x, nthreads=4)
    plan = ax.Plan(m, fmt, **kw)
    y = ax.spmv_host(plan, x)
    assert rel_err(ref, y) <= TOL, (mname, kname, rel_err(ref, y))
    # y is fully overwritten: a second call into a dirty buffer gives the same
    dx = ax.DeviceBuffer.from_array(x)
    dy = ax.DeviceBuffer.from_array(np.full(max(m.nrows, 1), np.nan))
    plan.spmv(dx, dy)
    y2 = dy.download(np.float64, m.nrows)
    atomic = fmt == ax.FMT_RCT or (kw.get("slices", 0) > 1 and not kw.get("flags", 0) & ax.PLAN_DETERMINISTIC)
    if atomic:  # the order of additions varies between runs
        assert rel_err(ref, y2) <= TOL
    else:
        assert np.array_equal(y, y2)


F32_KERNELS = [k for k in KERNELS if k[0] in ("csr_naive", "csr_vector_auto", "unc_h1_thw4", "unc_h4_thw64", "k1_c64", "axc_h8",
                                                "unc_h1_thw4_s4",
                                                "com_h1_thw64_bs256", "com_h8_thw64", "rct_r1024_s4",
                                                "rct_r1024_s4_atomic", "rct_r2048_s4_split")]


@pytest.mark.parametrize("mname", ["rand200", "powerlaw64k", "long_rows"])
@pytest.mark.parametrize("kname,fmt,kw", F32_KERNELS, ids=[k[0] for k in F32_KERNELS])
def test_spmv_f32_values(mname, kname, fmt, kw):
    m = MATS[mname]
    x = ax.uniform(12, m.ncols)
    ref = orc.ncsr_f32val(m, x)
    y = ax.spmv_host(ax.Plan(m, fmt, vtype=ax.F32, **kw), x)
    assert rel_err(ref, y) <= TOL
    # and the stated accuracy against the FP64 product
    assert rel_err(orc.ncsr(m, x), y) < 1e-6


def test_powerlaw_config2_full_size():
    m = ax.powerlaw(1 << 20, 2.0, 2048, 20.0, seed=1)
    x = ax.uniform(1, m.nrows)
    ref = orc.ncsr(m, x, nthreads=int(os.environ.get("OMP_NUM_THREADS", "8")))
    for fmt, kw in ((ax.FMT_AXT, dict(thw=4, th=1)), (ax.FMT_AXT, dict(thw=64, th=4)),
                    (ax.FMT_AXT, dict(thw=2, slices=4)), (ax.FMT_AXT, dict(thw=4, slices=4)),
                    (ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=1, bs=256)),
                    (ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=4)), (ax.FMT_K1, dict(chunk=64)),
                    (ax.FMT_AXC, dict(hbs=4)), (ax.FMT_RCT, dict()),
                    (ax.FMT_AXT, dict(thw=2, slices=4, flags=ax.PLAN_DETERMINISTIC))):
        y = ax.spmv_host(ax.Plan(m, fmt, **kw), x)
        assert rel_err(ref, y) <= TOL, kw
    # FP32 values: K1's and COM's streams still pass 128 MB (non-temporal loads)
    ref32 = orc.ncsr_f32val(m, x, nthreads=int(os.environ.get("OMP_NUM_THREADS", "8")))
    for fmt, kw in ((ax.FMT_K1, dict(chunk=64)), (ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=1, bs=256)),
                    (ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=4))):
        y = ax.spmv_host(ax.Plan(m, fmt, vtype=ax.F32, **kw), x)
        assert rel_err(ref32, y) <= TOL, kw


@pytest.mark.parametrize("gen,slices", [
    (lambda: ax.laplace2d(1024), 1),                          # 2^20 columns, 5-point band
    (lambda: ax.banded(1 << 19, 64, 0.5, seed=3), 1),         # 2^19 columns, band of 129
    (lambda: ax.stencil27(96, permute=False), 1),             # natural order: bands of about 2 planes
    (lambda: ax.stencil27(96, permute=True, seed=2), 4),      # 884,736 random columns
    (lambda: ax.powerlaw(1 << 19, 2.0, 2048, 20.0, seed=1), 2),
    (lambda: ax.random_sparse(200000, 1, 30, seed=4), 1),     # x of 1.6 MB fits one slice
    # 128 row blocks of 16,384 rows: 2 slices give the 256 CUs a tile each
    # (x slices of 8 MiB; 2 MB slices would take 8)
    (lambda: ax.stencil27(128, permute=True, seed=2), 2),
], ids=["laplace1024", "banded", "stencil96", "stencil96p", "powerlaw512k", "random200k", "stencil128p"])
def test_rct_auto_slices(gen, slices):
    """Default slicing: one slice when the row blocks read narrow column
    bands or x fits 2 MB; else the fewest slices for which the full-height
    row blocks times the slices cover the CUs, but no more than needed for
    x slices of 2 MB and no fewer than for x slices of 8 MiB.  The product
    is the same either way."""
    m = gen()
    p = ax.Plan(m, ax.FMT_RCT)
    assert p.info()["name"].split("_")[2] == f"s{slices}", p.info()["name"]
    x = ax.uniform(15, m.ncols)
    assert rel_err(orc.ncsr(m, x), ax.spmv_host(p, x)) <= TOL


@pytest.mark.parametrize("slices,flags,atomic", [(4, 0, False), (2, 0, True), (4, ax.PLAN_RCT_ATOMIC, True)])
def test_rct_slice_write_form(slices, flags, atomic):
    """Slices write their row sums as partials merged in slice order when
    slices x nrows >= nnz / 6 (here about 20 nonzeros per row: 4 slices
    yes, 2 no), else as FP64 adds into a cleared y; AXSPMV_PLAN_RCT_ATOMIC
    always adds.  Both give the oracle's product."""
    m = ax.powerlaw(1 << 16, 2.0, 2048, 20.0, seed=1)
    assert 6 * 4 * m.nrows >= m.nnz > 6 * 2 * m.nrows
    p = ax.Plan(m, ax.FMT_RCT, tile_rows=2048, slices=slices, flags=flags)
    assert p.info()["name"].endswith("_atomic") == atomic, p.info()["name"]
    x = ax.uniform(19, m.ncols)
    assert rel_err(orc.ncsr(m, x), ax.spmv_host(p, x)) <= TOL


def test_rct_rejects_invalid_block():
    m = MATS["lap16"]
    for b in (32, 96, 2048):
        with pytest.raises(ax.AxspmvError, match="multiple of 64"):
            ax.Plan(m, ax.FMT_RCT, block=b)


@pytest.mark.parametrize("gen,lanes", [(lambda: ax.laplace2d(128), 4),
                                       (lambda: ax.powerlaw(1 << 14, 2.0, 1024, 20.0, seed=1), 16),
                                       (lambda: ax.banded(20000, 32, 0.9, seed=4), 64)],
                         ids=["laplace_mean5", "powerlaw_mean20", "banded_mean58"])
def test_default_lanes_per_row(gen, lanes):
    """The CSR vector kernel's lanes per row and AXC's brick size default to
    the power of two g with 1.5 g >= the mean row length, within [2, 64]."""
    m = gen()
    x = ax.uniform(18, m.ncols)
    for p, name in ((ax.Plan(m, ax.FMT_CSR, kernel=ax.KER_STREAM), f"csr_vector_g{lanes}"),
                    (ax.Plan(m, ax.FMT_AXC), f"axc_hbs{lanes}")):
        assert p.info()["name"] == name, (m.nnz / m.nrows, p.info()["name"])
        assert rel_err(orc.ncsr(m, x), ax.spmv_host(p, x)) <= TOL


def wide_csr(nrows, ncols, seed):
    """One entry per row at a random column of a wide matrix."""
    rng = np.random.default_rng(seed)
    return ax.CsrMatrix(np.arange(nrows + 1, dtype=np.uint32), rng.integers(0, ncols, nrows).astype(np.uint32),
                        rng.uniform(-1, 1, nrows), ncols)


@pytest.mark.parametrize("gen,slices,flags,key", [
    # a band of 16 each side: row - column spans 33 values (6 bits), a run of
    # 128 entries at most 36 columns where it steps back into the next tile
    # (6 bits): the 16-bit key
    (lambda: ax.banded(300000, 16, 0.5, seed=9), 1, 0, "k16"),
    (lambda: ax.banded(300000, 16, 0.5, seed=9), 4, 0, "k16"),
    # ... unless asked for the 32-bit key: runs of 1024 entries step back a few
    # columns where they pass into the next row block: based on their
    # smallest column, 14 + 6 bits
    (lambda: ax.banded(300000, 16, 0.5, seed=9), 1, ax.PLAN_RCT_WIDE_KEY, ""),
    # the widest band of the 256-matrix batch: 129 values of row - column (8
    # bits) and runs stepping back about 128 columns (8 bits)
    (lambda: ax.banded(200000, 64, 0.3, seed=12), 1, 0, "k16"),
    # a band of 200: 401 values (9 bits) and runs stepping back about 400
    # columns (9 bits) do not fit 16 bits
    (lambda: ax.banded(100000, 200, 0.3, seed=13), 1, 0, ""),
    # a run covers about 1024 x 300,000 / 106,000 columns: 14 + 12 bits
    (lambda: ax.random_sparse(300000, 1, 12, seed=10), 1, 0, ""),
    # runs passing from the last slice of a row block into the first slice
    # of the next count round from their first column
    (lambda: ax.random_sparse(300000, 1, 12, seed=10), 4, 0, ""),
    # 16,384 entries per tile over 2^23 columns: a run covers ~2^19 columns,
    # 14 + 19 bits do not fit one word
    (lambda: wide_csr(32768, 1 << 23, seed=11), 1, 0, "split"),
], ids=["banded", "banded_s4", "banded_wide_key", "band64", "band200", "random", "random_s4", "wide"])
def test_rct_key_packing(gen, slices, flags, key):
    """Column and tile row of an entry: 16 bits (the column's offset from the
    smallest column of its run of 128 entries, and the row minus the column)
    on narrow bands; else one 32-bit word (the column as an offset from a
    base column per run of 1024 entries, and the row); where the offsets
    need more bits than the row leaves, separate column and row arrays.  All
    give the product of the oracle, and SpMM too."""
    m = gen()
    p = ax.Plan(m, ax.FMT_RCT, tile_rows=16384, slices=slices, flags=flags)
    name = p.info()["name"]
    assert ("_k16" in name) == (key == "k16") and ("_split" in name) == (key == "split"), name
    x = ax.uniform(16, m.ncols)
    assert rel_err(orc.ncsr(m, x), ax.spmv_host(p, x)) <= TOL
    if m.ncols > 1 << 20:
        return  # X of 8 columns would take 512 MB
    X = ax.uniform(17, m.ncols * 8).reshape(m.ncols, 8)
    q = ax.Plan(m, ax.FMT_RCT, tile_rows=2048, slices=slices, flags=flags)
    assert ("_k16" in q.info()["name"]) == (key == "k16"), q.info()["name"]
    dX, dY = ax.DeviceBuffer.from_array(X), ax.DeviceBuffer(8 * m.nrows * 8)
    q.spmm(8, dX, dY)
    assert rel_err(orc.spmm(m, X), dY.download(np.float64, m.nrows * 8).reshape(m.nrows, 8)) <= TOL


def test_rct_banded_tile_rows():
    """Bands (x wider than 2 MB) with rows of uneven length (two bands of 2
    and 60 each side, one after the other) take 512-row tiles; a band of one
    width, with rows of about one length, the tallest tiles that still give
    every CU one (300,000 / 256 rows, rounded up to 64)."""
    import scipy.sparse as sp
    a, b = ax.banded(200000, 2, 0.3, seed=21), ax.banded(200000, 60, 0.3, seed=22)
    bd = sp.block_diag([sp.csr_matrix((v.val, v.col, v.row), shape=(v.nrows, v.ncols)) for v in (a, b)]).tocsr()
    bd.sort_indices()
    m = ax.CsrMatrix(bd.indptr.astype(np.uint32), bd.indices.astype(np.uint32), bd.data, ncols=bd.shape[1])
    p = ax.Plan(m, ax.FMT_RCT)
    assert p.info()["name"] == "rct_r512_s1_k16", p.info()["name"]
    x = ax.uniform(24, m.ncols)
    assert rel_err(orc.ncsr(m, x), ax.spmv_host(p, x)) <= TOL
    c = ax.banded(300000, 60, 0.3, seed=23)
    assert ax.Plan(c, ax.FMT_RCT).info()["name"] == "rct_r1216_s1_k16"


def small_banded(n, hbw, seed):
    """Banded n x n matrix with about half its band filled, small-integer
    values (exact sums)."""
    m = ax.banded(n, hbw, 0.5, seed=seed)
    rng = np.random.default_rng(seed)
    return ax.CsrMatrix(m.row, m.col, rng.integers(-9, 10, m.nnz).astype(np.float64), ncols=n)


@pytest.mark.parametrize("n,hbw", [(1, 0), (2, 1), (3, 1), (5, 2), (70, 1), (129, 3), (700, 9), (4000, 64)])
@pytest.mark.parametrize("slices,block,vtype", [(1, 64, ax.F64), (2, 128, ax.F64), (8, 64, ax.F64), (4, 192, ax.F32)])
def test_rct_diag_key_pairs(n, hbw, slices, block, vtype):
    """The 16-bit key on bands: tiles starting or ending at odd entries, runs
    of 128 entries crossing tiles and slices, the spare entry of an odd
    count; the product equals the oracle bit for bit (integer values and x),
    and the value gradient and adjoint of the transposed plan too."""
    m = small_banded(n, hbw, seed=n + hbw)
    x = np.random.default_rng(n).integers(-9, 10, m.ncols).astype(np.float64)
    p = ax.Plan(m, ax.FMT_RCT, tile_rows=64, slices=slices, block=block, vtype=vtype,
                flags=ax.PLAN_TRANSPOSE if vtype == ax.F64 else 0)
    assert "_k16" in p.info()["name"], p.info()["name"]
    np.testing.assert_array_equal(ax.spmv_host(p, x), orc.ncsr(m, x))
    if vtype == ax.F32:
        return
    w = np.random.default_rng(n + 1).integers(-9, 10, m.nrows).astype(np.float64)
    dw, dx = ax.DeviceBuffer.from_array(w), ax.DeviceBuffer.from_array(x)
    dxb, dvb = ax.DeviceBuffer(8 * m.ncols), ax.DeviceBuffer(8 * max(m.nnz, 1))
    p.adjoint(dw, dxb)
    p.vbar(dw, dx, dvb)
    np.testing.assert_array_equal(dxb.download(np.float64, m.ncols), orc.ncsr_t(m, w, m.ncols))
    np.testing.assert_array_equal(p.vbar_csr(dvb.download(np.float64, m.nnz)), orc.vbar(m, w, x))


def small_csr(nrows, ncols, nnz, seed):
    """nnz entries at random (row, column) positions, rows sorted, values
    small integers (exact sums)."""
    rng = np.random.default_rng(seed)
    rows = np.sort(rng.integers(0, nrows, nnz))
    row = np.zeros(nrows + 1, np.uint32)
    np.add.at(row, rows + 1, 1)
    row = np.cumsum(row).astype(np.uint32)
    col = rng.integers(0, ncols, nnz).astype(np.uint32)
    return ax.CsrMatrix(row, col, rng.integers(-9, 10, nnz).astype(np.float64), ncols=ncols)


@pytest.mark.parametrize("nnz", [1, 2, 3, 5, 64, 127, 129, 1023, 4097])
@pytest.mark.parametrize("slices,block,vtype", [(1, 64, ax.F64), (2, 128, ax.F64), (8, 64, ax.F64), (4, 192, ax.F32)])
def test_rct_entry_pairs(nnz, slices, block, vtype):
    """The tile kernel loads entries in aligned pairs: tiles that start or end
    at odd entries, one-entry tiles, an odd entry count (the plan's spare
    entry is read and ignored), lanes past the end.  Integer values and x
    make every sum exact, so y must equal the oracle bit for bit."""
    m = small_csr(300, 700, nnz, seed=nnz)
    x = np.random.default_rng(nnz + 1).integers(-9, 10, m.ncols).astype(np.float64)
    p = ax.Plan(m, ax.FMT_RCT, tile_rows=64, slices=slices, block=block, vtype=vtype)
    np.testing.assert_array_equal(ax.spmv_host(p, x), orc.ncsr(m, x))


@pytest.mark.parametrize("kw", [dict(slices=4, tile_rows=512), dict(slices=1, tile_rows=512)])
def test_rct_unaligned_y(kw):
    """y at an 8-byte offset (the clear of y falls back to 8-byte stores)."""
    m = MATS["random_odd"]
    x = ax.uniform(13, m.ncols)
    p = ax.Plan(m, ax.FMT_RCT, **kw)
    dx = ax.DeviceBuffer.from_array(x)
    dy = ax.DeviceBuffer.from_array(np.full(m.nrows + 2, np.nan))
    p.spmv(dx, dy.ptr + 8)
    ax._lib.call("axspmv_device_sync")
    y = dy.download(np.float64, m.nrows + 2)
    assert np.isnan(y[0]) and np.isnan(y[-1])
    assert rel_err(orc.ncsr(m, x), y[1:-1]) <= TOL


@pytest.mark.parametrize("mname", ["laplace256", "band300", "banded_empty", "lap16", "long_rows", "powerlaw64k"])
@pytest.mark.parametrize("kw", [dict(thw=4), dict(thw=2), dict(thw=64, th=4), dict(thw=32, th=6)])
def test_lds_staging_is_exact(mname, kw):
    """x from the per-wave LDS window gives the same y, bit for bit, as x from
    HBM (same summation order); banded matrices stage, random ones do not."""
    m = MATS[mname]
    x = ax.uniform(14, m.ncols)
    staged = ax.Plan(m, ax.FMT_AXT, flags=ax.PLAN_STAGE_X, **kw)
    plain = ax.Plan(m, ax.FMT_AXT, **kw)
    assert plain.info()["staged_blocks"] == 0
    if mname in ("laplace256", "band300", "banded_empty", "lap16"):
        assert staged.info()["staged_blocks"] > 0
    if mname == "powerlaw64k":
        assert staged.info()["staged_blocks"] == 0
    y = ax.spmv_host(staged, x)
    assert np.array_equal(y, ax.spmv_host(plain, x))
    assert rel_err(orc.ncsr(m, x), y) <= TOL


EDGE_PLANS = [(ax.FMT_AXT, dict(thw=4)), (ax.FMT_AXT, dict(thw=64, th=4)), (ax.FMT_CSR, dict(kernel=ax.KER_NAIVE)),
              (ax.FMT_CSR, dict(kernel=ax.KER_STREAM)), (ax.FMT_CSR, dict(kernel=ax.KER_STREAM, hbs=64)),
              (ax.FMT_K1, dict(chunk=64)), (ax.FMT_AXC, dict(hbs=4)), (ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=1)),
              (ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=4)), (ax.FMT_RCT, dict()),
              (ax.FMT_RCT, dict(slices=4, tile_rows=64)), (ax.FMT_AXT, dict(thw=4, flags=ax.PLAN_AX_LAYOUT)),
              (ax.FMT_AXT, dict(thw=4, vtype=ax.F32)), (ax.FMT_RCT, dict(vtype=ax.F32, slices=2)),
              (ax.FMT_AXT, dict(thw=2, slices=4)), (ax.FMT_AXT, dict(thw=4, slices=8, vtype=ax.F32))]


@pytest.mark.parametrize("fmt,kw", EDGE_PLANS, ids=[f"{f}-{k}" for f, k in EDGE_PLANS])
def test_edge_shapes(fmt, kw):
    """One nonzero; no nonzeros at all (y = 0 over a dirty buffer); more
    columns than rows."""
    one = ax.CsrMatrix(np.array([0, 1], np.uint32), np.array([0], np.uint32), np.array([2.5]))
    assert ax.spmv_host(ax.Plan(one, fmt, **kw), np.array([3.0]))[0] == 7.5
    empty = ax.CsrMatrix(np.zeros(7, np.uint32), np.zeros(0, np.uint32), np.zeros(0), ncols=6)
    p = ax.Plan(empty, fmt, **kw)
    dx = ax.DeviceBuffer.from_array(np.ones(6))
    dy = ax.DeviceBuffer.from_array(np.full(6, np.nan))
    p.spmv(dx, dy)
    ax._lib.call("axspmv_device_sync")
    assert np.array_equal(dy.download(np.float64, 6), np.zeros(6))
    wide = ax.CsrMatrix(np.array([0, 2, 2, 5], np.uint32), np.array([1, 900, 0, 7, 999], np.uint32),
                        np.array([1.0, -2.0, 0.5, 4.0, 3.0]), ncols=1000)
    x = ax.uniform(15, 1000)
    y = ax.spmv_host(ax.Plan(wide, fmt, **kw), x)
    assert rel_err(orc.ncsr(wide, x), y) <= (TOL if kw.get("vtype") != ax.F32 else 1e-6)


def test_timing_modes():
    m = ax.laplace2d(128)
    x = ax.uniform(0, m.ncols)
    p = ax.Plan(m, ax.FMT_AXT, thw=4)
    dx, dy = ax.DeviceBuffer.from_array(x), ax.DeviceBuffer(8 * m.nrows)
    for mode in (ax.core.TIME_EVENTS, ax.core.TIME_GRAPH, ax.core.TIME_REGION, ax.core.TIME_COLD):
        t = p.time(dx, dy, iters=20, warmup=2, mode=mode, batch=10)
        assert t.mean_ms > 0 and t.min_ms <= t.median_ms <= t.max_ms and t.mode == mode
    assert rel_err(orc.ncsr(m, x), dy.download(np.float64, m.nrows)) <= TOL
    # the main-kernel split: plans with slices mark their main kernel, before
    # the merge of slice partials or after the clear of y
    t = p.time(dx, dy, iters=20, warmup=2, mode=ax.core.TIME_KERNEL)
    assert t.mode == ax.core.TIME_KERNEL and t.main_ms == pytest.approx(t.mean_ms)
    for fmt, kw in ((ax.FMT_RCT, dict(tile_rows=1024, slices=4)),
                    (ax.FMT_RCT, dict(tile_rows=1024, slices=4, flags=ax.PLAN_RCT_ATOMIC)),
                    (ax.FMT_AXT, dict(thw=2, slices=4)),
                    (ax.FMT_AXT, dict(thw=2, slices=4, flags=ax.PLAN_DETERMINISTIC))):
        r = ax.Plan(m, fmt, **kw)
        t = r.time(dx, dy, iters=20, warmup=2, mode=ax.core.TIME_KERNEL)
        assert 0 < t.main_ms < t.mean_ms, kw
        assert rel_err(orc.ncsr(m, x), dy.download(np.float64, m.nrows)) <= TOL
    r = ax.Plan(m, ax.FMT_RCT, tile_rows=1024, slices=4)
    with pytest.raises(ax.AxspmvError):
        r.time(dx, dy, iters=5000, warmup=0, mode=ax.core.TIME_KERNEL)


NONFINITE_KERNELS = [k for k in KERNELS if not k[0].startswith("csr_vector_g")] + [
    ("unc_h1_thw4_ax", ax.FMT_AXT, dict(thw=4, flags=ax.PLAN_AX_LAYOUT)),
    ("unc_h1_thw4_f32", ax.FMT_AXT, dict(thw=4, vtype=ax.F32)),
    ("unc_h4_thw64_staged", ax.FMT_AXT, dict(thw=64, th=4, flags=ax.PLAN_STAGE_X))]


@pytest.mark.parametrize("mname", ["band300", "random_odd", "long_rows"])
@pytest.mark.parametrize("kname,fmt,kw", NONFINITE_KERNELS, ids=[k[0] for k in NONFINITE_KERNELS])
def test_nonfinite_x_stays_in_its_rows(mname, kname, fmt, kw):
    """x[0] = inf and x[1] = nan: only the rows holding column 0 or 1 become
    non-finite, as in the CPU product.  Padding slots (value 0) carry the
    column 0xffffffff in the plans and read 0.0, so 0 * inf never reaches a
    row through padding."""
    m = MATS[mname]
    x = ax.uniform(19, m.ncols)
    x[0], x[1] = np.inf, np.nan
    if kw.get("vtype") == ax.F32:
        ref = orc.ncsr_f32val(m, x)
    else:
        ref = orc.ncsr(m, x)
    y = ax.spmv_host(ax.Plan(m, fmt, **kw), x)
    fin = np.isfinite(ref)
    assert not fin.all() and fin.any()
    np.testing.assert_array_equal(np.isfinite(y), fin)
    np.testing.assert_array_equal(y[~fin], ref[~fin])  # the same inf / nan
    assert rel_err(ref[fin], y[fin]) <= TOL


@pytest.mark.parametrize("slices", [2, 4, 8])
def test_axt_slices_layout_and_dispatch(slices):
    """AXT in column slices: each slice is the builders' AXT layout of its
    columns (the plan's slot count is the sum of the slices' layouts), the
    clear + block kernel pair is one SpMV, and the adjoint through the
    sliced plan of A^T matches the oracle."""
    m = MATS["powerlaw64k"]
    p = ax.Plan(m, ax.FMT_AXT, thw=4, slices=slices, flags=ax.PLAN_TRANSPOSE)
    info = p.info()
    assert info["name"] == f"axt_unc_h1_thw4_s{slices}" and info["launches"] == 2
    C = -(-m.ncols // slices)
    slots = 0
    for k in range(slices):
        keep = (m.col // C) == k
        rows = np.repeat(np.arange(m.nrows), np.diff(m.row))[keep]
        row = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=m.nrows))]).astype(np.uint32)
        sub = ax.CsrMatrix(row, m.col[keep], m.val[keep], ncols=m.ncols)
        slots += ax.build_axt(sub, None, 256, 4, 1, ax.AXT_UNC, ax.BUILD_PLANES)["lenAX"] // 2
    assert info["slots"] == slots
    x = ax.uniform(20, m.ncols)
    assert rel_err(orc.ncsr(m, x), ax.spmv_host(p, x)) <= TOL
    w = ax.uniform(21, m.nrows)
    dw, dxb = ax.DeviceBuffer.from_array(w), ax.DeviceBuffer(8 * m.ncols)
    p.adjoint(dw, dxb)
    assert rel_err(orc.ncsr_t(m, w, m.ncols), dxb.download(np.float64, m.ncols)) <= TOL
    with pytest.raises(ax.AxspmvError):
        p.vbar(dw, ax.DeviceBuffer.from_array(x), ax.DeviceBuffer(8 * m.nnz))
    with pytest.raises(ax.AxspmvError, match="column slices"):
        ax.Plan(m, ax.FMT_AXT, thw=64, th=4, slices=slices)
