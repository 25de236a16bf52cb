"""GPU: the benchmark configurations at their full sizes (the power-law
matrix of 2^20 rows is covered in test_gpu_spmv.py).

* 27-point stencil on a 128^3 grid, symmetric random permutation (seed 2):
  n = 2,097,152, nnz = 55,742,968 (the nonzero count is a closed form of the
  grid, checked first); every format against the oracle's ncsr, 1e-12.
* The same with AXSPMV_PLAN_REORDER (RCT, AXT UNC th=1, K1): x and y in the
  matrix's own numbering, the format's kernels on the renumbered matrix.
* Forward + adjoint on the same structure with values N(1, 0.1) (seed 4),
  b ~ U[-1, 1]: y, r = y - b, x_bar = A^T r against the oracle (1e-12 /
  1e-11), v_bar = r[row] * x[col] bit for bit against the oracle's product of
  the r the kernel wrote.
* Config 4A: k = 8, 32 and 64 right-hand sides on the power-law matrix, FP64
  and FP32 values.
* Config 4B: the 256 banded matrices (n ~ U{1000..200000}, 457,872,800
  nonzeros) in one launch, as the AXT persistent batch (uniform tiles and the
  autotuner's per-group tiles) and as the RCT block-diagonal plan; every
  member against the oracle at 1e-12.
"""
import os

import numpy as np
import pytest

import gpu_spmv_amd as ax
from gpu_spmv_amd import workloads as W
from conftest import rel_err
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
TOL = 1e-12
NT = int(os.environ.get("OMP_NUM_THREADS", "8"))


@pytest.fixture(scope="module")
def stencil():
    m = ax.stencil27(128, permute=True, seed=2)
    # interior 26 neighbours, fewer on faces: sum over the grid of the
    # product of the per-axis counts (1 + 2 interior, 2 at the ends)
    per_axis = 3 * 128 - 2
    assert (m.nrows, m.nnz) == (128 ** 3, per_axis ** 3) == (2097152, 55742968)
    x = ax.uniform(1, m.ncols)
    return m, x, orc.ncsr(m, x, nthreads=NT)


FORMATS = [
    ("csr_naive", ax.FMT_CSR, dict(kernel=ax.KER_NAIVE)),
    ("csr_vector", ax.FMT_CSR, dict(kernel=ax.KER_STREAM)),
    ("axt_unc_h1_thw4", ax.FMT_AXT, dict(thw=4, th=1)),
    ("axt_unc_h4_thw64", ax.FMT_AXT, dict(thw=64, th=4)),
    ("axt_com_h1_thw64_bs256", ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=1, bs=256)),
    ("axt_com_h8_thw64", ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=8)),
    ("k1_c64", ax.FMT_K1, dict(chunk=64)),
    ("axc_hbs4", ax.FMT_AXC, dict(hbs=4)),
    ("rct_auto", ax.FMT_RCT, dict()),
    ("rct_r8192_s1_split", ax.FMT_RCT, dict(tile_rows=8192, slices=1, flags=ax.PLAN_RCT_SPLIT)),
    # renumbered by reverse Cuthill-McKee inside the plan
    ("rct_auto_rcm", ax.FMT_RCT, dict(flags=ax.PLAN_REORDER)),
    ("axt_unc_h1_thw4_rcm", ax.FMT_AXT, dict(thw=4, flags=ax.PLAN_REORDER)),
    ("k1_c64_rcm", ax.FMT_K1, dict(chunk=64, flags=ax.PLAN_REORDER)),
]


@pytest.mark.timeout(300)
@pytest.mark.parametrize("name,fmt,kw", FORMATS, ids=[f[0] for f in FORMATS])
def test_stencil_config3_full_size(stencil, name, fmt, kw):
    m, x, ref = stencil
    p = ax.Plan(m, fmt, **kw)
    y = ax.spmv_host(p, x)
    assert rel_err(ref, y) <= TOL, p.info()["name"]


@pytest.mark.timeout(300)
def test_long_rows_full_size(stencil):
    """The permuted stencil (a stream of more than 512 MB gathering wide: the
    CSR vector and AXC plans read it with non-temporal loads) with 8 rows of
    5000 entries appended: their one-wave-per-row kernels take the same
    loads."""
    s, _, _ = stencil
    rng = np.random.default_rng(31)
    extra = [np.sort(rng.choice(s.ncols, 5000, replace=False)).astype(np.uint32) for _ in range(8)]
    col = np.concatenate([s.col] + extra)
    val = np.concatenate([s.val, rng.standard_normal(8 * 5000)])
    row = np.concatenate([s.row, s.row[-1] + 5000 * np.arange(1, 9, dtype=np.uint64)]).astype(np.uint32)
    m = ax.CsrMatrix(row, col, val, ncols=s.ncols)
    x = ax.uniform(1, m.ncols)
    ref = orc.ncsr(m, x, nthreads=NT)
    for fmt, kw in ((ax.FMT_CSR, dict(kernel=ax.KER_STREAM)), (ax.FMT_AXC, dict(hbs=4))):
        y = ax.spmv_host(ax.Plan(m, fmt, **kw), x)
        assert rel_err(ref, y) <= TOL, kw
        assert rel_err(ref[-8:], y[-8:]) <= TOL, kw


@pytest.mark.timeout(300)
@pytest.mark.parametrize("fmt,kw", [(ax.FMT_AXT, dict(thw=4)), (ax.FMT_RCT, dict()),
                                    (ax.FMT_AXT, dict(thw=4, flags=ax.PLAN_REORDER)),
                                    (ax.FMT_RCT, dict(flags=ax.PLAN_REORDER))],
                         ids=["axt", "rct", "axt_rcm", "rct_rcm"])
def test_fwd_adj_config5_full_size(stencil, fmt, kw):
    s, x, _ = stencil
    m = ax.CsrMatrix(s.row, s.col, s.val.copy(), ncols=s.ncols)
    ax.fill_values_normal(m, 1.0, 0.1, 4)
    b = ax.uniform(4, m.nrows, stream=ax.VEC_STREAM + 1)
    n, nc, nnz = m.nrows, m.ncols, m.nnz
    kw = dict(kw)
    flags = kw.pop("flags", 0)
    p = ax.Plan(m, fmt, flags=ax.PLAN_TRANSPOSE | flags, **kw)
    dx, db = ax.DeviceBuffer.from_array(x), ax.DeviceBuffer.from_array(b)
    outs = [ax.DeviceBuffer(8 * k) for k in (n, n, nc, nnz)]
    y_ref = orc.ncsr(m, x, nthreads=NT)
    xbar_ref = orc.ncsr_t(m, y_ref - b, nc)
    # the split form too where its kernels differ (the AXT value-gradient
    # kernel streams this large, wide matrix with non-temporal loads)
    for fused in ((True, False) if fmt == ax.FMT_AXT and not flags else (True,)):
        p.fwd_adj(dx, db, *outs, fused=fused)
        y, r, xbar, vbar = (o.download(np.float64, k) for o, k in zip(outs, (n, n, nc, nnz)))
        assert rel_err(y_ref, y) <= TOL
        assert rel_err(y_ref - b, r) <= TOL
        assert rel_err(xbar_ref, xbar) <= 1e-11
        assert np.array_equal(p.vbar_csr(vbar), orc.vbar(m, r, x))


@pytest.fixture(scope="module")
def powerlaw():
    return ax.powerlaw(1 << 20, 2.0, 2048, 20.0, seed=1)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("k", [8, 32, 64])
@pytest.mark.parametrize("vt", [ax.F64, ax.F32], ids=["f64", "f32val"])
@pytest.mark.parametrize("fmt,kw", [(ax.FMT_AXT, dict(thw=4)), (ax.FMT_RCT, dict(tile_rows=1024)),
                                    (ax.FMT_RCT, dict(tile_rows=256, slices=1))],
                         ids=["axt", "rct_r1024", "rct_r256_s1"])
def test_spmm_config4a_full_size(powerlaw, k, vt, fmt, kw):
    """Config 4A: k right-hand sides (x ~ U[-1, 1], seed 3) on the power-law
    matrix; FP32 values against the oracle on the float-rounded values."""
    m = powerlaw
    X = ax.uniform(3, m.ncols * k).reshape(m.ncols, k)
    mv = m if vt == ax.F64 else ax.CsrMatrix(m.row, m.col, m.val.astype(np.float32).astype(np.float64),
                                              ncols=m.ncols)
    ref = orc.spmm(mv, X, nthreads=NT)
    p = ax.Plan(m, fmt, vtype=vt, **kw)
    dX, dY = ax.DeviceBuffer.from_array(X), ax.DeviceBuffer(8 * m.nrows * k)
    p.spmm(k, dX, dY)
    assert rel_err(ref, dY.download(np.float64, m.nrows * k).reshape(m.nrows, k)) <= TOL


@pytest.fixture(scope="module")
def batch256():
    mats = W.config4b()
    assert sum(m.nnz for m in mats) == 457872800
    return mats


def check_members(b, mats, x):
    dx = ax.DeviceBuffer.from_array(x)
    ny = int(b.yoff[-1])
    dy = ax.DeviceBuffer.from_array(np.full(ny, np.nan))
    b.spmv(dx, dy)
    y = dy.download(np.float64, ny)
    worst = 0.0
    for i, m in enumerate(mats):
        e = rel_err(orc.ncsr(m, x[b.xoff[i]:b.xoff[i + 1]], nthreads=NT), y[b.yoff[i]:b.yoff[i + 1]])
        assert e <= TOL, (i, m.nrows, e)
        worst = max(worst, e)
    return worst


@pytest.mark.timeout(600)
@pytest.mark.parametrize("how", ["axt_thw4", "axt_autotuned", "rct"])
def test_batch_config4b_full_size(batch256, how):
    mats = batch256
    if how == "axt_thw4":
        b = ax.Batch(mats, thw=4)
    elif how == "axt_autotuned":
        tuned, _ = ax.autotune_batch(mats, iters=3)
        b = ax.Batch(mats, opts_list=tuned)
    else:
        b = ax.Batch(mats, fmt=ax.FMT_RCT)
    info = b.info()
    assert info["count"] == 256 and info["nnz"] == 457872800
    x = ax.uniform(6, int(b.xoff[-1]))
    check_members(b, mats, x)
    b.close()
