"""GPU: plans with AXSPMV_PLAN_REORDER (the plan of B = Q A Q^T, Q the reverse
Cuthill-McKee order, run in A's numbering) against the oracle on A.

Bar: y, x_bar, r against ncsr / ncsr_t on A at 1e-12 (FP64; B's rows sum in
a different order); v_bar mapped to A's CSR positions is r[row] * x[col]
bit for bit.  Matrices: permuted stencils, a shuffled Laplacian and a
shuffled banded matrix with empty rows (where the new order applies), and a
power-law matrix, where it does not (the plan is A's own).
"""
import numpy as np
import pytest

import gpu_spmv_amd as ax
from conftest import rel_err
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
TOL = 1e-12

KERNELS = [
    ("csr_naive", ax.FMT_CSR, dict(kernel=ax.KER_NAIVE)),
    ("csr_vector", ax.FMT_CSR, dict(kernel=ax.KER_STREAM)),
    ("unc_h1_thw4", ax.FMT_AXT, dict(thw=4, th=1)),
    ("unc_h1_thw2_s4", ax.FMT_AXT, dict(thw=2, slices=4)),
    ("unc_h4_thw64", ax.FMT_AXT, dict(thw=64, th=4)),
    ("unc_h1_thw4_ax", ax.FMT_AXT, dict(thw=4, flags=ax.PLAN_AX_LAYOUT)),
    ("com_h1_thw64_bs256", ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=1, bs=256)),
    ("com_h8_thw64", ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=8)),
    ("k1_c64", ax.FMT_K1, dict(chunk=64)),
    ("axc_h8", ax.FMT_AXC, dict(hbs=8)),
    ("rct_auto", ax.FMT_RCT, dict()),
    ("rct_r1024_s4", ax.FMT_RCT, dict(tile_rows=1024, slices=4)),
    ("rct_r512_s2_atomic", ax.FMT_RCT, dict(tile_rows=512, slices=2, flags=ax.PLAN_RCT_ATOMIC)),
]


def shuffled(m, seed):
    return m.permute(np.random.default_rng(seed).permutation(m.nrows).astype(np.uint32))


def mats():
    # 9261 rows: two chunks of the 8192-position permutation passes, the second short
    yield "stencil21p", ax.stencil27(21, permute=True, seed=2), True
    yield "stencil9p_odd", ax.stencil27(9, permute=True, seed=3), True
    yield "laplace64_shuffled", shuffled(ax.laplace2d(64), 5), True
    yield "banded_empty_shuffled", shuffled(ax.banded(5000, 3, 0.3, seed=6), 7), True
    yield "powerlaw16k", ax.powerlaw(1 << 14, 2.0, 1024, 20.0, seed=1), False
    yield "stencil2p_tiny", ax.stencil27(2, permute=True, seed=1), False  # 8 rows, all adjacent: nothing to gain
    yield "one_row", ax.CsrMatrix(np.array([0, 1], np.uint32), np.array([0], np.uint32), np.array([3.0])), True


MATS = {name: (m, applies) for name, m, applies in mats()}


@pytest.mark.parametrize("mname", list(MATS))
@pytest.mark.parametrize("kname,fmt,kw", KERNELS, ids=[k[0] for k in KERNELS])
def test_reordered_spmv(mname, kname, fmt, kw):
    m, applies = MATS[mname]
    kw = dict(kw)
    p = ax.Plan(m, fmt, flags=ax.PLAN_REORDER | kw.pop("flags", 0), **kw)
    name = p.info()["name"]
    assert name.endswith("_rcm") == applies, name
    x = ax.uniform(11, m.ncols)
    y_ref = orc.ncsr(m, x)
    for _ in range(2):  # the plan's vectors are reused from call to call
        y = ax.spmv_host(p, x)
        assert rel_err(y_ref, y) <= TOL, (mname, kname)
    print(f"{mname} {name}: rel err {rel_err(y_ref, y):.2e}", flush=True)


GRAD_KERNELS = [
    ("csr_naive", ax.FMT_CSR, dict(kernel=ax.KER_NAIVE)),
    ("unc_h1_thw4", ax.FMT_AXT, dict(thw=4, th=1)),
    ("rct_auto", ax.FMT_RCT, dict()),
    ("rct_r512_s4_det", ax.FMT_RCT, dict(tile_rows=512, slices=4, flags=ax.PLAN_DETERMINISTIC)),
]


@pytest.mark.parametrize("mname", ["stencil21p", "banded_empty_shuffled", "powerlaw16k"])
@pytest.mark.parametrize("kname,fmt,kw", GRAD_KERNELS, ids=[k[0] for k in GRAD_KERNELS])
def test_reordered_adjoint_vbar_fwd_adj(mname, kname, fmt, kw):
    m, _ = MATS[mname]
    n, nc = m.nrows, m.ncols
    kw = dict(kw)
    p = ax.Plan(m, fmt, flags=ax.PLAN_REORDER | ax.PLAN_TRANSPOSE | kw.pop("flags", 0), **kw)
    x = ax.uniform(21, nc)
    b = ax.uniform(22, n, stream=ax.VEC_STREAM + 1)
    w = ax.uniform(23, n)
    dx, db, dw = (ax.DeviceBuffer.from_array(v) for v in (x, b, w))
    dxb, dvb = ax.DeviceBuffer(8 * nc), ax.DeviceBuffer(8 * m.nnz)
    p.adjoint(dw, dxb)
    assert rel_err(orc.ncsr_t(m, w, nc), dxb.download(np.float64, nc)) <= TOL
    p.vbar(dw, dx, dvb)
    pos = p.vbar_positions()
    assert np.array_equal(np.sort(pos), np.arange(m.nnz, dtype=np.uint32))
    assert np.array_equal(p.vbar_csr(dvb.download(np.float64, m.nnz)), orc.vbar(m, w, x))
    y_ref = orc.ncsr(m, x)
    r_ref = y_ref - b
    for fused in (True, False):
        outs = [ax.DeviceBuffer.from_array(np.full(s, np.nan)) for s in (n, n, nc, m.nnz)]
        p.fwd_adj(dx, db, *outs, fused=fused)
        y, r, xbar, vbar = (o.download(np.float64, s) for o, s in zip(outs, (n, n, nc, m.nnz)))
        vbar = p.vbar_csr(vbar)
        assert rel_err(y_ref, y) <= TOL
        assert rel_err(r_ref, r) <= TOL
        assert rel_err(orc.ncsr_t(m, r_ref, nc), xbar) <= 1e-11
        assert np.array_equal(vbar, orc.vbar(m, r, x))


def test_reordered_graph_replays_and_timing():
    m, _ = MATS["stencil21p"]
    p = ax.Plan(m, ax.FMT_RCT, flags=ax.PLAN_REORDER)
    x = ax.uniform(31, m.ncols)
    y_ref = orc.ncsr(m, x)
    dx, dy = ax.DeviceBuffer.from_array(x), ax.DeviceBuffer(8 * m.nrows)
    g = ax.Graph.of_plan(p, ax._lib.OP_SPMV, repeats=2, x=dx, y=dy)
    for _ in range(3):
        dy.upload(np.full(m.nrows, np.nan))
        g.launch()
        assert rel_err(y_ref, dy.download(np.float64, m.nrows)) <= TOL
        p.spmv(dx, dy)  # eager work between replays
    t = p.time(dx, dy, iters=5, warmup=1, mode=ax.core.TIME_KERNEL)
    assert t.mean_ms > 0 and t.main_ms > 0
    assert rel_err(y_ref, dy.download(np.float64, m.nrows)) <= TOL


SPMM_KERNELS = [
    ("csr_vector", ax.FMT_CSR, dict(kernel=ax.KER_STREAM)),
    ("unc_h1_thw4", ax.FMT_AXT, dict(thw=4, th=1)),
    ("rct_r1024_s1", ax.FMT_RCT, dict(tile_rows=1024, slices=1)),
    ("rct_r512_s2", ax.FMT_RCT, dict(tile_rows=512, slices=2)),
]


@pytest.mark.parametrize("mname", ["stencil21p", "banded_empty_shuffled"])
@pytest.mark.parametrize("kname,fmt,kw", SPMM_KERNELS, ids=[k[0] for k in SPMM_KERNELS])
def test_reordered_spmm(mname, kname, fmt, kw):
    m, _ = MATS[mname]
    p = ax.Plan(m, fmt, flags=ax.PLAN_REORDER, **kw)
    assert p.info()["name"].endswith("_rcm")
    # k = 8 after k = 24: the larger operand buffer serves the smaller k
    for k in (24, 8, 40):
        X = ax.uniform(41 + k, m.ncols * k).reshape(m.ncols, k)
        dX, dY = ax.DeviceBuffer.from_array(X), ax.DeviceBuffer.from_array(np.full(m.nrows * k, np.nan))
        p.spmm(k, dX, dY)
        Y = dY.download(np.float64, m.nrows * k).reshape(m.nrows, k)
        assert rel_err(orc.spmm(m, X), Y) <= TOL, (mname, kname, k)


def test_reordered_spmm_graph_and_x_refresh():
    m, _ = MATS["stencil21p"]
    p = ax.Plan(m, ax.FMT_AXT, thw=4, flags=ax.PLAN_REORDER)
    k = 16
    X = ax.uniform(51, m.ncols * k).reshape(m.ncols, k)
    dX, dY = ax.DeviceBuffer.from_array(X), ax.DeviceBuffer(8 * m.nrows * k)
    g = ax.Graph.of_plan(p, ax._lib.OP_SPMM, repeats=2, k=k, X=dX, Y=dY)
    ref = orc.spmm(m, X)
    for _ in range(3):
        dY.upload(np.full(m.nrows * k, np.nan))
        g.launch()
        assert rel_err(ref, dY.download(np.float64, m.nrows * k).reshape(m.nrows, k)) <= TOL
        p.spmm(32, ax.DeviceBuffer(8 * m.ncols * 32), ax.DeviceBuffer(8 * m.nrows * 32))  # grows the buffer
    # the AX layout's x refresh through a reordered plan, then its SpMV
    pa = ax.Plan(m, ax.FMT_AXT, thw=4, flags=ax.PLAN_REORDER | ax.PLAN_AX_LAYOUT)
    x = ax.uniform(52, m.ncols)
    dx = ax.DeviceBuffer.from_array(x)
    t = pa.time_op(ax._lib.OP_XREFRESH, iters=3, warmup=1, x=dx)
    assert t.mean_ms > 0
    assert rel_err(orc.ncsr(m, x), ax.spmv_host(pa, x)) <= TOL


COMBOS = [
    ("rct_f32val", ax.FMT_RCT, dict(vtype=ax.F32)),
    ("unc_h1_thw4_f32val", ax.FMT_AXT, dict(thw=4, vtype=ax.F32)),
    ("k1_c64_f32val", ax.FMT_K1, dict(chunk=64, vtype=ax.F32)),
    ("csr_vector_f32val", ax.FMT_CSR, dict(kernel=ax.KER_STREAM, vtype=ax.F32)),
    ("unc_h1_thw4_stage_x", ax.FMT_AXT, dict(thw=4, flags=ax.PLAN_STAGE_X)),
    ("unc_h1_thw2_s4_det", ax.FMT_AXT, dict(thw=2, slices=4, flags=ax.PLAN_DETERMINISTIC)),
    ("rct_r256_split", ax.FMT_RCT, dict(tile_rows=256, flags=ax.PLAN_RCT_SPLIT)),
]


@pytest.mark.parametrize("kname,fmt,kw", COMBOS, ids=[k[0] for k in COMBOS])
def test_reordered_with_other_options(kname, fmt, kw):
    """The flag combines with FP32 values (against the oracle on the
    float-rounded values), x staging, deterministic slices (bit-identical
    over repeated calls) and split RCT keys."""
    m, _ = MATS["stencil21p"]
    kw = dict(kw)
    p = ax.Plan(m, fmt, flags=ax.PLAN_REORDER | kw.pop("flags", 0), **kw)
    assert p.info()["name"].endswith("_rcm"), p.info()["name"]
    x = ax.uniform(61, m.ncols)
    ref = orc.ncsr_f32val(m, x) if kw.get("vtype") == ax.F32 else orc.ncsr(m, x)
    ys = [ax.spmv_host(p, x) for _ in range(3)]
    for y in ys:
        assert rel_err(ref, y) <= TOL, kname
    if "det" in kname:
        assert all(np.array_equal(ys[0], y) for y in ys[1:])


def test_reordered_rejects():
    m, _ = MATS["stencil21p"]
    rect = ax.CsrMatrix(np.array([0, 1, 2], np.uint32), np.array([0, 2], np.uint32), np.ones(2), ncols=3)
    with pytest.raises(ax.AxspmvError):
        ax.Plan(rect, ax.FMT_CSR, flags=ax.PLAN_REORDER)
    with pytest.raises(ax.AxspmvError):
        ax.Batch([m, m], fmt=ax.FMT_AXT, thw=4, flags=ax.PLAN_REORDER)
