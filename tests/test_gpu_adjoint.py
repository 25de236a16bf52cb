"""GPU: adjoint x_bar = A^T y_bar, value gradient v_bar, the fused forward +
adjoint pass, SpMM, and the autograd wrappers.

Checks: oracle (ncsr on A^T, vbar, spmm) at 1e-12; v_bar is one product per
entry, so bit-exact; <A x, w> = <x, A^T w>; central finite differences
(h = 1e-6) and a dense FP64 gradient on a 4,096-row random matrix of about
10 entries per row, values N(0,1).
"""
import numpy as np
import pytest

import gpu_spmv_amd as ax
from conftest import rel_err
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
TOL = 1e-12

ADJ_KERNELS = [
    ("unc_h1_thw4", ax.FMT_AXT, dict(thw=4, th=1)),
    ("unc_h1_thw8", ax.FMT_AXT, dict(thw=8, th=1)),
    ("unc_h1_thw2", ax.FMT_AXT, dict(thw=2, th=1)),
    ("csr_naive", ax.FMT_CSR, dict(kernel=ax.KER_NAIVE)),
    ("rct_auto", ax.FMT_RCT, dict()),
    ("rct_r512_s4", ax.FMT_RCT, dict(tile_rows=512, slices=4)),
    ("rct_r512_s4_atomic", ax.FMT_RCT, dict(tile_rows=512, slices=4, flags=ax.PLAN_RCT_ATOMIC)),
    ("rct_r1024_s2_split", ax.FMT_RCT, dict(tile_rows=1024, slices=2, flags=ax.PLAN_RCT_SPLIT)),
    # taller than the transposed plan's tiles may be (x_bar sums and x values
    # of the fused gradient: 16 B of LDS per row): that plan is cut to fit
    ("rct_r16384_s4", ax.FMT_RCT, dict(tile_rows=16384, slices=4)),
    # A^T also as a deterministic AXT plan in column slices for the adjoint
    ("rct_r512_s4_det", ax.FMT_RCT, dict(tile_rows=512, slices=4, flags=ax.PLAN_DETERMINISTIC)),
]


def mats():
    rng = np.random.default_rng(5)
    yield "random4096", ax.random_sparse(4096, 5, 15, seed=8)
    yield "banded_empty", ax.banded(3000, 3, 0.3, seed=6)
    yield "stencil16p", ax.stencil27(16, permute=True, seed=4)
    # an odd row count: the slice merge (and its residual) in 8-byte steps
    yield "random_odd", ax.random_sparse(3001, 2, 30, seed=9)
    # long rows (more than 64 units: the fused kernel's second sweep)
    n = 2000
    rl = rng.integers(0, 8, n)
    rl[[7, 900, 1999]] = [700, 1500, 300]
    row = np.concatenate([[0], np.cumsum(rl)]).astype(np.uint32)
    col = np.concatenate([np.sort(rng.choice(n, k, replace=False)) for k in rl]).astype(np.uint32)
    yield "long_rows", ax.CsrMatrix(row, col, rng.standard_normal(col.size), ncols=n)


MATS = dict(mats())


@pytest.mark.parametrize("mname", list(MATS))
@pytest.mark.parametrize("kname,fmt,kw", ADJ_KERNELS, ids=[k[0] for k in ADJ_KERNELS])
def test_adjoint_vbar_fused(mname, kname, fmt, kw):
    m = MATS[mname]
    n, nc = m.nrows, m.ncols
    x = ax.uniform(21, nc)
    b = ax.uniform(22, n, stream=ax.VEC_STREAM + 1)
    w = ax.uniform(23, n)
    kw = dict(kw)
    p = ax.Plan(m, fmt, flags=ax.PLAN_TRANSPOSE | kw.pop("flags", 0), **kw)
    dx, db, dw = (ax.DeviceBuffer.from_array(v) for v in (x, b, w))
    dxb, dvb = ax.DeviceBuffer(8 * nc), ax.DeviceBuffer(8 * max(m.nnz, 1))
    p.adjoint(dw, dxb)
    p.vbar(dw, dx, dvb)
    xb = dxb.download(np.float64, nc)
    assert rel_err(orc.ncsr_t(m, w, nc), xb) <= TOL
    # v_bar in CSR order (RCT writes entry order; its positions map it back)
    pos = p.vbar_positions()
    assert np.array_equal(np.sort(pos), np.arange(m.nnz, dtype=np.uint32))
    assert np.array_equal(p.vbar_csr(dvb.download(np.float64, m.nnz)), orc.vbar(m, w, x))
    y_ref = orc.ncsr(m, x)
    assert abs(np.dot(y_ref, w) - np.dot(x, xb)) <= 1e-11 * (np.abs(y_ref).sum() * np.abs(w).max() + 1)
    # fused and split forward + adjoint
    r_ref = y_ref - b
    for fused in (True, False):
        outs = [ax.DeviceBuffer.from_array(np.full(s, np.nan)) for s in (n, n, nc, max(m.nnz, 1))]
        p.fwd_adj(dx, db, *outs, fused=fused)
        y, r, xbar, vbar = (o.download(np.float64, s) for o, s in zip(outs, (n, n, nc, m.nnz)))
        vbar = p.vbar_csr(vbar)
        assert rel_err(y_ref, y) <= TOL
        assert rel_err(r_ref, r) <= TOL
        assert rel_err(orc.ncsr_t(m, r_ref, nc), xbar) <= 1e-11
        assert rel_err(orc.vbar(m, r_ref, x), vbar) <= 1e-11
        assert np.array_equal(vbar, orc.vbar(m, r, x))  # v_bar is r[row] * x[col] of the r it wrote


def test_gradients_finite_differences_and_dense():
    """Config 5 gradient check: 4,096 rows, ~10 entries per row, N(0,1)."""
    m = ax.random_sparse(4096, 5, 15, seed=8)
    x = ax.uniform(31, m.ncols)
    b = ax.uniform(32, m.nrows, stream=ax.VEC_STREAM + 1)
    p = ax.Plan(m, ax.FMT_AXT, thw=4, flags=ax.PLAN_TRANSPOSE)
    dx, db = ax.DeviceBuffer.from_array(x), ax.DeviceBuffer.from_array(b)
    n, nc, nnz = m.nrows, m.ncols, m.nnz
    outs = [ax.DeviceBuffer(8 * s) for s in (n, n, nc, nnz)]
    p.fwd_adj(dx, db, *outs, fused=True)
    _, r, xbar, vbar = (o.download(np.float64, s) for o, s in zip(outs, (n, n, nc, nnz)))
    # dense FP64 reference gradient
    A = m.to_dense()
    rd = A @ x - b
    assert rel_err(A.T @ rd, xbar) < 1e-12
    assert rel_err(rd[m.row_ids()] * x[m.col], vbar) < 1e-12

    def loss(xv, vals):
        mm = ax.CsrMatrix(m.row, m.col, vals, ncols=nc)
        y = orc.ncsr(mm, xv)
        return 0.5 * float(np.dot(y - b, y - b))

    h = 1e-6
    rng = np.random.default_rng(0)
    for i in rng.choice(nc, 24, replace=False):
        e = np.zeros(nc)
        e[i] = h
        fd = (loss(x + e, m.val) - loss(x - e, m.val)) / (2 * h)
        assert abs(fd - xbar[i]) <= 2e-5 * max(1.0, abs(xbar[i])), (i, fd, xbar[i])
    for k in rng.choice(nnz, 24, replace=False):
        e = np.zeros(nnz)
        e[k] = h
        fd = (loss(x, m.val + e) - loss(x, m.val - e)) / (2 * h)
        assert abs(fd - vbar[k]) <= 2e-5 * max(1.0, abs(vbar[k])), (k, fd, vbar[k])


@pytest.mark.timeout(900)  # the first import of torch on a fresh machine is slow
@pytest.mark.parametrize("fmt,kw", [(ax.FMT_AXT, dict(thw=4)), (ax.FMT_RCT, dict(tile_rows=128, slices=2))],
                         ids=["axt", "rct"])
def test_autograd_wrappers(fmt, kw):
    torch = pytest.importorskip("torch")
    from gpu_spmv_amd import autograd as ag
    m = ax.random_sparse(500, 2, 9, seed=3)
    vals = torch.tensor(m.val, requires_grad=True)
    x = torch.tensor(ax.uniform(41, m.ncols), requires_grad=True)
    b = torch.tensor(ax.uniform(42, m.nrows))
    # spmv: gradients against a dense torch reference
    y = ag.spmv(vals, x, m, fmt, **kw)
    g = torch.tensor(ax.uniform(43, m.nrows))
    (y * g).sum().backward()
    A = torch.tensor(m.to_dense())
    assert torch.allclose(y, A @ x.detach(), rtol=0, atol=1e-12)
    assert torch.allclose(x.grad, A.T @ g, rtol=0, atol=1e-12)
    rows = torch.tensor(m.row_ids().astype(np.int64))
    cols = torch.tensor(m.col.astype(np.int64))
    assert torch.allclose(vals.grad, g[rows] * x.detach()[cols], rtol=0, atol=1e-13)
    # least squares, fused and split
    for fused in (True, False):
        vals.grad = None
        x.grad = None
        loss = ag.least_squares(vals, x, b, m, fmt, fused=fused, **kw)
        loss.backward()
        r = A @ x.detach() - b
        assert abs(loss.item() - 0.5 * float(r @ r)) <= 1e-12 * (1 + loss.item())
        assert torch.allclose(x.grad, A.T @ r, rtol=0, atol=1e-11)
        assert torch.allclose(vals.grad, r[rows] * x.detach()[cols], rtol=0, atol=1e-12)


SPMM_CASES = [(1, 0), (3, 0), (8, 8), (8, 0), (32, 32), (32, 8), (64, 64), (64, 16), (20, 0), (96, 64)]


@pytest.mark.parametrize("mname", ["random4096", "long_rows", "banded_empty"])
@pytest.mark.parametrize("k,kt", SPMM_CASES)
@pytest.mark.parametrize("fmt,kw", [(ax.FMT_AXT, dict(thw=4)), (ax.FMT_CSR, dict(kernel=ax.KER_NAIVE)),
                                    (ax.FMT_AXC, dict(hbs=8)), (ax.FMT_RCT, dict(tile_rows=256)),
                                    (ax.FMT_RCT, dict(tile_rows=200, slices=4))],
                         ids=["axt", "csr", "axc", "rct_r256", "rct_r200_s4"])
def test_spmm(mname, k, kt, fmt, kw):
    m = MATS[mname]
    X = ax.uniform(50 + k, m.ncols * k).reshape(m.ncols, k)
    ref = orc.spmm(m, X)
    p = ax.Plan(m, fmt, ktile=kt, **kw)
    dX = ax.DeviceBuffer.from_array(X)
    dY = ax.DeviceBuffer.from_array(np.full(m.nrows * k, np.nan))
    p.spmm(k, dX, dY)
    Y = dY.download(np.float64, m.nrows * k).reshape(m.nrows, k)
    assert rel_err(ref, Y) <= TOL


@pytest.mark.parametrize("fmt,kw", [(ax.FMT_AXT, dict(thw=4)), (ax.FMT_RCT, dict(tile_rows=512, slices=2))],
                         ids=["axt", "rct"])
def test_spmm_f32_values(fmt, kw):
    m = MATS["random4096"]
    k = 32
    X = ax.uniform(60, m.ncols * k).reshape(m.ncols, k)
    m32 = ax.CsrMatrix(m.row, m.col, m.val.astype(np.float32).astype(np.float64), ncols=m.ncols)
    p = ax.Plan(m, fmt, vtype=ax.F32, **kw)
    dX, dY = ax.DeviceBuffer.from_array(X), ax.DeviceBuffer(8 * m.nrows * k)
    p.spmm(k, dX, dY)
    assert rel_err(orc.spmm(m32, X), dY.download(np.float64, m.nrows * k).reshape(m.nrows, k)) <= TOL


@pytest.mark.parametrize("k", [8, 5])
def test_rct_spmm_unaligned_y(k):
    """Y at an 8-byte offset with column slices (the clear of Y falls back to
    8-byte stores); the words either side stay untouched."""
    m = MATS["long_rows"]
    X = ax.uniform(61, m.ncols * k).reshape(m.ncols, k)
    p = ax.Plan(m, ax.FMT_RCT, tile_rows=1024, slices=8)
    n = m.nrows * k
    dX, dY = ax.DeviceBuffer.from_array(X), ax.DeviceBuffer.from_array(np.full(n + 2, np.nan))
    p.spmm(k, dX, dY.ptr + 8)
    ax._lib.call("axspmv_device_sync")
    Y = dY.download(np.float64, n + 2)
    assert np.isnan(Y[0]) and np.isnan(Y[-1])
    assert rel_err(orc.spmm(m, X), Y[1:-1].reshape(m.nrows, k)) <= TOL


def test_rct_spmm_lds_limit():
    """RCT SpMM keeps tile_rows x ktile sums in LDS: the default k-tile is cut
    to fit, an explicit one that does not fit is an error."""
    m = MATS["random4096"]
    k = 64
    X = ax.uniform(62, m.ncols * k).reshape(m.ncols, k)
    dX, dY = ax.DeviceBuffer.from_array(X), ax.DeviceBuffer(8 * m.nrows * k)
    p = ax.Plan(m, ax.FMT_RCT, tile_rows=2048)  # ktile 64 -> 8 (2048 x 8 x 8 B = 128 KB)
    p.spmm(k, dX, dY)
    assert rel_err(orc.spmm(m, X), dY.download(np.float64, m.nrows * k).reshape(m.nrows, k)) <= TOL
    with pytest.raises(ax.AxspmvError, match="LDS"):
        ax.Plan(m, ax.FMT_RCT, tile_rows=2048, ktile=16).spmm(k, dX, dY)
    with pytest.raises(ax.AxspmvError, match="LDS"):
        ax.Plan(m, ax.FMT_RCT, tile_rows=4096).spmm(k, dX, dY)


def test_unsupported_combinations_raise():
    m = MATS["random4096"]
    p = ax.Plan(m, ax.FMT_AXT, thw=4)  # no transpose
    d = ax.DeviceBuffer(8 * m.nrows)
    with pytest.raises(ax.AxspmvError):
        p.adjoint(d, d)
    pk = ax.Plan(m, ax.FMT_K1, chunk=64)
    with pytest.raises(ax.AxspmvError):
        pk.spmm(8, d, d)
    with pytest.raises(ax.AxspmvError):
        pk.vbar(d, d, d)
    with pytest.raises(ax.AxspmvError):
        pk.vbar_positions()
    pr = ax.Plan(m, ax.FMT_RCT)  # no transpose: no plan of A^T for the value gradient
    with pytest.raises(ax.AxspmvError):
        pr.vbar_positions()
    with pytest.raises(ax.AxspmvError, match="TRANSPOSE"):
        pr.vbar(d, d, d)
    assert np.array_equal(ax.Plan(m, ax.FMT_AXT, thw=4, flags=ax.PLAN_TRANSPOSE).vbar_positions(),
                          np.arange(m.nnz, dtype=np.uint32))


@pytest.mark.parametrize("kw", [dict(tile_rows=16384, slices=1), dict(tile_rows=16384, slices=4), dict()])
def test_rct_fused_gradient_tall_tiles(kw):
    """30,000 rows with 16384-row tiles: the plan of A^T keeps its rows' x_bar
    sums and x values in LDS (16 B a row), so its tiles are cut to fit; the
    fused sweep (x_bar and v_bar together) equals the separate kernels' v_bar
    bit for bit, and both match the oracle."""
    m = ax.random_sparse(30000, 1, 12, seed=33)
    n, nc, nnz = m.nrows, m.ncols, m.nnz
    x, b = ax.uniform(34, nc), ax.uniform(35, n, stream=ax.VEC_STREAM + 1)
    p = ax.Plan(m, ax.FMT_RCT, flags=ax.PLAN_TRANSPOSE, **kw)
    dx, db = ax.DeviceBuffer.from_array(x), ax.DeviceBuffer.from_array(b)
    res = {}
    for fused in (True, False):
        outs = [ax.DeviceBuffer.from_array(np.full(k, np.nan)) for k in (n, n, nc, nnz)]
        p.fwd_adj(dx, db, *outs, fused=fused)
        res[fused] = [o.download(np.float64, k) for o, k in zip(outs, (n, n, nc, nnz))]
    y, r, xbar, vbar = res[True]
    r_ref = orc.ncsr(m, x) - b
    assert rel_err(r_ref, r) <= TOL
    assert rel_err(orc.ncsr_t(m, r_ref, nc), xbar) <= 1e-11
    assert np.array_equal(p.vbar_csr(vbar), orc.vbar(m, r, x))
    # the separate kernels: the same entry order, each v_bar from its own r
    assert np.array_equal(p.vbar_csr(res[False][3]), orc.vbar(m, res[False][1], x))


@pytest.mark.parametrize("fused", [True, False])
def test_rct_vbar_at_8_byte_offset(fused):
    """The fused sweep writes v_bar pairs with 16-byte stores when v_bar is
    16-byte aligned; at an 8-byte offset it falls back to 8-byte stores and
    writes the same values, and nothing outside v_bar[0 .. nnz)."""
    m = ax.random_sparse(20000, 1, 12, seed=36)
    n, nc, nnz = m.nrows, m.ncols, m.nnz
    x, b = ax.uniform(37, nc), ax.uniform(38, n, stream=ax.VEC_STREAM + 1)
    p = ax.Plan(m, ax.FMT_RCT, flags=ax.PLAN_TRANSPOSE)
    dx, db = ax.DeviceBuffer.from_array(x), ax.DeviceBuffer.from_array(b)
    outs = [ax.DeviceBuffer.from_array(np.full(k, np.nan)) for k in (n, n, nc)]
    big = ax.DeviceBuffer.from_array(np.full(nnz + 2, np.nan))
    p.fwd_adj(dx, db, *outs, big.ptr + 8, fused=fused)
    r = outs[1].download(np.float64, n)
    full = big.download(np.float64, nnz + 2)
    assert np.isnan(full[0]) and np.isnan(full[-1])
    assert np.array_equal(p.vbar_csr(full[1:-1]), orc.vbar(m, r, x))
    assert rel_err(orc.ncsr(m, x) - b, r) <= TOL


@pytest.mark.parametrize("gen", [lambda: ax.powerlaw(1 << 18, 2.0, 2048, 20.0, seed=1),
                                 lambda: ax.stencil27(48, permute=True, seed=2),
                                 lambda: MATS["long_rows"]], ids=["powerlaw256k", "stencil48p", "long_rows"])
def test_rct_adjoint_deterministic(gen):
    """RCT plans with PLAN_DETERMINISTIC: x_bar = A^T y_bar runs on an AXT plan
    of A^T in column slices whose slice partials are summed in slice order --
    bit-identical over repeated calls (into dirty buffers), and within 1e-12
    of the oracle; without the flag the same plan adds with atomics."""
    m = gen()
    w = ax.uniform(31, m.nrows)
    ref = orc.ncsr_t(m, w, m.ncols)
    p = ax.Plan(m, ax.FMT_RCT, flags=ax.PLAN_TRANSPOSE | ax.PLAN_DETERMINISTIC)
    dw = ax.DeviceBuffer.from_array(w)
    outs = []
    for i in range(4):
        dxb = ax.DeviceBuffer.from_array(np.full(m.ncols, np.nan if i % 2 else 7.0))
        p.adjoint(dw, dxb)
        outs.append(dxb.download(np.float64, m.ncols))
    assert rel_err(ref, outs[0]) <= TOL
    for o in outs[1:]:
        assert o.tobytes() == outs[0].tobytes()
    # the forward SpMV, v_bar and the fused pass of the same plan still work
    x = ax.uniform(32, m.ncols)
    assert rel_err(orc.ncsr(m, x), ax.spmv_host(p, x)) <= TOL


@pytest.mark.parametrize("slices", [2, 4, 8])
def test_axt_slices_deterministic(slices):
    """AXT in column slices with PLAN_DETERMINISTIC: y bit-identical over
    repeated SpMVs (slice partials summed in slice order), and the adjoint
    through its sliced plan of A^T too."""
    m = ax.powerlaw(1 << 17, 2.0, 2048, 20.0, seed=3)
    p = ax.Plan(m, ax.FMT_AXT, thw=2, slices=slices, flags=ax.PLAN_DETERMINISTIC | ax.PLAN_TRANSPOSE)
    assert p.info()["name"].endswith("_det")
    x = ax.uniform(33, m.ncols)
    w = ax.uniform(34, m.nrows)
    ys = [ax.spmv_host(p, x) for _ in range(3)]
    assert rel_err(orc.ncsr(m, x), ys[0]) <= TOL
    assert all(y.tobytes() == ys[0].tobytes() for y in ys)
    dw = ax.DeviceBuffer.from_array(w)
    xs = []
    for _ in range(3):
        dxb = ax.DeviceBuffer.from_array(np.full(m.ncols, np.nan))
        p.adjoint(dw, dxb)
        xs.append(dxb.download(np.float64, m.ncols))
    assert rel_err(orc.ncsr_t(m, w, m.ncols), xs[0]) <= TOL
    assert all(v.tobytes() == xs[0].tobytes() for v in xs)
