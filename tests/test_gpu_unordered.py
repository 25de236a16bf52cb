"""GPU: CSR rows in any column order, with repeated columns.  The reference
sums every stored entry whatever its order; so does every kernel here.
Checked against the oracle's ncsr (which sums in CSR order) at 1e-12 for
every format, x staging (whose windows must span each row's true smallest
and largest column), SpMM, the adjoint and v_bar, a batch, and the device
builders (bit for bit against the host builders, which keep CSR order)."""
import numpy as np
import pytest

import gpu_spmv_amd as ax
from conftest import rel_err
from oracle import oracle as orc
from test_gpu_spmv import KERNELS

pytestmark = pytest.mark.gpu
TOL = 1e-12


def scramble(m, seed):
    """Shuffle every row's entries and repeat about one entry in ten (the
    copy gets its own value)."""
    rng = np.random.default_rng(seed)
    rows, cols, vals = [0], [], []
    for r in range(m.nrows):
        b, e = int(m.row[r]), int(m.row[r + 1])
        c, v = list(m.col[b:e]), list(m.val[b:e])
        for k in rng.choice(e - b, size=(e - b) // 10, replace=True) if e > b else []:
            c.append(c[k])
            v.append(float(rng.normal()))
        p = rng.permutation(len(c))
        cols += [c[i] for i in p]
        vals += [v[i] for i in p]
        rows.append(len(cols))
    return ax.CsrMatrix(np.array(rows), np.array(cols, dtype=np.uint32), np.array(vals), ncols=m.ncols)


MATS = {
    "random": scramble(ax.random_sparse(3000, 1, 40, seed=21), 1),
    "banded": scramble(ax.banded(4000, 24, 0.5, seed=22), 2),     # narrow windows: x staging kicks in
    "long_rows": scramble(ax.powerlaw(4096, 2.0, 1500, 30.0, seed=23), 3),
}


def test_scrambled_matrices_are_unordered():
    for m in MATS.values():
        with pytest.raises(ax.AxspmvError, match="increasing"):
            m.check()


@pytest.mark.parametrize("mname", sorted(MATS))
@pytest.mark.parametrize("kname,fmt,kw", KERNELS, ids=[k[0] for k in KERNELS])
def test_spmv_unordered(mname, kname, fmt, kw):
    m = MATS[mname]
    x = ax.uniform(90, m.ncols)
    assert rel_err(orc.ncsr(m, x), ax.spmv_host(ax.Plan(m, fmt, **kw), x)) <= TOL


@pytest.mark.parametrize("mname", sorted(MATS))
@pytest.mark.parametrize("kw", [dict(thw=4), dict(thw=64, th=4)])
def test_x_staging_unordered(mname, kw):
    """Staged and unstaged give the same y bit for bit (same order)."""
    m = MATS[mname]
    x = ax.uniform(91, m.ncols)
    y0 = ax.spmv_host(ax.Plan(m, ax.FMT_AXT, **kw), x)
    y1 = ax.spmv_host(ax.Plan(m, ax.FMT_AXT, flags=ax.PLAN_STAGE_X, **kw), x)
    assert np.array_equal(y0, y1)
    assert rel_err(orc.ncsr(m, x), y1) <= TOL


@pytest.mark.parametrize("mname", sorted(MATS))
@pytest.mark.parametrize("fmt,kw", [(ax.FMT_AXT, dict(thw=4)), (ax.FMT_RCT, dict(tile_rows=512, slices=2)),
                                    (ax.FMT_CSR, dict(kernel=ax.KER_NAIVE))], ids=["axt", "rct", "csr"])
def test_spmm_and_adjoint_unordered(mname, fmt, kw):
    m = MATS[mname]
    n, nc = m.nrows, m.ncols
    k = 16
    X = ax.uniform(92, nc * k).reshape(nc, k)
    p = ax.Plan(m, fmt, flags=ax.PLAN_TRANSPOSE, **kw)
    dX, dY = ax.DeviceBuffer.from_array(X), ax.DeviceBuffer(8 * n * k)
    p.spmm(k, dX, dY)
    assert rel_err(orc.spmm(m, X), dY.download(np.float64, n * k).reshape(n, k)) <= TOL
    x, w = ax.uniform(93, nc), ax.uniform(94, n)
    dx, dw = ax.DeviceBuffer.from_array(x), ax.DeviceBuffer.from_array(w)
    dxb, dvb = ax.DeviceBuffer(8 * nc), ax.DeviceBuffer(8 * m.nnz)
    p.adjoint(dw, dxb)
    assert rel_err(orc.ncsr_t(m, w, nc), dxb.download(np.float64, nc)) <= TOL
    p.vbar(dw, dx, dvb)
    assert np.array_equal(p.vbar_csr(dvb.download(np.float64, m.nnz)), orc.vbar(m, w, x))


@pytest.mark.parametrize("fmt", [ax.FMT_AXT, ax.FMT_RCT], ids=["axt", "rct"])
def test_batch_unordered(fmt):
    mats = list(MATS.values())
    b = ax.Batch(mats, fmt=fmt)
    x = ax.uniform(95, int(b.xoff[-1]))
    dx, dy = ax.DeviceBuffer.from_array(x), ax.DeviceBuffer(8 * int(b.yoff[-1]))
    b.spmv(dx, dy)
    y = dy.download(np.float64, int(b.yoff[-1]))
    for i, m in enumerate(mats):
        assert rel_err(orc.ncsr(m, x[b.xoff[i]:b.xoff[i + 1]]), y[b.yoff[i]:b.yoff[i + 1]]) <= TOL, i


@pytest.mark.parametrize("mname", sorted(MATS))
def test_device_builders_unordered(mname):
    m = MATS[mname]
    x = ax.uniform(96, m.ncols)
    for args in ((256, 4, 1, ax.AXT_UNC), (256, 64, 4, ax.AXT_UNC), (256, 64, 1, ax.AXT_COM), (256, 64, 4, ax.AXT_COM)):
        h, d = ax.build_axt(m, x, *args), ax.dev_build_axt(m, x, *args)
        assert np.array_equal(d["ax"], h["ax"]) and np.array_equal(d["sec"], h["sec"]), args
    h, d = ax.build_axc(m, x, 8), ax.dev_build_axc(m, x, 8)
    assert np.array_equal(d["ax"], h["ax"]) and np.array_equal(d["plane"], h["mapx"])
