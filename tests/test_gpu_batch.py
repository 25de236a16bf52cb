"""GPU: batched persistent SpMV over many uneven banded matrices (the batched
configuration at reduced sizes), per-matrix autotuning, and hipGraph-captured
repeat loops checked replay by replay against the eager result."""
import ctypes as C

import numpy as np
import pytest

import gpu_spmv_amd as ax
from conftest import rel_err
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
TOL = 1e-12


def batch_mats(count, nmax, seed=7):
    """count banded matrices: n ~ U{1000..nmax}, half-bandwidth ~ U{1..64},
    in-band fill 0.3 (diagonal included in the draw), values N(0,1)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(1000, nmax + 1))
        hbw = int(rng.integers(1, 65))
        out.append(ax.banded(n, hbw, 0.3, seed=1000 + i))
    return out


MATS = batch_mats(48, 12000)


def check_batch(b, mats, x):
    dx = ax.DeviceBuffer.from_array(x)
    ny = int(b.yoff[-1])
    dy = ax.DeviceBuffer.from_array(np.full(ny, np.nan))
    b.spmv(dx, dy)
    y = dy.download(np.float64, ny)
    for i, m in enumerate(mats):
        ref = orc.ncsr(m, x[b.xoff[i]:b.xoff[i + 1]])
        assert rel_err(ref, y[b.yoff[i]:b.yoff[i + 1]]) <= TOL, i
    return dx, dy, y


@pytest.mark.parametrize("kw", [dict(thw=4), dict(thw=2), dict(thw=64, th=4), dict(thw=8, budget=64),
                                dict(thw=4, vtype=ax.F32)])
def test_batch_uniform_options(kw):
    b = ax.Batch(MATS, **kw)
    x = ax.uniform(77, int(b.xoff[-1]))
    if kw.get("vtype") == ax.F32:
        mats = [ax.CsrMatrix(m.row, m.col, m.val.astype(np.float32).astype(np.float64), ncols=m.ncols) for m in MATS]
    else:
        mats = MATS
    check_batch(b, mats, x)
    info = b.info()
    assert info["count"] == len(MATS) and info["nnz"] == sum(m.nnz for m in MATS)


@pytest.mark.parametrize("kw", [dict(), dict(tile_rows=1024), dict(tile_rows=256, slices=2), dict(vtype=ax.F32)])
def test_rct_batch(kw):
    """RCT batch: one RCT plan of the block-diagonal matrix; row blocks cross
    matrix boundaries; empty matrices and empty rows give zero rows."""
    e = ax.CsrMatrix(np.zeros(51, np.uint32), np.zeros(0, np.uint32), np.zeros(0), ncols=70)
    mats = MATS[:20] + [e] + MATS[20:30] + [ax.banded(3000, 3, 0.3, seed=6)]
    b = ax.Batch(mats, fmt=ax.FMT_RCT, **kw)
    if kw.get("vtype") == ax.F32:
        ref = [ax.CsrMatrix(m.row, m.col, m.val.astype(np.float32).astype(np.float64), ncols=m.ncols) for m in mats]
    else:
        ref = mats
    dx, dy, y = check_batch(b, ref, ax.uniform(85, int(b.xoff[-1])))
    info = b.info()
    assert info["count"] == len(mats) and info["nnz"] == sum(m.nnz for m in mats) and info["nitems"] >= 1
    # graph replays, eager work in between (the LDS sums are unordered: equal
    # to the eager result within the tolerance, not bit for bit)
    g = ax.Graph.of_batch(b, dx, dy, repeats=3)
    for rep in range(3):
        dy.upload(np.full(y.size, np.nan))
        g.launch()
        ax._lib.call("axspmv_device_sync")
        assert rel_err(y, dy.download(np.float64, y.size)) <= TOL, rep


def test_rct_batch_options():
    """An RCT batch takes one option set: equal per-matrix options are fine,
    different ones are an error."""
    mats = MATS[:4]
    same = [ax.opts(ax.FMT_RCT, tile_rows=512)] * 4
    b = ax.Batch(mats, opts_list=same)
    check_batch(b, mats, ax.uniform(86, int(b.xoff[-1])))
    with pytest.raises(ax.AxspmvError, match="one option set"):
        ax.Batch(mats, opts_list=same[:3] + [ax.opts(ax.FMT_RCT, tile_rows=1024)])


def test_batch_repeated_matrices():
    """One matrix listed several times (built once per tile shape, as the
    tuner's group batches do), also with different options per listing."""
    mats = [MATS[0], MATS[1], MATS[0], MATS[0], MATS[2], MATS[1]]
    b = ax.Batch(mats, thw=4)
    check_batch(b, mats, ax.uniform(82, int(b.xoff[-1])))
    ol = [ax.opts(ax.FMT_AXT, thw=t, th=h) for t, h in ((4, 1), (2, 1), (2, 1), (64, 4), (4, 1), (4, 1))]
    b = ax.Batch(mats, opts_list=ol)
    check_batch(b, mats, ax.uniform(83, int(b.xoff[-1])))


def test_autotune_batch_groups():
    """Each shape group is tuned on a batch of its repeated members; a group
    of empty matrices is tuned once and still gets options."""
    e = ax.CsrMatrix(np.zeros(101, np.uint32), np.zeros(0, np.uint32), np.zeros(0), ncols=100)
    mats = [MATS[3], e, MATS[4], e, MATS[3]]
    cands = [ax.opts(ax.FMT_AXT, thw=t) for t in (2, 4)] + [ax.opts(ax.FMT_AXT, thw=64, th=4)]
    ol, ng = ax.autotune_batch(mats, candidates=cands, iters=3)
    assert len(ol) == len(mats) and ng >= 2
    assert (ol[0].thw, ol[0].th) == (ol[4].thw, ol[4].th) and (ol[1].thw, ol[1].th) == (ol[3].thw, ol[3].th)
    b = ax.Batch(mats, opts_list=ol)
    check_batch(b, mats, ax.uniform(84, int(b.xoff[-1])))


def test_autotune_batch_whole_batch_candidates():
    """RCT candidates compete for the whole batch against the per-group AXT
    choice; the result is one format, and the batch it builds is correct."""
    mats = MATS[:12]
    ol, ng = ax.autotune_batch(mats, candidates=[ax.opts(ax.FMT_AXT, thw=4), ax.opts(ax.FMT_RCT)], iters=3)
    assert len({o.format for o in ol}) == 1
    if ol[0].format == ax.FMT_RCT:
        assert ng == 1
    b = ax.Batch(mats, opts_list=ol)
    check_batch(b, mats, ax.uniform(87, int(b.xoff[-1])))
    # RCT candidates only: one of them for every matrix
    rc = [ax.opts(ax.FMT_RCT, tile_rows=512), ax.opts(ax.FMT_RCT)]
    ol, ng = ax.autotune_batch(mats, candidates=rc, iters=3)
    assert ng == 1 and ol[0].format == ax.FMT_RCT and ol[0].tile_rows in (0, 512)
    assert all(o.tile_rows == ol[0].tile_rows for o in ol)
    # the built-in set (AXT per group + default RCT)
    ol, ng = ax.autotune_batch(mats, iters=3)
    b = ax.Batch(mats, opts_list=ol)
    check_batch(b, mats, ax.uniform(88, int(b.xoff[-1])))
    with pytest.raises(ax.AxspmvError, match="AXT or RCT"):
        ax.autotune_batch(mats, candidates=[ax.opts(ax.FMT_K1)], iters=3)


def test_batch_mixed_per_matrix_options():
    choices = [dict(thw=4), dict(thw=2), dict(thw=16), dict(thw=64, th=2), dict(thw=64, th=4), dict(thw=32, th=8),
               dict(thw=1)]
    ol = [ax.opts(ax.FMT_AXT, **choices[i % len(choices)]) for i in range(len(MATS))]
    b = ax.Batch(MATS, opts_list=ol)
    check_batch(b, MATS, ax.uniform(78, int(b.xoff[-1])))


def test_batch_lds_staging_is_exact():
    """Items that stage x in LDS give the same y, bit for bit, as without."""
    for kw in (dict(thw=4), dict(thw=64, th=4)):
        staged = ax.Batch(MATS, flags=ax.PLAN_STAGE_X, **kw)
        plain = ax.Batch(MATS, **kw)
        x = ax.uniform(80, int(staged.xoff[-1]))
        dx = ax.DeviceBuffer.from_array(x)
        ny = int(staged.yoff[-1])
        ys = []
        for b in (staged, plain):
            dy = ax.DeviceBuffer(8 * ny)
            b.spmv(dx, dy)
            ax._lib.call("axspmv_device_sync")
            ys.append(dy.download(np.float64, ny))
        assert np.array_equal(ys[0], ys[1]), kw
        check_batch(staged, MATS, x)


def test_autotune_and_tuned_batch():
    defaults = ax.autotune_defaults()
    assert {o.format for o in defaults} == {ax.FMT_AXT, ax.FMT_K1, ax.FMT_RCT}
    best, times = ax.autotune(MATS[0])
    assert times.size == len(defaults) and 0 <= best < times.size and times[best] > 0
    assert times[best] == times[times >= 0].min()
    # the built-in set through the C API: times_ms must hold the whole set
    with pytest.raises(ax.AxspmvError):
        b = C.c_uint32()
        t = np.zeros(4)
        ax._lib.call("axspmv_autotune", MATS[0].ref, ax._lib.OP_SPMV, 0, None, 4, 3, None, C.byref(b),
                     t.ctypes.data_as(C.POINTER(C.c_double)))
    cands = [ax.opts(ax.FMT_AXT, thw=t) for t in (2, 4, 8)] + [ax.opts(ax.FMT_AXT, thw=64, th=4)]
    ol, ngroups = ax.autotune_batch(MATS, candidates=cands, iters=5)
    assert 1 <= ngroups <= len(MATS)
    b = ax.Batch(MATS, opts_list=ol)
    check_batch(b, MATS, ax.uniform(79, int(b.xoff[-1])))
    # SpMM k-tile tuning
    m = ax.random_sparse(3000, 5, 30, seed=2)
    kt_best, kt_times = ax.autotune(m, op=ax._lib.OP_SPMM, k=32,
                                    candidates=[ax.opts(ax.FMT_AXT, thw=4, ktile=k) for k in (8, 16, 32, 64)])
    assert 0 <= kt_best < 4 and (kt_times >= 0).all()
    # the built-in SpMM set (AXT k-tiles, RCT tiles sized to the LDS): every
    # candidate runs, and the winner computes the product
    sd = ax.autotune_defaults(ax._lib.OP_SPMM)
    assert {o.format for o in sd} == {ax.FMT_AXT, ax.FMT_RCT}
    s_best, s_times = ax.autotune(m, op=ax._lib.OP_SPMM, k=24)
    assert s_times.size == len(sd) and (s_times > 0).all() and s_times[s_best] == s_times.min()
    c = sd[s_best]
    p = ax.Plan(m, c.format, ktile=c.ktile, thw=c.thw, tile_rows=c.tile_rows, slices=c.slices)
    X = ax.uniform(81, m.ncols * 24).reshape(m.ncols, 24)
    dX, dY = ax.DeviceBuffer.from_array(X), ax.DeviceBuffer(8 * m.nrows * 24)
    p.spmm(24, dX, dY)
    assert rel_err(orc.spmm(m, X), dY.download(np.float64, m.nrows * 24).reshape(m.nrows, 24)) <= 1e-12


def test_batch_graph_replays():
    b = ax.Batch(MATS, thw=4)
    x = ax.uniform(80, int(b.xoff[-1]))
    dx, dy, y_eager = check_batch(b, MATS, x)
    g = ax.Graph.of_batch(b, dx, dy, repeats=5)
    ny = int(b.yoff[-1])
    for rep in range(4):
        dy.upload(np.full(ny, np.nan))  # eager work between replays
        g.launch()
        ax._lib.call("axspmv_device_sync")
        assert np.array_equal(dy.download(np.float64, ny), y_eager), rep
    t = b.time(dx, dy, iters=5, mode=ax.core.TIME_GRAPH, batch=4)
    assert t.mean_ms > 0
    assert np.array_equal(dy.download(np.float64, ny), y_eager)


def test_plan_graph_replays():
    m = ax.powerlaw(1 << 15, 2.0, 2048, 20.0, seed=3)
    x = ax.uniform(81, m.ncols)
    for fmt, kw in ((ax.FMT_AXT, dict(thw=4)), (ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=4)),
                    (ax.FMT_K1, dict(chunk=64)), (ax.FMT_RCT, dict(tile_rows=2048, slices=4)),
                    (ax.FMT_RCT, dict(tile_rows=1024, slices=1))):
        p = ax.Plan(m, fmt, **kw)
        dx = ax.DeviceBuffer.from_array(x)
        dy = ax.DeviceBuffer(8 * m.nrows)
        p.spmv(dx, dy)
        ax._lib.call("axspmv_device_sync")
        y_eager = dy.download(np.float64, m.nrows)
        assert rel_err(orc.ncsr(m, x), y_eager) <= TOL
        g = ax.Graph.of_plan(p, ax._lib.OP_SPMV, repeats=3, x=dx, y=dy)
        for rep in range(3):
            dy.upload(np.full(m.nrows, np.nan))
            g.launch()
            ax._lib.call("axspmv_device_sync")
            y = dy.download(np.float64, m.nrows)
            if fmt == ax.FMT_RCT:  # LDS and slice atomics: rounding order varies
                assert rel_err(y_eager, y) <= TOL, (fmt, rep)
            else:
                assert np.array_equal(y, y_eager), (fmt, rep)


@pytest.mark.parametrize("fmt,kw", [(ax.FMT_AXT, dict(thw=4)), (ax.FMT_RCT, dict(tile_rows=2048, slices=4)),
                                    (ax.FMT_RCT, dict(tile_rows=1024, slices=1))], ids=["axt", "rct_s4", "rct_s1"])
def test_plan_graph_spmm_and_fwd_adj(fmt, kw):
    """Graphs of SpMM and of the forward + adjoint pass (fused and split):
    every replay, with eager work in between, gives the eager result (RCT
    within the tolerance: unordered sums; v_bar exactly r[row] * x[col] of
    the r that replay wrote)."""
    m = ax.powerlaw(1 << 14, 2.0, 1024, 20.0, seed=7)
    n, nc, nnz = m.nrows, m.ncols, m.nnz
    p = ax.Plan(m, fmt, flags=ax.PLAN_TRANSPOSE, **kw)
    k = 16
    X = ax.uniform(89, nc * k).reshape(nc, k)
    dX, dY = ax.DeviceBuffer.from_array(X), ax.DeviceBuffer(8 * n * k)
    Y_ref = orc.spmm(m, X)
    g = ax.Graph.of_plan(p, ax._lib.OP_SPMM, repeats=2, k=k, X=dX, Y=dY)
    for rep in range(3):
        dY.upload(np.full(n * k, np.nan))
        g.launch()
        ax._lib.call("axspmv_device_sync")
        assert rel_err(Y_ref, dY.download(np.float64, n * k).reshape(n, k)) <= TOL, rep
    x, b = ax.uniform(90, nc), ax.uniform(91, n, stream=ax.VEC_STREAM + 1)
    dx, db = ax.DeviceBuffer.from_array(x), ax.DeviceBuffer.from_array(b)
    outs = dict(y=ax.DeviceBuffer(8 * n), r=ax.DeviceBuffer(8 * n), xbar=ax.DeviceBuffer(8 * nc),
                vbar=ax.DeviceBuffer(8 * nnz))
    r_ref = orc.ncsr(m, x) - b
    xbar_ref = orc.ncsr_t(m, r_ref, nc)
    for op in (ax._lib.OP_FWD_ADJ, ax._lib.OP_FWD_ADJ_SPLIT):
        g = ax.Graph.of_plan(p, op, repeats=2, x=dx, b=db, **outs)
        for rep in range(3):
            for name, size in (("y", n), ("r", n), ("xbar", nc), ("vbar", nnz)):
                outs[name].upload(np.full(size, np.nan))
            g.launch()
            ax._lib.call("axspmv_device_sync")
            r = outs["r"].download(np.float64, n)
            assert rel_err(r_ref, r) <= TOL, (op, rep)
            assert rel_err(xbar_ref, outs["xbar"].download(np.float64, nc)) <= 1e-11, (op, rep)
            assert np.array_equal(p.vbar_csr(outs["vbar"].download(np.float64, nnz)), orc.vbar(m, r, x)), (op, rep)
