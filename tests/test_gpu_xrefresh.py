"""GPU: the x-refresh step of the interleaved layouts and the AX-layout plan.

The reference keeps x inside its AXC / AXT arrays (ax = tiles of hw values
then hw x-values; get_axAXC cudaSpmv.cu:952-979, getArraysAxSecAXT_*
cudaSpmv.cu:1239-1387), so a new x means refilling every x-half.
axspmv_x_refresh does that on the device from the column of each value slot
(AXC mapx, get_mapxAXC cudaSpmv.cu:983-1000, or an AXT column plane).

Bars: the refreshed array equals the host builder's ax bit for bit (the
builders are pinned to the reference's own builders in test_builders.py)
on every value slot and on the x-half of every real entry; a padding slot
gets x[map] = x[0] (map 0, as the reference's mapx) or 0.0 (map >= nx).
The AX-layout SpMV meets the FP64 parity bar of the other kernels.
"""
import numpy as np
import pytest

import gpu_spmv_amd as ax
from gpu_spmv_amd import _lib as L
from conftest import FIXTURES, FixtureMatrix, rel_err
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
TOL = 1e-12
PAD = 0xFFFFFFFF


def mats():
    out = {name: FixtureMatrix(name).csr() for name in FIXTURES}
    out["powerlaw16k"] = ax.powerlaw(1 << 14, 2.0, 512, 12.0, seed=3)
    out["banded_empty"] = ax.banded(3000, 2, 0.3, seed=6)
    return out


MATS = mats()


def split(arr, hw):
    """(value halves, x halves) of an interleaved array, each in slot order."""
    g = arr.reshape(-1, 2, hw)
    return g[:, 0, :].reshape(-1), g[:, 1, :].reshape(-1)


def refresh(map_, hw, nx, x, ax_in):
    dmap = ax.DeviceBuffer.from_array(np.ascontiguousarray(map_, np.uint32))
    dx = ax.DeviceBuffer.from_array(x)
    dax = ax.DeviceBuffer.from_array(ax_in)
    ax.x_refresh(dmap, map_.size, hw, nx, dx, dax)
    L.call("axspmv_device_sync")
    return dax.download(np.float64, ax_in.size)


def zero_x_halves(a, hw):
    g = a.copy().reshape(-1, 2, hw)
    g[:, 1, :] = 0.0
    return g.reshape(-1)


@pytest.mark.parametrize("mname", list(MATS))
@pytest.mark.parametrize("thw,th,mode", [(4, 1, ax.AXT_UNC), (16, 1, ax.AXT_UNC), (2, 1, ax.AXT_UNC),
                                         (32, 4, ax.AXT_UNC), (64, 8, ax.AXT_UNC), (32, 1, ax.AXT_COM),
                                         (64, 4, ax.AXT_COM)])
def test_refresh_axt_equals_builder(mname, thw, th, mode):
    m = MATS[mname]
    x = ax.uniform(21, m.ncols)
    f = ax.build_axt(m, x, 256, thw, th, mode)
    ref = f["ax"]
    real = f["vplane"] != 0.0  # values are N(0,1) draws: exact zeros are padding
    # padding marked: the refreshed array is the builder's, bit for bit
    marked = np.where(real, f["cplane"], PAD).astype(np.uint32)
    out = refresh(marked, thw, m.ncols, x, zero_x_halves(ref, thw))
    assert np.array_equal(out, ref)
    # the builder's own column plane (padding column 0): padding gets x[0]
    out0 = refresh(f["cplane"], thw, m.ncols, x, zero_x_halves(ref, thw))
    v0, x0 = split(out0, thw)
    vr, xr = split(ref, thw)
    assert np.array_equal(v0, vr)
    assert np.array_equal(x0[real], xr[real])
    assert np.all(x0[~real] == x[0])


@pytest.mark.parametrize("mname", list(MATS))
@pytest.mark.parametrize("hbs", [4, 32, 64])
def test_refresh_axc_from_mapx(mname, hbs):
    m = MATS[mname]
    x = ax.uniform(22, m.ncols)
    f = ax.build_axc(m, x, hbs)
    ref = f["ax"]
    nslots = f["lenAX"] // 2
    out = refresh(f["mapx"][:nslots], hbs, m.ncols, x, zero_x_halves(ref, hbs))
    v, xh = split(out, hbs)
    vr, xr = split(ref, hbs)
    real = f["vplane"] != 0.0
    assert np.array_equal(v, vr)
    assert np.array_equal(xh[real], xr[real])
    assert np.all(xh[~real] == x[0])


def test_refresh_rejects_bad_width():
    d = ax.DeviceBuffer(64)
    with pytest.raises(ax.AxspmvError):
        ax.x_refresh(d, 4, 3, 1, d, d)


AX_PLANS = [dict(thw=4), dict(thw=8), dict(thw=16), dict(thw=64), dict(thw=2), dict(thw=4, budget=64),
            dict(fmt=ax.FMT_AXC, hbs=4), dict(fmt=ax.FMT_AXC, hbs=32)]


@pytest.mark.parametrize("mname", list(MATS) + ["long_rows"])
@pytest.mark.parametrize("kw", AX_PLANS, ids=[str(k) for k in AX_PLANS])
def test_ax_layout_spmv(mname, kw):
    if mname == "long_rows":
        from test_gpu_spmv import long_rows
        m = long_rows()
    else:
        m = MATS[mname]
    kw = dict(kw)
    p = ax.Plan(m, kw.pop("fmt", ax.FMT_AXT), flags=ax.PLAN_AX_LAYOUT, **kw)
    info = p.info()
    assert info["name"].endswith("_ax") and info["launches"] == 2
    dy = ax.DeviceBuffer(8 * max(m.nrows, 1))
    for seed in (31, 32):  # a new x each call: the refresh must replace the old one
        x = ax.uniform(seed, m.ncols)
        dx = ax.DeviceBuffer.from_array(x)
        p.spmv(dx, dy)
        L.call("axspmv_device_sync")
        y = dy.download(np.float64, m.nrows)
        assert rel_err(orc.ncsr(m, x), y) <= TOL, (mname, kw, seed)
    t = p.time_op(L.OP_XREFRESH, iters=5, warmup=1, x=dx)
    assert t.mean_ms > 0


def test_ax_layout_unsupported():
    m = MATS["powerlaw16k"]
    for kw in (dict(thw=64, th=4), dict(thw=4, vtype=ax.F32), dict(mode=ax.AXT_COM, thw=64, th=1)):
        with pytest.raises(ax.AxspmvError):
            ax.Plan(m, ax.FMT_AXT, flags=ax.PLAN_AX_LAYOUT, **kw)
    p = ax.Plan(m, ax.FMT_AXT, thw=4)
    dx = ax.DeviceBuffer(8 * m.ncols)
    with pytest.raises(ax.AxspmvError):
        p.time_op(L.OP_XREFRESH, iters=1, warmup=0, x=dx)
