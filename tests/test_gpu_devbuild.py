"""GPU: format builders on the device against the host builders.

axspmv_dev_build_axt / axspmv_dev_build_axc build the reference's AXT and
AXC arrays (getFormatDataAXT cudaSpmv.cu:1391-1566, getFormatDataAXC
cudaSpmv.cu:1004-1062) with kernels.  Bar: every array equal, bit for bit,
to the host builder's (which test_builders.py pins to the reference's own
builders compiled from its sources, and to the oracle's restatement).
"""
import numpy as np
import pytest

import gpu_spmv_amd as ax
from conftest import FIXTURES, FixtureMatrix

pytestmark = pytest.mark.gpu


def mats():
    out = {name: FixtureMatrix(name).csr() for name in FIXTURES}
    out["powerlaw64k"] = ax.powerlaw(1 << 16, 2.0, 2048, 20.0, seed=1)
    out["stencil12p"] = ax.stencil27(12, permute=True, seed=2)
    out["banded_empty"] = ax.banded(3000, 2, 0.3, seed=6)
    out["empty"] = ax.CsrMatrix(np.zeros(6, np.uint32), np.zeros(0, np.uint32), np.zeros(0), ncols=5)
    return out


MATS = mats()
AXT = [(256, 4, 1, ax.AXT_UNC), (256, 16, 1, ax.AXT_UNC), (256, 32, 1, ax.AXT_UNC), (256, 3, 1, ax.AXT_UNC),
       (256, 32, 4, ax.AXT_UNC), (256, 64, 8, ax.AXT_UNC), (256, 32, 6, ax.AXT_UNC),
       (256, 64, 1, ax.AXT_COM), (64, 32, 1, ax.AXT_COM), (2, 4, 1, ax.AXT_COM),
       (256, 64, 4, ax.AXT_COM), (256, 32, 16, ax.AXT_COM)]


@pytest.mark.parametrize("mname", list(MATS))
@pytest.mark.parametrize("bs,thw,th,mode", AXT, ids=[f"bs{a}_thw{b}_th{c}_m{d}" for a, b, c, d in AXT])
def test_dev_build_axt_equals_host(mname, bs, thw, th, mode):
    m = MATS[mname]
    x = ax.uniform(41, max(m.ncols, 1))[:m.ncols]
    h = ax.build_axt(m, x, bs, thw, th, mode)
    d = ax.dev_build_axt(m, x, bs, thw, th, mode)
    assert d["name"] == h["name"]
    for k in ("tileN", "lenAX", "lenSEC", "lenCON", "log"):
        assert d[k] == h[k], k
    assert np.array_equal(d["ax"], h["ax"])
    assert np.array_equal(d["sec"], h["sec"])
    assert np.array_equal(d["con"], h["con"])
    assert np.array_equal(d["plane"], h["cplane"])
    assert d["ms"] > 0 or m.nrows == 0


@pytest.mark.parametrize("mname", list(MATS))
@pytest.mark.parametrize("hbs", [1, 4, 32, 64, 24])
def test_dev_build_axc_equals_host(mname, hbs):
    m = MATS[mname]
    x = ax.uniform(42, max(m.ncols, 1))[:m.ncols]
    h = ax.build_axc(m, x, hbs)
    d = ax.dev_build_axc(m, x, hbs)
    assert d["name"] == "faxc" and d["lenAX"] == h["lenAX"] and d["lenMAPX"] == h["lenMAPX"]
    assert np.array_equal(d["ax"], h["ax"])
    assert np.array_equal(d["con"], h["brp"])
    assert np.array_equal(d["plane"], h["mapx"])


def test_dev_build_without_x_leaves_x_halves_zero():
    m = MATS["plaw512"]
    x = ax.uniform(43, m.ncols)
    h = ax.build_axt(m, x, 256, 8, 1, ax.AXT_UNC)
    d = ax.dev_build_axt(m, None, 256, 8, 1, ax.AXT_UNC)
    g, gh = d["ax"].reshape(-1, 2, 8), h["ax"].reshape(-1, 2, 8)
    assert np.array_equal(g[:, 0], gh[:, 0]) and not np.any(g[:, 1])


def test_dev_build_rejects_what_the_host_rejects():
    m = MATS["toy12"]
    for args in ((256, 0, 1, ax.AXT_UNC), (256, 4, 0, ax.AXT_UNC), (3, 4, 1, ax.AXT_COM), (256, 4, 3, ax.AXT_COM),
                 (256, 4, 1, 7)):
        with pytest.raises(ax.AxspmvError):
            ax.build_axt(m, None, *args, flags=ax.BUILD_PLANES)
        with pytest.raises(ax.AxspmvError):
            ax.dev_build_axt(m, None, *args)
    with pytest.raises(ax.AxspmvError):
        ax.dev_build_axc(m, None, 0)


@pytest.mark.parametrize("case,row,col,ncols,msg", [
    ("column >= ncols", [0, 2, 3, 5], [0, 2, 1, 0, 3], 3, "column >= ncols"),
    ("offsets decrease", [0, 2, 1, 5], [0, 2, 1, 0, 2], 3, "row offsets decrease"),
    ("row[nrows] != nnz", [0, 2, 3, 4], [0, 2, 1, 0, 2], 3, "row\\[nrows\\] != nnz"),
    ("row[0] != 0", [1, 2, 3, 5], [0, 2, 1, 0, 2], 3, "row\\[0\\] != 0"),
])
def test_dev_build_checks_the_device_csr(case, row, col, ncols, msg):
    """A device CSR is checked on the device (offsets, column bounds) before
    the fill reads through its offsets: an error, nothing is built."""
    m = ax.CsrMatrix(np.array(row), np.array(col), np.ones(len(col)), ncols=ncols)
    with pytest.raises(ax.AxspmvError, match="invalid device CSR:.*" + msg):
        ax.dev_build_axt(m, None, 256, 4, 1, ax.AXT_UNC)
    with pytest.raises(ax.AxspmvError, match="invalid device CSR:.*" + msg):
        ax.dev_build_axt(m, None, 256, 64, 1, ax.AXT_COM)
    with pytest.raises(ax.AxspmvError, match="invalid device CSR:.*" + msg):
        ax.dev_build_axc(m, None, 4)
