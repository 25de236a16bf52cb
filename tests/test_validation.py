"""CPU: every entry point that takes a host CSR checks its offsets and
column bounds before it builds or uploads anything (the kernels gather
x[col] unchecked), so a bad matrix is an error code with a message, not a
GPU access fault.  Plans and batches reject it before their first device
call, which is why this runs without a GPU.  Column order is not required
(axspmv_csr_check alone insists on it): unsorted rows and repeated columns
are summed entry by entry, as by the reference's kernels."""
import numpy as np
import pytest

import gpu_spmv_amd as ax
from gpu_spmv_amd import _lib as L
from oracle import oracle as orc

GOOD = dict(row=[0, 2, 3, 5], col=[0, 2, 1, 0, 2], val=[1.0, 2.0, 3.0, 4.0, 5.0], ncols=3)
UNSAFE = {
    "column >= ncols": (dict(GOOD, col=[0, 3, 1, 0, 2]), "ncols"),
    "row offsets decrease": (dict(GOOD, row=[0, 2, 1, 5]), "decrease"),
    "row[0] != 0": (dict(GOOD, row=[1, 2, 3, 5]), "row\\[0\\]"),
    "row[nrows] != nnz": (dict(GOOD, row=[0, 2, 3, 4]), "nnz"),
}
UNORDERED = {
    "unsorted row": dict(GOOD, col=[2, 0, 1, 0, 2]),
    "repeated column": dict(GOOD, col=[1, 1, 1, 0, 2]),
}


def make(d):
    return ax.CsrMatrix(np.array(d["row"]), np.array(d["col"]), np.array(d["val"]), ncols=d["ncols"])


@pytest.mark.parametrize("case", sorted(UNSAFE))
def test_unsafe_csr_is_rejected_everywhere(case):
    d, msg = UNSAFE[case]
    m = make(d)
    with pytest.raises(ax.AxspmvError, match=msg):
        m.check()
    for fmt in (ax.FMT_CSR, ax.FMT_AXT, ax.FMT_AXC, ax.FMT_K1, ax.FMT_RCT):
        with pytest.raises(ax.AxspmvError, match="plan_create: .*" + msg):
            ax.Plan(m, fmt)
    for fmt in (ax.FMT_AXT, ax.FMT_RCT):
        with pytest.raises(ax.AxspmvError, match="matrix 1: .*" + msg):
            ax.Batch([make(GOOD), m], fmt=fmt)
    with pytest.raises(ax.AxspmvError, match="build_axt: .*" + msg):
        ax.build_axt(m, None, 256, 4, 1, L.AXT_UNC, L.BUILD_PLANES)
    with pytest.raises(ax.AxspmvError, match="build_axc: .*" + msg):
        ax.build_axc(m, None, 4, L.BUILD_PLANES)
    with pytest.raises(ax.AxspmvError, match="build_k1: .*" + msg):
        ax.build_k1(m, 32)
    with pytest.raises(ax.AxspmvError, match="csr_transpose: .*" + msg):
        m.transpose()


@pytest.mark.parametrize("case", sorted(UNORDERED))
def test_unordered_rows_are_accepted_by_the_builders(case):
    m = make(UNORDERED[case])
    with pytest.raises(ax.AxspmvError, match="increasing"):
        m.check()
    x = np.array([1.0, 10.0, 100.0])
    axt = ax.build_axt(m, x, 256, 4, 1, L.AXT_UNC)
    # ax holds every value and the gathered x of its column, padding zero
    assert axt["ax"].sum() == m.val.sum() + x[m.col].sum()
    t = m.transpose()
    w = np.array([1.0, -2.0, 3.0])
    assert t.nnz == m.nnz and np.allclose(orc.ncsr(t, w), orc.ncsr_t(m, w, m.ncols))


def test_good_csr_passes_the_host_side():
    m = make(GOOD)
    m.check()
    assert ax.build_axt(m, None, 256, 4, 1, L.AXT_UNC, L.BUILD_PLANES)
    t = m.transpose()
    assert t.nrows == 3 and t.nnz == 5
