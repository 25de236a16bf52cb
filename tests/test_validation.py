"""CPU: every entry point that takes a host CSR checks it whole before it
builds or uploads anything (the kernels gather x[col] unchecked), so a bad
matrix is an error code with a message, not a GPU access fault.  Plans and
batches reject it before their first device call, which is why this runs
without a GPU."""
import numpy as np
import pytest

import gpu_spmv_amd as ax
from gpu_spmv_amd import _lib as L

GOOD = dict(row=[0, 2, 3, 5], col=[0, 2, 1, 0, 2], val=[1.0, 2.0, 3.0, 4.0, 5.0], ncols=3)
BAD = {
    "column >= ncols": (dict(GOOD, col=[0, 3, 1, 0, 2]), "ncols"),
    "unsorted row": (dict(GOOD, col=[2, 0, 1, 0, 2]), "increasing"),
    "repeated column": (dict(GOOD, col=[1, 1, 1, 0, 2]), "increasing"),
    "row offsets decrease": (dict(GOOD, row=[0, 2, 1, 5]), "decrease"),
    "row[0] != 0": (dict(GOOD, row=[1, 2, 3, 5]), "row\\[0\\]"),
}


def make(d):
    return ax.CsrMatrix(np.array(d["row"]), np.array(d["col"]), np.array(d["val"]), ncols=d["ncols"])


@pytest.mark.parametrize("case", sorted(BAD))
def test_bad_csr_is_rejected_everywhere(case):
    d, msg = BAD[case]
    m = make(d)
    with pytest.raises(ax.AxspmvError, match=msg):
        m.check()
    for fmt in (ax.FMT_CSR, ax.FMT_AXT, ax.FMT_AXC, ax.FMT_K1, ax.FMT_RCT):
        with pytest.raises(ax.AxspmvError, match=msg):
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
    with pytest.raises(ax.AxspmvError, match=msg):
        m.transpose()


def test_good_csr_passes_the_host_side():
    m = make(GOOD)
    m.check()
    assert ax.build_axt(m, None, 256, 4, 1, L.AXT_UNC, L.BUILD_PLANES)
    t = m.transpose()
    assert t.nrows == 3 and t.nnz == 5
