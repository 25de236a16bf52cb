"""GPU: the benchmark's rocSPARSE comparator (librocsparse_cmp.so, not part
of the library) computes the same product as the oracle with each of its CSR
algorithms, so the numbers it reports are for the right answer."""
import os

import numpy as np
import pytest

import bench
import gpu_spmv_amd as ax
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(not os.path.exists(bench.CMP_LIB), reason="librocsparse_cmp.so not built")
@pytest.mark.parametrize("gen", [lambda: ax.laplace2d(64), lambda: ax.powerlaw(1 << 14, 2.0, 1024, 20.0, seed=5),
                                 lambda: ax.random_sparse(5000, 0, 30, seed=6)], ids=["laplace64", "powerlaw16k", "random"])
def test_rocsparse_comparator_matches_oracle(gen):
    m = gen()
    x = ax.uniform(7, m.ncols)
    rows = bench.rocsparse_rows(m, ax.DeviceBuffer.from_array(x), orc.ncsr(m, x), iters=5)
    assert [r["label"] for r in rows] == ["rocsparse_" + a for a in bench.ROCSPARSE_ALGS]
    for r in rows:
        assert "error" not in r, r
        assert r["rel_err"] <= 1e-12 and r["us"] > 0, r
