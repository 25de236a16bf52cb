"""GPU: the benchmark's rocSPARSE comparator (librocsparse_cmp.so, not part
of the library) computes the same product as the oracle with each of its CSR
algorithms, so the numbers it reports are for the right answer.

It runs in a child process, as in the benchmark (which never imports
torch): the torch wheel ships its own rocSPARSE build, and in a process that
has imported torch (the autograd tests do) librocsparse.so.1 can resolve to
that copy instead of the ROCm one the comparator was built against."""
import json
import os
import subprocess
import sys

import pytest

import bench
from conftest import ROOT

pytestmark = pytest.mark.gpu

CHILD = r"""
import json, sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import bench
import gpu_spmv_amd as ax
from oracle import oracle as orc
gens = {"laplace64": lambda: ax.laplace2d(64), "powerlaw16k": lambda: ax.powerlaw(1 << 14, 2.0, 1024, 20.0, seed=5),
        "random": lambda: ax.random_sparse(5000, 0, 30, seed=6)}
out = {}
for name, gen in gens.items():
    m = gen()
    x = ax.uniform(7, m.ncols)
    out[name] = bench.rocsparse_rows(m, ax.DeviceBuffer.from_array(x), orc.ncsr(m, x), iters=5)
print(json.dumps(out))
"""


@pytest.mark.skipif(not os.path.exists(bench.CMP_LIB), reason="librocsparse_cmp.so not built")
def test_rocsparse_comparator_matches_oracle():
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert sorted(out) == ["laplace64", "powerlaw16k", "random"]
    for name, rows in out.items():
        assert [x["label"] for x in rows] == ["rocsparse_" + a for a in bench.ROCSPARSE_ALGS], name
        for x in rows:
            assert "error" not in x, (name, x)
            assert x["rel_err"] <= 1e-12 and x["us"] > 0, (name, x)
