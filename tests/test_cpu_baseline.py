"""CPU: the benchmark's CPU baseline -- the reference's own ncsr from
oracle/_ref where it is built (its output must equal the oracle's ncsr bit
for bit: same left-to-right row sums), else the oracle's restatement."""
import os

import numpy as np
import pytest

import bench
import gpu_spmv_amd as ax
from oracle import oracle as orc


@pytest.mark.skipif(not os.path.exists(bench.REF_BIN), reason="oracle/_ref not built (needs the reference sources)")
def test_reference_baseline_equals_oracle():
    m = ax.powerlaw(4096, 2.0, 256, 20.0, seed=3)
    x = ax.uniform(5, m.ncols)
    y, info = bench.cpu_baseline(m, x, 0.2)
    assert info["kind"] == "reference" and info["cores"] >= 1 and info["value"] > 0
    assert np.array_equal(y, orc.ncsr(m, x))


def test_port_baseline_when_reference_absent(monkeypatch):
    monkeypatch.setattr(bench, "REF_BIN", "/nonexistent/ref_dump_c64")
    m = ax.powerlaw(4096, 2.0, 256, 20.0, seed=3)
    x = ax.uniform(5, m.ncols)
    y, info = bench.cpu_baseline(m, x, 0.2)
    assert info["kind"] == "port" and info["value"] > 0
    assert np.array_equal(y, orc.ncsr(m, x))
