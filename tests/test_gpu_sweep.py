"""GPU: the sweep driver over formats x block sizes, checked against the oracle."""
import json

import pytest

import gpu_spmv_amd as ax
from gpu_spmv_amd import sweep
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("spec", ["laplace2d:64", "powerlaw:20000", "banded:5000"])
def test_sweep_matrix_rows(spec):
    name, m = sweep.load_matrix(spec)
    r = sweep.sweep_matrix(name, m, blocks=(64, 256, 1024), iters=5, warmup=1, check=lambda a, x: orc.ncsr(a, x))
    ok = [row for row in r["rows"] if "us" in row]
    # every candidate ran; block-sensitive ones once per block size
    n_expected = sum(3 if c[3] else 1 for c in sweep.CANDIDATES)
    assert len(ok) == n_expected, [row for row in r["rows"] if "error" in row]
    assert max(row["rel_err"] for row in ok) <= 1e-12
    assert r["best"]["us"] == min(row["us"] for row in ok)


def test_sweep_cli_writes_database(tmp_path):
    m = ax.random_sparse(3000, 1, 40, seed=3)
    f = tmp_path / "M98_rand.bin"
    m.write(str(f))
    out = tmp_path / "db.json"
    assert sweep.main([str(f), "--generate", "laplace2d:32", "--blocks", "128,512", "--iters", "3",
                       "--out", str(out), "--sta"]) == 0
    db = json.loads(out.read_text())
    assert set(db["matrices"]) == {"M98_rand", "laplace2d_32"}
    assert (tmp_path / "M98_rand.sta").exists()
    fmt, opts = sweep.best_opts(str(out), "M98_rand")
    p = ax.Plan(m, fmt, **opts)  # the recorded choice builds a plan
    x = ax.uniform(5, m.ncols)
    assert float(abs(ax.spmv_host(p, x) - orc.ncsr(m, x)).max()) <= 1e-12 * float(abs(orc.ncsr(m, x)).max())
