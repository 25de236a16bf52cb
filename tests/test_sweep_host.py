"""Sweep driver: matrix specs and the tuning database (host side only)."""
import json

import numpy as np

import gpu_spmv_amd as ax
from gpu_spmv_amd import sweep


def test_load_matrix_generators_and_files(tmp_path):
    name, m = sweep.load_matrix("laplace2d:16")
    assert name == "laplace2d_16" and m.nrows == 256 and m.nnz == 5 * 256 - 4 * 16
    p = tmp_path / "M99_toy.bin"
    m.write(str(p))
    name2, m2 = sweep.load_matrix(str(p))
    assert name2 == "M99_toy" and np.array_equal(m2.col, m.col) and np.array_equal(m2.val, m.val)


def test_best_opts_reads_database(tmp_path):
    db = dict(matrices=dict(a=dict(best=dict(label="rct_auto", format=ax.FMT_RCT, opts=dict(block=512), us=1.0)),
                            b=dict(best=None)))
    path = tmp_path / "db.json"
    path.write_text(json.dumps(db))
    assert sweep.best_opts(str(path), "a") == (ax.FMT_RCT, dict(block=512))
    assert sweep.best_opts(db, "b") is None and sweep.best_opts(db, "missing") is None


def test_candidates_cover_every_format():
    fmts = {c[1] for c in sweep.CANDIDATES}
    assert fmts == {ax.FMT_CSR, ax.FMT_K1, ax.FMT_AXC, ax.FMT_AXT, ax.FMT_RCT}
    assert sweep.BLOCKS == (32, 64, 128, 256, 512, 1024)
