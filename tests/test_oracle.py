"""Pins every function of the oracle (oracle/oracle.c).

  ncsr         bit-exact vs the reference's ncsr (golden fixtures), vs a
               pure-Python left-to-right sum, and closed forms on the 2D
               Laplacian (x = 1: 0/1/2 by boundary class; x_i = i: 0 inside)
  get_errors   hand cases (sign blindness of the reference, signed variant)
  row_lengths  vs the reference's rl[] (fixtures)
  k1/axc/axt   bit-exact vs the reference's builders (fixtures, 32- and
               64-lane parameters)
  spmm         = k independent ncsr calls;  ncsr_f32val = ncsr on rounded
               values;  ncsr_t = ncsr on the explicit transpose and the
               identity <Ax, w> = <x, A^T w>;  vbar = brute force
"""
import numpy as np
import pytest

from conftest import parse_axt_key, rel_err
from oracle import oracle as orc


def py_ncsr(row, col, val, x):
    y = []
    for i in range(row.size - 1):
        acc = 0.0
        for j in range(int(row[i]), int(row[i + 1])):
            acc = acc + float(val[j]) * float(x[col[j]])
        y.append(acc)
    return np.array(y)


def test_ncsr_matches_reference(fixture_matrix):
    f = fixture_matrix
    y = orc.ncsr(f, f.x)
    assert np.array_equal(y, f.z["ncsr.y"])
    assert np.array_equal(y, orc.ncsr(f, f.x, nthreads=4))


def test_ncsr_matches_python_loop(fixture_matrix):
    f = fixture_matrix
    assert np.array_equal(orc.ncsr(f, f.x), py_ncsr(f.row, f.col, f.val, f.x))


def test_row_lengths_match_reference(fixture_matrix):
    f = fixture_matrix
    assert np.array_equal(orc.row_lengths(f), f.z["ncsr.rl"])


def test_ncsr_closed_forms_laplace():
    import gpu_spmv_amd as ax
    m = ax.laplace2d(64)
    y = orc.ncsr(m, np.ones(m.nrows))
    # interior rows sum to 0, edge rows to 1, the 4 corners to 2
    assert sorted(np.unique(y)) == [0.0, 1.0, 2.0]
    assert (y == 0).sum() == 62 * 62 and (y == 2).sum() == 4 and (y == 1).sum() == 4 * 62
    x = np.arange(m.nrows, dtype=np.float64)
    y = orc.ncsr(m, x).reshape(64, 64)
    assert np.all(y[1:-1, 1:-1] == 0.0)
    # a dropped term or a wrong column index breaks these: e.g. without the
    # diagonal the interior would be -4*i
    assert y[0, 0] == 4 * 0 - 1 - 64


def test_get_errors_hand_cases():
    ref = np.array([1.0, -2.0, 4.0])
    res = np.array([1.0, 2.0, 4.0])  # sign flip in entry 1
    a, r, p = orc.get_errors(ref, res, sign_blind=True)
    assert a == 0.0 and r == 0.0  # the reference's metric cannot see it
    a, r, p = orc.get_errors(ref, res, sign_blind=False)
    assert a == 4.0 and r == 1.0 and p == 1
    a, r, p = orc.get_errors(ref, np.array([1.0, -2.0, 4.5]))
    assert a == 0.5 and r == 0.125 and p == 2


def test_k1_matches_reference(fixture_matrix):
    f = fixture_matrix
    for c in (32, 64):
        g = f.arrays(f"k1_c{c}")
        o = orc.k1(f, c, reorder=True)
        for k in ("permi", "nmc", "chp", "val", "col"):
            assert np.array_equal(o[k], g[k]), (c, k)
        assert o["lenVC"] == f.meta[f"k1_c{c}"]["lenVC"]


def test_axc_matches_reference(fixture_matrix):
    f = fixture_matrix
    for key in f.requests("axc_h"):
        g = f.arrays(key)
        o = orc.axc(f, f.x, f.meta[key]["hbs"])
        assert o["lenAX"] == f.meta[key]["lenAX"]
        for k in ("ax", "brp", "mapx"):
            assert np.array_equal(o[k], g[k]), (key, k)


def test_axt_matches_reference(fixture_matrix):
    f = fixture_matrix
    keys = f.requests("axt_")
    assert keys
    for key in keys:
        mode, bs, thw, th = parse_axt_key(key)
        g, meta = f.arrays(key), f.meta[key]
        o = orc.axt(f, f.x, bs, thw, th, mode)
        assert (o["tileN"], o["lenAX"], o["lenSEC"]) == (meta["tileN"], meta["lenAX"], meta["lenSEC"]), key
        if mode == 1:
            assert o["log"] == meta["log"], key
        for k in ("ax", "sec", "con"):
            assert np.array_equal(o[k], g[k]), (key, k)


def test_spmm_equals_columnwise_ncsr(fixture_matrix):
    f = fixture_matrix
    rng = np.random.default_rng(0)
    X = rng.uniform(-1, 1, (f.row.size - 1, 5))
    Y = orc.spmm(f, X, nthreads=2)
    for c in range(5):
        assert np.array_equal(Y[:, c], orc.ncsr(f, X[:, c]))


def test_ncsr_f32val(fixture_matrix):
    f = fixture_matrix

    class M:
        row, col = f.row, f.col
        val = f.val.astype(np.float32).astype(np.float64)

    assert np.array_equal(orc.ncsr_f32val(f, f.x), orc.ncsr(M, f.x))


def test_transpose_product_and_vbar(fixture_matrix):
    f = fixture_matrix
    m = f.csr()
    n = m.nrows
    rng = np.random.default_rng(1)
    w = rng.uniform(-1, 1, n)
    xt = orc.ncsr_t(f, w, n)
    assert np.array_equal(xt, orc.ncsr(m.transpose(), w))
    d = m.to_dense()
    assert rel_err(d.T @ w, xt) < 1e-13
    assert abs(np.dot(orc.ncsr(f, f.x), w) - np.dot(f.x, xt)) < 1e-10 * (1 + abs(np.dot(f.x, xt)))
    vb = orc.vbar(f, w, f.x)
    rows = m.row_ids()
    assert np.array_equal(vb, np.array([w[r] * f.x[c] for r, c in zip(rows, f.col)]))
