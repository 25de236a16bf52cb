"""The seeded RNG and the synthetic benchmark matrices: published RNG values,
an independent Python restatement of xoshiro256**, closed-form sizes."""
import hashlib

import numpy as np
import pytest

import gpu_spmv_amd as ax
from gpu_spmv_amd import workloads as W
from oracle import oracle as orc

M64 = (1 << 64) - 1


def py_splitmix(s):
    s = (s + 0x9E3779B97F4A7C15) & M64
    z = s
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return s, z ^ (z >> 31)


def py_xoshiro(seed, stream, n):
    sm = seed ^ ((stream * 0x9E3779B97F4A7C15) & M64)
    st = []
    for _ in range(4):
        sm, v = py_splitmix(sm)
        st.append(v)
    rotl = lambda x, k: ((x << k) | (x >> (64 - k))) & M64
    out = []
    for _ in range(n):
        r = (rotl((st[1] * 5) & M64, 7) * 9) & M64
        t = (st[1] << 17) & M64
        st[2] ^= st[0]
        st[3] ^= st[1]
        st[1] ^= st[2]
        st[0] ^= st[3]
        st[2] ^= t
        st[3] = rotl(st[3], 45)
        out.append(r)
    return out


def test_splitmix_published_value():
    # splitmix64 seeded with 0: first output 0xE220A8397B1DCDAF (Vigna)
    assert py_splitmix(0)[1] == 0xE220A8397B1DCDAF


@pytest.mark.parametrize("seed,stream", [(0, 0), (1, 5), (12345, 1 << 40)])
def test_xoshiro_matches_restatement(seed, stream):
    assert ax.rng_u64(seed, stream, 50).tolist() == py_xoshiro(seed, stream, 50)


def test_uniform_and_normal():
    u = ax.uniform(3, 200000)
    assert u.min() >= -1 and u.max() < 1 and abs(u.mean()) < 0.01
    raw = ax.rng_u64(3, ax.VEC_STREAM, 4)
    assert np.allclose(ax.uniform(3, 4), 2 * (raw >> np.uint64(11)).astype(np.float64) * 2.0 ** -53 - 1, rtol=0, atol=0)
    g = ax.normal(4, 200000)
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1) < 0.01


@pytest.mark.parametrize("grid", [1, 2, 5, 256])
def test_laplace2d_closed_form(grid):
    m = ax.laplace2d(grid)
    assert m.nrows == grid * grid and m.nnz == 5 * grid * grid - 4 * grid
    m.check()
    rl = np.diff(m.row)
    if grid >= 2:
        assert (rl == 5).sum() == (grid - 2) ** 2 and (rl == 3).sum() == 4
    if grid == 256:
        assert m.nnz == 326656


@pytest.mark.parametrize("grid,permute", [(3, False), (6, True), (12, True)])
def test_stencil27_closed_form(grid, permute):
    m = ax.stencil27(grid, permute=permute, seed=2)
    assert m.nrows == grid ** 3 and m.nnz == (3 * grid - 2) ** 3
    m.check()
    y = orc.ncsr(m, np.ones(m.nrows))
    assert np.array_equal(y, 27.0 - np.diff(m.row))  # diag 26, off -1: y = 27 - rl
    if permute:
        base = ax.stencil27(grid, permute=False)
        # A' = P A P^T: same multiset of (row length, value) and symmetric structure
        d = m.to_dense()
        assert np.array_equal(d, d.T)
        assert sorted(np.diff(m.row)) == sorted(np.diff(base.row))
        # the permutation is a relabelling: A' and A have the same spectrum
        assert np.allclose(np.sort(np.linalg.eigvalsh(d)), np.sort(np.linalg.eigvalsh(base.to_dense())))


# Fingerprints of the benchmark inputs: a change to a generator, the RNG or
# the seeds changes the benchmark and fails here first.  nnz of configs 1
# and 3 are closed forms of the grid (5 g^2 - 4 g, (3 g - 2)^3); config 2's
# count and every digest are those of the inputs the benchmark ran on.
def test_config1_laplace_pinned():
    m, x = W.config1()
    assert (m.nrows, m.nnz) == (65536, 5 * 256 ** 2 - 4 * 256) == (65536, 326656)
    assert W.digest(m) == "25f9449b4f39ef36"


def test_config2_powerlaw_pinned():
    m, x = W.config2()
    assert (m.nrows, m.nnz) == (1 << 20, 20970465)
    assert W.digest(m) == "dc05322526159495"
    assert hashlib.sha256(x.tobytes()).hexdigest()[:16] == "a38ec121af095151"


def test_config3_stencil_pinned():
    m, _ = W.config3()
    assert (m.nrows, m.nnz) == (128 ** 3, (3 * 128 - 2) ** 3) == (2097152, 55742968)
    assert W.digest(m) == "5fad6478cc6e7f4f"


def test_config4b_batch_pinned():
    """256 banded members, generated one at a time: the total nonzero count
    and a digest of the members' digests."""
    shapes = W.config4b_shapes()
    assert len(shapes) == 256
    assert all(1000 <= n <= 200000 and 1 <= h <= 64 for n, h, _ in shapes)
    tot, h = 0, hashlib.sha256()
    for s in shapes:
        m = W.config4b_matrix(*s)
        tot += m.nnz
        h.update(W.digest(m).encode())
    assert tot == 457872800
    assert h.hexdigest()[:16] == "66096c682f8cd455"


def test_powerlaw_properties():
    m = ax.powerlaw(1 << 16, 2.0, 2048, 20.0, seed=1)
    m.check()
    rl = np.diff(m.row)
    assert rl.min() >= 1 and rl.max() <= 2048
    assert abs(rl.mean() - 20.0) < 0.5
    # Zipf(2) head: about 61% of rows come from z = 1
    assert 0.55 < np.mean(rl == rl.min()) < 0.67
    again = ax.powerlaw(1 << 16, 2.0, 2048, 20.0, seed=1)
    assert np.array_equal(m.col, again.col) and np.array_equal(m.val, again.val)


def test_banded_and_random():
    m = ax.banded(5000, 7, 0.3, seed=3)
    m.check()
    rows = m.row_ids().astype(np.int64)
    assert np.all(np.abs(rows - m.col.astype(np.int64)) <= 7)
    assert abs(m.nnz / (5000 * 15) - 0.3) < 0.02
    assert (np.diff(m.row) == 0).any()  # empty rows occur
    r = ax.random_sparse(4096, 5, 15, seed=8)
    r.check()
    assert np.diff(r.row).min() >= 5 and np.diff(r.row).max() <= 15


def test_io_roundtrip(tmp_path):
    m = ax.random_sparse(300, 0, 20, seed=2)
    for ext in (".bin", ".csr"):
        p = str(tmp_path / ("m" + ext))
        m.write(p)
        r = ax.CsrMatrix.read(p)
        assert np.array_equal(r.row, m.row) and np.array_equal(r.col, m.col) and np.array_equal(r.val, m.val)
    with pytest.raises(ax.AxspmvError):
        ax.CsrMatrix.read(str(tmp_path / "m.txt"))
    with pytest.raises(ax.AxspmvError):
        ax.CsrMatrix.read(str(tmp_path / "missing.bin"))


@pytest.mark.parametrize("name", ["varied", "writer"])
def test_csr_text_reader_matches_reference(name, tmp_path):
    """The .csr text reader against the reference's matrixReading
    (cudaSpmv.cu:268-293, fscanf "%lf %d"): the arrays it parsed from a file
    with varied number formats and whitespace (denormals, -0.0, 17- and
    21-digit mantissas, tabs, blank lines), and from a file the library's
    writer produced, are the library's parse bit for bit."""
    import os
    from conftest import GOLDEN
    z = np.load(os.path.join(GOLDEN, "csrtext.npz"), allow_pickle=False)
    p = str(tmp_path / "m.csr")
    with open(p, "w") as f:
        f.write(str(z[f"{name}.text"]))
    m = ax.CsrMatrix.read(p)
    np.testing.assert_array_equal(m.row, z[f"{name}.row"])
    np.testing.assert_array_equal(m.col, z[f"{name}.col"])
    assert m.val.tobytes() == z[f"{name}.val"].tobytes()  # -0.0 and denormals included
    if name == "writer":
        # the writer's text is what the reference parsed back to the original arrays
        w = ax.random_sparse(40, 0, 8, seed=21)
        q = str(tmp_path / "w.csr")
        w.write(q)
        assert open(q).read() == str(z["writer.text"])
        assert w.val.tobytes() == z["writer.val"].tobytes()
        np.testing.assert_array_equal(w.col, z["writer.col"])


def test_stats_and_sta_match_reference(fixture_matrix, tmp_path):
    import gpu_spmv_amd._lib as L
    f = fixture_matrix
    m = f.csr()
    s = m.stats()
    meta = f.meta["ncsr"]
    assert (s["rmin"], s["rmax"], s["bw"]) == (meta["rmin"], meta["rmax"], meta["bw"])
    assert s["rave"] == pytest.approx(meta["rave"], rel=1e-15) and s["rsd"] == pytest.approx(meta["rsd"], rel=1e-12)
    p = str(tmp_path / "m.sta")
    ref_text = str(f.z["ncsr.sta"])
    name = ref_text.splitlines()[3].split()[1]
    L.call("axspmv_csr_write_sta", m.ref, name.encode(), p.encode())
    assert open(p).read() == ref_text
