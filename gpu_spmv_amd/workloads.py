"""The five benchmark configurations as concrete, seeded inputs.

Every matrix comes from the library's own generators (C, exported through
the ABI) and every vector from its xoshiro256** streams, so the benchmark,
the tests and a C caller see the same inputs bit for bit.  `digest` is the
fingerprint the CPU tests pin for the large matrices.

  config1  2D 5-point Laplacian 256^2 (n = 65,536, nnz = 326,656), diagonal
           4, off-diagonals -1; x ~ U[-1, 1], seed 0
  config2  power-law, n = 2^20, row lengths Zipf(2) on {1..2048} rescaled to
           mean 20 (nnz = 20,970,465), values N(0, 1); x ~ U[-1, 1], seed 1
  config3  27-point stencil on 128^3 (n = 2,097,152, nnz = 55,742,968),
           diagonal 26, off -1, symmetric random permutation (seed 2);
           x ~ U[-1, 1], seed 2
  config4a k right-hand sides X (ncols x k, interleaved) on the config-2
           matrix, X ~ U[-1, 1], seed 3
  config4b 256 banded matrices: n ~ U{1000..200000}, half-bandwidth ~
           U{1..64}, in-band fill 0.3 (the diagonal drawn like every other
           in-band position), values N(0, 1); sizes from numpy's PCG64 with
           seed 5, matrix i from seed 5000 + i; total nnz 457,872,800
  config5  the config-3 structure with values N(1, 0.1) (seed 4), x of
           config 3, b ~ U[-1, 1] (seed 4, the stream after the vectors')
"""
from __future__ import annotations

import hashlib

import numpy as np

from .core import VEC_STREAM, CsrMatrix, banded, fill_values_normal, laplace2d, powerlaw, stencil27, uniform

C4B_COUNT = 256


def config1():
    m = laplace2d(256)
    return m, uniform(0, m.ncols)


def config2():
    m = powerlaw(1 << 20, 2.0, 2048, 20.0, seed=1)
    return m, uniform(1, m.ncols)


def config3():
    m = stencil27(128, permute=True, seed=2)
    return m, uniform(2, m.ncols)


def config4a_rhs(ncols: int, k: int) -> np.ndarray:
    return uniform(3, ncols * k).reshape(ncols, k)


def config4b_shapes(count: int = C4B_COUNT):
    """(n, half-bandwidth, seed) of every member."""
    rng = np.random.default_rng(5)
    out = []
    for i in range(count):
        n = int(rng.integers(1000, 200001))
        hbw = int(rng.integers(1, 65))
        out.append((n, hbw, 5000 + i))
    return out


def config4b_matrix(n: int, hbw: int, seed: int) -> CsrMatrix:
    return banded(n, hbw, 0.3, seed=seed)


def config4b(count: int = C4B_COUNT):
    return [config4b_matrix(*s) for s in config4b_shapes(count)]


def config5(structure: CsrMatrix | None = None):
    """(A with values N(1, 0.1), x, b); `structure` reuses a config-3 matrix
    (its values are replaced in a copy)."""
    s = structure if structure is not None else stencil27(128, permute=True, seed=2)
    m = CsrMatrix(s.row, s.col, s.val.copy(), ncols=s.ncols)
    fill_values_normal(m, 1.0, 0.1, 4)
    return m, uniform(2, m.ncols), uniform(4, m.nrows, stream=VEC_STREAM + 1)


def digest(m: CsrMatrix) -> str:
    """First 16 hex digits of SHA-256 over row, col and val (native byte order)."""
    h = hashlib.sha256()
    for a in (m.row, m.col, m.val):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()[:16]
