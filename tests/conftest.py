import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
FIXTURES = ["toy12", "rand200", "lap16", "band300", "plaw512"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and runs HIP kernels")
    # A fresh checkout has no built libraries (they are not versioned): build
    # them once, here, before any test loads them.  A tree that has them
    # (built beforehand, as on the GPU machine) is used as it is.
    libs = [os.path.join(ROOT, "gpu_spmv_amd", "libaxspmv.so"), os.path.join(ROOT, "oracle", "liboracle.so")]
    if not all(os.path.exists(p) for p in libs):
        import subprocess
        subprocess.run(["make", "-j8", "-C", os.path.join(ROOT, "gpu_spmv_amd", "csrc")], check=True)
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True)


class FixtureMatrix:
    """A golden fixture: the matrix, x, and the reference's outputs."""

    def __init__(self, name):
        import json
        self.name = name
        self.z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
        self.row, self.col, self.val, self.x = (self.z[k] for k in ("row", "col", "val", "x"))
        self.meta = json.loads(str(self.z["meta"]))

    def arrays(self, key):
        pre = key + "."
        return {k[len(pre):]: self.z[k] for k in self.z.files if k.startswith(pre)}

    def requests(self, prefix):
        return sorted(k for k in self.meta if k.startswith(prefix))

    def csr(self):
        from gpu_spmv_amd import CsrMatrix
        return CsrMatrix(self.row, self.col, self.val, ncols=self.row.size - 1)


@pytest.fixture(scope="session", params=FIXTURES)
def fixture_matrix(request):
    return FixtureMatrix(request.param)


def parse_axt_key(key):
    """'axt_UNC_bs256_hw4_h1' -> (mode, bs, thw, th)."""
    _, mode, bs, hw, h = key.split("_")
    return (0 if mode == "UNC" else 1), int(bs[2:]), int(hw[2:]), int(h[1:])


def rel_err(ref, res):
    """Signed norm-wise relative error max|res-ref| / max|ref|."""
    ref = np.asarray(ref, np.float64)
    res = np.asarray(res, np.float64)
    den = np.max(np.abs(ref)) if ref.size else 0.0
    num = np.max(np.abs(res - ref)) if ref.size else 0.0
    return num / den if den > 0 else num
