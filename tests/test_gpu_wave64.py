"""GPU: the wave64 cross-lane primitives behind every segmented kernel.

They replace the reference's 32-lane __shfl_down trees (cudaSpmv.cu:1086-1090,
1586-1594) and Hillis-Steele __shfl_up scans (cudaSpmv.cu:2061-2086,
scan.cu:239-414) with DPP and ds_swizzle on 64 lanes.  axspmv_wave_selftest
runs one primitive lane by lane; each result is compared with the
primitive's definition restated below as plain loops over the 64 lanes of
each wave.  Values are integers of magnitude <= 1000, so every sum is exact
whatever the association order and the bar is bit-exact: a lane paired
with the wrong partner, a segment that leaks across a head, or a shift by
the wrong distance changes the output.
"""
import numpy as np
import pytest

import gpu_spmv_amd as ax
from gpu_spmv_amd import _lib as L

pytestmark = pytest.mark.gpu

SEG_SCAN, SCAN_MAX, SHIFT_UP, SHIFT_DOWN, GROUP_SUM_1 = 0, 1, 2, 3, 4
FILL_U32 = 0xFFFFFFFF


def head_patterns(rng):
    """Rows of 64 segment-head words: random, none, all, lane 0 only,
    lane 63 only, alternating, one head per 16-lane row boundary and one
    just after it (the DPP row boundaries), and arbitrary nonzero words."""
    rows = [rng.random(64) < p for p in (0.05, 0.2, 0.5, 0.8) for _ in range(4)]
    rows.append(np.zeros(64, bool))
    rows.append(np.ones(64, bool))
    for lanes in ([0], [63], list(range(0, 64, 2)), [16, 32, 48], [15, 31, 47], [17, 33, 49], [31], [32]):
        r = np.zeros(64, bool)
        r[lanes] = True
        rows.append(r)
    h = np.concatenate(rows).astype(np.uint32)
    words = rng.integers(1, 1 << 32, h.size, dtype=np.uint64).astype(np.uint32)
    return np.where(h != 0, words, 0).astype(np.uint32)


def run(op, v, h):
    n = v.size
    dv, dh = ax.DeviceBuffer.from_array(v), ax.DeviceBuffer.from_array(h)
    ov, oh = ax.DeviceBuffer(8 * n), ax.DeviceBuffer(4 * n)
    L.call("axspmv_wave_selftest", op, dv.ptr, dh.ptr, ov.ptr, oh.ptr, n, None)
    L.call("axspmv_device_sync")
    return ov.download(np.float64, n), oh.download(np.uint32, n)


@pytest.fixture(scope="module")
def data():
    rng = np.random.default_rng(64)
    h = head_patterns(rng)
    v = rng.integers(-1000, 1001, h.size).astype(np.float64)
    return v, h


def test_seg_scan(data):
    v, h = data
    ov, oh = run(SEG_SCAN, v, h)
    want_v, want_h = np.empty_like(v), np.zeros(v.size, bool)
    for w in range(0, v.size, 64):
        acc, seen = 0.0, False
        for i in range(w, w + 64):
            if h[i]:
                acc, seen = 0.0, True
            acc += v[i]
            want_v[i], want_h[i] = acc, seen
    np.testing.assert_array_equal(ov, want_v)
    np.testing.assert_array_equal(oh != 0, want_h)


def test_scan_max(data):
    v, _ = data
    rng = np.random.default_rng(7)
    h = rng.integers(0, 1 << 32, v.size, dtype=np.uint64).astype(np.uint32)
    h[::5] = 0
    ov, oh = run(SCAN_MAX, v, h)
    want = np.empty_like(h)
    for w in range(0, h.size, 64):
        m = 0
        for i in range(w, w + 64):
            m = max(m, int(h[i]))
            want[i] = m
    np.testing.assert_array_equal(oh, want)
    np.testing.assert_array_equal(ov, v)


def test_shift_up(data):
    v, h = data
    ov, oh = run(SHIFT_UP, v, h)
    for w in range(0, v.size, 64):
        assert ov[w] == -1.0 and oh[w] == FILL_U32
        for i in range(w + 1, w + 64):
            assert ov[i] == v[i - 1] and oh[i] == h[i - 1], i


def test_shift_down(data):
    v, h = data
    ov, oh = run(SHIFT_DOWN, v, h)
    for w in range(0, v.size, 64):
        for i in range(w, w + 63):
            assert oh[i] == h[i + 1], i
        assert oh[w + 63] == FILL_U32
    np.testing.assert_array_equal(ov, v)


@pytest.mark.parametrize("log_g", range(7))
def test_group_sum(data, log_g):
    v, h = data
    g = 1 << log_g
    ov, oh = run(GROUP_SUM_1 + log_g, v, h)
    want = np.empty_like(v)
    for s in range(0, v.size, g):
        want[s:s + g] = sum(v[s:s + g])
    np.testing.assert_array_equal(ov, want)
    np.testing.assert_array_equal(oh, h)


def test_group_sum_beyond_one_grid():
    """More elements than one launch grid covers (the grid is capped at
    16,384 workgroups of 256 lanes): the kernel strides over whole waves."""
    rng = np.random.default_rng(65)
    n = 64 * 70000  # 4,480,000 > 16,384 * 256
    v = rng.integers(-1000, 1001, n).astype(np.float64)
    h = np.zeros(n, np.uint32)
    ov, oh = run(GROUP_SUM_1 + 6, v, h)
    np.testing.assert_array_equal(ov, np.repeat(v.reshape(-1, 64).sum(1), 64))
    ov, _ = run(SHIFT_UP, v, h)
    want = np.concatenate([np.full((n // 64, 1), -1.0), v.reshape(-1, 64)[:, :-1]], axis=1).ravel()
    np.testing.assert_array_equal(ov, want)
