"""GPU: h2w_dense_matvec_noalloc (the dense product with device pointers and no allocation) against
h2w_dense_matvec on the same inputs, through the C ABI: two calls in a row, the second with a smaller
input (another handle of fewer points, fewer rows), bit for bit; and its argument checks."""
import ctypes

import numpy as np
import pytest

import h2weak_amd as h2w

pytestmark = pytest.mark.gpu


class _Dev:
    """Device buffers through the HIP runtime (test plumbing)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.bufs = []

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(arr.nbytes, 8))) == 0
        self.bufs.append(p)
        assert self.hip.hipMemcpy(p, arr.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(arr.nbytes), 1) == 0
        return p

    def get(self, p, n, dtype=np.float64):
        out = np.empty(n, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def _handle(N, seed):
    P = np.random.default_rng(seed).uniform(-1, 1, (3, N))
    return h2w.Initialisation(3, N, 64, 1.0, 10, h2w.CHOICE_RBF, h2w.NESTED, points=P)


def _both(dev, g, x, rows):
    """(no-allocation result, h2w_dense_matvec result) for x and rows (None: every row)."""
    L = h2w.lib()
    n = g.N if rows is None else len(rows)
    dx, dy, dy2 = dev.put(x), dev.put(np.full(n, np.nan)), dev.put(np.full(n, np.nan))
    drows = None if rows is None else dev.put(np.asarray(rows, np.int64))
    rc = L.h2w_dense_matvec_noalloc(g._h, dx, dy, drows, n)
    assert rc == h2w.H2W_OK, L.h2w_last_error()
    hrows = None if rows is None else np.ascontiguousarray(rows, np.int64)
    rc = L.h2w_dense_matvec(g._h, dx, dy2, None if hrows is None else hrows.ctypes.data_as(ctypes.c_void_p),
                            n, h2w.DEVICE)
    assert rc == h2w.H2W_OK, L.h2w_last_error()
    return dev.get(dy, n), dev.get(dy2, n)


def test_noalloc_matches_dense_matvec():
    dev = _Dev()
    try:
        a = _handle(6000, 1)
        x = np.random.default_rng(2).normal(size=a.N)
        y, ref = _both(dev, a, x, None)  # every row
        assert np.all(np.isfinite(ref)) and np.array_equal(y, ref)
        b = _handle(2500, 3)  # second call: fewer points, fewer rows
        xb = np.random.default_rng(4).normal(size=b.N)
        rows = np.random.default_rng(5).choice(b.N, 333, replace=False)
        y, ref = _both(dev, b, xb, rows)
        assert np.all(np.isfinite(ref)) and np.array_equal(y, ref)
        # and the first handle again, on a few rows with repeats
        y, ref = _both(dev, a, x, [5, 5, 0, a.N - 1])
        assert np.array_equal(y, ref)
    finally:
        dev.free()


def test_noalloc_argument_checks():
    dev = _Dev()
    L = h2w.lib()
    try:
        g = _handle(3000, 6)
        x, y = dev.put(np.ones(g.N)), dev.put(np.zeros(4))
        bad = dev.put(np.array([0, 1, g.N, 2], np.int64))
        assert L.h2w_dense_matvec_noalloc(g._h, x, y, bad, 4) == h2w.H2W_EINVAL
        assert "out of range" in L.h2w_last_error().decode()
        assert L.h2w_dense_matvec_noalloc(g._h, x, y, bad, -1) == h2w.H2W_EINVAL
        assert L.h2w_dense_matvec_noalloc(g._h, None, y, bad, 4) == h2w.H2W_EINVAL
        # a valid call after the refused ones (the bad-row count starts from zero again)
        good = dev.put(np.array([0, 1, 2, 3], np.int64))
        assert L.h2w_dense_matvec_noalloc(g._h, x, y, good, 4) == h2w.H2W_OK
        # sizes beyond the staging limit are refused before any device work
        big = h2w.Initialisation(1, (1 << 20) + 8, 64, 1.0, 10, h2w.CHOICE_RBF, h2w.NESTED,
                                 points=np.linspace(-1, 1, (1 << 20) + 8)[None, :])
        assert L.h2w_dense_matvec_noalloc(big._h, x, y, good, 4) == h2w.H2W_EINVAL
        assert "staging limit" in L.h2w_last_error().decode()
    finally:
        dev.free()
