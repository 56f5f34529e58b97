"""CPU FP64 oracle for the H2* / (H2+H)* hot path -- TEST INFRASTRUCTURE ONLY.

A restatement of RiteshKhan/H2weak's CPU algorithm (kernel entries, GL-16
diagonal quadrature, partial-pivot ACA, tree and interaction lists, nested and
semi-nested assembly and matvec, storage count, dense product, GMRES), see the
header of ``h2w_oracle.cpp`` for the file:line each part follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker / CPU baseline.
The product (``h2weak_amd``) never calls it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_d = ctypes.c_double
_i = ctypes.c_int
_i64 = ctypes.c_int64
_p = ctypes.c_void_p
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int)
_pi64 = ctypes.POINTER(ctypes.c_int64)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.orc_create.restype = _p
        L.orc_create.argtypes = [_i, _i, _i64, _pd, _i, _d, _i, _i, _i, _i]
        L.orc_destroy.argtypes = [_p]
        for name in ("orc_nlevels",):
            getattr(L, name).restype = _i
            getattr(L, name).argtypes = [_p]
        L.orc_assemble.argtypes = [_p]
        for name in ("orc_memory", "orc_diag", "orc_kii"):
            getattr(L, name).restype = _d
            getattr(L, name).argtypes = [_p]
        L.orc_entry.restype = _d
        L.orc_entry.argtypes = [_p, _i, _i]
        L.orc_points.argtypes = [_p, _pd]
        L.orc_matvec.argtypes = [_p, _pd, _pd]
        L.orc_dense_matvec.argtypes = [_p, _pd, _pd]
        L.orc_dense_rows.argtypes = [_p, _pi64, _i64, _pd, _pd]
        L.orc_gmres.restype = _i
        L.orc_gmres.argtypes = [_p, _pd, _pd, _i, _d, _pd, _pi]
        L.orc_gl16.argtypes = [_pd, _pd]
        L.orc_tree_order.argtypes = [_p, _pi]
        L.orc_box_count.restype = _i
        L.orc_box_count.argtypes = [_p, _i, _i]
        L.orc_il.restype = _i
        L.orc_il.argtypes = [_p, _i, _i, _pi, _pi, _i]
        L.orc_near.restype = _i
        L.orc_near.argtypes = [_p, _i, _i, _pi, _i]
        L.orc_rank.restype = _i
        L.orc_rank.argtypes = [_p, _i, _i, _i]
        L.orc_pivots.restype = _i
        L.orc_pivots.argtypes = [_p, _i, _i, _i, _pi, _pi]
        L.orc_threads.restype = _i
        L.orc_threads.argtypes = []
        L.orc_aca.restype = _i
        L.orc_aca.argtypes = [_p, _pi, _i, _pi, _i, _i, _pi, _pi, _pd, _pd]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_pd)


def _ip(a):
    return a.ctypes.data_as(_pi)


def threads():
    """OpenMP threads the oracle's parallel regions use."""
    return lib().orc_threads()


def gl16():
    n = np.zeros(16)
    w = np.zeros(16)
    lib().orc_gl16(_dp(n), _dp(w))
    return n, w


class Oracle:
    """Mirror of the reference's ``initialisation`` facade (*/codes/matvec.hpp)."""

    def __init__(self, dim, N, n_max, L=1.0, tol_pow=12, choice=1, variant=0, points=None, nlevels=-1, compat=1):
        self.dim, self.N = dim, int(N)
        self._pts = None
        pp = None
        if points is not None:
            self._pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(dim, N))
            pp = _dp(self._pts)
        self.h = lib().orc_create(dim, variant, int(N), pp, n_max, float(L), tol_pow, choice, nlevels, compat)
        if not self.h:
            raise ValueError("invalid oracle configuration")
        self.nlevels = lib().orc_nlevels(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def tree_initialisation(self):
        lib().orc_assemble(self.h)

    @property
    def memory(self):
        return lib().orc_memory(self.h)

    @property
    def diag(self):
        return lib().orc_diag(self.h)

    @property
    def kii(self):
        return lib().orc_kii(self.h)

    def entry(self, i, j):
        return lib().orc_entry(self.h, int(i), int(j))

    def points(self):
        out = np.zeros(self.dim * self.N)
        lib().orc_points(self.h, _dp(out))
        return out.reshape(self.dim, self.N)

    def mat_vec_prod(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.N)
        lib().orc_matvec(self.h, _dp(x), _dp(y))
        return y

    def dense_matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.N)
        lib().orc_dense_matvec(self.h, _dp(x), _dp(y))
        return y

    def dense_rows(self, rows, x):
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(len(rows))
        lib().orc_dense_rows(self.h, rows.ctypes.data_as(_pi64), len(rows), _dp(x), _dp(y))
        return y

    def gmres(self, b, x0=None, max_iter=500, tol=1e-12):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.ones(self.N) if x0 is None else np.array(x0, dtype=np.float64)
        resid = ctypes.c_double(0.0)
        nmv = ctypes.c_int(0)
        it = lib().orc_gmres(self.h, _dp(b), _dp(x), max_iter, tol, ctypes.byref(resid), ctypes.byref(nmv))
        return x, it, resid.value, nmv.value

    def tree_order(self):
        p = np.zeros(self.N, dtype=np.int32)
        lib().orc_tree_order(self.h, _ip(p))
        return p

    def box_count(self, level, k):
        return lib().orc_box_count(self.h, level, k)

    def interaction_list(self, level, k):
        out = np.zeros(512, dtype=np.int32)
        vf = np.zeros(512, dtype=np.int32)
        n = lib().orc_il(self.h, level, k, _ip(out), _ip(vf), 512)
        return out[:n].tolist(), vf[:n].tolist()

    def near(self, level, k):
        out = np.zeros(64, dtype=np.int32)
        n = lib().orc_near(self.h, level, k, _ip(out), 64)
        return out[:n].tolist()

    def rank(self, family, level, k):
        return lib().orc_rank(self.h, family, level, k)

    def pivots(self, family, level, k):
        r = self.rank(family, level, k)
        ir = np.zeros(max(r, 1), dtype=np.int32)
        ic = np.zeros(max(r, 1), dtype=np.int32)
        lib().orc_pivots(self.h, family, level, k, _ip(ir), _ip(ic))
        return ir[:r].tolist(), ic[:r].tolist()

    def aca(self, rows, cols, tol_pow):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        cap = min(len(rows), len(cols))
        rb = np.zeros(max(cap, 1), dtype=np.int32)
        cb = np.zeros(max(cap, 1), dtype=np.int32)
        Lm = np.zeros(max(cap * cap, 1))
        Rm = np.zeros(max(cap * cap, 1))
        r = lib().orc_aca(self.h, _ip(rows), len(rows), _ip(cols), len(cols), tol_pow, _ip(rb), _ip(cb), _dp(Lm), _dp(Rm))
        L = Lm[: r * r].reshape(r, r, order="F")
        R = Rm[: r * r].reshape(r, r, order="F")
        return r, rb[:r].tolist(), cb[:r].tolist(), L, R
