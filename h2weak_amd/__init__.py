"""h2weak_amd -- MI355X-native H2* / (H2+H)* fast matrix-vector product and
GMRES driver, a drop-in for RiteshKhan/H2weak's operator facade.

The compute path is libh2w.so (hand-written HIP kernels for gfx950 behind the
C ABI in include/h2w.h).  This module is a thin ctypes mirror of the
reference's host interface:

  * ``Initialisation`` mirrors ``class initialisation``
    ({1,2,3}D/codes/matvec.hpp): constructor, ``tree_initialisation()``,
    ``mat_vec_prod(charge, output)``;
  * ``IterativeSolver`` mirrors ``class iterative_solver``
    (1D/codes/gmres.hpp:14-201): ``GMRES(A, x, b)`` with ``resid`` and
    ``max_iter`` (the iteration count) set on return.

There is no CPU fallback: importing works without a GPU, but constructing an
operator raises ``H2WError`` when libh2w.so is missing or no GPU is visible.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libh2w.so")

H2W_OK, H2W_EINVAL, H2W_ESTATE, H2W_EHIP, H2W_ENOMEM, H2W_ENOCONV, H2W_ENOLIB = range(7)
NESTED, SEMI_NESTED, NON_NESTED = 0, 1, 2
CHOICE_IE, CHOICE_RBF, CHOICE_HELMHOLTZ, CHOICE_HANKEL = 0, 1, 2, 10
HOST, DEVICE = 0, 1
FP64, FP32, FP32_OTF, FP64_OTF, FP32_INNER = 0, 1, 2, 3, 4
ORDER_INPUT, ORDER_TREE = 0, 1
GMRES_FP64, GMRES_FP32_IR, GMRES_FP32_OTF_IR, GMRES_FP32_INNER_IR = 0, 1, 2, 3

# every symbol include/h2w.h declares
EXPORTS = ("h2w_default_opts", "h2w_create", "h2w_assemble", "h2w_matvec", "h2w_matvec_ex", "h2w_gmres",
           "h2w_dense_matvec", "h2w_get_stats", "h2w_tree_order", "h2w_box_skeleton", "h2w_interaction_list",
           "h2w_near_list", "h2w_destroy", "h2w_last_error", "h2w_bench_matvec", "h2w_set_inner_tolerance", "h2w_set_deflation", "h2w_assemble_many",
           "h2w_autotune", "h2w_gmres_multi", "h2w_bench_matvec_multi", "h2w_batch_create", "h2w_batch_matvec",
           "h2w_batch_gmres", "h2w_batch_bench_matvec", "h2w_batch_destroy", "h2w_set_rank_model",
           "h2w_rank_samples", "h2w_rank_predict", "h2w_save", "h2w_load", "h2w_bench_stream_read",
           "h2w_dense_matvec_noalloc")
RANK_FEATURES, RANK_PARAMS = 16, 9473
TUNE_UNROLL, TUNE_INNER, TUNE_DEFLATION = 1, 2, 4


class H2WError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"h2w error {code}: {msg}")
        self.code = code


class Opts(ctypes.Structure):
    _fields_ = [("compat", ctypes.c_int), ("nlevels", ctypes.c_int), ("fp32_shadow", ctypes.c_int),
                ("device", ctypes.c_int), ("rank_cap", ctypes.c_int), ("workspace_bytes", ctypes.c_int64),
                ("plan_only", ctypes.c_int), ("ir_tol_pow", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("variant", ctypes.c_int), ("choice", ctypes.c_int), ("nlevels", ctypes.c_int),
                ("N", ctypes.c_int64), ("memory", ctypes.c_double), ("diag", ctypes.c_double),
                ("assembly_seconds", ctypes.c_double), ("max_rank_far", ctypes.c_int),
                ("max_rank_vertex", ctypes.c_int), ("mean_rank_far", ctypes.c_double),
                ("mean_rank_vertex", ctypes.c_double), ("stored_bytes_fp64", ctypes.c_int64),
                ("stored_bytes_fp32", ctypes.c_int64), ("n_launches_matvec", ctypes.c_int),
                ("aca_reruns", ctypes.c_int), ("stored_bytes_inner_fp32", ctypes.c_int64),
                ("aca_batches", ctypes.c_int), ("aca_pool_bytes", ctypes.c_int64), ("rank_pred_blocks", ctypes.c_int),
                ("rank_pred_under", ctypes.c_int), ("rank_pred_mae_log2", ctypes.c_double),
                ("inner_tol_exp", ctypes.c_double), ("rank_fixed_blocks", ctypes.c_int),
                ("rank_fixed_resumed", ctypes.c_int), ("rank_fixed_extra_steps", ctypes.c_int64),
                ("rank_fixed_sampled_err", ctypes.c_double), ("diag_im", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class GmresStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("outer_iterations", ctypes.c_int), ("matvecs_fp64", ctypes.c_int),
                ("matvecs_fp32", ctypes.c_int), ("resid", ctypes.c_double), ("seconds", ctypes.c_double),
                ("converged", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BenchResult(ctypes.Structure):
    _fields_ = [("ms_per_matvec", ctypes.c_double), ("launch_ms", ctypes.c_double * 64),
                ("n_launch", ctypes.c_int), ("launch_bytes", ctypes.c_int64 * 64), ("launch_role", ctypes.c_int * 64)]


class Tune(ctypes.Structure):
    _fields_ = [("ir_tol_exp", ctypes.c_double), ("n_cand", ctypes.c_int), ("cand_ir_tol_exp", ctypes.c_double * 12),
                ("cand_seconds", ctypes.c_double * 12), ("cand_iterations", ctypes.c_int * 12),
                ("cand_outer", ctypes.c_int * 12), ("cand_converged", ctypes.c_int * 12),
                ("cand_inner_bytes", ctypes.c_int64 * 12), ("matvec_ms_before", ctypes.c_double * 3),
                ("matvec_ms_after", ctypes.c_double * 3), ("defl_m", ctypes.c_int), ("defl_k", ctypes.c_int),
                ("defl_build_seconds", ctypes.c_double), ("n_defl_cand", ctypes.c_int),
                ("cand_defl_m", ctypes.c_int * 12), ("cand_defl_k", ctypes.c_int * 12),
                ("cand_defl_ir_tol_exp", ctypes.c_double * 12), ("cand_defl_seconds", ctypes.c_double * 12),
                ("cand_defl_iterations", ctypes.c_int * 12), ("cand_defl_outer", ctypes.c_int * 12),
                ("cand_defl_converged", ctypes.c_int * 12)]


_lib = None


def lib():
    """Load libh2w.so (in-tree).  Raises H2WError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise H2WError(H2W_ENOLIB, f"{LIB_PATH} is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    p, i, i64, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
    L.h2w_default_opts.argtypes = [ctypes.POINTER(Opts)]
    L.h2w_create.argtypes = [i, i, i64, p, i, d, i, i, ctypes.POINTER(Opts), ctypes.POINTER(p)]
    L.h2w_assemble.argtypes = [p]
    L.h2w_matvec.argtypes = [p, p, p, i]
    L.h2w_matvec_ex.argtypes = [p, p, p, i, i64, i, i, i]
    L.h2w_gmres.argtypes = [p, p, p, i, d, i, i, ctypes.POINTER(GmresStats)]
    L.h2w_dense_matvec.argtypes = [p, p, p, p, i64, i]
    L.h2w_dense_matvec_noalloc.argtypes = [p, p, p, p, i64]
    L.h2w_get_stats.argtypes = [p, ctypes.POINTER(Stats)]
    L.h2w_tree_order.argtypes = [p, p]
    L.h2w_box_skeleton.argtypes = [p, i, i, i, p, p, i]
    L.h2w_interaction_list.argtypes = [p, i, i, p, p, i]
    L.h2w_near_list.argtypes = [p, i, i, p, i]
    L.h2w_destroy.argtypes = [p]
    L.h2w_last_error.restype = ctypes.c_char_p
    L.h2w_bench_matvec.argtypes = [p, i, i, ctypes.POINTER(BenchResult)]
    L.h2w_set_inner_tolerance.argtypes = [p, d]
    L.h2w_set_deflation.argtypes = [p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.h2w_assemble_many.argtypes = [p, ctypes.c_int, ctypes.c_int]
    L.h2w_bench_matvec_multi.argtypes = [p, i, i, ctypes.POINTER(BenchResult)]
    L.h2w_batch_create.argtypes = [ctypes.POINTER(p), i, ctypes.POINTER(p)]
    L.h2w_batch_matvec.argtypes = [p, p, p, i, i, i]
    L.h2w_batch_gmres.argtypes = [p, p, p, i, i, d, i, i, ctypes.POINTER(GmresStats)]
    L.h2w_batch_bench_matvec.argtypes = [p, i, i, ctypes.POINTER(BenchResult)]
    L.h2w_batch_destroy.argtypes = [p]
    L.h2w_set_rank_model.argtypes = [p, p, i64]
    L.h2w_rank_samples.argtypes = [p, p, p, i64]
    L.h2w_rank_samples.restype = i64
    L.h2w_rank_predict.argtypes = [p, p, i64, p]
    L.h2w_save.argtypes = [p, ctypes.c_char_p]
    L.h2w_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(Opts), ctypes.POINTER(p)]
    L.h2w_gmres_multi.argtypes = [p, p, i64, p, i64, i, i, d, i, i, ctypes.POINTER(GmresStats)]
    L.h2w_autotune.argtypes = [p, p, i, d, i, i, ctypes.POINTER(Tune)]
    L.h2w_bench_stream_read.argtypes = [i, i64, i, ctypes.POINTER(d)]
    _lib = L
    return L


def _check(rc):
    if rc != H2W_OK:
        raise H2WError(rc, lib().h2w_last_error().decode())


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def bench_stream_read(nbytes=4 << 30, nrep=5, device=0):
    """Read bandwidth (GB/s) of a plain read-only stream over nbytes of device memory: the
    achievable rate a streaming kernel is compared with."""
    g = ctypes.c_double()
    _check(lib().h2w_bench_stream_read(int(device), int(nbytes), int(nrep), ctypes.byref(g)))
    return g.value


class Initialisation:
    """Mirror of the reference's ``initialisation`` facade ({1,2,3}D/codes/matvec.hpp).

    dim, N, n_max, L, tol_pow (ACA/NCA tolerance 10^-tol_pow) and choice
    (0 = integral equation, 1 = RBF) have the reference's meaning; ``variant``
    picks H2* (NESTED) or (H2+H)* (SEMI_NESTED), which the reference selects
    at compile time.  ``points`` (dim x N) defaults to the reference's grid.
    """

    def __init__(self, dim, N, n_max, L=1.0, tol_pow=12, choice=CHOICE_RBF, variant=NESTED, points=None,
                 nlevels=0, compat=1, fp32_shadow=1, rank_cap=256, device=0, workspace_bytes=0,
                 plan_only=0, ir_tol_pow=6):
        self.dim, self.N = int(dim), int(N)
        o = Opts()
        lib().h2w_default_opts(ctypes.byref(o))
        o.compat, o.nlevels, o.fp32_shadow, o.device = compat, nlevels, fp32_shadow, device
        o.rank_cap, o.workspace_bytes, o.plan_only = rank_cap, workspace_bytes, plan_only
        o.ir_tol_pow = ir_tol_pow
        self._pts = None
        pp = None
        if points is not None:
            self._pts = _f64(np.asarray(points).reshape(dim, N))
            pp = _ptr(self._pts)
        h = ctypes.c_void_p()
        _check(lib().h2w_create(dim, variant, int(N), pp, int(n_max), float(L), int(tol_pow), int(choice),
                                ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        # complex kernels (choice 2 / 10): vectors are complex128 (interleaved re, im in the C ABI)
        self.cplx = int(choice) in (CHOICE_HELMHOLTZ, CHOICE_HANKEL)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.h2w_destroy(h)
            self._h = None

    def tree_initialisation(self):
        _check(lib().h2w_assemble(self._h))

    def mat_vec_prod(self, charge, output=None, precision=FP64):
        """output = A charge (input order).  numpy arrays (host) or objects with
        ``data_ptr()`` (device pointers on the handle's GPU)."""
        if hasattr(charge, "data_ptr"):
            _check(lib().h2w_matvec_ex(self._h, ctypes.c_void_p(charge.data_ptr()),
                                       ctypes.c_void_p(output.data_ptr()), 1, self.N, precision, ORDER_INPUT, DEVICE))
            return output
        if self.cplx:
            x = np.ascontiguousarray(charge, dtype=np.complex128)
            y = np.empty(self.N, dtype=np.complex128)
            _check(lib().h2w_matvec_ex(self._h, _ptr(x), _ptr(y), 1, self.N, precision, ORDER_INPUT, HOST))
            return y
        x = _f64(charge)
        y = np.empty(self.N) if output is None else output
        _check(lib().h2w_matvec_ex(self._h, _ptr(x), _ptr(y), 1, self.N, precision, ORDER_INPUT, HOST))
        return y

    def matvec_multi(self, X, precision=FP64):
        X = np.ascontiguousarray(X, dtype=np.complex128) if self.cplx else _f64(X)
        nrhs = X.shape[0]
        Y = np.empty_like(X)
        _check(lib().h2w_matvec_ex(self._h, _ptr(X), _ptr(Y), nrhs, self.N, precision, ORDER_INPUT, HOST))
        return Y

    def dense_rows(self, rows, x):
        x = np.ascontiguousarray(x, dtype=np.complex128) if self.cplx else _f64(x)
        rows = None if rows is None else np.ascontiguousarray(rows, dtype=np.int64)
        n = self.N if rows is None else len(rows)
        y = np.empty(n, dtype=x.dtype)
        _check(lib().h2w_dense_matvec(self._h, _ptr(x), _ptr(y), None if rows is None else _ptr(rows), n, HOST))
        return y

    def gmres(self, b, x0=None, max_iter=500, tol=1e-12, mode=GMRES_FP64):
        dt = np.complex128 if self.cplx else np.float64
        b = np.ascontiguousarray(b, dtype=dt)
        x = np.ones(self.N, dtype=dt) if x0 is None else np.array(x0, dtype=dt)
        st = GmresStats()
        rc = lib().h2w_gmres(self._h, _ptr(b), _ptr(x), int(max_iter), float(tol), mode, HOST, ctypes.byref(st))
        if rc not in (H2W_OK, H2W_ENOCONV):
            _check(rc)
        return x, st.as_dict()

    def gmres_multi(self, B, X0=None, max_iter=500, tol=1e-12, mode=GMRES_FP64):
        """nrhs independent solves (rows of B), 16 at a time in lockstep on the matrix cores."""
        B = _f64(np.atleast_2d(B))
        nrhs = B.shape[0]
        X = np.ones_like(B) if X0 is None else np.array(np.atleast_2d(X0), dtype=np.float64)
        st = (GmresStats * nrhs)()
        rc = lib().h2w_gmres_multi(self._h, _ptr(B), self.N, _ptr(X), self.N, nrhs, int(max_iter), float(tol), mode,
                                   HOST, st)
        if rc not in (H2W_OK, H2W_ENOCONV):
            _check(rc)
        return X, [st[c].as_dict() for c in range(nrhs)]

    def stats(self):
        st = Stats()
        _check(lib().h2w_get_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    @property
    def memory(self):
        return self.stats()["memory"]

    def tree_order(self):
        p = np.empty(self.N, dtype=np.int32)
        _check(lib().h2w_tree_order(self._h, _ptr(p)))
        return p

    def skeleton(self, family, level, box, cap=4096):
        ir = np.empty(cap, dtype=np.int32)
        ic = np.empty(cap, dtype=np.int32)
        r = lib().h2w_box_skeleton(self._h, family, level, box, _ptr(ir), _ptr(ic), cap)
        if r < 0:
            _check(-r)
        return ir[:r].tolist(), ic[:r].tolist()

    def interaction_list(self, level, box):
        il = np.empty(512, dtype=np.int32)
        v = np.empty(512, dtype=np.int32)
        n = lib().h2w_interaction_list(self._h, level, box, _ptr(il), _ptr(v), 512)
        if n < 0:
            _check(-n)
        return il[:n].tolist(), v[:n].tolist()

    def near_list(self, level, box):
        nb = np.empty(64, dtype=np.int32)
        n = lib().h2w_near_list(self._h, level, box, _ptr(nb), 64)
        if n < 0:
            _check(-n)
        return nb[:n].tolist()

    def bench_matvec(self, nrep=10, precision=FP64):
        r = BenchResult()
        _check(lib().h2w_bench_matvec(self._h, int(nrep), precision, ctypes.byref(r)))
        n = r.n_launch
        return {"ms_per_matvec": r.ms_per_matvec, "launch_ms": list(r.launch_ms[:n]),
                "launch_bytes": list(r.launch_bytes[:n]), "launch_role": list(r.launch_role[:n])}


    def save(self, path):
        """Write the assembled operator's skeletons (h2w_save)."""
        _check(lib().h2w_save(self._h, os.fsencode(path)))

    @classmethod
    def load(cls, path, **opts):
        """An assembled Initialisation from a file written by save(); no cross approximation runs."""
        o = Opts()
        lib().h2w_default_opts(ctypes.byref(o))
        for k, v in opts.items():
            setattr(o, k, v)
        h = ctypes.c_void_p()
        _check(lib().h2w_load(os.fsencode(path), ctypes.byref(o), ctypes.byref(h)))
        self = cls.__new__(cls)
        self._h = h
        self._pts = None
        st = Stats()
        _check(lib().h2w_get_stats(h, ctypes.byref(st)))
        self.dim, self.N = st.dim, st.N
        self.cplx = st.choice in (CHOICE_HELMHOLTZ, CHOICE_HANKEL)
        self.nLevels = st.nlevels
        return self

    def set_rank_model(self, params):
        """BF16 parameters (uint16 array of RANK_PARAMS) of the rank predictor; None removes it.
        Takes effect at the next tree_initialisation()."""
        if params is None:
            _check(lib().h2w_set_rank_model(self._h, None, 0))
            return
        w = np.ascontiguousarray(params, dtype=np.uint16)
        _check(lib().h2w_set_rank_model(self._h, _ptr(w), w.size))

    def rank_samples(self):
        """(features [n, 16] float32, ranks [n] int32) of every ACA block of the last assembly."""
        n = lib().h2w_rank_samples(self._h, None, None, 0)
        if n < 0:
            _check(-n)
        F = np.empty((n, RANK_FEATURES), dtype=np.float32)
        R = np.empty(n, dtype=np.int32)
        lib().h2w_rank_samples(self._h, _ptr(F), _ptr(R), n)
        return F, R

    def rank_predict(self, feat):
        """log2(1 + rank) predicted by the GPU kernel for feature rows [n, 16]."""
        F = np.ascontiguousarray(feat, dtype=np.float32)
        out = np.empty(F.shape[0], dtype=np.float32)
        _check(lib().h2w_rank_predict(self._h, _ptr(F), F.shape[0], _ptr(out)))
        return out

    def set_inner_tolerance(self, ir_tol_pow):
        """Rebuild the truncated FP32 inner operator for tolerance 10^-ir_tol_pow (0 removes it;
        fractional exponents as the autotuner's half decades)."""
        _check(lib().h2w_set_inner_tolerance(self._h, float(ir_tol_pow)))

    def set_deflation(self, m, k=0, precision=None):
        """Deflation space of the refinement: m Arnoldi steps of the inner operator (default the truncated
        inner operator, H2W_FP32_INNER) from a fixed pseudo-random vector, the whole space (k = 0) or its k
        Ritz vectors farthest from the diagonal value; every later correction in that refinement mode starts
        from it (single solves, gmres_multi, and batches whose every system has one).  m = 0 drops it."""
        _check(lib().h2w_set_deflation(self._h, int(m), int(k), FP32_INNER if precision is None else int(precision)))

    def autotune(self, b=None, tol=1e-10, max_iter=1000, flags=TUNE_UNROLL | TUNE_INNER):
        """GEMV unroll depth per launch, inner tolerance of the refinement and (TUNE_DEFLATION) its deflation
        space (b in input order: a right-hand side other than the ones to be timed)."""
        t = Tune()
        bb = None
        if b is not None:
            bb = np.ascontiguousarray(b, dtype=np.float64)
            if bb.shape != (self.N,):
                raise ValueError("b must have N entries")
        _check(lib().h2w_autotune(self._h, _ptr(bb) if bb is not None else None, HOST, float(tol),
                                  int(max_iter), int(flags), ctypes.byref(t)))
        n = t.n_cand
        return {"ir_tol_exp": t.ir_tol_exp,
                "candidates": [{"ir_tol_exp": t.cand_ir_tol_exp[q], "seconds": t.cand_seconds[q],
                                "iterations": t.cand_iterations[q], "outer": t.cand_outer[q],
                                "converged": bool(t.cand_converged[q]), "inner_bytes": t.cand_inner_bytes[q]}
                               for q in range(n)],
                "matvec_ms_before": list(t.matvec_ms_before), "matvec_ms_after": list(t.matvec_ms_after),
                "deflation": {"m": t.defl_m, "k": t.defl_k, "build_seconds": t.defl_build_seconds,
                              "candidates": [{"m": t.cand_defl_m[q], "k": t.cand_defl_k[q],
                                              "ir_tol_exp": t.cand_defl_ir_tol_exp[q],
                                              "seconds": t.cand_defl_seconds[q],
                                              "iterations": t.cand_defl_iterations[q], "outer": t.cand_defl_outer[q],
                                              "converged": bool(t.cand_defl_converged[q])}
                                             for q in range(t.n_defl_cand)]}}


    def bench_matvec_multi(self, nrep=10, precision=FP64):
        """HIP-event time of one 16-wide product (16 right-hand sides per pass) and the operator bytes
        it multiplies (launch_bytes[0]; symmetric blocks count twice)."""
        r = BenchResult()
        _check(lib().h2w_bench_matvec_multi(self._h, int(nrep), precision, ctypes.byref(r)))
        n = r.n_launch
        return {"ms_per_pass": r.ms_per_matvec, "nrhs": 16, "launch_bytes": list(r.launch_bytes[:n]),
                "launch_ms": list(r.launch_ms[:n]), "launch_role": list(r.launch_role[:n])}


class Batch:
    """Independent systems (assembled Initialisation objects on one device) multiplied and
    solved together: one launch per phase for all of them, 16 right-hand sides per pass.
    Vectors are lists with one (nrhs, N_s) array per system."""

    def __init__(self, systems):
        self.systems = list(systems)  # keeps the handles alive as long as the batch
        arr = (ctypes.c_void_p * len(self.systems))(*[s._h for s in self.systems])
        h = ctypes.c_void_p()
        _check(lib().h2w_batch_create(arr, len(self.systems), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().h2w_batch_destroy(h)
            self._h = None

    def _pack(self, Xs):
        Xs = [np.atleast_2d(np.asarray(x, dtype=np.float64)) for x in Xs]
        if len(Xs) != len(self.systems) or any(x.shape[1] != s.N for x, s in zip(Xs, self.systems)):
            raise ValueError("one (nrhs, N_s) array per system")
        nrhs = Xs[0].shape[0]
        if any(x.shape[0] != nrhs for x in Xs):
            raise ValueError("every system takes the same number of right-hand sides")
        return np.ascontiguousarray(np.concatenate([x.ravel() for x in Xs])), nrhs

    def _unpack(self, flat, nrhs):
        out, o = [], 0
        for s in self.systems:
            out.append(flat[o:o + nrhs * s.N].reshape(nrhs, s.N))
            o += nrhs * s.N
        return out

    def matvec(self, Xs, precision=FP64):
        X, nrhs = self._pack(Xs)
        Y = np.empty_like(X)
        _check(lib().h2w_batch_matvec(self._h, _ptr(X), _ptr(Y), nrhs, precision, HOST))
        return self._unpack(Y, nrhs)

    def gmres(self, Bs, X0s=None, max_iter=500, tol=1e-10, mode=GMRES_FP64):
        B, nrhs = self._pack(Bs)
        X = np.ones_like(B) if X0s is None else self._pack(X0s)[0].copy()
        st = (GmresStats * (len(self.systems) * nrhs))()
        rc = lib().h2w_batch_gmres(self._h, _ptr(B), _ptr(X), nrhs, int(max_iter), float(tol), mode, HOST, st)
        if rc not in (H2W_OK, H2W_ENOCONV):
            _check(rc)
        stats = [[st[s * nrhs + r].as_dict() for r in range(nrhs)] for s in range(len(self.systems))]
        return self._unpack(X, nrhs), stats

    def _recreate(self):
        """A new native batch over the same systems (after their inner operators were rebuilt)."""
        lib().h2w_batch_destroy(self._h)
        self._h = None
        arr = (ctypes.c_void_p * len(self.systems))(*[s._h for s in self.systems])
        h = ctypes.c_void_p()
        _check(lib().h2w_batch_create(arr, len(self.systems), ctypes.byref(h)))
        self._h = h

    def autotune(self, Bs, tol=1e-8, max_iter=300, inner=(6.0, 5.5, 5.0, 4.5), spaces=((128, 48), (128, 64))):
        """Choose the refinement's inner tolerance (10^-inner, the same for every system) and every system's
        deflation space (m Arnoldi steps, k Ritz vectors) by timing lockstep FP32-inner refined solves of Bs,
        right-hand sides that should not be the ones timed afterwards; keeps the fastest candidate whose
        columns all converged.  Returns the choice and every candidate's time and step count."""
        cands, best = [], None
        for itol in inner:
            for g in self.systems:
                g.set_inner_tolerance(itol)
            self._recreate()
            for m, k in spaces:
                for g in self.systems:
                    g.set_deflation(m, k)
                self.gmres([b[:1] for b in Bs], None, max_iter, tol, GMRES_FP32_INNER_IR)  # warm-up
                _, st = self.gmres(Bs, None, max_iter, tol, GMRES_FP32_INNER_IR)
                cols = [c for row in st for c in row]
                c = {"inner_tol_exp": float(itol), "m": int(m), "k": int(k), "seconds": cols[0]["seconds"],
                     "max_iterations": max(c["iterations"] for c in cols),
                     "converged": all(c["converged"] for c in cols)}
                cands.append(c)
                if c["converged"] and (best is None or c["seconds"] < best["seconds"]):
                    best = c
        if best is None:
            raise H2WError(H2W_ENOCONV, "no candidate converged")
        for g in self.systems:
            g.set_inner_tolerance(best["inner_tol_exp"])
        self._recreate()
        for g in self.systems:
            g.set_deflation(best["m"], best["k"])
        return {"inner_tol_exp": best["inner_tol_exp"], "m": best["m"], "k": best["k"], "candidates": cands}

    def bench_matvec(self, nrep=10, precision=FP64):
        r = BenchResult()
        _check(lib().h2w_batch_bench_matvec(self._h, int(nrep), precision, ctypes.byref(r)))
        n = r.n_launch
        return {"ms_per_pass": r.ms_per_matvec, "nrhs": 16, "launch_bytes": list(r.launch_bytes[:n]),
                "launch_role": list(r.launch_role[:n]), "launch_ms": list(r.launch_ms[:n])}


def assemble_many(systems, nthreads=4):
    """tree_initialisation() of several Initialisation objects at once: up to nthreads host threads, each
    system on its own stream (h2w_assemble_many); the operators equal one-by-one assembly's."""
    systems = list(systems)
    arr = (ctypes.c_void_p * len(systems))(*[g._h for g in systems])
    _check(lib().h2w_assemble_many(arr, len(systems), int(nthreads)))


class IterativeSolver:
    """Mirror of ``iterative_solver`` (1D/codes/gmres.hpp:14-201)."""

    def __init__(self, max_iter, gmres_tol_pow, mode=GMRES_FP64):
        self.restart = self.max_iter = int(max_iter)
        self.gmres_tol = 10.0 ** (-gmres_tol_pow)
        self.resid = 0.0
        self.mode = mode
        self.stats = None

    def GMRES(self, A, x, b):
        """Solves A x = b in place (x holds x0 on entry), as the reference's GMRES."""
        xs, st = A.gmres(b, x, self.max_iter, self.gmres_tol, self.mode)
        x[:] = xs
        self.resid = st["resid"]
        self.max_iter = st["iterations"]
        self.stats = st
        return x
