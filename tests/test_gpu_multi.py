"""GPU: 16 right-hand sides per pass on the matrix cores.  The 16-wide product
equals 16 single-vector products (to rounding) for every stored-operator
precision and format, including a column count that is not a multiple of 16;
lockstep GMRES reaches the single-vector solve's accuracy in every column."""
import numpy as np
import pytest

import h2weak_amd as h2w
import oracle

pytestmark = pytest.mark.gpu

# dim, N, n_max, tol_pow, choice, variant, points
CASES = [
    (3, 16 ** 3, 27, 10, 1, 0, "grid"),   # H2*, RBF
    (3, 4000, 32, 8, 0, 1, "rand"),       # (H2+H)*, IE: CUR column factors (atomic items)
    (2, 48 * 48, 36, 10, 0, 2, "grid"),   # H*
    (1, 4096, 64, 12, 1, 0, "rand"),      # 1D
    # the transposed products' two slab shapes: blocks of few rows (narrow slabs in every
    # precision) and 256-point leaves (wide slabs, output boxes of two row tiles)
    (3, 16 ** 3, 27, 4, 1, 0, "grid"),
    (2, 64 * 64, 256, 12, 1, 0, "grid"),
]


def _ids(c):
    return f"{c[0]}d-N{c[1]}-tol{c[3]}-ch{c[4]}-{['nested', 'semi', 'hstar'][c[5]]}-{c[6]}"


@pytest.fixture(scope="module")
def built():
    cache = {}

    def get(case):
        if case not in cache:
            dim, N, nmax, tp, ch, var, kind = case
            P = None if kind == "grid" else np.random.default_rng(17).uniform(-1, 1, (dim, N))
            g = h2w.Initialisation(dim, N, nmax, 1.0, tp, ch, var, points=P)
            g.tree_initialisation()
            cache[case] = (g, P)
        return cache[case]

    return get


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_multi_rhs_product(built, case):
    g, _ = built(case)
    N = g.N
    X = np.random.default_rng(3).normal(size=(21, N))  # one full pass of 16 and a partial one of 5
    for prec, tol in ((h2w.FP64, 1e-13), (h2w.FP32, 2e-6), (h2w.FP32_INNER, 2e-6)):
        if prec == h2w.FP32_INNER and case[3] <= 6:
            continue
        Y = g.matvec_multi(X, precision=prec)
        for c in (0, 7, 15, 16, 20):
            y1 = g.mat_vec_prod(X[c], precision=prec)
            assert _rel(Y[c], y1) <= tol, (prec, c)


@pytest.mark.parametrize("case", CASES[:2], ids=_ids)
def test_multi_rhs_gmres(built, case):
    g, P = built(case)
    dim, N, nmax, tp, ch, var, kind = case
    o = oracle.Oracle(dim, N, nmax, 1.0, tp, ch, var, points=P)
    o.tree_initialisation()
    rng = np.random.default_rng(11)
    Xt = rng.normal(size=(5, N))
    B = np.stack([o.dense_matvec(x) for x in Xt])
    gtol = 1e-10 if tp >= 10 else 1e-8
    single64 = [g.gmres(B[c], None, 300, gtol, h2w.GMRES_FP64) for c in range(5)]
    for mode in (h2w.GMRES_FP64, h2w.GMRES_FP32_IR, h2w.GMRES_FP32_INNER_IR):
        if mode == h2w.GMRES_FP32_INNER_IR and tp <= 6:
            continue
        X, st = g.gmres_multi(B, None, 300, gtol, mode)
        for c in range(5):
            x1, s1 = single64[c]
            assert st[c]["converged"], (mode, c, st[c])
            # the true residual of every column meets the tolerance
            assert _rel(g.mat_vec_prod(X[c]), B[c]) <= 2 * gtol
            # and the solution is as accurate as the single-vector FP64 solve's
            e1, em = _rel(x1, Xt[c]), _rel(X[c], Xt[c])
            assert em <= 10 * e1 + 1e-13, (mode, c, em, e1)
            if mode == h2w.GMRES_FP64:
                assert abs(st[c]["iterations"] - s1["iterations"]) <= 1


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_single_read_launch_matches_two_launches(built, case, monkeypatch):
    """The fused launch that applies every stored block directly and transposed in one read equals the
    direct launch followed by the separate transposed launches (a handle built with H2W_NO_SYM16 set),
    in every stored precision, to the rounding of the atomics' summation order."""
    g, P = built(case)
    dim, N, nmax, tp, ch, var, kind = case
    monkeypatch.setenv("H2W_NO_SYM16", "1")
    g2 = h2w.Initialisation(dim, N, nmax, 1.0, tp, ch, var, points=P)
    g2.tree_initialisation()
    monkeypatch.delenv("H2W_NO_SYM16")
    X = np.random.default_rng(5).normal(size=(16, N))
    for prec, tol in ((h2w.FP64, 1e-14), (h2w.FP32, 2e-6), (h2w.FP32_INNER, 2e-6)):
        if prec == h2w.FP32_INNER and case[3] <= 6:
            continue
        Y1, Y2 = g.matvec_multi(X, precision=prec), g2.matvec_multi(X, precision=prec)
        for c in range(16):
            assert _rel(Y1[c], Y2[c]) <= tol, (prec, c)
