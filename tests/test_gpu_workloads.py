"""GPU at the benchmark workloads themselves: the 32-system batch (3D H2* RBF, N
log-uniform in [4096, 65536], 16 right-hand sides each, tolerance 1e-8) and
the learned-rank assembly at N = 262144 (uniform and Gaussian points).  Every
right-hand side is the dense FP64 product A x_true (the reference's
getTruePoten, 3D/codes/HODLR3Dtree.hpp:2630-2643, on the GPU and itself
parity-tested against the oracle), so the solution errors are against x_true
and the operators are checked against the dense operator, not against the
library's own compressed products (the reference's check, 3D/codes/main.cpp:27-38,86)."""
import numpy as np
import pytest

import h2weak_amd as h2w
from h2weak_amd import rank_mlp

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _sampled_rows_error(A, x, nrows=1024, seed=4):
    rows = np.sort(np.random.default_rng(seed).choice(A.N, min(nrows, A.N), replace=False))
    return _rel(A.mat_vec_prod(x)[rows], A.dense_rows(rows, x))


def c4_sizes(S=32, seed=5):
    rng = np.random.default_rng(seed)
    return [int(v) for v in np.rint(np.exp(rng.uniform(np.log(4096), np.log(65536), S)))]


@pytest.fixture(scope="module")
def c4():
    """The batch workload: 32 systems, points U([-1,1]^3) (seed 100 + s), n_max 64, tol 1e-8,
    16 right-hand sides each with X_true ~ N(0, 1) (seed 200 + s) and B = A X_true (dense FP64)."""
    tol_pow = 8
    systems, Xt, Bs = [], [], []
    for q, N in enumerate(c4_sizes()):
        P = np.random.default_rng(100 + q).uniform(-1.0, 1.0, (3, N))
        g = h2w.Initialisation(3, N, 64, 1.0, tol_pow, h2w.CHOICE_RBF, h2w.NESTED, points=P)
        g.tree_initialisation()
        systems.append(g)
        X = np.random.default_rng(200 + q).normal(size=(16, N))
        Xt.append(X)
        Bs.append(np.stack([g.dense_rows(None, X[c]) for c in range(16)]))
    print(f"[c4] 32 systems, N {min(g.N for g in systems)}..{max(g.N for g in systems)}", flush=True)
    return systems, Xt, Bs, 10.0 ** -tol_pow


@pytest.mark.timeout(900)
def test_c4_products_against_dense_rows(c4):
    """1024 sampled rows of every system's product (and the batched product) against the dense FP64 rows."""
    systems, Xt, Bs, tol = c4
    B = h2w.Batch(systems)
    Ys = B.matvec([X[:2] for X in Xt])
    worst = 0.0
    for q, g in enumerate(systems):
        rows = np.sort(np.random.default_rng(4).choice(g.N, 1024, replace=False))
        for c in range(2):
            e = _rel(Ys[q][c][rows], Bs[q][c][rows])
            worst = max(worst, e)
            assert e <= 10 * tol, (q, c, e)
        assert _sampled_rows_error(g, Xt[q][3]) <= 10 * tol
    print(f"[c4] worst sampled-row error of the batched product {worst:.2e} (tol {tol:g})", flush=True)


@pytest.mark.timeout(900)
def test_c4_batched_solves_against_x_true(c4):
    """Lockstep solves of all 512 columns (FP64, and FP32 inner operator inside FP64 refinement):
    every column converges, its FP64 residual meets the tolerance, and its error against x_true is
    within 10x of the same column's own single-right-hand-side FP64 solve."""
    systems, Xt, Bs, tol = c4
    e64 = [[_rel(g.gmres(Bs[q][c], None, 300, tol, h2w.GMRES_FP64)[0], Xt[q][c]) for c in range(16)]
           for q, g in enumerate(systems)]
    B = h2w.Batch(systems)
    for mode in (h2w.GMRES_FP64, h2w.GMRES_FP32_INNER_IR):
        Xs, st = B.gmres(Bs, None, 300, tol, mode)
        worst = 0.0
        for q, g in enumerate(systems):
            R = g.matvec_multi(Xs[q])
            for c in range(16):
                assert st[q][c]["converged"], (mode, q, c, st[q][c])
                assert _rel(R[c], Bs[q][c]) <= 2 * tol, (mode, q, c)
                e = _rel(Xs[q][c], Xt[q][c])
                worst = max(worst, e / e64[q][c])
                assert e <= 10 * e64[q][c] + 1e-12, (mode, q, c, e, e64[q][c])
        print(f"[c4] mode {mode}: worst error / single FP64 solve error {worst:.2f}", flush=True)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("kind", ["uniform", "gauss"])
def test_c5_rank_model_full_size(kind):
    """N = 262144, 3D H2* RBF, n_max 64, tol 1e-10 with the shipped rank model: the operator equals the
    one assembled with the heuristic estimates, 1024 sampled rows match the dense FP64 rows at the
    tolerance, and the refined solve reaches 1e-10 in the FP64 residual with the solution error of
    the plain FP64 solve."""
    N, tol_pow = 262144, 10
    P = rank_mlp.points(kind, N, 7)
    params, _ = rank_mlp.load(rank_mlp.default_model_path(3, 1, tol_pow))
    a = h2w.Initialisation(3, N, 64, 1.0, tol_pow, h2w.CHOICE_RBF, h2w.NESTED, points=P)
    a.tree_initialisation()
    mem_heuristic = a.stats()["memory"]
    del a
    g = h2w.Initialisation(3, N, 64, 1.0, tol_pow, h2w.CHOICE_RBF, h2w.NESTED, points=P)
    g.set_rank_model(params)
    g.tree_initialisation()
    st = g.stats()
    print(f"[c5 {kind}] assembly {st['assembly_seconds']:.1f} s, model blocks {st['rank_pred_blocks']}, "
          f"under-estimates {st['rank_pred_under']}, mae(log2) {st['rank_pred_mae_log2']:.3f}", flush=True)
    assert st["rank_pred_blocks"] > 0
    assert st["memory"] == mem_heuristic
    xt = np.random.default_rng(2).uniform(-1.0, 1.0, N)
    e_rows = _sampled_rows_error(g, xt)
    assert e_rows <= 10 * 10.0 ** -tol_pow, e_rows
    b = g.dense_rows(None, xt)
    x64, s64 = g.gmres(b, None, 500, 1e-10, h2w.GMRES_FP64)
    x, s = g.gmres(b, None, 500, 1e-10, h2w.GMRES_FP32_INNER_IR)
    r = _rel(g.mat_vec_prod(x), b)
    e, e64 = _rel(x, xt), _rel(x64, xt)
    print(f"[c5 {kind}] sampled rows {e_rows:.2e}; FP64 GMRES {s64['iterations']} its error {e64:.2e}; "
          f"refined {s['iterations']}+{s['outer_iterations']} its, resid {r:.2e}, error {e:.2e}", flush=True)
    assert s["converged"] and r <= 2e-10
    assert e <= 10 * e64 + 1e-12 and e <= 1e-9
