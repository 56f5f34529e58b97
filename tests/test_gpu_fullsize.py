"""GPU at full size: the reference's eight published runs and the benchmark
configurations, where the oracle would take minutes to hours.  Checked through
what does not depend on size: the storage count against the published one and
the oracle's (tests/golden/transcripts_oracle.json, the oracle run at these
configurations), iteration counts, linearity, symmetry of the nested operator,
sampled rows against the dense FP64 product (itself parity-tested against the
oracle), and the solution error against x_true."""
import json
import os

import numpy as np
import pytest

import h2weak_amd as h2w

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
RUNS = json.load(open(os.path.join(HERE, "golden", "transcripts_oracle.json")))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _resid(A, x, b):
    """FP64 residual of x with the stored FP64 operator, relative to b."""
    return float(np.linalg.norm(b - A.mat_vec_prod(x)) / np.linalg.norm(b))


def _sampled_rows_error(A, x, nrows=1024, seed=4):
    """Fast product against the dense FP64 product on rows sampled without replacement."""
    rows = np.sort(np.random.default_rng(seed).choice(A.N, nrows, replace=False))
    return _rel(A.mat_vec_prod(x)[rows], A.dense_rows(rows, x))


def _linearity_error(A, x, z):
    return _rel(A.mat_vec_prod(2 * x - 3 * z), 2 * A.mat_vec_prod(x) - 3 * A.mat_vec_prod(z))


def _symmetry_error(A, x, z):
    Ax, Az = A.mat_vec_prod(x), A.mat_vec_prod(z)
    return abs(z @ Ax - x @ Az) / (np.linalg.norm(z) * np.linalg.norm(Ax))


@pytest.mark.parametrize("name", sorted(RUNS))
def test_published_runs(name):
    """Storage equal to the published value (to its printed digits) and to the
    oracle's count; FP64 GMRES iterations within one of the published run and
    the oracle's, residual below the tolerance, error within 10x; the FP32-inner
    refinement reaching the same tolerance in the FP64 residual."""
    r = RUNS[name]
    p = r["published"]
    A = h2w.Initialisation(p["dim"], p["N"], p["n_max"], 1.0, p["tol_pow"], p["choice"], p["variant"])
    A.tree_initialisation()
    st = A.stats()
    gb = st["memory"] * 8e-9
    assert abs(gb - p["storage_gb"]) <= 1e-5 * p["storage_gb"] + 1e-6, (gb, p["storage_gb"])
    assert st["memory"] == pytest.approx(r["memory"], rel=1e-6)
    assert st["nlevels"] == r["nlevels"]
    xt = np.random.default_rng(2).uniform(-1.0, 1.0, p["N"])
    b = A.dense_rows(None, xt)
    x, s = A.gmres(b, None, p["max_iter"], p["gtol"], h2w.GMRES_FP64)
    err = _rel(x, xt)
    print(f"{name}: {gb:.6f} GB, {s['iterations']} its (published {p['iters']}, oracle {r['iters']}), "
          f"resid {s['resid']:.2e}, error {err:.2e} (published {p['err']:.2e}, oracle {r['err']:.2e})")
    assert s["converged"] and s["resid"] < p["gtol"]
    assert abs(s["iterations"] - p["iters"]) <= 1 and abs(s["iterations"] - r["iters"]) <= 1
    assert err <= 10 * max(p["err"], r["err"])
    x32, s32 = A.gmres(b, None, p["max_iter"], p["gtol"], h2w.GMRES_FP32_INNER_IR)
    assert s32["converged"] and s32["resid"] <= p["gtol"]
    assert _resid(A, x32, b) <= 2 * p["gtol"]
    assert _rel(x32, xt) <= 10 * max(p["err"], r["err"])


def test_1d_rbf_65536():
    """1D RBF interpolation, N = 65536 uniform random points (seed 1), n_max 100,
    ACA tolerance 1e-12, GMRES to 1e-12 from x0 = 1, x_true ~ U(-1, 1) (seed 2)."""
    N = 65536
    P = np.random.default_rng(1).uniform(-1, 1, (1, N))
    A = h2w.Initialisation(1, N, 100, 1.0, 12, h2w.CHOICE_RBF, h2w.NESTED, points=P)
    A.tree_initialisation()
    assert A.stats()["nlevels"] == 10
    rng = np.random.default_rng(2)
    xt = rng.uniform(-1.0, 1.0, N)
    e_rows = _sampled_rows_error(A, xt)
    e_lin = _linearity_error(A, xt, rng.normal(size=N))
    e_sym = _symmetry_error(A, xt, rng.normal(size=N))
    b = A.dense_rows(None, xt)
    x, s = A.gmres(b, None, 500, 1e-12, h2w.GMRES_FP64)
    err = _rel(x, xt)
    print(f"rows {e_rows:.2e}, linearity {e_lin:.2e}, symmetry {e_sym:.2e}, {s['iterations']} its, error {err:.2e}")
    assert e_rows <= 1e-10 and e_lin <= 1e-13 and e_sym <= 1e-13
    assert s["converged"] and s["resid"] < 1e-12 and err <= 1e-10


@pytest.mark.parametrize("choice", [h2w.CHOICE_IE, h2w.CHOICE_RBF])
def test_1d_250000_random(choice):
    """1D H2*, N = 250000 uniform random points, n_max 100, tolerance 1e-12, both
    kernels (the integral equation with its diagonal rule, and the RBF)."""
    N = 250000
    P = np.random.default_rng(1).uniform(-1, 1, (1, N))
    A = h2w.Initialisation(1, N, 100, 1.0, 12, choice, h2w.NESTED, points=P)
    A.tree_initialisation()
    assert A.stats()["nlevels"] == 12
    xt = np.random.default_rng(2).uniform(-1.0, 1.0, N)
    e_rows = _sampled_rows_error(A, xt)
    b = A.dense_rows(None, xt)
    x, s = A.gmres(b, None, 500, 1e-12, h2w.GMRES_FP64)
    err = _rel(x, xt)
    print(f"choice {choice}: rows {e_rows:.2e}, {s['iterations']} its, error {err:.2e}")
    assert e_rows <= 1e-10
    # (the stopping test is GMRES's recurrence estimate, as in the reference; the residual
    # recomputed with the product is bounded by that product's FP64 rounding, which at
    # N = 250000 lies above 1e-12: measured 1.4e-11 and 7e-12, and the oracle's GMRES
    # leaves the same gap, 7.9e-14 estimated against 2.9e-11 recomputed, for the first kernel)
    assert s["converged"] and s["resid"] < 1e-12 and _resid(A, x, b) <= 1e-10
    # (interpolation on random points: nearby points make the RBF system far worse
    # conditioned than on the grid, so its error stays well above the residual)
    assert err <= (1e-10 if choice == h2w.CHOICE_IE else 1e-6)


def test_2d_semi_nested_ie_400x400():
    """2D (H2+H)* integral equation on the reference's 400 x 400 grid, n_max 64,
    tolerance 1e-10; x_true ~ N(0, 1) (seed 3); the FP32 refinements reach 1e-10
    in the FP64 residual with the FP64 solve's error."""
    N = 400 * 400
    A = h2w.Initialisation(2, N, 64, 1.0, 10, h2w.CHOICE_IE, h2w.SEMI_NESTED)
    A.tree_initialisation()
    assert A.stats()["nlevels"] == 6
    rng = np.random.default_rng(3)
    xt = rng.normal(size=N)
    e_rows = _sampled_rows_error(A, xt)
    e_sym = _symmetry_error(A, xt, rng.normal(size=N))
    b = A.dense_rows(None, xt)
    x64, s64 = A.gmres(b, None, 500, 1e-10, h2w.GMRES_FP64)
    e64 = _rel(x64, xt)
    print(f"rows {e_rows:.2e}, symmetry {e_sym:.2e}, FP64 {s64['iterations']} its, error {e64:.2e}")
    assert e_rows <= 1e-8 and e_sym <= 1e-9
    assert s64["converged"] and s64["resid"] < 1e-10
    for mode in (h2w.GMRES_FP32_IR, h2w.GMRES_FP32_INNER_IR):
        x, s = A.gmres(b, None, 500, 1e-10, mode)
        e = _rel(x, xt)
        print(f"mode {mode}: {s['iterations']} inner its, {s['outer_iterations']} refinements, error {e:.2e}")
        assert s["converged"] and _resid(A, x, b) <= 2e-10
        assert e <= 10 * e64 + 1e-12


def test_3d_rbf_grid64():
    """3D H2* RBF on the 64^3 grid (N = 262144), n_max 64, tolerance 1e-10: the
    nested operator is symmetric to rounding, and the FP32-inner refinement
    reaches a relative solution error of 1e-10 or better, as FP64 GMRES does."""
    N = 64 ** 3
    A = h2w.Initialisation(3, N, 64, 1.0, 10, h2w.CHOICE_RBF, h2w.NESTED)
    A.tree_initialisation()
    assert A.stats()["nlevels"] == 4
    rng = np.random.default_rng(2)
    xt = rng.uniform(-1.0, 1.0, N)
    z = rng.normal(size=N)
    e_rows = _sampled_rows_error(A, xt)
    e_lin = _linearity_error(A, xt, z)
    e_sym = _symmetry_error(A, xt, z)
    b = A.dense_rows(None, xt)
    x64, s64 = A.gmres(b, None, 500, 1e-10, h2w.GMRES_FP64)
    x32, s32 = A.gmres(b, None, 500, 1e-10, h2w.GMRES_FP32_INNER_IR)
    e64, e32 = _rel(x64, xt), _rel(x32, xt)
    print(f"rows {e_rows:.2e}, linearity {e_lin:.2e}, symmetry {e_sym:.2e}; FP64 {s64['iterations']} its "
          f"error {e64:.2e}; FP32-inner IR {s32['iterations']} + {s32['matvecs_fp64']} FP64 products error {e32:.2e}")
    assert e_rows <= 1e-8 and e_lin <= 1e-13 and e_sym <= 1e-13
    assert s64["converged"] and s64["resid"] < 1e-10 and e64 <= 1e-10
    assert s32["converged"] and _resid(A, x32, b) <= 2e-10 and e32 <= 1e-10
