"""GPU: GMRES edge cases of the reference's loop (1D/codes/gmres.hpp:91-200).

* A right-hand side inside an invariant subspace of dimension 3: the Arnoldi
  process breaks down (H(4,3) is zero up to rounding) and GMRES must return
  the exact solution after 3 steps without dividing by the vanishing
  subdiagonal (the reference's gap, 1D/codes/gmres.hpp:153-154); the
  refinement modes, which keep the first correction's Krylov space for the
  next ones, must stay finite.
* A zero right-hand side: converged at once with x = 0.
"""
import numpy as np
import pytest

import h2weak_amd as h2w

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    N = 500
    P = np.random.default_rng(9).uniform(-1, 1, (3, N))
    g = h2w.Initialisation(3, N, 200, 1.0, 10, h2w.CHOICE_RBF, h2w.NESTED, points=P)
    g.tree_initialisation()
    # the compressed operator as a dense matrix, 16 columns per product
    A = np.empty((N, N))
    for c0 in range(0, N, 16):
        E = np.zeros((min(16, N - c0), N))
        E[np.arange(E.shape[0]), c0 + np.arange(E.shape[0])] = 1.0
        A[:, c0:c0 + E.shape[0]] = g.matvec_multi(E).T
    return g, A


def test_invariant_subspace_breakdown(small):
    g, A = small
    N = A.shape[0]
    # the nested operator is symmetric up to rounding
    assert np.linalg.norm(A - A.T) <= 1e-13 * np.linalg.norm(A)
    lam, V = np.linalg.eigh(0.5 * (A + A.T))
    idx = [N - 1, N - 2, N // 2]
    coef = np.array([1.0, -2.0, 0.5])
    b = V[:, idx] @ coef
    x_exact = V[:, idx] @ (coef / lam[idx])
    x, st = g.gmres(b, np.zeros(N), 50, 1e-12, h2w.GMRES_FP64)
    assert np.all(np.isfinite(x))
    assert st["converged"] and st["iterations"] <= 4, st
    assert np.linalg.norm(x - x_exact) <= 1e-11 * np.linalg.norm(x_exact)
    for mode in (h2w.GMRES_FP32_IR, h2w.GMRES_FP32_INNER_IR):
        x, st = g.gmres(b, np.zeros(N), 50, 1e-10, mode)
        assert np.all(np.isfinite(x)), mode
        assert st["converged"], (mode, st)
        assert np.linalg.norm(A @ x - b) <= 2e-10 * np.linalg.norm(b)
        # a second solve reuses nothing from the first (no state leaks between solves)
        x2, st2 = g.gmres(2.0 * b, np.zeros(N), 50, 1e-10, mode)
        assert st2["converged"] and np.linalg.norm(x2 - 2.0 * x) <= 1e-8 * np.linalg.norm(x)


def test_zero_rhs(small):
    g, A = small
    N = A.shape[0]
    for mode in (h2w.GMRES_FP64, h2w.GMRES_FP32_IR, h2w.GMRES_FP32_INNER_IR):
        x, st = g.gmres(np.zeros(N), np.zeros(N), 20, 1e-10, mode)
        assert st["converged"], (mode, st)
        assert np.all(x == 0.0), mode
