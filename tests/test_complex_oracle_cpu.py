"""CPU: the complex part of the oracle (oracle/complex_ref.py) pinned before it
checks the GPU -- kernel entries against scipy's Hankel function and closed
forms, the complex diagonal quadrature against adaptive quadrature, the complex
cross's invariants by brute force, complex GMRES against a direct solve -- and
the library's argument checks for the complex choices (no GPU needed)."""
import numpy as np
import pytest
from scipy import integrate, special

import h2weak_amd as h2w
from oracle import complex_ref as cr


def test_kernel_entries():
    R = np.array([1e-3, 0.05, 0.3, 1.0, 2.7])
    # 2D: (i/4) H0^(1)(R), against scipy's Hankel function of the first kind
    np.testing.assert_allclose(cr.koff(cr.HANKEL, R), 0.25j * special.hankel1(0, R), rtol=1e-13)
    # small-argument limits: J0 -> 1, Y0 ~ (2/pi)(log(R/2) + gamma)
    v = cr.koff(cr.HANKEL, np.array([1e-8]))[0]
    assert abs(v.imag - 0.25) < 1e-12
    assert abs(v.real + (2 / np.pi) * (np.log(0.5e-8) + np.euler_gamma) / 4) < 1e-12
    # 3D: e^{iR} / (4 pi) (no 1/R: the reference's Helmholtz3D), modulus 1/(4 pi)
    z = cr.koff(cr.HELMHOLTZ, R)
    np.testing.assert_allclose(np.abs(z), 1 / (4 * np.pi), rtol=1e-14)
    np.testing.assert_allclose(np.angle(z), R, rtol=1e-14)


def test_hankel_diagonal_quadrature():
    """Kii = 4 GL16^2 of (i/4) H0^(1) on [0, h/2]^2: the reference's 16-point rule on a log-singular
    integrand, close to the adaptive integral (the imaginary part, J0 / 4, is smooth: near exact)."""
    for N in (2304, 160000):
        h = 1 / np.sqrt(N)
        k = cr.kii_hankel(N)
        re = integrate.dblquad(lambda y, x: -special.y0(np.hypot(x, y)) / 4, 0, h / 2, 0, h / 2, epsabs=0,
                               epsrel=1e-12)[0] * 4
        im = integrate.dblquad(lambda y, x: special.j0(np.hypot(x, y)) / 4, 0, h / 2, 0, h / 2, epsabs=0,
                               epsrel=1e-13)[0] * 4
        assert abs(k.imag - im) <= 1e-13 * abs(im)
        assert abs(k.real - re) <= 2e-3 * abs(re)
        d = cr.diag(2, cr.HANKEL, N)
        np.testing.assert_allclose(d, (1 + 15 * np.log(N) + k) / h ** 2, rtol=1e-15)
    assert cr.diag(3, cr.HELMHOLTZ, 4096) == 500 * 64


def test_complex_cross_invariants():
    rng = np.random.default_rng(3)
    # exact rank on a complex rank-3 block; L R reproduces the pivot block
    A = (rng.normal(size=(60, 3)) + 1j * rng.normal(size=(60, 3))) @ (rng.normal(size=(3, 80))
                                                                      + 1j * rng.normal(size=(3, 80)))
    r, rb, cb, L, R, _ = cr.aca(A, 1e-12)
    # (the stopping test runs after a step is accepted, 2D/codes/aca.hpp:363,394-399: one rounding-level
    # term may follow the exact ones)
    assert r in (3, 4)
    np.testing.assert_allclose(L @ R, A[np.ix_(rb, cb)], rtol=1e-12, atol=1e-12 * np.abs(A).max())
    approx = A[:, cb[:3]] @ np.linalg.solve(A[np.ix_(rb[:3], cb[:3])], A[rb[:3], :])
    assert np.linalg.norm(A - approx) <= 1e-12 * np.linalg.norm(A)
    assert np.allclose(np.triu(L, 1), 0) and np.allclose(np.diag(L), 1) and np.allclose(np.tril(R, -1), 0)
    # a zero block has rank 0
    assert cr.aca(np.zeros((5, 7), complex), 1e-10)[0] == 0
    # well-separated Helmholtz and Hankel blocks: the cross meets the tolerance (brute force)
    for choice, dim in ((cr.HELMHOLTZ, 3), (cr.HANKEL, 2)):
        P = np.concatenate([rng.uniform(-1, -0.5, (dim, 150)), rng.uniform(0.5, 1, (dim, 170))], axis=1)
        B = cr.block(P, choice, np.arange(150), np.arange(150, 320))
        for tol in (1e-6, 1e-10):
            r, rb, cb, L, R, hist = cr.aca(B, tol)
            Ac, Ar = B[:, cb], B[rb, :]
            approx = Ac @ np.linalg.solve(L @ R, Ar)
            err = np.linalg.norm(B - approx) / np.linalg.norm(B)
            assert 0 < r < 150 and err < 50 * tol, (choice, tol, r, err)
            # the stop: the last accepted step is the first with |u||v| <= tol normS
            assert hist[-1][0] <= tol * hist[-1][1] or r == 150
            assert all(h[0] > tol * h[1] for h in hist[:-1])


def test_complex_gmres():
    rng = np.random.default_rng(4)
    n = 40
    A = np.eye(n) * 8 + (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    xt = rng.normal(size=n) + 1j * rng.normal(size=n)
    b = A @ xt
    x, it, res = cr.gmres(A, b, np.zeros(n), 200, 1e-12)
    assert res < 1e-12 and it <= n
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9)
    # the identity converges in one step
    x, it, res = cr.gmres(np.eye(n, dtype=complex), b, np.zeros(n), 10, 1e-12)
    assert it == 1 and np.allclose(x, b)
    # the recurrence residual is the true residual
    x, it, res = cr.gmres(A, b, None, 10, 1e-14)
    assert abs(res - np.linalg.norm(b - A @ x) / np.linalg.norm(b)) <= 1e-10


def test_library_refuses_complex_choices_in_the_wrong_dimension():
    with pytest.raises(h2w.H2WError) as e:
        h2w.Initialisation(2, 2304, 36, 1.0, 8, h2w.CHOICE_HELMHOLTZ, h2w.NESTED, plan_only=1)
    assert e.value.code == h2w.H2W_EINVAL
    with pytest.raises(h2w.H2WError) as e:
        h2w.Initialisation(3, 4096, 27, 1.0, 8, h2w.CHOICE_HANKEL, h2w.NESTED, plan_only=1)
    assert e.value.code == h2w.H2W_EINVAL
    g = h2w.Initialisation(2, 2304, 36, 1.0, 8, h2w.CHOICE_HANKEL, h2w.NESTED, plan_only=1)
    st = g.stats()
    d = cr.diag(2, cr.HANKEL, 2304)
    assert st["diag"] == pytest.approx(d.real, rel=1e-13) and st["diag_im"] == pytest.approx(d.imag, rel=1e-12)
    g = h2w.Initialisation(3, 4096, 27, 1.0, 8, h2w.CHOICE_HELMHOLTZ, h2w.NESTED, plan_only=1)
    assert g.stats()["diag"] == 500 * 64 and g.stats()["diag_im"] == 0.0
