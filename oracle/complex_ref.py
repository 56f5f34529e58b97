"""Complex-kernel part of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

A numpy restatement of the reference's compile-time complex builds, used as the
checker for the HIP library's complex path (only ``tests/`` may import it):

  3D Helmholtz (-DUSE_Helm, ker_choice 2) ... 3D/codes/kernel.hpp:132-144: e^{iR} / (4 pi),
                                              diagonal 500 sqrt(N)
  2D Hankel (-DUSE_Hankel, ker_choice 10) ... 2D/codes/kernel.hpp:126-145: (i/4) H0^(1)(R) =
                                              (i J0(R) - Y0(R)) / 4, diagonal (1 + 15 log N + Kii) / h^2,
                                              Kii = 4 GL16^2 of the same integrand on [0, h/2]^2
                                              (2D/codes/integral2d.hpp:16-57,60-90)
  complex partial-pivot cross ............... 2D/codes/aca.hpp:23-38,296-446 with complex entries
                                              (pivots by modulus, Eigen's conjugate-linear dot)
  complex GMRES .............................. 1D/codes/gmres.hpp:52-67,91-200, with the Hermitian
                                              inner product v_k^H w and the rotation
                                              [c, s; -conj(s), c] (the reference's complex branch
                                              forms w.dot(v_k) = w^H v_k, the conjugate of v_k^H w,
                                              and applies -s where the unitary rotation has
                                              -conj(s), 1D/codes/gmres.hpp:70-74,150; this is the
                                              textbook complex GMRES instead)

The Bessel functions come from scipy.special (the reference uses Boost.Math,
which is not available here); their values are pinned against
scipy.special.hankel1 and closed-form limits in the CPU tests.
"""
import numpy as np
from scipy import special

import oracle

HELMHOLTZ, HANKEL = 2, 10


def grid_points(dim, N):
    """The reference's uniform grid (3D/codes/HODLR3Dtree.hpp:82-102, 2D/codes/HODLR2Dtree.hpp:75-91),
    dim x N in input order."""
    n = int(round(N ** (1.0 / dim)))
    nd = -1.0 + 2.0 * (np.arange(n) + 1.0) / (n + 1.0)
    if dim == 2:
        return np.stack([np.tile(nd, n), np.repeat(nd, n)])
    J, K, I = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return np.stack([nd[K.ravel()], nd[J.ravel()], nd[I.ravel()]])


def kii_hankel(N):
    """4 * GL16^2 of (i/4) H0^(1)(R) on [0, h/2]^2, h = N^(-1/2)."""
    nodes, weights = oracle.gl16()
    h = 1.0 / np.sqrt(N)
    cx = 0.25 * h
    t = cx * nodes + cx
    X, Y = np.meshgrid(t, t, indexing="ij")
    R = np.sqrt(X * X + Y * Y)
    W = np.outer(weights, weights)
    f = (-special.y0(R) + 1j * special.j0(R)) / 4.0
    return 4.0 * cx * cx * np.sum(W * f)


def diag(dim, choice, N):
    if choice == HELMHOLTZ:
        return complex(500.0 * np.sqrt(N))
    h = 1.0 / np.sqrt(N)
    return (1.0 + 15.0 * np.log(N) + kii_hankel(N)) / h ** 2


def koff(choice, R):
    """Off-diagonal entries as functions of the distance R."""
    if choice == HELMHOLTZ:
        return np.exp(1j * R) / (4.0 * np.pi)
    return (-special.y0(R) + 1j * special.j0(R)) / 4.0


def block(P, choice, rows, cols):
    """K(rows, cols) (input-order indices), diagonal rule where the indices coincide."""
    rows, cols = np.asarray(rows), np.asarray(cols)
    D = P[:, rows][:, :, None] - P[:, cols][:, None, :]
    R = np.sqrt(np.sum(D * D, axis=0))
    same = rows[:, None] == cols[None, :]
    Rs = np.where(same, 1.0, R)
    A = koff(choice, Rs)
    if np.any(same):
        A = np.where(same, diag(P.shape[0], choice, P.shape[1]), A)
    return A


def dense(P, choice):
    N = P.shape[1]
    idx = np.arange(N)
    return block(P, choice, idx, idx)


def aca(A, tol):
    """Partial-pivot cross of the dense block A (2D/codes/aca.hpp:296-446): returns rank, local row and
    column pivots, L (unit lower), R (upper) with A[rb][:, cb] = L R, and the per-step (|u||v|, normS)."""
    N1, N2 = A.shape
    mn = min(N1, N2)
    rrem = np.ones(N1, bool)
    crem = np.ones(N2, bool)

    def first_argmax(v, rem):
        a = np.where(rem, np.abs(v), -1.0)
        return int(np.argmax(a)) if rem.any() else -1

    U, V, rb, cb, hist = [], [], [], [], []
    l = 0
    while l < N1:
        row = A[l].copy()
        ci = first_argmax(row, crem)
        if abs(row[ci]) > 1e-36:
            break
        l += 1
    if l == N1:
        return 0, [], [], np.zeros((0, 0), complex), np.zeros((0, 0), complex), []
    ri = l
    v = row
    col = A[:, ci].copy()
    u = col / row[ci]
    U.append(u)
    V.append(v)
    rb.append(ri)
    cb.append(ci)
    crem[ci] = False
    rrem[ri] = False
    rank = 1
    normS = 0.0
    hist.append((np.linalg.norm(u) * np.linalg.norm(v), 0.0))
    ri = first_argmax(col, rrem)
    while np.linalg.norm(u) * np.linalg.norm(v) > tol * normS and rank < mn:
        if ri < 0:
            break
        row = A[ri].copy()
        for q in range(rank):
            row = row - U[q][ri] * V[q]
        ci = first_argmax(row, crem)
        if ci < 0:
            break
        v = row
        col = A[:, ci].copy()
        for q in range(rank):
            col = col - V[q][ci] * U[q]
        u = col / row[ci]
        un, vn = np.linalg.norm(u), np.linalg.norm(v)
        if un < 1e-36 or vn < 1e-36:
            break
        U.append(u)
        V.append(v)
        rb.append(ri)
        cb.append(ci)
        rank += 1
        crem[ci] = False
        rrem[ri] = False
        normS = normS ** 2 + (un * vn) ** 2
        for q in range(rank):
            normS += 2 * abs(np.vdot(U[q], u) * np.vdot(V[q], v))
        normS = np.sqrt(normS)
        hist.append((un * vn, normS))
        ri = first_argmax(col, rrem)
    L = np.eye(rank, dtype=complex)
    R = np.zeros((rank, rank), dtype=complex)
    for i in range(rank):
        for j in range(i):
            L[i, j] = U[j][rb[i]]
        for j in range(i + 1):
            R[j, i] = V[j][cb[i]]
    return rank, rb, cb, L, R, hist


def gmres(A, b, x0=None, max_iter=500, tol=1e-10):
    """Unrestarted GMRES (1D/codes/gmres.hpp:91-200) on a complex dense matrix: modified Gram-Schmidt with
    h = v_k^H w, complex Givens rotations, resid = |s_{i+1}| / ||b||.  Returns x, iterations, resid."""
    N = b.size
    x = np.ones(N, complex) if x0 is None else np.array(x0, complex)
    normb = np.linalg.norm(b) or 1.0
    r = b - A @ x
    beta = np.linalg.norm(r)
    if beta / normb <= tol:
        return x, 0, beta / normb
    m = max_iter
    Vb = [r / beta]
    H = np.zeros((m + 1, m), complex)
    s = np.zeros(m + 1, complex)
    cs = np.zeros(m + 1, complex)
    sn = np.zeros(m + 1, complex)
    s[0] = beta

    def rot(dx, dy, c, sg):
        return c * dx + sg * dy, -np.conj(sg) * dx + c * dy

    resid = beta / normb
    for i in range(m):
        w = A @ Vb[i]
        for k in range(i + 1):
            H[k, i] = np.vdot(Vb[k], w)
            w = w - H[k, i] * Vb[k]
        H[i + 1, i] = np.linalg.norm(w)
        Vb.append(w / H[i + 1, i] if H[i + 1, i] != 0 else np.zeros(N, complex))
        for k in range(i):
            H[k, i], H[k + 1, i] = rot(H[k, i], H[k + 1, i], cs[k], sn[k])
        dx, dy = H[i, i], H[i + 1, i]
        if dx == 0:
            cs[i], sn[i] = 0.0, 1.0
        else:
            nrm = np.hypot(abs(dx), abs(dy))
            cs[i] = abs(dx) / nrm
            sn[i] = (dx / abs(dx)) * np.conj(dy) / nrm
        H[i, i], H[i + 1, i] = rot(H[i, i], H[i + 1, i], cs[i], sn[i])
        s[i], s[i + 1] = rot(s[i], s[i + 1], cs[i], sn[i])
        resid = abs(s[i + 1]) / normb
        if resid < tol or i == m - 1:
            y = np.linalg.solve(np.triu(H[: i + 1, : i + 1]), s[: i + 1])
            x = x + np.stack(Vb[: i + 1], axis=1) @ y
            return x, i + 1, resid
    return x, m, resid
