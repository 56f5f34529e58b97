"""GPU: the complex kernels (3D Helmholtz e^{iR}/(4 pi), choice 2; 2D Hankel (i/4) H0^(1)(R), choice 10)
against the complex oracle (oracle/complex_ref.py): dense rows, the skeletons of the complex cross
(leaf-level far field and level-1 vertex field, where the row and column sets are the boxes' own
points), the compressed product against the dense product, complex symmetry of the operator, and
complex GMRES against the oracle's GMRES on the dense matrix and x_true."""
import numpy as np
import pytest

import h2weak_amd as h2w
import oracle
from oracle import complex_ref as cr

pytestmark = pytest.mark.gpu

# dim, N, n_max, tol_pow, choice, variant, points
CASES = [
    (3, 16 ** 3, 27, 8, cr.HELMHOLTZ, 0, "grid"),
    (3, 16 ** 3, 27, 8, cr.HELMHOLTZ, 1, "grid"),
    (3, 16 ** 3, 27, 8, cr.HELMHOLTZ, 2, "grid"),
    (3, 3000, 32, 10, cr.HELMHOLTZ, 0, "rand"),
    (2, 48 * 48, 36, 10, cr.HANKEL, 0, "grid"),
    (2, 48 * 48, 36, 10, cr.HANKEL, 1, "grid"),
    (2, 2500, 40, 12, cr.HANKEL, 0, "rand"),
]


def _ids(c):
    return (f"{c[0]}d-N{c[1]}-tol{c[3]}-{'helmholtz' if c[4] == cr.HELMHOLTZ else 'hankel'}-"
            f"{['nested', 'semi', 'hstar'][c[5]]}-{c[6]}")


@pytest.fixture(scope="module")
def built():
    cache = {}

    def get(case):
        if case not in cache:
            dim, N, nmax, tp, ch, var, kind = case
            P = (cr.grid_points(dim, N) if kind == "grid"
                 else np.random.default_rng(11).uniform(-1, 1, (dim, N)))
            g = h2w.Initialisation(dim, N, nmax, 1.0, tp, ch, var, points=P)
            g.tree_initialisation()
            A = cr.dense(P, ch)
            cache[case] = (g, A, P)
        return cache[case]

    return get


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("case", CASES[::2], ids=_ids)
def test_dense_rows(built, case):
    g, A, P = built(case)
    x = np.random.default_rng(1).normal(size=g.N) + 1j * np.random.default_rng(2).normal(size=g.N)
    rows = np.random.default_rng(3).choice(g.N, 300, replace=False)
    np.testing.assert_allclose(g.dense_rows(rows, x), A[rows] @ x, rtol=1e-12, atol=1e-12 * np.abs(A @ x).max())
    assert _rel(g.dense_rows(None, x), A @ x) <= 1e-13


@pytest.mark.parametrize("case", [c for c in CASES if c[5] == 0], ids=_ids)
def test_skeletons_match_the_complex_cross(built, case):
    """Leaf-level far field and level-1 vertex field, where rows and columns are the boxes' own points:
    the GPU's complex cross against the oracle's on the same block -- the same rank (to within a step or
    two), a cross approximation as accurate as the oracle's, and identical pivots where the pivot choice
    is well defined.  (3D Helmholtz entries all have modulus 1/(4 pi), so the first pivot of every
    block is an argmax over entries equal up to rounding, decided by the last bits of cos and sin; grids
    add exact ties.  Only random 2D points make every argmax unique.)"""
    g, A, P = built(case)
    dim, N, nmax, tp, ch, var, kind = case
    o = oracle.Oracle(dim, N, nmax, 1.0, tp, 1, var, points=P)  # (its tree and lists: kernel-independent)
    nl = o.nlevels
    perm = o.tree_order()
    tol = 10.0 ** -tp

    def box_points(level):
        cnt = [o.box_count(level, k) for k in range((1 << dim) ** level)]
        start = np.concatenate([[0], np.cumsum(cnt)])
        return [perm[start[k]:start[k + 1]] for k in range(len(cnt))]

    def cross_err(B, rb, cb):
        approx = B[:, cb] @ np.linalg.solve(B[np.ix_(rb, cb)], B[rb, :])
        return np.linalg.norm(B - approx) / np.linalg.norm(B)

    blocks = []
    leaves = box_points(nl)
    for b in range(0, len(leaves), 7):
        il, vf = o.interaction_list(nl, b)
        far = [leaves[f] for f, v in zip(il, vf) if not v]
        if len(leaves[b]) and far:
            blocks.append((0, nl, b, leaves[b], np.concatenate(far).astype(int)))
    lev1 = box_points(1)
    for b in range(len(lev1)):
        il, vf = o.interaction_list(1, b)
        blocks.append((1, 1, b, lev1[b], np.concatenate([lev1[f] for f, v in zip(il, vf) if v]).astype(int)))
    same = 0
    for fam, lev, b, rows, cols in blocks:
        B = A[np.ix_(rows, cols)]
        r, rb, cb, L, R, _ = cr.aca(B, tol)
        gi, gc = g.skeleton(fam, lev, b)
        same += (list(rows[rb]) == gi and list(cols[cb]) == gc)
        pos_r = {v: i for i, v in enumerate(rows)}
        pos_c = {v: i for i, v in enumerate(cols)}
        grb, gcb = [pos_r[v] for v in gi], [pos_c[v] for v in gc]
        assert abs(len(gi) - r) <= max(2, r // 10), (fam, lev, b, len(gi), r)
        e_o, e_g = cross_err(B, rb, cb), cross_err(B, grb, gcb)
        assert e_g <= 10 * e_o + 10 * tol, (fam, lev, b, e_g, e_o)
    print(f"{same}/{len(blocks)} complex skeletons equal the oracle's")
    if dim == 2 and kind == "rand":
        assert same / len(blocks) >= 0.9


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_product_and_symmetry(built, case):
    g, A, P = built(case)
    tol = 10.0 ** -case[3]
    rng = np.random.default_rng(4)
    x = rng.normal(size=g.N) + 1j * rng.normal(size=g.N)
    z = rng.normal(size=g.N) + 1j * rng.normal(size=g.N)
    y = g.mat_vec_prod(x)
    e = _rel(y, A @ x)
    print(f"{_ids(case)}: product error against the dense operator {e:.2e}")
    assert e <= 30 * tol
    # linearity, and the several-right-hand-side form column by column
    assert _rel(g.mat_vec_prod(2 * x - 3j * z), 2 * y - 3j * g.mat_vec_prod(z)) <= 1e-13
    Y = g.matvec_multi(np.stack([x, z]))
    assert _rel(Y[0], y) <= 1e-14
    if case[5] == 0:
        # complex symmetric (not Hermitian): z^T A x = x^T A z
        assert abs(z @ y - x @ g.mat_vec_prod(z)) <= 1e-12 * np.linalg.norm(z) * np.linalg.norm(y)


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_complex_gmres(built, case):
    """GMRES on the compressed complex operator: the FP64 residual meets the tolerance, and the error
    against x_true stays within 10x of the oracle GMRES on the dense matrix."""
    g, A, P = built(case)
    gtol = 10.0 ** -min(case[3], 10)
    rng = np.random.default_rng(5)
    xt = rng.normal(size=g.N) + 1j * rng.normal(size=g.N)
    b = A @ xt
    xo, ito, ro = cr.gmres(A, b, None, 300, gtol)
    x, st = g.gmres(b, None, 300, gtol)
    eo, eg = _rel(xo, xt), _rel(x, xt)
    print(f"{_ids(case)}: GPU {st['iterations']} its resid {st['resid']:.2e} error {eg:.2e}; oracle {ito} its "
          f"error {eo:.2e}")
    assert st["converged"] and st["resid"] < gtol
    assert _rel(g.mat_vec_prod(x), b) <= 5 * gtol + 30 * 10.0 ** -case[3]
    assert abs(st["iterations"] - ito) <= max(2, ito // 5)
    assert eg <= 10 * eo + 50 * 10.0 ** -case[3]


def test_complex_refusals(built):
    g, A, P = built(CASES[0])
    b = np.ones(g.N, dtype=complex)
    for mode in (h2w.GMRES_FP32_IR, h2w.GMRES_FP32_INNER_IR):
        with pytest.raises(h2w.H2WError) as e:
            g.gmres(b, None, 10, 1e-8, mode)
        assert e.value.code == h2w.H2W_EINVAL
    with pytest.raises(h2w.H2WError) as e:
        g.mat_vec_prod(b, precision=h2w.FP32)
    assert e.value.code == h2w.H2W_EINVAL
    with pytest.raises(h2w.H2WError):
        g.set_inner_tolerance(4)
    with pytest.raises(h2w.H2WError):
        h2w.Batch([g])
