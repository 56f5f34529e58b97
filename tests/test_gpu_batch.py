"""GPU: batches of independent systems.  A batched product (one launch per
phase for every system, 16 right-hand sides per pass) equals each system's own
16-wide product and matches the oracle's product of the same system; lockstep
solves over all (system, right-hand side) columns of right-hand sides from the
dense FP64 product reach each column's tolerance with the single-system solve's
accuracy.  The systems differ in size, depth, kernel and format."""
import numpy as np
import pytest

import h2weak_amd as h2w

pytestmark = pytest.mark.gpu

# dim, N, n_max, tol_pow, choice, variant, seed
SYSTEMS = [
    (3, 2000, 64, 8, 1, 0, 1),
    (3, 9000, 64, 8, 1, 0, 2),    # one level deeper
    (3, 5000, 64, 8, 0, 1, 3),    # semi-nested IE
    (2, 3000, 40, 8, 1, 0, 4),
]


@pytest.fixture(scope="module")
def batch():
    sys_ = []
    for dim, N, nmax, tp, ch, var, seed in SYSTEMS:
        P = np.random.default_rng(seed).uniform(-1, 1, (dim, N))
        g = h2w.Initialisation(dim, N, nmax, 1.0, tp, ch, var, points=P)
        g.tree_initialisation()
        sys_.append(g)
    return h2w.Batch(sys_), sys_


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_depths_differ(batch):
    _, sys_ = batch
    assert len({g.stats()["nlevels"] for g in sys_}) >= 2


def test_batched_product_against_oracle(batch):
    """The batched FP64 product of every system against the oracle's product of the same system
    (same skeletons: the pivots are the oracle's) and against the dense FP64 product."""
    import oracle
    B, sys_ = batch
    Xs = [np.random.default_rng(11).normal(size=(2, g.N)) for g in sys_]
    Ys = B.matvec(Xs)
    for (dim, N, nmax, tp, ch, var, seed), X, Y, g in zip(SYSTEMS, Xs, Ys, sys_):
        P = np.random.default_rng(seed).uniform(-1, 1, (dim, N))
        o = oracle.Oracle(dim, N, nmax, 1.0, tp, ch, var, points=P)
        o.tree_initialisation()
        for c in range(2):
            yo, yd = o.mat_vec_prod(X[c]), g.dense_rows(None, X[c])
            assert _rel(Y[c], yo) <= 1e-11, (N, c, _rel(Y[c], yo))
            assert _rel(Y[c], yd) <= 10 * _rel(yo, yd) + 1e-14


@pytest.mark.parametrize("nrhs", [3, 16, 19])
def test_batched_product(batch, nrhs):
    B, sys_ = batch
    rng = np.random.default_rng(nrhs)
    Xs = [rng.normal(size=(nrhs, g.N)) for g in sys_]
    for prec, tol in ((h2w.FP64, 1e-13), (h2w.FP32, 1e-6), (h2w.FP32_INNER, 1e-6)):
        Ys = B.matvec(Xs, precision=prec)
        for X, Y, g in zip(Xs, Ys, sys_):
            ref = g.matvec_multi(X, precision=prec)
            for c in range(nrhs):
                assert _rel(Y[c], ref[c]) <= tol, (prec, c)


def test_batched_solves(batch):
    B, sys_ = batch
    rng = np.random.default_rng(7)
    Xt = [rng.normal(size=(4, g.N)) for g in sys_]
    # right-hand sides from the dense FP64 product (not the compressed one), so errors are against x_true
    Bs = [np.stack([g.dense_rows(None, x[c]) for c in range(4)]) for g, x in zip(sys_, Xt)]
    gtol = 1e-8
    # accuracy reference: each column's own single FP64 solve to the same tolerance
    e64 = [[_rel(g.gmres(Bs[s][c], None, 300, gtol, h2w.GMRES_FP64)[0], Xt[s][c]) for c in range(4)]
           for s, g in enumerate(sys_)]
    for mode in (h2w.GMRES_FP64, h2w.GMRES_FP32_IR, h2w.GMRES_FP32_INNER_IR):
        Xs, st = B.gmres(Bs, None, 300, gtol, mode)
        for s, g in enumerate(sys_):
            for c in range(4):
                assert st[s][c]["converged"], (mode, s, c, st[s][c])
                assert _rel(g.mat_vec_prod(Xs[s][c]), Bs[s][c]) <= 2 * gtol
                assert _rel(Xs[s][c], Xt[s][c]) <= 10 * e64[s][c] + 1e-12, (mode, s, c)


def test_assemble_many_equals_one_by_one():
    """Systems assembled together from several host threads (one stream each) have the operators of
    one-by-one assembly: the same skeletons and the same products."""
    rng = np.random.default_rng(44)
    pts = [rng.uniform(-1, 1, (3, n)) for n in (2500, 4096, 3100, 5000, 2000, 3600)]
    one = []
    for P in pts:
        g = h2w.Initialisation(3, P.shape[1], 40, 1.0, 8, h2w.CHOICE_RBF, h2w.NESTED, points=P)
        g.tree_initialisation()
        one.append(g)
    many = [h2w.Initialisation(3, P.shape[1], 40, 1.0, 8, h2w.CHOICE_RBF, h2w.NESTED, points=P) for P in pts]
    h2w.assemble_many(many, 3)
    for a, b in zip(one, many):
        assert a.stats()["memory"] == b.stats()["memory"]
        nl = a.stats()["nlevels"]
        for fam, lev in ((0, nl), (1, 1)):
            for box in range(0, 8 ** lev, 5):
                assert a.skeleton(fam, lev, box) == b.skeleton(fam, lev, box)
        x = np.random.default_rng(2).normal(size=a.N)
        ya, yb = a.mat_vec_prod(x), b.mat_vec_prod(x)
        assert np.linalg.norm(ya - yb) <= 1e-13 * np.linalg.norm(ya)
    with pytest.raises(h2w.H2WError) as e:
        h2w.assemble_many([many[0], many[0]], 2)
    assert e.value.code == h2w.H2W_EINVAL
