"""GPU: the refinement's deflation space (h2w_set_deflation, the autotuner's TUNE_DEFLATION).
The space changes how many inner steps a correction needs, never what the refinement converges
to: every solve still ends on an FP64 residual below tol, with the accuracy of plain FP64 GMRES
(and of the oracle's GMRES on the same b)."""
import numpy as np
import pytest

import h2weak_amd as h2w
import oracle

pytestmark = pytest.mark.gpu

SIDE, TP = 24, 10  # 3D RBF grid 24^3, n_max 64, ACA and GMRES tol 1e-10
N = SIDE ** 3
TOL = 10.0 ** -TP


@pytest.fixture(scope="module")
def sys3d():
    g = h2w.Initialisation(3, N, 64, 1.0, TP, h2w.CHOICE_RBF, h2w.NESTED)
    g.tree_initialisation()
    g.set_inner_tolerance(6.5)
    xt = np.random.default_rng(2).uniform(-1, 1, N)
    b = g.dense_rows(None, xt)
    return g, xt, b


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_deflated_refinement_converges_with_fp64_accuracy(sys3d):
    g, xt, b = sys3d
    x64, s64 = g.gmres(b, None, 500, TOL, h2w.GMRES_FP64)
    g.set_deflation(0)
    x0, s0 = g.gmres(b, None, 500, TOL, h2w.GMRES_FP32_INNER_IR)
    e64, e0 = _rel(x64, xt), _rel(x0, xt)
    for m, k in ((32, 0), (64, 0), (64, 24), (96, 32)):
        g.set_deflation(m, k)
        x, s = g.gmres(b, None, 500, TOL, h2w.GMRES_FP32_INNER_IR)
        e = _rel(x, xt)
        print(f"m={m} k={k}: {s['iterations']}+{s['outer_iterations']} its (none: {s0['iterations']}+"
              f"{s0['outer_iterations']}), resid {s['resid']:.2e}, error {e:.2e} (FP64 GMRES {e64:.2e})")
        assert s["converged"] and s["resid"] < TOL
        # the FP64 residual, recomputed from x with the FP64 product
        assert _rel(g.mat_vec_prod(x), b) < TOL
        assert e <= 10 * e64
        assert s["iterations"] <= s0["iterations"]
    g.set_deflation(0)


def test_deflation_matches_the_oracle_solution(sys3d):
    g, xt, b = sys3d
    o = oracle.Oracle(3, N, 64, 1.0, TP, h2w.CHOICE_RBF, h2w.NESTED)
    o.tree_initialisation()
    xo, ito, ro, _ = o.gmres(b, None, 500, TOL)
    g.set_deflation(64, 24)
    x, s = g.gmres(b, None, 500, TOL, h2w.GMRES_FP32_INNER_IR)
    assert _rel(x, xt) <= 10 * _rel(xo, xt)
    assert _rel(x, xo) <= 1e-8  # both solve the same operator to 1e-10
    g.set_deflation(0)


def test_space_does_not_depend_on_the_right_hand_side(sys3d):
    """The space is built from the operator only: solves of other right-hand sides (and from another x0)
    converge as well, with the same kind of step count."""
    g, xt, b = sys3d
    g.set_deflation(64, 24)
    its = []
    for seed in (5, 6, 7):
        x2 = np.random.default_rng(seed).normal(size=N)
        b2 = g.dense_rows(None, x2)
        for x0 in (None, np.zeros(N)):
            x, s = g.gmres(b2, x0, 500, TOL, h2w.GMRES_FP32_INNER_IR)
            assert s["converged"] and _rel(g.mat_vec_prod(x), b2) < TOL
            assert _rel(x, x2) <= 1e-8
            its.append(s["iterations"])
    assert max(its) - min(its) <= 6, its
    g.set_deflation(0)


def test_rebuilding_the_inner_operator_invalidates_the_space(sys3d):
    g, xt, b = sys3d
    g.set_deflation(0)
    g.set_inner_tolerance(6)
    _, s_plain = g.gmres(b, None, 500, TOL, h2w.GMRES_FP32_INNER_IR)
    g.set_inner_tolerance(6.5)
    g.set_deflation(64, 24)
    _, s_defl = g.gmres(b, None, 500, TOL, h2w.GMRES_FP32_INNER_IR)
    g.set_inner_tolerance(6)  # the space belonged to the 10^-6.5 operator: not used any more
    x, s = g.gmres(b, None, 500, TOL, h2w.GMRES_FP32_INNER_IR)
    # (equal up to a step: the atomics' summation order may move a marginal residual)
    assert abs(s["iterations"] - s_plain["iterations"]) <= 1, (s, s_plain, s_defl)
    assert s["converged"] and _rel(x, xt) <= 1e-8
    g.set_inner_tolerance(6.5)


def test_stored_fp32_operator_space(sys3d):
    g, xt, b = sys3d
    _, s0 = g.gmres(b, None, 500, TOL, h2w.GMRES_FP32_IR)
    g.set_deflation(64, 0, h2w.FP32)
    x, s = g.gmres(b, None, 500, TOL, h2w.GMRES_FP32_IR)
    assert s["converged"] and s["iterations"] <= s0["iterations"] and _rel(x, xt) <= 1e-8
    # plain FP64 GMRES ignores it
    x64, s64 = g.gmres(b, None, 500, TOL, h2w.GMRES_FP64)
    assert s64["converged"] and s64["outer_iterations"] == 0
    g.set_deflation(0, 0, h2w.FP32)


def test_argument_checks(sys3d):
    g, xt, b = sys3d
    for m, k, p in ((1, 0, h2w.FP32_INNER), (513, 0, h2w.FP32_INNER), (-1, 0, h2w.FP32_INNER),
                    (16, 17, h2w.FP32_INNER), (16, -1, h2w.FP32_INNER), (16, 0, h2w.FP64),
                    (16, 0, h2w.FP64_OTF)):
        with pytest.raises(h2w.H2WError) as e:
            g.set_deflation(m, k, p)
        assert e.value.code == h2w.H2W_EINVAL, (m, k, p)
    small = h2w.Initialisation(3, 512, 27, 1.0, 8, h2w.CHOICE_RBF, h2w.NESTED)  # 8^3 grid, two levels
    small.tree_initialisation()
    small.set_inner_tolerance(6)
    with pytest.raises(h2w.H2WError) as e:
        small.set_deflation(512)  # m must be below N
    assert e.value.code == h2w.H2W_EINVAL
    small.set_deflation(16, 4)
    xs = np.random.default_rng(1).uniform(-1, 1, 512)
    bs = small.dense_rows(None, xs)
    x, s = small.gmres(bs, None, 200, 1e-8, h2w.GMRES_FP32_INNER_IR)
    assert s["converged"] and _rel(x, xs) <= 1e-6


def test_autotune_with_deflation(sys3d):
    g, xt, b = sys3d
    bt = g.dense_rows(None, np.random.default_rng(12).uniform(-1, 1, N))
    t = g.autotune(bt, tol=TOL, max_iter=500, flags=h2w.TUNE_UNROLL | h2w.TUNE_INNER | h2w.TUNE_DEFLATION)
    d = t["deflation"]
    print("deflation candidates", [(c["m"], c["k"], c["ir_tol_exp"], round(c["seconds"] * 1e3, 2), c["iterations"])
                                   for c in d["candidates"]], "chosen", d["m"], d["k"], t["ir_tol_exp"])
    assert len(d["candidates"]) >= 6 and all(c["converged"] for c in d["candidates"])
    assert (d["m"], d["k"]) in {(c["m"], c["k"]) for c in d["candidates"]}
    assert g.stats()["inner_tol_exp"] == t["ir_tol_exp"]
    x, s = g.gmres(b, None, 500, TOL, h2w.GMRES_FP32_INNER_IR)
    assert s["converged"] and _rel(g.mat_vec_prod(x), b) < TOL and _rel(x, xt) <= 1e-8
    g.set_deflation(0)
    g.set_inner_tolerance(6.5)


def test_lockstep_solves_share_the_space(sys3d):
    """16 right-hand sides refined in lockstep start every correction from the same deflation space (one
    row-major copy shared by the columns): every column converges to the single solve's accuracy, in no more
    steps than without the space."""
    g, xt, b = sys3d
    rng = np.random.default_rng(8)
    XT = rng.normal(size=(16, N))
    B = np.stack([g.dense_rows(None, x) for x in XT])
    g.set_deflation(0)
    _, st0 = g.gmres_multi(B, None, 500, TOL, h2w.GMRES_FP32_INNER_IR)
    g.set_deflation(64, 24)
    X, st = g.gmres_multi(B, None, 500, TOL, h2w.GMRES_FP32_INNER_IR)
    its0, its = max(s["iterations"] for s in st0), max(s["iterations"] for s in st)
    print(f"lockstep: {its} inner steps with the space, {its0} without")
    for c in range(16):
        assert st[c]["converged"] and st[c]["resid"] < TOL, (c, st[c])
        assert _rel(g.mat_vec_prod(X[c]), B[c]) < TOL
        assert _rel(X[c], XT[c]) <= 1e-8
    assert its <= its0
    # 5 columns (a partial pass) and the stored FP32 operator's space
    X5, st5 = g.gmres_multi(B[:5], None, 500, TOL, h2w.GMRES_FP32_INNER_IR)
    assert all(s["converged"] for s in st5) and max(_rel(X5[c], XT[c]) for c in range(5)) <= 1e-8
    g.set_deflation(64, 0, h2w.FP32)
    X, st = g.gmres_multi(B, None, 500, TOL, h2w.GMRES_FP32_IR)
    assert all(s["converged"] for s in st) and max(_rel(X[c], XT[c]) for c in range(16)) <= 1e-8
    g.set_deflation(0, 0, h2w.FP32)
    g.set_deflation(0)


def test_batched_solves_use_each_systems_space():
    """A batch whose systems each carry a deflation space: every system's lockstep columns start from its
    own space (zero-padded to the largest k), every column converges to the single FP64 solve's accuracy, in
    no more steps than without the spaces; a system without a space turns them off for the batch."""
    rng = np.random.default_rng(21)
    systems = []
    for n in (3000, 5200, 4100):
        P = rng.uniform(-1, 1, (3, n))
        g = h2w.Initialisation(3, n, 64, 1.0, 8, h2w.CHOICE_RBF, h2w.NESTED, points=P)
        g.tree_initialisation()
        g.set_inner_tolerance(5)
        systems.append(g)
    XT = [np.random.default_rng(30 + q).normal(size=(16, g.N)) for q, g in enumerate(systems)]
    Bs = [np.stack([g.dense_rows(None, x) for x in X]) for g, X in zip(systems, XT)]
    bt = h2w.Batch(systems)
    _, st0 = bt.gmres(Bs, None, 300, 1e-8, h2w.GMRES_FP32_INNER_IR)
    for g, (m, k) in zip(systems, ((48, 16), (64, 24), (32, 0))):
        g.set_deflation(m, k)
    Xs, st = bt.gmres(Bs, None, 300, 1e-8, h2w.GMRES_FP32_INNER_IR)
    its0 = max(c["iterations"] for row in st0 for c in row)
    its = max(c["iterations"] for row in st for c in row)
    print(f"batch: {its} inner steps with the spaces, {its0} without")
    for q, g in enumerate(systems):
        for c in range(16):
            assert st[q][c]["converged"], (q, c, st[q][c])
            e1 = _rel(g.gmres(Bs[q][c], None, 300, 1e-8, h2w.GMRES_FP64)[0], XT[q][c])
            assert _rel(Xs[q][c], XT[q][c]) <= 10 * e1 + 1e-12
    assert its <= its0
    # one system without a space: none is used, the batch still converges
    systems[1].set_deflation(0)
    _, st2 = bt.gmres(Bs, None, 300, 1e-8, h2w.GMRES_FP32_INNER_IR)
    assert all(c["converged"] for row in st2 for c in row)
    assert abs(max(c["iterations"] for row in st2 for c in row) - its0) <= 1


def test_batch_autotune_on_held_out_right_hand_sides():
    """Batch.autotune times lockstep refined solves of held-out right-hand sides over inner tolerances and
    deflation spaces, keeps the fastest converged candidate (applied to every system) and rebuilds the native
    batch: later solves of other right-hand sides converge to the single FP64 solve's accuracy."""
    rng = np.random.default_rng(23)
    systems = []
    for n in (2600, 4700, 3900):
        P = rng.uniform(-1, 1, (3, n))
        g = h2w.Initialisation(3, n, 64, 1.0, 8, h2w.CHOICE_RBF, h2w.NESTED, points=P)
        g.tree_initialisation()
        systems.append(g)
    bt = h2w.Batch(systems)
    Bh = bt.matvec([np.random.default_rng(40 + q).normal(size=(16, g.N)) for q, g in enumerate(systems)])
    t = bt.autotune(Bh, 1e-8, 300, inner=(6.0, 5.0), spaces=((32, 8), (64, 16)))
    conv = [c for c in t["candidates"] if c["converged"]]
    assert len(t["candidates"]) == 4 and conv
    best = min(conv, key=lambda c: c["seconds"])
    assert (t["inner_tol_exp"], t["m"], t["k"]) == (best["inner_tol_exp"], best["m"], best["k"])
    assert all(g.stats()["inner_tol_exp"] == t["inner_tol_exp"] for g in systems)
    XT = [np.random.default_rng(50 + q).normal(size=(16, g.N)) for q, g in enumerate(systems)]
    Bs = [np.stack([g.dense_rows(None, x) for x in X]) for g, X in zip(systems, XT)]
    Xs, st = bt.gmres(Bs, None, 300, 1e-8, h2w.GMRES_FP32_INNER_IR)
    for q, g in enumerate(systems):
        for c in range(16):
            assert st[q][c]["converged"], (q, c, st[q][c])
            e1 = _rel(g.gmres(Bs[q][c], None, 300, 1e-8, h2w.GMRES_FP64)[0], XT[q][c])
            assert _rel(Xs[q][c], XT[q][c]) <= 10 * e1 + 1e-12


@pytest.mark.parametrize("m,k", [(64, 24), (300, 200)])
def test_lockstep_columns_take_the_single_solves_steps(sys3d, m, k):
    """The lockstep solves project on the same space as the single solve (its row-major copy, on the matrix
    cores; k = 200 spans two workgroup tiles of the dot-product kernel): every column takes the single solve's
    inner steps on the same right-hand side, up to one step of summation-order noise."""
    g, xt, b = sys3d
    g.set_deflation(m, k)
    rng = np.random.default_rng(31)
    XT = rng.normal(size=(5, N))
    B = np.stack([g.dense_rows(None, x) for x in XT])
    X, st = g.gmres_multi(B, None, 500, TOL, h2w.GMRES_FP32_INNER_IR)
    for c in range(5):
        x1, s1 = g.gmres(B[c], None, 500, TOL, h2w.GMRES_FP32_INNER_IR)
        assert st[c]["converged"] and abs(st[c]["iterations"] - s1["iterations"]) <= 1, (c, st[c], s1)
        assert st[c]["outer_iterations"] == s1["outer_iterations"]
        assert _rel(X[c], x1) <= 1e-9
    g.set_deflation(0)
