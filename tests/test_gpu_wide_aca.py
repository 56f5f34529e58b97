"""GPU: the wide cross (one block over many workgroups, one launch per phase of a
pivot step), forced onto every block with m + n > 1500 through H2W_WIDE_MIN
(read by the library at each assembly).  Its rows and columns are the same
numbers as the one-workgroup cross, so skeletons equal the oracle's; only the
norms are summed in another order."""
import os

import numpy as np
import pytest

import h2weak_amd as h2w
import oracle
from h2weak_amd import rank_mlp

pytestmark = pytest.mark.gpu

CASES = [(3, 4096, 27, 10, 1, 0), (2, 2304, 36, 10, 0, 1), (3, 3000, 40, 8, 0, 2)]


@pytest.fixture
def wide():
    old = os.environ.get("H2W_WIDE_MIN")
    os.environ["H2W_WIDE_MIN"] = "1500"
    yield
    if old is None:
        del os.environ["H2W_WIDE_MIN"]
    else:
        os.environ["H2W_WIDE_MIN"] = old


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}d-N{c[1]}-tol{c[3]}-ch{c[4]}-v{c[5]}")
def test_wide_cross_matches_oracle(wide, case):
    dim, N, nmax, tp, ch, var = case
    P = None if var == 0 else np.random.default_rng(3).uniform(-1, 1, (dim, N))
    o = oracle.Oracle(dim, N, nmax, 1.0, tp, ch, var, points=P)
    o.tree_initialisation()
    g = h2w.Initialisation(dim, N, nmax, 1.0, tp, ch, var, points=P)
    g.tree_initialisation()
    fams = ([0] if dim > 1 and var != 2 else []) + ([1] if var == 0 else [])
    tot = same = 0
    for f in fams:
        for j in range(2 if f == 0 else 1, o.nlevels + 1):
            for b in range((1 << dim) ** j):
                tot += 1
                same += g.skeleton(f, j, b) == o.pivots(f, j, b)
    if tot:
        assert same / tot >= 0.98, (same, tot)
    x = np.random.default_rng(1).normal(size=N)
    yo, yg, yd = o.mat_vec_prod(x), g.mat_vec_prod(x), o.dense_matvec(x)
    err_o = np.linalg.norm(yo - yd) / np.linalg.norm(yd)
    if same == tot:
        # every pivot equal: the reference's count exactly and the same operator to rounding
        assert g.stats()["memory"] == o.memory
        assert np.linalg.norm(yg - yo) / np.linalg.norm(yd) <= 1e-12
    else:
        assert g.stats()["memory"] == pytest.approx(o.memory, rel=2e-3)
        assert np.linalg.norm(yg - yo) / np.linalg.norm(yd) <= max(1e-12, 5 * err_o)


def test_wide_cross_with_empty_boxes(wide):
    # clustered points: many empty boxes, so some wide blocks have no rows or no columns
    P = np.random.default_rng(4).normal(0, 0.15, (3, 12000))
    P = P[:, np.all(np.abs(P) <= 1, axis=0)][:, :6000]
    N = P.shape[1]
    o = oracle.Oracle(3, N, 32, 1.0, 8, 1, 0, points=P)
    o.tree_initialisation()
    g = h2w.Initialisation(3, N, 32, 1.0, 8, 1, 0, points=P)
    g.tree_initialisation()
    assert g.stats()["memory"] == pytest.approx(o.memory, rel=2e-3)
    x = np.random.default_rng(1).normal(size=N)
    yo, yg = o.mat_vec_prod(x), g.mat_vec_prod(x)
    assert np.linalg.norm(yg - yo) <= 1e-6 * np.linalg.norm(yo)


@pytest.mark.parametrize("N", [4096, 12000])
def test_fixed_rank_phase_matches_oracle(wide, N):
    """With the rank model set, every wide block first takes a fixed number of pivot steps (1.03 x its
    predicted rank + 2) without the stopping test, then the reference's stopping rule is applied to all
    of them at once from their Gram products: the block ends where the adaptive cross would have, or
    continues adaptively when the model under-predicted.  Either way the skeletons are the oracle's."""
    P = np.random.default_rng(7).uniform(-1, 1, (3, N))
    o = oracle.Oracle(3, N, 64, 1.0, 10, 1, 0, points=P)
    o.tree_initialisation()
    g = h2w.Initialisation(3, N, 64, 1.0, 10, h2w.CHOICE_RBF, h2w.NESTED, points=P)
    g.set_rank_model(rank_mlp.load(rank_mlp.default_model_path(3, 1, 10))[0])
    g.tree_initialisation()
    st = g.stats()
    assert st["rank_fixed_blocks"] > 0
    tot = same = 0
    for f in (0, 1):
        for j in range(2 if f == 0 else 1, o.nlevels + 1):
            for b in range(8 ** j):
                tot += 1
                same += g.skeleton(f, j, b) == o.pivots(f, j, b)
    print(f"N={N}: {st['rank_fixed_blocks']} fixed-rank blocks, {st['rank_fixed_resumed']} resumed, "
          f"{st['rank_fixed_extra_steps']} extra steps, sampled error {st['rank_fixed_sampled_err']:.2e}; "
          f"{same}/{tot} skeletons equal")
    assert same / tot >= 0.98
    assert st["rank_fixed_sampled_err"] <= 1e-8
    x = np.random.default_rng(1).normal(size=N)
    yo, yg = o.mat_vec_prod(x), g.mat_vec_prod(x)
    if same == tot:
        assert st["memory"] == o.memory
        assert np.linalg.norm(yg - yo) <= 1e-12 * np.linalg.norm(yo)
    else:
        assert st["memory"] == pytest.approx(o.memory, rel=2e-3)
        assert np.linalg.norm(yg - yo) <= 1e-9 * np.linalg.norm(yo)
