"""GPU: the learned rank predictor.  The BF16 perceptron kernel reproduces the
numpy forward pass of the same parameters; with the model set the assembled
operator is unchanged (the adaptive cross still decides every pivot) while the
workspace estimates come from the model; the harvest returns one sample per
ACA block."""
import numpy as np
import pytest

import h2weak_amd as h2w
from h2weak_amd import rank_mlp

pytestmark = pytest.mark.gpu

MODEL = rank_mlp.default_model_path(3, 1, 10)


@pytest.fixture(scope="module")
def params():
    assert MODEL is not None, "shipped model missing"
    p, meta = rank_mlp.load(MODEL)
    assert meta["tol_pow"] == 10 and meta["dim"] == 3
    return p


def test_kernel_matches_numpy_forward(params):
    g = h2w.Initialisation(3, 4096, 64, 1.0, 10, h2w.CHOICE_RBF, h2w.NESTED,
                           points=rank_mlp.points("uniform", 4096, 3))
    g.set_rank_model(params)
    F = np.random.default_rng(0).uniform(-0.2, 1.2, (777, 16)).astype(np.float32)
    out = g.rank_predict(F)
    ref = rank_mlp.forward(params, F)
    np.testing.assert_allclose(out, ref, rtol=1e-4, atol=1e-4)
    # random parameters as well (every layer exercised away from the trained regime)
    rp = rank_mlp.to_bf16(np.random.default_rng(1).normal(0, 0.3, rank_mlp.N_PARAMS))
    g.set_rank_model(rp)
    np.testing.assert_allclose(g.rank_predict(F), rank_mlp.forward(rp, F), rtol=1e-4, atol=1e-4)
    with pytest.raises(h2w.H2WError):
        g.set_rank_model(rp[:100])


@pytest.mark.parametrize("kind", ["uniform", "gauss"])
def test_model_does_not_change_the_operator(params, kind):
    N = 20000
    P = rank_mlp.points(kind, N, 11)
    a = h2w.Initialisation(3, N, 64, 1.0, 10, h2w.CHOICE_RBF, h2w.NESTED, points=P)
    a.tree_initialisation()
    b = h2w.Initialisation(3, N, 64, 1.0, 10, h2w.CHOICE_RBF, h2w.NESTED, points=P)
    b.set_rank_model(params)
    b.tree_initialisation()
    sa, sb = a.stats(), b.stats()
    assert sa["memory"] == sb["memory"]
    assert sa["rank_pred_blocks"] == 0 and sb["rank_pred_blocks"] > 0
    assert sb["rank_pred_under"] <= 0.05 * sb["rank_pred_blocks"]
    assert sb["rank_pred_mae_log2"] < 0.2
    x = np.random.default_rng(2).normal(size=N)
    ya, yb = a.mat_vec_prod(x), b.mat_vec_prod(x)
    np.testing.assert_allclose(yb, ya, rtol=0, atol=1e-14 * np.abs(ya).max())
    Fa, Ra = a.rank_samples()
    Fb, Rb = b.rank_samples()
    assert len(Ra) == len(Rb) > 0 and Fa.shape == (len(Ra), 16)
    # same blocks, same ranks (the model only reorders the batches)
    ka = sorted(zip(map(tuple, Fa.round(6)), Ra))
    kb = sorted(zip(map(tuple, Fb.round(6)), Rb))
    assert ka == kb
