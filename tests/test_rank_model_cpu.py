"""CPU: the rank predictor's host side -- BF16 rounding, parameter layout,
the shipped model file and its recorded held-out accuracy."""
import numpy as np

from h2weak_amd import rank_mlp


def test_bf16_rounding():
    # exactly representable values survive; halfway cases round to even
    x = np.array([1.0, -2.5, 0.0, 3.140625, 1e-3], dtype=np.float32)
    rt = rank_mlp.from_bf16(rank_mlp.to_bf16(x))
    np.testing.assert_allclose(rt, x, rtol=2 ** -8)
    assert rank_mlp.from_bf16(rank_mlp.to_bf16(np.float32(3.140625)))[()] == np.float32(3.140625)
    one_plus_half_ulp = np.frombuffer(np.uint32(0x3F808000).tobytes(), dtype=np.float32)  # 1 + 2^-8
    assert rank_mlp.to_bf16(one_plus_half_ulp)[0] == 0x3F80        # tie -> even (down)
    three_halves_ulp = np.frombuffer(np.uint32(0x3F818000).tobytes(), dtype=np.float32)
    assert rank_mlp.to_bf16(three_halves_ulp)[0] == 0x3F82         # tie -> even (up)


def test_parameter_layout_and_forward():
    assert rank_mlp.N_PARAMS == 9473
    # a network whose only path is feature 4 -> unit 0 -> ... -> output reproduces GELU chains
    p = np.zeros(rank_mlp.N_PARAMS, dtype=np.float32)
    F, H = rank_mlp.F_DIM, rank_mlp.H_DIM
    o = 0
    p[o + 4 * H + 0] = 1.0            # W1[4, 0]
    o += F * H + H
    p[o + 0 * H + 0] = 1.0            # W2[0, 0]
    o += H * H + H
    p[o + 0 * H + 0] = 1.0            # W3[0, 0]
    o += H * H + H
    p[o + 0] = 2.0                    # w4[0]
    p[-1] = 0.5                       # b4
    params = rank_mlp.to_bf16(p)
    from scipy.special import erf
    g = lambda x: 0.5 * x * (1 + erf(x / np.sqrt(2)))
    X = np.zeros((3, F), dtype=np.float32)
    X[:, 4] = [0.0, 1.0, -2.0]
    np.testing.assert_allclose(rank_mlp.forward(params, X), 2.0 * g(g(g(X[:, 4]))) + 0.5, rtol=1e-6, atol=1e-7)


def test_shipped_model():
    path = rank_mlp.default_model_path(3, 1, 10)
    assert path is not None
    params, meta = rank_mlp.load(path)
    assert params.dtype == np.uint16 and params.size == rank_mlp.N_PARAMS
    assert meta["tol_pow"] == 10 and meta["features"] == 16
    # held-out N = 262144 sets, recorded at training time
    for k, v in meta["test"].items():
        assert v["mae_log2"] < 0.1, k
        assert v["under_frac"] < 0.02, k
    y = rank_mlp.forward(params, np.random.default_rng(0).uniform(0, 1, (50, 16)).astype(np.float32))
    assert np.all(np.isfinite(y))
