"""GPU: the autotuner and the rebuildable inner operator of the refinement.
Tuning changes launch shapes and the truncation of the FP32 inner operator only;
the FP64 product stays the oracle's to rounding and the refined solve keeps FP64
accuracy."""
import numpy as np
import pytest

import h2weak_amd as h2w
import oracle

pytestmark = pytest.mark.gpu

CASE = (3, 16 ** 3, 27, 10, 1, 0)  # dim, N, n_max, tol_pow, choice, variant (3D RBF grid, H2*)


@pytest.fixture(scope="module")
def pair():
    dim, N, nmax, tp, ch, var = CASE
    o = oracle.Oracle(dim, N, nmax, 1.0, tp, ch, var)
    o.tree_initialisation()
    g = h2w.Initialisation(dim, N, nmax, 1.0, tp, ch, var)
    g.tree_initialisation()
    return o, g, N


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_inner_tolerance_controls_truncation(pair):
    o, g, N = pair
    x = np.random.default_rng(4).uniform(-1, 1, N)
    y64 = g.mat_vec_prod(x)
    assert _rel(y64, o.mat_vec_prod(x)) <= 1e-12
    sizes = []
    for p in (4, 6, 6.5, 8):  # (fractional exponents: the tuner's half decades)
        g.set_inner_tolerance(p)
        yin = g.mat_vec_prod(x, precision=h2w.FP32_INNER)
        assert _rel(yin, y64) <= 30 * 10.0 ** (-p) + 2e-6, p
        sizes.append(g.stats()["stored_bytes_inner_fp32"])
    # looser tolerance, fewer pivots kept
    assert sizes[0] <= sizes[1] <= sizes[2] <= sizes[3]
    g.set_inner_tolerance(0)
    with pytest.raises(h2w.H2WError) as e:
        g.mat_vec_prod(x, precision=h2w.FP32_INNER)
    assert e.value.code == h2w.H2W_ESTATE
    for bad in (CASE[3], CASE[3] + 0.5, -1.0):
        with pytest.raises(h2w.H2WError) as e:
            g.set_inner_tolerance(bad)
        assert e.value.code == h2w.H2W_EINVAL
    g.set_inner_tolerance(6)


def test_autotune_keeps_parity(pair):
    o, g, N = pair
    rng = np.random.default_rng(9)
    x = rng.normal(size=N)
    ref = o.mat_vec_prod(x)
    y64, y32 = g.mat_vec_prod(x), g.mat_vec_prod(x, precision=h2w.FP32)
    xt = rng.uniform(-1, 1, N)
    b = o.dense_matvec(xt)
    with pytest.raises(h2w.H2WError):
        g.autotune(None, flags=h2w.TUNE_INNER)
    t = g.autotune(b, tol=1e-10, max_iter=300)
    assert [c["ir_tol_exp"] for c in t["candidates"]] == [4, 4.5, 5, 5.5, 6, 6.5, 7, 7.5, 8]
    assert all(c["converged"] for c in t["candidates"])
    assert t["ir_tol_exp"] in [c["ir_tol_exp"] for c in t["candidates"]]
    assert g.stats()["inner_tol_exp"] == t["ir_tol_exp"]
    assert all(v > 0 for v in t["matvec_ms_after"])
    # products after tuning: same operator, other summation order at most
    assert _rel(g.mat_vec_prod(x), ref) <= 1e-12
    assert _rel(g.mat_vec_prod(x), y64) <= 1e-13
    assert _rel(g.mat_vec_prod(x, precision=h2w.FP32), y32) <= 2e-6
    # the tuned refinement reaches FP64 accuracy
    x64, s64 = g.gmres(b, None, 300, 1e-10, h2w.GMRES_FP64)
    xin, sin = g.gmres(b, None, 300, 1e-10, h2w.GMRES_FP32_INNER_IR)
    assert sin["converged"] and _rel(g.mat_vec_prod(xin), b) <= 2e-10
    assert _rel(xin, xt) <= 10 * _rel(x64, xt) + 1e-12
