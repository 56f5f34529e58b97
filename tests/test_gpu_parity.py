"""GPU parity: libh2w (HIP, gfx950) against the CPU oracle on the same seeded
inputs -- skeleton ranks, the reference's storage count, the product itself,
the dense rows, GMRES, and the FP32 + refinement path."""
import numpy as np
import pytest

import h2weak_amd as h2w
import oracle

pytestmark = pytest.mark.gpu


def _points(kind, dim, N, seed):
    rng = np.random.default_rng(seed)
    if kind == "grid":
        return None
    if kind == "rand":
        return rng.uniform(-1, 1, (dim, N))
    if kind == "gauss":  # clustered: leaves near the boundary stay empty
        P = rng.normal(0, 0.2, (dim, 4 * N))
        P = P[:, np.all(np.abs(P) <= 1, axis=0)][:, :N]
        return P
    raise ValueError(kind)


# dim, N, n_max, tol_pow, choice, variant, points
CASES = [
    (1, 4096, 64, 12, 1, 0, "rand"),
    (1, 4096, 64, 12, 0, 1, "rand"),
    (1, 3000, 50, 12, 1, 1, "grid"),
    (2, 48 * 48, 36, 10, 0, 0, "grid"),
    (2, 48 * 48, 36, 12, 1, 1, "grid"),
    (2, 3000, 40, 10, 1, 0, "rand"),
    (3, 16 ** 3, 27, 8, 1, 0, "grid"),
    (3, 16 ** 3, 27, 7, 0, 1, "grid"),
    (3, 5000, 40, 10, 1, 0, "rand"),
    (3, 4000, 32, 8, 0, 0, "gauss"),
    (3, 4000, 32, 8, 1, 1, "gauss"),
    (2, 48 * 48, 36, 10, 1, 2, "grid"),
    (3, 16 ** 3, 27, 7, 0, 2, "grid"),
    # leaves of 256 points: near-field blocks and output boxes of two 128-row tiles
    (2, 64 * 64, 256, 10, 1, 0, "grid"),
    (2, 64 * 64, 256, 11, 0, 1, "rand"),
]


def _ids(c):
    return f"{c[0]}d-N{c[1]}-nmax{c[2]}-tol{c[3]}-ch{c[4]}-{['nested', 'semi', 'hstar'][c[5]]}-{c[6]}"


@pytest.fixture(scope="module")
def built():
    cache = {}

    def get(case):
        if case not in cache:
            dim, N, nmax, tp, ch, var, kind = case
            P = _points(kind, dim, N, 17)
            if P is not None:
                N = P.shape[1]
            o = oracle.Oracle(dim, N, nmax, 1.0, tp, ch, var, points=P)
            o.tree_initialisation()
            g = h2w.Initialisation(dim, N, nmax, 1.0, tp, ch, var, points=P)
            g.tree_initialisation()
            cache[case] = (o, g, N)
        return cache[case]

    return get


def _same_pivots(o, g, case):
    """(boxes with the oracle's pivots, boxes) over the nested families."""
    dim, var = case[0], case[5]
    fams = ([0] if dim > 1 and var != 2 else []) + ([1] if var == 0 else [])
    total = same = 0
    for f in fams:
        for j in range(2 if f == 0 else 1, o.nlevels + 1):
            for b in range((1 << dim) ** j):
                total += 1
                gi, gc = g.skeleton(f, j, b)
                oi, oc = o.pivots(f, j, b)
                same += (gi == oi and gc == oc)
    return same, total


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_skeletons_and_storage(built, case):
    o, g, N = built(case)
    same, total = _same_pivots(o, g, case)
    if total:
        assert same / total >= 0.98, f"{same}/{total} boxes with identical pivots"
    st = g.stats()
    if same == total:
        # every pivot equal: the same ranks everywhere, hence the reference's count exactly
        assert st["memory"] == o.memory
    else:
        assert st["memory"] == pytest.approx(o.memory, rel=2e-3)


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_matvec_parity(built, case):
    o, g, N = built(case)
    tol = 10.0 ** (-case[3])
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, N)
    yg = g.mat_vec_prod(x)
    yo = o.mat_vec_prod(x)
    yd = o.dense_matvec(x)
    err_o = np.linalg.norm(yo - yd) / np.linalg.norm(yd)
    err_g = np.linalg.norm(yg - yd) / np.linalg.norm(yd)
    same, total = _same_pivots(o, g, case)
    if same == total:
        # every pivot equal: the same compressed operator, so the products agree to rounding
        assert np.linalg.norm(yg - yo) <= 1e-12 * np.linalg.norm(yd)
    else:
        # same compressed operator up to the few boxes whose pivots differ, same compression error
        assert np.linalg.norm(yg - yo) <= max(1e-12, 0.5 * err_o) * np.linalg.norm(yd) * 10
    assert err_g <= 10 * err_o + 1e-14
    assert err_g <= 100 * tol
    # FP32 shadow: FP32 operator and accumulation
    y32 = g.mat_vec_prod(x, precision=h2w.FP32)
    assert np.linalg.norm(y32 - yg) <= 2e-5 * np.linalg.norm(yg)
    # near field and M2L evaluated on the fly: FP64 equals the stored product to rounding,
    # FP32 stays within FP32 accuracy of it
    y64o = g.mat_vec_prod(x, precision=h2w.FP64_OTF)
    assert np.linalg.norm(y64o - yg) <= 1e-12 * np.linalg.norm(yg)
    y32o = g.mat_vec_prod(x, precision=h2w.FP32_OTF)
    assert np.linalg.norm(y32o - yg) <= 2e-5 * np.linalg.norm(yg)
    # truncated inner operator of the refinement (ranks cut where the cross meets 1e-6)
    if case[3] > 6:
        yin = g.mat_vec_prod(x, precision=h2w.FP32_INNER)
        assert np.linalg.norm(yin - yg) <= 1e-4 * np.linalg.norm(yg)
        assert g.stats()["stored_bytes_inner_fp32"] <= g.stats()["stored_bytes_fp32"]
    # several right-hand sides at once
    X = np.stack([x, 2 * x, rng.normal(size=N)])
    Y = g.matvec_multi(X)
    np.testing.assert_allclose(Y[1], 2 * yg, rtol=0, atol=1e-13 * np.abs(yg).max())


@pytest.mark.parametrize("case", CASES[::3], ids=_ids)
def test_dense_rows_parity(built, case):
    o, g, N = built(case)
    rng = np.random.default_rng(8)
    x = rng.normal(size=N)
    rows = rng.choice(N, size=min(N, 200), replace=False)
    np.testing.assert_allclose(g.dense_rows(rows, x), o.dense_rows(rows, x), rtol=1e-12,
                               atol=1e-12 * np.abs(o.dense_rows(rows, x)).max())


@pytest.mark.parametrize("case", [c for c in CASES if c[4] == 0 or c[0] != 3], ids=_ids)
def test_gmres_parity(built, case):
    """Relative solution error and final residual within 10x of the oracle's
    (the reference's own correctness check, */codes/main.cpp), iterations within one."""
    o, g, N = built(case)
    gtol = 10.0 ** (-min(case[3], 12))
    xt = np.random.default_rng(2).uniform(-1, 1, N)
    b = o.dense_matvec(xt)
    xo, ito, ro, _ = o.gmres(b, None, 300, gtol)
    xg, st = g.gmres(b, None, 300, gtol)
    eo = np.linalg.norm(xo - xt) / np.linalg.norm(xt)
    eg = np.linalg.norm(xg - xt) / np.linalg.norm(xt)
    assert ito > 0 and st["converged"]
    assert abs(st["iterations"] - ito) <= 1
    assert st["resid"] <= 10 * ro and st["resid"] < gtol
    assert eg <= 10 * eo + 1e-15


@pytest.mark.parametrize("case", [CASES[3], CASES[6], CASES[9]], ids=_ids)
def test_fp32_refinement_reaches_fp64_accuracy(built, case):
    o, g, N = built(case)
    gtol = 1e-10
    xt = np.random.default_rng(3).normal(size=N)
    b = o.dense_matvec(xt)
    x64, s64 = g.gmres(b, None, 300, gtol, h2w.GMRES_FP64)
    e64 = np.linalg.norm(x64 - xt) / np.linalg.norm(xt)
    for mode in (h2w.GMRES_FP32_IR, h2w.GMRES_FP32_OTF_IR, h2w.GMRES_FP32_INNER_IR):
        x32, s32 = g.gmres(b, None, 300, gtol, mode)
        assert s32["converged"] and s32["resid"] <= gtol
        # the true FP64 residual of the refined solution
        r = np.linalg.norm(b - g.mat_vec_prod(x32)) / np.linalg.norm(b)
        assert r <= 2 * gtol
        e32 = np.linalg.norm(x32 - xt) / np.linalg.norm(xt)
        assert e32 <= 10 * e64 + 1e-12


def test_single_level_and_errors():
    # nLevels = 1: no far field, one vertex level, the near field is the whole block structure
    P = np.random.default_rng(1).uniform(-1, 1, (3, 600))
    o = oracle.Oracle(3, 600, 200, 1.0, 10, 1, 0, points=P)
    o.tree_initialisation()
    g = h2w.Initialisation(3, 600, 200, 1.0, 10, 1, 0, points=P)
    assert g.stats()["nlevels"] == 1
    with pytest.raises(h2w.H2WError) as e:
        g.mat_vec_prod(np.ones(600))
    assert e.value.code == h2w.H2W_ESTATE
    g.tree_initialisation()
    x = np.random.default_rng(2).normal(size=600)
    np.testing.assert_allclose(g.mat_vec_prod(x), o.mat_vec_prod(x), rtol=1e-10, atol=1e-12)


def test_graph_replays_match(built):
    """The matvec is a captured HIP graph replayed on every call: replays with
    different inputs, interleaved with other work, each match the oracle."""
    o, g, N = built(CASES[6])
    rng = np.random.default_rng(21)
    xs = [rng.normal(size=N) for _ in range(3)]
    ref = [o.mat_vec_prod(x) for x in xs]
    for rep in range(3):
        for x, r in zip(xs, ref):
            y = g.mat_vec_prod(x)
            assert np.linalg.norm(y - r) <= 1e-12 * np.linalg.norm(r)
            g.mat_vec_prod(x, precision=h2w.FP32)
        g.dense_rows([0, 1, 2], xs[0])


def test_stream_read_bandwidth():
    """The read-only stream helper the bench compares the products with: a positive rate no
    higher than the published HBM peak, and argument checks."""
    g = h2w.bench_stream_read(1 << 30, 3)
    assert 500.0 < g < 8000.0
    with pytest.raises(h2w.H2WError) as e:
        h2w.bench_stream_read(1024, 1)
    assert e.value.code == h2w.H2W_EINVAL


def test_bench_helpers(built):
    """The timing entry points (single, 16-wide, batch): positive times, one fused launch
    per product with its operator bytes, roles as documented in include/h2w.h."""
    o, g, N = built(CASES[6])
    for prec in (h2w.FP64, h2w.FP32, h2w.FP32_INNER):
        r = g.bench_matvec(2, prec)
        assert r["ms_per_matvec"] > 0 and r["launch_role"][0] == -1
        assert r["launch_role"].count(1) == 1 and r["launch_bytes"][r["launch_role"].index(1)] > 0
        assert all(t >= 0 for t in r["launch_ms"])
        m = g.bench_matvec_multi(2, prec)
        # (the single-read launch applies the transposes too: no separate role-4 launch; entry 0 counts the
        # symmetric blocks twice, as multiplied)
        assert m["ms_per_pass"] > 0 and m["launch_role"][0] == -2 and 4 not in m["launch_role"]
        assert m["launch_role"].count(1) == 1
        q = m["launch_role"].index(1)
        assert m["launch_bytes"][q] == r["launch_bytes"][r["launch_role"].index(1)]
        assert m["launch_bytes"][0] > m["launch_bytes"][q]
    b = h2w.Batch([g]).bench_matvec(2, h2w.FP64)
    assert b["ms_per_pass"] > 0 and len(b["launch_ms"]) == len(b["launch_role"]) > 0


@pytest.mark.parametrize("mode", [h2w.GMRES_FP32_IR, h2w.GMRES_FP32_INNER_IR])
def test_refinement_reports_fp64_residuals(built, mode):
    """The refinement's first correction starts from the FP32 operator's residual, but what it
    reports and decides on is an FP64 residual: with too few iterations allowed, and when x0
    already solves the system (max_iter 0 is refused, as for FP64 GMRES)."""
    o, g, N = built(CASES[6])
    xt = np.random.default_rng(5).normal(size=N)
    b = g.mat_vec_prod(xt)
    x0 = np.ones(N)

    def true_resid(x):
        return np.linalg.norm(b - g.mat_vec_prod(x)) / np.linalg.norm(b)

    with pytest.raises(h2w.H2WError) as e:
        g.gmres(b, x0, 0, 1e-10, mode)
    assert e.value.code == h2w.H2W_EINVAL
    for its in (1, 3):
        x, s = g.gmres(b, x0, its, 1e-10, mode)
        assert s["iterations"] <= its and not s["converged"] and s["matvecs_fp64"] >= 1
        assert s["resid"] == pytest.approx(true_resid(x), rel=1e-6)
    x, s = g.gmres(b, xt, 300, 1e-10, mode)
    assert s["converged"] and true_resid(x) <= 2e-10
    assert np.linalg.norm(x - xt) <= 1e-8 * np.linalg.norm(xt)
