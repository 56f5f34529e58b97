"""GPU: saving and loading an assembled operator.  The loaded operator (plan
rebuilt from the saved points, skeletons put back, blocks re-evaluated) equals
the saved one (storage count, pivots and products to the last bit up to atomic
ordering, for every format) and the oracle's assembly of the same problem
(pivots, the reference's storage count, the product to rounding); altered or
truncated files are refused."""
import os

import numpy as np
import pytest

import h2weak_amd as h2w
import oracle

pytestmark = pytest.mark.gpu

# dim, N, n_max, tol_pow, choice, variant, points
CASES = [
    (3, 16 ** 3, 27, 10, 1, 0, "grid"),
    (3, 5000, 40, 8, 0, 1, "rand"),
    (2, 48 * 48, 36, 10, 1, 2, "grid"),
    (1, 4096, 64, 12, 1, 0, "rand"),
    (1, 3000, 50, 12, 0, 1, "rand"),
]


def _ids(c):
    return f"{c[0]}d-N{c[1]}-tol{c[3]}-ch{c[4]}-{['nested', 'semi', 'hstar'][c[5]]}-{c[6]}"


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_save_load_roundtrip(case, tmp_path):
    dim, N, nmax, tp, ch, var, kind = case
    P = None if kind == "grid" else np.random.default_rng(5).uniform(-1, 1, (dim, N))
    a = h2w.Initialisation(dim, N, nmax, 1.0, tp, ch, var, points=P)
    a.tree_initialisation()
    path = str(tmp_path / "op.h2w")
    a.save(path)
    b = h2w.Initialisation.load(path)
    sa, sb = a.stats(), b.stats()
    for k in ("memory", "nlevels", "stored_bytes_fp64", "stored_bytes_fp32", "stored_bytes_inner_fp32",
              "max_rank_far", "max_rank_vertex", "diag"):
        assert sa[k] == sb[k], k
    assert a.tree_order().tolist() == b.tree_order().tolist()
    fams = ([0] if dim > 1 and var != 2 else []) + ([1] if var == 0 else [])
    for f in fams:
        for j in range(2 if f == 0 else 1, sa["nlevels"] + 1):
            for box in range(0, (1 << dim) ** j, 7):
                assert a.skeleton(f, j, box) == b.skeleton(f, j, box)
    x = np.random.default_rng(1).normal(size=N)
    # the loaded operator against the oracle's assembly of the same problem: its skeletons are the
    # oracle's pivots, and its product is the oracle's product to rounding
    o = oracle.Oracle(dim, N, nmax, 1.0, tp, ch, var, points=P)
    o.tree_initialisation()
    same = tot = 0
    for f in fams:
        for j in range(2 if f == 0 else 1, sa["nlevels"] + 1):
            for box in range((1 << dim) ** j):
                tot += 1
                same += b.skeleton(f, j, box) == o.pivots(f, j, box)
    assert same == tot, (same, tot)
    assert sb["memory"] == o.memory
    yo = o.mat_vec_prod(x)
    assert np.linalg.norm(b.mat_vec_prod(x) - yo) <= 1e-12 * np.linalg.norm(yo)
    for prec in (h2w.FP64, h2w.FP32, h2w.FP32_INNER):
        if prec == h2w.FP32_INNER and tp <= 6:
            continue
        ya, yb = a.mat_vec_prod(x, precision=prec), b.mat_vec_prod(x, precision=prec)
        np.testing.assert_allclose(yb, ya, rtol=0, atol=1e-14 * np.abs(ya).max() if prec == h2w.FP64
                                   else 1e-6 * np.abs(ya).max())
    # an altered byte, and a truncated file, are refused
    raw = open(path, "rb").read()
    bad = bytearray(raw)
    bad[len(raw) // 2] ^= 0x40
    open(path, "wb").write(bytes(bad))
    with pytest.raises(h2w.H2WError) as e:
        h2w.Initialisation.load(path)
    assert e.value.code == h2w.H2W_EINVAL
    open(path, "wb").write(raw[: len(raw) - 9])
    with pytest.raises(h2w.H2WError) as e:
        h2w.Initialisation.load(path)
    assert e.value.code == h2w.H2W_EINVAL
