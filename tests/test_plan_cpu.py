"""Host side of libh2w on the CPU: the library loads and exports every symbol
include/h2w.h declares, refuses to compute without a GPU, and its tree /
interaction lists equal the oracle's restatement of the reference tables."""
import os
import re

import numpy as np
import pytest

import h2weak_amd as h2w
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_header_symbols():
    hdr = open(os.path.join(ROOT, "include", "h2w.h")).read()
    declared = sorted(set(re.findall(r"\b(h2w_[a-z0-9_]+)\s*\(", hdr)))
    L = h2w.lib()
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    assert set(declared) == set(h2w.EXPORTS)


def test_no_cpu_fallback_without_gpu():
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("GPU present")
    with pytest.raises(h2w.H2WError) as e:
        h2w.Initialisation(3, 4096, 8)
    assert e.value.code == h2w.H2W_ENOLIB


@pytest.mark.parametrize("args", [dict(dim=4, N=100, n_max=8), dict(dim=3, N=100, n_max=8, choice=2),
                                  dict(dim=3, N=10, n_max=64), dict(dim=3, N=1000, n_max=8, variant=5),
                                  dict(dim=2, N=1000, n_max=8)])
def test_invalid_arguments_return_errors(args):
    # the reference exits on a bad choice / nLevels < 1; the ABI returns H2W_EINVAL
    with pytest.raises(h2w.H2WError) as e:
        h2w.Initialisation(plan_only=1, **args)
    assert e.value.code == h2w.H2W_EINVAL


PLANS = [(1, 3000, 40, None), (2, 40 * 40, 16, None), (3, 16 ** 3, 8, None), (2, 3000, 20, "rand"),
         (3, 6000, 20, "rand"), (3, 4000, 16, "gauss"), (3, 27 ** 3, 27, None)]


@pytest.mark.parametrize("dim,N,nmax,kind", PLANS)
def test_plan_equals_oracle_tree(dim, N, nmax, kind):
    rng = np.random.default_rng(N)
    P = None
    if kind == "rand":
        P = rng.uniform(-1, 1, (dim, N))
    elif kind == "gauss":
        P = np.clip(rng.normal(0, 0.2, (dim, N)), -1, 1)
    o = oracle.Oracle(dim, N, nmax, 1.0, 10, 1, 0, points=P)
    g = h2w.Initialisation(dim, N, nmax, 1.0, 10, 1, 0, points=P, plan_only=1)
    assert np.array_equal(g.tree_order(), o.tree_order())
    nl = o.nlevels
    for j in range(1, nl + 1):
        for b in range((1 << dim) ** j):
            assert g.interaction_list(j, b) == o.interaction_list(j, b), (j, b)
    for b in range((1 << dim) ** nl):
        assert g.near_list(nl, b) == o.near(nl, b), b


def test_interaction_counts_interior():
    """Interior boxes: 2D 15 = 3 vertex + 12 far, 3D 133 = 7 vertex + 126 far;
    near: 2D 4, 3D 18 (counts of the reference tables)."""
    for dim, n, count, nv, nn in [(2, 64, 15, 3, 4), (3, 32, 133, 7, 18)]:
        g = h2w.Initialisation(dim, n ** dim, (n // 16) ** dim, plan_only=1)
        nl = 4
        side = 1 << nl
        # an interior box: integer coordinates away from the boundary
        for b in range((1 << dim) ** nl):
            il, vf = g.interaction_list(nl, b)
            if len(il) == count:
                assert sum(vf) == nv and len(g.near_list(nl, b)) == nn
                break
        else:
            pytest.fail("no interior box found")
        assert side > 4


def test_lists_brute_force_classification():
    """IL(B) = children of near(parent) u {parent} that are neither B nor near(B),
    checked by brute force over all box pairs with integer coordinates."""
    dim, nl = 3, 3
    g = h2w.Initialisation(dim, 24 ** 3, 27, plan_only=1, nlevels=nl)
    nch = 1 << dim

    def coords(j, b):
        c = [0, 0, 0]
        digits = []
        for _ in range(j):
            digits.append(b % nch)
            b //= nch
        for ch in reversed(digits):
            bx = int(ch % 4 in (1, 2))
            by = int(ch % 4 in (2, 3))
            bz = int(ch >= 4)
            c = [2 * c[0] + bx, 2 * c[1] + by, 2 * c[2] + bz]
        return c

    def rel(a, b):
        d = [abs(x - y) for x, y in zip(a, b)]
        nz = sum(1 for x in d if x)
        if nz == 0:
            return "self"
        if max(d) > 1:
            return "far"
        return "vertex" if nz == dim else "near"

    for j in range(2, nl + 1):
        nb = nch ** j
        C = [coords(j, b) for b in range(nb)]
        Cp = [coords(j - 1, b) for b in range(nch ** (j - 1))]
        for b in range(0, nb, 7):
            pb = b // nch
            expect = set()
            for c in range(nb):
                pc = c // nch
                if rel(Cp[pb], Cp[pc]) in ("self", "near") and rel(C[b], C[c]) in ("far", "vertex"):
                    expect.add(c)
            il, vf = g.interaction_list(j, b)
            assert set(il) == expect
            for c, v in zip(il, vf):
                assert bool(v) == (rel(C[b], C[c]) == "vertex")
