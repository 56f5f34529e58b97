"""CPU: the operator-file reader refuses what is not a complete h2w operator
file before touching any device."""
import numpy as np
import pytest

import h2weak_amd as h2w


def test_refuses_foreign_and_truncated_files(tmp_path):
    p = tmp_path / "x.h2w"
    p.write_bytes(b"NOTANOPS" + bytes(100))
    with pytest.raises(h2w.H2WError) as e:
        h2w.Initialisation.load(str(p))
    assert e.value.code == h2w.H2W_EINVAL
    # right magic, header cut short
    p.write_bytes(b"H2WSKEL1" + np.array([3, 0, 1, 10], dtype=np.int32).tobytes())
    with pytest.raises(h2w.H2WError) as e:
        h2w.Initialisation.load(str(p))
    assert e.value.code == h2w.H2W_EINVAL
    # impossible header values
    hdr = np.array([7, 0, 1, 10, 64, 3, 1, 0], dtype=np.int32).tobytes() + np.int64(1000).tobytes()
    p.write_bytes(b"H2WSKEL1" + hdr + np.float64(1.0).tobytes())
    with pytest.raises(h2w.H2WError) as e:
        h2w.Initialisation.load(str(p))
    assert e.value.code == h2w.H2W_EINVAL
    with pytest.raises(h2w.H2WError):
        h2w.Initialisation.load(str(tmp_path / "missing.h2w"))


# ---- crafted files: well-formed (right layout, recomputed hash) but holding indices or counts that do
# not fit the plan rebuilt from the file's points; the hash detects accidents only, so the reader
# checks every index against the plan before any device work.

def _fnv1a(data):
    h = 1469598103934665603
    for c in data:
        h = ((h ^ c) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def _craft(path, dim, variant, nL, N, fams, pairs, n_max=2, tol_pow=8, choice=1, extra=b""):
    """fams[f][j]: list of (rank, Ir, Ic) per box; pairs: list of (level, b, v, rank, Ir, Ic)."""
    grid = -1.0 + 2.0 * (np.arange(N) + 1.0) / (N + 1.0)
    pts = np.tile(grid, dim) if dim == 1 else None
    if dim == 2:
        n = int(round(N ** 0.5))
        nd = -1.0 + 2.0 * (np.arange(n) + 1.0) / (n + 1.0)
        pts = np.concatenate([np.tile(nd, n), np.repeat(nd, n)])
    out = bytearray(b"H2WSKEL1")
    out += np.array([dim, variant, choice, tol_pow, n_max, nL, 1, 0], np.int32).tobytes()
    out += np.int64(N).tobytes() + np.float64(1.0).tobytes() + np.asarray(pts, np.float64).tobytes()
    for f in range(2):
        for j in range(nL + 1):
            boxes = fams[f][j]
            out += np.int32(len(boxes)).tobytes()
            for rk, Ir, Ic in boxes:
                out += np.int32(rk).tobytes() + np.asarray(Ir, np.int32).tobytes()
                out += np.asarray(Ic, np.int32).tobytes() + np.ones(2 * rk * rk).tobytes()
                out += np.int32(0).tobytes()
    out += np.int64(len(pairs)).tobytes()
    for lv, b, v, rk, Ir, Ic in pairs:
        out += np.array([lv, b, v, rk], np.int32).tobytes()
        out += np.asarray(Ir, np.int32).tobytes() + np.asarray(Ic, np.int32).tobytes() + np.ones(2 * rk * rk).tobytes()
    out += np.uint64(_fnv1a(bytes(out))).tobytes()
    path.write_bytes(bytes(out) + extra)


def _semi_1d_pairs(N, nL):
    """The directed vertex pairs of a 1D semi-nested plan, in the assembly's order, from a host-only plan."""
    g = h2w.Initialisation(1, N, 2, 1.0, 8, 1, h2w.SEMI_NESTED, nlevels=nL, plan_only=1)
    pairs = []
    for j in range(1, nL + 1):
        for b in range(2 ** j):
            il, isv = g.interaction_list(j, b)
            pairs += [(j, b, int(v)) for v, f in zip(il, isv) if f]
    return pairs


def _load_code(path, **opts):
    try:
        h2w.Initialisation.load(str(path), **opts)
    except h2w.H2WError as e:
        return e.code
    return h2w.H2W_OK


def test_crafted_files_are_checked_against_the_plan(tmp_path):
    N, nL = 16, 2
    empty = [[[] for _ in range(nL + 1)] for _ in range(2)]
    pl = _semi_1d_pairs(N, nL)
    assert len(pl) == 6
    box = lambda j, b: (N >> j) * b  # first point of box b at level j (uniform grid, tree order = input order)
    good = [(j, b, v, 1, [box(j, b)], [box(j, v)]) for j, b, v in pl]
    p = tmp_path / "c.h2w"
    _craft(p, 1, h2w.SEMI_NESTED, nL, N, empty, good)
    # well-formed: accepted by the checks (checked-only load), and on a host without a GPU the device
    # step is what refuses it
    assert _load_code(p, plan_only=1) == h2w.H2W_OK
    # a pair pivot outside [0, N)
    bad = list(good)
    bad[3] = bad[3][:4] + ([N + 3],) + bad[3][5:]
    _craft(p, 1, h2w.SEMI_NESTED, nL, N, empty, bad)
    assert _load_code(p, plan_only=1) == h2w.H2W_EINVAL
    assert _load_code(p) == h2w.H2W_EINVAL  # refused before any device is looked for
    bad = list(good)
    bad[2] = bad[2][:5] + ([-1],)
    _craft(p, 1, h2w.SEMI_NESTED, nL, N, empty, bad)
    assert _load_code(p, plan_only=1) == h2w.H2W_EINVAL
    # pairs out of the assembly's order, a missing pair, a rank above the box size
    _craft(p, 1, h2w.SEMI_NESTED, nL, N, empty, [good[1], good[0]] + good[2:])
    assert _load_code(p, plan_only=1) == h2w.H2W_EINVAL
    _craft(p, 1, h2w.SEMI_NESTED, nL, N, empty, good[:-1])
    assert _load_code(p, plan_only=1) == h2w.H2W_EINVAL
    big = list(good)
    big[0] = (1, 0, 1, 9, list(range(9)), list(range(9)))
    _craft(p, 1, h2w.SEMI_NESTED, nL, N, empty, big)
    assert _load_code(p, plan_only=1) == h2w.H2W_EINVAL
    # skeletons given for a family the variant does not use
    fams = [[[] for _ in range(nL + 1)] for _ in range(2)]
    fams[1][1] = [(0, [], [])] * 2
    _craft(p, 1, h2w.SEMI_NESTED, nL, N, fams, good)
    assert _load_code(p, plan_only=1) == h2w.H2W_EINVAL
    # bytes after the hash
    _craft(p, 1, h2w.SEMI_NESTED, nL, N, empty, good, extra=b"\0")
    assert _load_code(p, plan_only=1) == h2w.H2W_EINVAL


def test_crafted_nested_files(tmp_path):
    """2D H2*: a level the far and vertex families use must hold every box; pivots must be points."""
    N, nL = 16, 2  # 4 x 4 grid, one point per leaf
    p = tmp_path / "n.h2w"
    full = [[[] for _ in range(nL + 1)] for _ in range(2)]
    full[0][2] = [(1, [b], [15 - b]) for b in range(16)]
    full[1][1] = [(1, [4 * b], [(4 * b + 8) % 16]) for b in range(4)]
    full[1][2] = [(1, [b], [15 - b]) for b in range(16)]
    _craft(p, 2, h2w.NESTED, nL, N, full, [], n_max=1)
    assert _load_code(p, plan_only=1) == h2w.H2W_OK
    holes = [[list(l) for l in f] for f in full]
    holes[0][2] = []  # nb = 0 at a level the far family uses
    _craft(p, 2, h2w.NESTED, nL, N, holes, [], n_max=1)
    assert _load_code(p, plan_only=1) == h2w.H2W_EINVAL
    oob = [[list(l) for l in f] for f in full]
    oob[1][2][5] = (1, [5], [16])
    _craft(p, 2, h2w.NESTED, nL, N, oob, [], n_max=1)
    assert _load_code(p, plan_only=1) == h2w.H2W_EINVAL
    # pairs in a nested file
    _craft(p, 2, h2w.NESTED, nL, N, full, [(1, 0, 2, 0, [], [])], n_max=1)
    assert _load_code(p, plan_only=1) == h2w.H2W_EINVAL
