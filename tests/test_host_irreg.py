"""The irregular particle tree of the lat-lon path (initialize_cubed_sphere.cpp:
446-592, fmm_funcs.cpp:125-231): the product's host build equals the oracle's
literal restatement (panels, leaf point lists, interaction list), and the
invariants fix what the restatement could share with it: leaves hold at most
leaf_size points, every point lies in its leaf, and every ordered pair of
point-holding leaves is covered by exactly one interaction."""
import math

import numpy as np
import pytest


def random_points(n, seed):
    x = np.random.default_rng(seed).normal(size=(3, n))
    return x / np.linalg.norm(x, axis=0)


def latlon_points(spacing_deg):
    """Cell midpoints of a regular lat-lon grid (as the lat-lon driver places its points,
    swe_time_evo_ll.cpp:72)."""
    lat = np.radians(-90.0 + spacing_deg * (np.arange(int(180 / spacing_deg)) + 0.5))
    lon = np.radians(spacing_deg * (np.arange(int(360 / spacing_deg)) + 0.5))
    la, lo = np.meshgrid(lat, lon, indexing="ij")
    return np.stack([np.cos(la) * np.cos(lo), np.cos(la) * np.sin(lo), np.sin(la)]).reshape(3, -1)


def edge_points():
    """Points on face edges, corners and mid-lines (ties of the face rule and of the splits)."""
    pts = [(1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1), (-1, 1, -1), (1, 0, 0), (0, 0, -1), (1, -1, 0)]
    p = np.array(pts, dtype=float).T
    return p / np.linalg.norm(p, axis=0)


CASES = {
    "random3000": (lambda: random_points(3000, 0), 40, 6),
    "latlon5deg": (lambda: latlon_points(5.0), 30, 8),
    "random+edges": (lambda: np.concatenate([random_points(1500, 1), edge_points()], 1), 25, 5),
}


@pytest.fixture(params=sorted(CASES))
def case(request, sb, orc):
    make, leaf, deg = CASES[request.param]
    X = make()
    return X, leaf, deg, sb.HostIrregTree(*X, leaf, deg, 0.7), orc.OracleIrreg(*X, leaf, deg, 0.7)


def test_tree_equals_oracle(case):
    X, leaf, deg, t, o = case
    ints, dbls = t.panels()
    oi, od, opts = o.panels()
    assert (t.P, t.levels) == (o.P, o.levels)
    np.testing.assert_array_equal(ints[:, :9], oi)
    np.testing.assert_array_equal(dbls, od)
    np.testing.assert_array_equal(t.points(), opts)
    np.testing.assert_array_equal(t.interaction_list(), o.interactions())


def test_tree_invariants(case, orc):
    X, leaf, deg, t, o = case
    ints, dbls = t.panels()
    perm = t.points()
    leaves = np.nonzero(ints[:, 7])[0]
    assert ints[leaves, 8].max() <= leaf and ints[leaves, 8].sum() == X.shape[1]
    assert np.array_equal(np.sort(perm), np.arange(X.shape[1]))
    # each point inside its leaf's label bounds
    owner = np.zeros(X.shape[1], dtype=int)
    for p in leaves:
        first, cnt = ints[p, 9], ints[p, 8]
        owner[perm[first:first + cnt]] = p
    for i in range(0, X.shape[1], max(1, X.shape[1] // 400)):
        xe = orc.xieta_from_xyz(*X[:, i])
        p = owner[i]
        assert orc.face_from_xyz(*X[:, i]) == ints[p, 0]
        assert dbls[p, 0] - 1e-15 <= xe[0] <= dbls[p, 1] + 1e-15 and dbls[p, 2] - 1e-15 <= xe[1] <= dbls[p, 3] + 1e-15
    # every ordered pair of point-holding leaves covered exactly once
    parent = ints[:, 2]

    def chain(p):
        out = [p]
        while parent[out[-1]] >= 0:
            out.append(parent[out[-1]])
        return out

    il = t.interaction_list()
    cover = {}
    for tg, s, _ in il:
        cover.setdefault(tg, set()).add(s)
    full = [p for p in leaves if ints[p, 8] > 0]
    rng = np.random.default_rng(0)
    for tl in rng.choice(full, size=min(len(full), 40), replace=False):
        for sl in full:
            n = sum(1 for a in chain(tl) for b in chain(sl) if b in cover.get(a, ()))
            assert n == 1, (tl, sl, n)
    # interaction types: particle sides only at the deepest level or for adjacent leaves
    for tg, s, typ in il:
        if typ in (0, 1):
            assert ints[tg, 7] == 1
        if typ in (0, 2):
            assert ints[s, 7] == 1


def test_bad_input(sb):
    X = random_points(100, 2)
    with pytest.raises(sb.SbslpmError):
        sb.HostIrregTree(*X, 0, 6)
    X[0, 3] = math.nan
    with pytest.raises(sb.SbslpmError):
        sb.HostIrregTree(*X, 10, 6)


@pytest.mark.parametrize("levels", [0, 2, 4])
def test_midpoints(sb, orc, levels):
    x, y, z, a = sb.cubed_sphere_midpoints(levels)
    side = 2 ** levels
    assert x.size == 6 * side * side
    np.testing.assert_allclose(x * x + y * y + z * z, 1.0, rtol=0, atol=1e-15)
    assert abs(a.sum() - 4 * math.pi) < 1e-12 and np.all(a > 0)
    # each centre sits at the centre of its cell in label coordinates
    h = math.pi / (2 * side)
    for f in range(6):
        for j in (0, side - 1):
            for k in (0, side // 2):
                o = (f * side + j) * side + k
                assert orc.face_from_xyz(x[o], y[o], z[o]) == f
                xe, _ = orc.xieta_from_xyz(x[o], y[o], z[o], f)
                assert abs(xe[0] - (-math.pi / 4 + h / 2 + h * j)) < 1e-14
                assert abs(xe[1] - (-math.pi / 4 + h / 2 + h * k)) < 1e-14
