"""CPU checks of the product library: it loads, exports every symbol the C
header declares, and its host set-up (tree, particle grid, weights,
interaction list, transfer matrices) equals the oracle's."""
import os
import re

import numpy as np
import pytest

from conftest import REPO

CONFIGS = [
    # levels, d, leaf degree, theta, weights
    (3, 8, 8, 0.7, "reference"),
    (4, 8, 7, 0.7, "clenshaw_curtis"),
    (4, 8, 7, 0.7, "equal_area"),
    (3, 10, 7, 0.6, "clenshaw_curtis"),
    (2, 4, 6, 0.5, "reference"),
]
RULE = {"reference": 0, "clenshaw_curtis": 1, "equal_area": 2}


def test_library_exports_every_declared_symbol(sb):
    lib = sb.load()
    header = open(os.path.join(REPO, "include", "sbslpm.h")).read()
    declared = sorted(set(re.findall(r"\b(sbslpm_[a-z0-9_]+)\s*\(", header)))
    assert len(declared) >= 30
    for name in declared:
        assert hasattr(lib, name), name
    assert set(sb.exported_symbols()) == set(declared)


@pytest.mark.parametrize("L,d,md,theta,w", CONFIGS)
def test_host_tree_matches_oracle(sb, orc, L, d, md, theta, w):
    h = sb.HostTree(sb.RunConfig(levels=L, interp_degree=d, leaf_degree=md, fmm_theta=theta, weights=w))
    o = orc.OracleFMM(L, d, md, theta, RULE[w])
    assert h.sizes.P == o.P and h.sizes.Pl == o.Pl and h.sizes.I == o.interaction_count
    hi, hd = h.panels()
    oi, od = o.panels()
    np.testing.assert_array_equal(hi, oi)
    np.testing.assert_array_equal(hd, od)
    np.testing.assert_array_equal(h.interaction_list(), o.interactions())
    for a, b in zip(h.spec_points(), o.spec_points()):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-17)


@pytest.mark.parametrize("d", [4, 8, 10])
def test_transfer_matrices_equal_per_node_bli(sb, orc, d):
    """The fixed child<->parent 1-D matrices reproduce the reference's per-node
    basis evaluation (fmm_funcs_impl.hpp:61-64) on every panel of a level."""
    md = 7
    h = sb.HostTree(sb.RunConfig(levels=3, interp_degree=d, leaf_degree=md))
    hi, lo, p2m = h.transfer_matrices()
    o = orc.OracleFMM(3, d, md, 0.7, 0)
    ints, dbls = o.panels()
    tab = orc.interp_init(d)
    tabm = orc.interp_init(md)
    for par in range(6, 30):
        pb = np.pi / 4 * dbls[par, :4]
        for c in range(4):
            ch = ints[par, 3 + c]
            q = dbls[ch]
            xo, xs = np.pi / 4 * 0.5 * (q[1] + q[0]), np.pi / 4 * 0.5 * (q[1] - q[0])
            eo, es = np.pi / 4 * 0.5 * (q[3] + q[2]), np.pi / 4 * 0.5 * (q[3] - q[2])
            A = hi if c in (0, 3) else lo
            B = hi if c in (0, 1) else lo
            for i in range(0, (d + 1) ** 2, 7):
                basis, err = orc.interp_vals_bli(tab[i, 0] * xs + xo, tab[i, 1] * es + eo, *pb, d)
                assert err == 0
                k, l = divmod(i, d + 1)
                np.testing.assert_allclose(basis, np.outer(A[:, k], B[:, l]).ravel(), rtol=0, atol=1e-14)
    # proxy basis at the particle grid of a leaf
    leaf = o.offset + 5
    q = dbls[leaf]
    xo, xs = np.pi / 4 * 0.5 * (q[1] + q[0]), np.pi / 4 * 0.5 * (q[1] - q[0])
    eo, es = np.pi / 4 * 0.5 * (q[3] + q[2]), np.pi / 4 * 0.5 * (q[3] - q[2])
    for j in range((md + 1) ** 2):
        basis, _ = orc.interp_vals_bli(tabm[j, 0] * xs + xo, tabm[j, 1] * es + eo, *(np.pi / 4 * q[:4]), d)
        k, l = divmod(j, md + 1)
        np.testing.assert_allclose(basis, np.outer(p2m[:, k], p2m[:, l]).ravel(), rtol=0, atol=1e-14)


@pytest.mark.parametrize("L", [1, 2, 4, 6])
def test_host_locate_matches_oracle(sb, orc, L):
    """The arithmetic descent against the oracle's panel-reading descent, on
    random points and on points built on every mid-line, edge and corner of
    the level-(L-1) grid of each face (the tie rules)."""
    h = sb.HostTree(sb.RunConfig(levels=L, interp_degree=4))
    o = orc.OracleFMM(L, 4, 4, 0.7, 0)
    rng = np.random.default_rng(8)
    v = rng.normal(size=(300, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    g = np.pi / 4 * np.linspace(-1, 1, 2 ** (L - 1) + 1)
    for face in range(6):
        for xi in g:
            for eta in g[:: max(1, len(g) // 5)]:
                v = np.vstack([v, orc.xyz_from_xieta(xi, eta, face)[0]])
    for p in v:
        assert h.locate(*p) == o.point_locate(*p)
    if L > 1:
        assert h.locate(float("nan"), 0.0, 0.0) == -1


def test_invalid_parameters_are_rejected(sb):
    for kw in [dict(levels=0, interp_degree=4), dict(levels=3, interp_degree=11),
               dict(levels=3, interp_degree=4, leaf_degree=0), dict(levels=3, interp_degree=4, fmm_theta=0.0)]:
        with pytest.raises(sb.SbslpmError):
            sb.HostTree(sb.RunConfig(**kw))
