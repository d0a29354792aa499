"""Pins the oracle's tree, traversal, passes and pair sums against invariants,
brute force and closed forms (initialize_cubed_sphere.cpp:64-288,
fmm_funcs.cpp:17-123, 341-392, fmm_funcs_impl.hpp:17-188,
bve_vel_impl.hpp:34-70, inverse_poisson_impl.hpp:23-52, direct_sum_impl.hpp:18-27)."""
import numpy as np
import pytest

W = None


@pytest.fixture(scope="module", autouse=True)
def _w(sb):
    global W
    W = sb.workloads


@pytest.mark.parametrize("L", [1, 2, 3, 4])
def test_tree_counts_areas_and_quadrants(orc, L):
    o = orc.OracleFMM(L, 4, 4, 0.7, 0)
    assert o.P == 2 * (4 ** L - 1) and o.Pl == 6 * 4 ** (L - 1)
    ints, dbls = o.panels()
    start = 0
    for lev in range(L):
        cnt = 6 * 4 ** lev
        assert np.all(ints[start:start + cnt, 1] == lev)
        assert abs(dbls[start:start + cnt, 5].sum() - 4 * np.pi) < 1e-12  # areas tile the sphere
        start += cnt
    for p in range(o.offset):
        ch = ints[p, 3:7]
        assert np.all(ints[ch, 2] == p) and np.all(ints[ch, 0] == ints[p, 0])
        mx = 0.5 * (dbls[p, 0] + dbls[p, 1])
        me = 0.5 * (dbls[p, 2] + dbls[p, 3])
        # children 1..4 = (hi,hi), (lo,hi), (lo,lo), (hi,lo)
        assert dbls[ch[0], 0] == mx and dbls[ch[0], 2] == me
        assert dbls[ch[1], 1] == mx and dbls[ch[1], 2] == me
        assert dbls[ch[2], 1] == mx and dbls[ch[2], 3] == me
        assert dbls[ch[3], 0] == mx and dbls[ch[3], 3] == me
    assert np.all(ints[o.offset:, 7] == 1) and np.all(ints[:o.offset, 7] == 0)


@pytest.mark.parametrize("L,theta", [(2, 0.7), (3, 0.7), (3, 0.4), (4, 0.9)])
def test_traversal_covers_every_leaf_pair_once(orc, L, theta):
    o = orc.OracleFMM(L, 4, 4, theta, 0)
    ints, dbls = o.panels()
    il = o.interactions()

    def leaves(p):
        if ints[p, 7]:
            return [p]
        return [q for c in ints[p, 3:7] for q in leaves(c)]

    cover = np.zeros((o.P, o.P), dtype=np.int32)
    centre = {}
    for (t, s, typ) in il:
        for a in leaves(t):
            for b in leaves(s):
                cover[a, b] += 1
        tl = 2 * (ints[t, 1] < L - 1) + (ints[s, 1] < L - 1)
        if typ != 0 or not (ints[t, 7] and ints[s, 7]):
            assert typ == tl
    lv = np.arange(o.offset, o.P)
    assert np.all(cover[np.ix_(lv, lv)] == 1)
    # well-separated pairs satisfy the MAC
    for (t, s, typ) in il:
        for p in (t, s):
            if p not in centre:
                xi = np.pi / 4 * 0.5 * (dbls[p, 0] + dbls[p, 1])
                eta = np.pi / 4 * 0.5 * (dbls[p, 2] + dbls[p, 3])
                centre[p] = orc.xyz_from_xieta(xi, eta, ints[p, 0])[0]
        dist = np.arccos(np.clip(centre[t] @ centre[s], -1, 1))
        if typ != 0:
            assert (dbls[t, 4] + dbls[s, 4]) / dist < theta


def test_spec_point_weights_integrate(orc, sb):
    W = sb.workloads
    for (L, d, md, rule) in [(3, 8, 8, 0), (3, 8, 7, 1), (3, 8, 7, 2)]:
        o = orc.OracleFMM(L, d, md, 0.7, rule)
        x, y, z, a = o.spec_points()
        assert abs(a.sum() - 4 * np.pi) < 1e-12
        assert np.allclose(x * x + y * y + z * z, 1.0, atol=1e-15)
        if rule != 2:  # spectral-quality weights: integral of z^2 is 4 pi/3 to O(h^2)
            assert abs(np.sum(a * z * z) - 4 * np.pi / 3) < 2e-3


def test_point_locate_contains_point(orc):
    o = orc.OracleFMM(4, 4, 4, 0.7, 0)
    ints, dbls = o.panels()
    rng = np.random.default_rng(3)
    v = rng.normal(size=(400, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for p in v:
        leaf = o.point_locate(*p)
        assert ints[leaf, 7] == 1
        face = orc.face_from_xyz(*p)
        assert ints[leaf, 0] == face
        xe, _ = orc.xieta_from_xyz(*p, face=face)
        q = dbls[leaf]
        assert np.pi / 4 * q[0] - 1e-15 <= xe[0] <= np.pi / 4 * q[1] + 1e-15
        assert np.pi / 4 * q[2] - 1e-15 <= xe[1] <= np.pi / 4 * q[3] + 1e-15
    assert o.point_locate(float("nan"), 0.0, 1.0) == -1


@pytest.mark.parametrize("md", [8, 6])
def test_upward_pass_preserves_total_charge(orc, md):
    o = orc.OracleFMM(3, 8, md, 0.7, 1)
    x, y, z, a = o.spec_points()
    f = np.cos(3 * x) + y * z
    qp, qx = o.upward_pass(a, f)
    total = np.sum(a * f)
    np.testing.assert_allclose(qp, a * f)
    start = 0
    for lev in range(3):  # partition of unity: each level's proxies carry the total charge
        cnt = 6 * 4 ** lev
        assert abs(qx[start:start + cnt].sum() - total) < 1e-12 * max(1, abs(total))
        start += cnt


@pytest.mark.parametrize("md", [8, 6])
def test_downward_pass_exact_for_polynomials(orc, md):
    """L2L/L2P reproduce a field that is a polynomial of degree <= min(d, m-1)
    in each label coordinate of a face."""
    d = 8
    o = orc.OracleFMM(3, d, md, 0.7, 0)
    ints, dbls = o.panels()
    tab_d = orc.interp_init(d)
    tab_m = orc.interp_init(md)
    deg = min(d, md)
    rng = np.random.default_rng(4)
    coef = rng.normal(size=(deg + 1, deg + 1))
    poly = lambda a, b: sum(coef[i, j] * a ** i * b ** j for i in range(deg + 1) for j in range(deg + 1))
    phi = np.zeros((o.P, o.n))
    for p in range(6):  # only the face panels hold the field; L2L carries it down
        xo, xs = 0.5 * (dbls[p, 1] + dbls[p, 0]), 0.5 * (dbls[p, 1] - dbls[p, 0])
        eo, es = 0.5 * (dbls[p, 3] + dbls[p, 2]), 0.5 * (dbls[p, 3] - dbls[p, 2])
        phi[p] = poly(tab_d[:, 0] * xs + xo, tab_d[:, 1] * es + eo)
    sols = o.downward_pass([phi])[0]
    for leaf in range(o.Pl):
        p = o.offset + leaf
        xo, xs = 0.5 * (dbls[p, 1] + dbls[p, 0]), 0.5 * (dbls[p, 1] - dbls[p, 0])
        eo, es = 0.5 * (dbls[p, 3] + dbls[p, 2]), 0.5 * (dbls[p, 3] - dbls[p, 2])
        expect = poly(tab_m[:, 0] * xs + xo, tab_m[:, 1] * es + eo)
        np.testing.assert_allclose(sols[leaf], expect, rtol=0, atol=1e-11)


def test_poisson_closed_form_converges(orc, sb):
    """sh43-type field (degree-4 harmonic): psi = f/20 (the reference's sign, Lap psi = -f)."""
    W = sb.workloads
    errs = []
    for L in (2, 3, 4):
        o = orc.OracleFMM(L, 8, 8, 0.7, 0)
        x, y, z, a = o.spec_points()
        f = 1.770130769779931 * x * (x * x - 3.0 * y * y) * z  # initial_conditions.cpp:71
        psi = o.poisson_solve(a, f)
        errs.append(W.rel_l2(psi, f / 20.0))
    assert errs[-1] < 0.015
    assert errs[0] / errs[1] > 2.5 and errs[1] / errs[2] > 2.5


def test_rh4_velocity_closed_form_converges(orc, sb):
    W = sb.workloads
    errs = []
    for L in (2, 3, 4):
        o = orc.OracleFMM(L, 8, 8, 0.7, 0)
        x, y, z, a = o.spec_points()
        vel = o.bve_velocity(a, W.rh4_vorticity(x, y, z), 1e-10)
        errs.append(W.vec_rel_l2(vel, W.rh4_velocity(x, y, z)))
    assert errs[-1] < 0.01
    assert errs[0] / errs[1] > 2.5 and errs[1] / errs[2] > 2.5


@pytest.mark.parametrize("md,theta", [(8, 0.7), (7, 0.7), (7, 0.5)])
def test_fmm_matches_direct_sum(orc, sb, md, theta):
    W = sb.workloads
    o = orc.OracleFMM(3, 8, md, theta, 1)
    x, y, z, a = o.spec_points()
    vor = W.gaussian_vortices(x, y, z, a, seed=5)
    flat = [v.ravel().copy() for v in (x, y, z, a, vor)]
    ds = orc.direct_sum_bve_vel(*flat, 1e-4)
    vel = o.bve_velocity(a, vor, 1e-4)
    assert W.vec_rel_l2([v.ravel() for v in vel], ds) < (5e-7 if theta == 0.7 else 5e-8)
    # log kernel: the FMM skips coincident (duplicated edge) points, so use the guarded direct sum
    psi = o.poisson_solve(a, vor)
    dl = orc.direct_sum_inv_lap_targets(flat[0], flat[1], flat[2], *flat)
    assert W.rel_l2(psi.ravel(), dl) < 1e-6


def test_direct_sum_inv_lap_skips_only_self(orc):
    rng = np.random.default_rng(6)
    v = rng.normal(size=(50, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    x, y, z = [v[:, i].copy() for i in range(3)]
    area, pots = rng.uniform(0.1, 1, 50), rng.normal(size=50)
    out = orc.direct_sum_inv_lap(x, y, z, area, pots, 10, 30)
    expect = np.zeros(50)
    for i in range(10, 30):
        for j in range(50):
            if i != j:
                expect[i] += -1 / (4 * np.pi) * np.log(1 - v[i] @ v[j]) * area[j] * pots[j]
    np.testing.assert_allclose(out, expect, rtol=1e-13, atol=1e-15)
