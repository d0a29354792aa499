"""Pins the oracle's interpolation, quadrature and cubed-sphere maps against
closed forms and brute force (interp_funcs.cpp:11-73, interp_funcs_impl.hpp:11-112,
cubed_sphere_transforms_impl.hpp:8-123, general_utils.cpp:5-21)."""
import numpy as np
import pytest


@pytest.mark.parametrize("d", [2, 4, 6, 8, 10])
def test_interp_init_nodes_and_weights(orc, d):
    tab = orc.interp_init(d)
    k = np.arange(d + 1)
    nodes = np.cos(np.pi * k / d)
    nodes[np.abs(nodes) < 1e-16] = 0.0
    # node index = xi_index*(d+1) + eta_index
    np.testing.assert_allclose(tab[:, 0], np.repeat(nodes, d + 1), rtol=0, atol=3e-16)
    np.testing.assert_allclose(tab[:, 1], np.tile(nodes, d + 1), rtol=0, atol=3e-16)
    # second-kind barycentric weights (Berrut-Trefethen): 1/2, -1, 1, ..., (-1)^d/2
    w = (-1.0) ** k
    w[0] *= 0.5
    w[-1] *= 0.5
    np.testing.assert_array_equal(tab[:, 2], np.outer(w, w).ravel())


@pytest.mark.parametrize("d", [2, 4, 6, 8, 10])
def test_reference_quadrature_even_degree(orc, d):
    """The reference formula sums to 2 and is exact for degree <= d-1 (not d:
    the j = d/2 cosine term carries factor 2 instead of 1)."""
    tab = orc.interp_init(d)
    cc = np.sqrt(tab[:: d + 2, 3])  # diagonal entries cc_i^2
    x = np.cos(np.pi * np.arange(d + 1) / d)
    assert abs(cc.sum() - 2.0) < 1e-14
    for p in range(d):
        exact = 0.0 if p % 2 else 2.0 / (p + 1)
        assert abs(np.dot(cc, x ** p) - exact) < 1e-13, p
    exact = 2.0 / (d + 1)
    assert abs(np.dot(cc, x ** d) - exact) > 1e-6  # degree d is not integrated exactly


@pytest.mark.parametrize("d", [3, 5, 7])
def test_reference_quadrature_odd_degree_is_broken(orc, d):
    tab = orc.interp_init(d)
    cc = np.array([tab[i * (d + 2), 3] for i in range(d + 1)])
    assert abs(np.sqrt(cc).sum() - 2.0) > 1e-3 or np.any(cc == 0)


@pytest.mark.parametrize("d", [1, 2, 3, 5, 7, 8, 9, 10])
def test_textbook_clenshaw_curtis(orc, d):
    w = orc.cc_weights_textbook(d)
    x = np.cos(np.pi * np.arange(d + 1) / d)
    for p in range(d + 1):
        exact = 0.0 if p % 2 else 2.0 / (p + 1)
        assert abs(np.dot(w, x ** p) - exact) < 1e-13


def _lagrange_brute(nodes, t):
    out = np.ones(len(nodes))
    for i in range(len(nodes)):
        for j in range(len(nodes)):
            if i != j:
                out[i] *= (t - nodes[j]) / (nodes[i] - nodes[j])
    return out


@pytest.mark.parametrize("d", [1, 4, 7, 8, 10])
def test_bli_against_brute_force_lagrange(orc, d):
    rng = np.random.default_rng(d)
    lo, hi, elo, ehi = -0.3, 0.2, 0.1, 0.45
    xn = np.cos(np.pi * np.arange(d + 1) / d) * 0.5 * (hi - lo) + 0.5 * (hi + lo)
    en = np.cos(np.pi * np.arange(d + 1) / d) * 0.5 * (ehi - elo) + 0.5 * (ehi + elo)
    for _ in range(20):
        xi, eta = rng.uniform(lo, hi), rng.uniform(elo, ehi)
        basis, err = orc.interp_vals_bli(xi, eta, lo, hi, elo, ehi, d)
        assert err == 0
        expect = np.outer(_lagrange_brute(xn, xi), _lagrange_brute(en, eta)).ravel()
        np.testing.assert_allclose(basis, expect, rtol=0, atol=1e-12)
        assert abs(basis.sum() - 1.0) < 1e-13  # partition of unity


def test_bli_reproduces_polynomials_and_snaps_to_nodes(orc):
    d = 8
    lo, hi = -0.5, 0.7
    xn = np.cos(np.pi * np.arange(d + 1) / d) * 0.6 + 0.1
    rng = np.random.default_rng(1)
    coef = rng.normal(size=(d + 1, d + 1))
    f = lambda a, b: sum(coef[i, j] * a ** i * b ** j for i in range(d + 1) for j in range(d + 1))
    vals = np.array([f(a, b) for a in xn for b in xn])
    for xi, eta in rng.uniform(lo, hi, size=(10, 2)):
        basis, _ = orc.interp_vals_bli(xi, eta, lo, hi, lo, hi, d)
        assert abs(basis @ vals - f(xi, eta)) < 1e-9 * max(1.0, abs(f(xi, eta)))
    # exactly at a node (and within 1e-15 of it): a delta
    basis, _ = orc.interp_vals_bli(xn[3], xn[5] + 5e-16, lo, hi, lo, hi, d)
    e = np.zeros((d + 1) ** 2)
    e[3 * (d + 1) + 5] = 1.0
    np.testing.assert_array_equal(basis, e)


def test_bli_out_of_panel_is_an_error(orc):
    _, err = orc.interp_vals_bli(0.5 + 2e-15, 0.0, -0.5, 0.5, -0.5, 0.5, 4)
    assert err != 0
    _, err = orc.interp_vals_bli(0.5 + 5e-16, 0.0, -0.5, 0.5, -0.5, 0.5, 4)
    assert err == 0
    _, err = orc.interp_vals_bli(float("nan"), 0.0, -0.5, 0.5, -0.5, 0.5, 4)
    assert err != 0


def test_maps_roundtrip_and_face_rule(orc):
    rng = np.random.default_rng(2)
    for face in range(6):
        for xi, eta in rng.uniform(-np.pi / 4, np.pi / 4, size=(25, 2)):
            xyz, err = orc.xyz_from_xieta(xi, eta, face)
            assert err == 0
            assert abs(np.linalg.norm(xyz) - 1.0) < 1e-15
            assert orc.face_from_xyz(*xyz) == face
            xe, _ = orc.xieta_from_xyz(*xyz, face=face)
            assert abs(xe[0] - xi) < 1e-15 and abs(xe[1] - eta) < 1e-15
            np.testing.assert_allclose(orc.xieta_from_xyz(*xyz), xe, rtol=0, atol=0)
    # face centres
    centres = {0: (1, 0, 0), 1: (0, 1, 0), 2: (-1, 0, 0), 3: (0, -1, 0), 4: (0, 0, 1), 5: (0, 0, -1)}
    for f, c in centres.items():
        xyz, _ = orc.xyz_from_xieta(0.0, 0.0, f)
        np.testing.assert_allclose(xyz, c, atol=1e-16)
    _, err = orc.xyz_from_xieta(0.0, 0.0, 6)
    assert err != 0  # cubed_sphere_transforms.cpp:63-65 throws for a bad face


def test_sphere_tri_area_octant(orc):
    """Girard: the octant triangle has area 4 pi / 8."""
    a = orc.sphere_tri_area([1, 0, 0], [0, 1, 0], [0, 0, 1])
    assert abs(a - np.pi / 2) < 1e-14
