"""Pins the oracle's shallow-water step pieces (swe_time_step.cpp, deriv_funcs_impl.hpp,
laplacian_impl.hpp, topography.cpp) with checks that do not restate them:

  * the panel Laplacian against Delta Y_lm = -l(l+1) Y_lm, converging spectrally;
  * the leaf-rule velocity gradient against the surface gradients of the coordinate
    functions and of Y_31 (great-circle differences), pole nodes included;
  * the tendencies of a solid-body rotation against their hand-derived closed form;
  * the cone topography against a latitude/longitude formula written differently;
  * the regularised-Laplacian kernel: K * 1 = 8/eps and the low-degree eigenvalues
    near -l(l+1), and its FMM against an O(N^2) sum with displaced sources;
  * the height correction at rest (interpolation at the nodes is the identity);
  * the full step at rest leaves the state unchanged.
The reference rule of the velocity gradient (leaf-local face index, metric applied
twice, pole shortcut) is pinned through its parts only: the differentiation and the
vector conversion above, and equality with the leaf rule where the two coincide."""
import numpy as np
import pytest

import bench


def sh(W, l, m, x, y, z):
    return W.real_sph_harm_all(l, x, y, z)[(l, m)]


@pytest.mark.parametrize("L,d,tol", [(2, 8, 5e-4), (3, 8, 2e-5), (3, 10, 1e-7)])
def test_panel_laplacian_spherical_harmonics(orc, sb, L, d, tol):
    W = sb.workloads
    o = orc.OracleFMM(L, d, d, 0.7, 0)
    x, y, z, _ = o.spec_points()
    f = sh(W, 3, 2, x, y, z) + 0.5 * sh(W, 4, -1, x, y, z)
    expect = -12 * sh(W, 3, 2, x, y, z) - 10 * sh(W, 4, -1, x, y, z)
    assert W.rel_l2(o.panel_laplacian(f), expect) < tol


@pytest.mark.parametrize("L,d,tol_c,tol_y", [(2, 8, 2e-6, 5e-5), (3, 10, 1e-9, 1e-8)])
def test_leaf_rule_gradient_is_the_surface_gradient(orc, sb, L, d, tol_c, tol_y):
    W = sb.workloads
    o = orc.OracleFMM(L, d, d, 0.7, 0)
    x, y, z, _ = o.spec_points()
    g = o.velocity_gradient([x, y, z], rule=1)
    expect = [1 - x * x, -x * y, -x * z, -x * y, 1 - y * y, -y * z, -x * z, -y * z, 1 - z * z]
    assert max(np.abs(g[k] - expect[k]).max() for k in range(9)) < tol_c
    f = sh(W, 3, 1, x, y, z)
    G = bench.tangent_gradient(lambda a, b, c: sh(W, 3, 1, a, b, c), x.ravel(), y.ravel(), z.ravel())
    g = o.velocity_gradient([f, f, f], rule=1)
    assert max(np.abs(g[k].ravel() - G[:, k]).max() for k in range(3)) < tol_y


def test_reference_rule_gradient_differs_where_the_reference_does(orc, sb):
    """With one level (leaves = faces, so panel i is leaf i) the two rules differ only by the
    doubled metric and the pole shortcut; the leaf rule is exact for linear functions, the
    reference rule is not (its doubled metric scales the gradient off the panel centre)."""
    o = orc.OracleFMM(1, 8, 8, 0.7, 0)
    x, y, z, _ = o.spec_points()
    g0 = o.velocity_gradient([x, y, z], rule=0)
    g1 = o.velocity_gradient([x, y, z], rule=1)
    # at the face centres of faces 0..3 (X = Y = 0) the metric is the identity: both agree
    centre = (np.abs(np.abs(x) - 1) < 1e-14) | (np.abs(np.abs(y) - 1) < 1e-14)
    assert centre.sum() == 4
    for k in range(9):
        assert np.allclose(g0[k][centre], g1[k][centre], atol=1e-12)
    assert max(np.abs(g0[k] - g1[k]).max() for k in range(9)) > 0.1


@pytest.mark.parametrize("rule", [0, 1])
def test_tendency_of_solid_body_rotation(orc, sb, rule):
    """u = W0 z^ x r, vor = 2 W0 z, div = 0: vor_t = 0, h_t = 0 and
    div_t = 4 Om W0 z^2 - tr(grad u grad u) - lap - |u|^2 - 2 Om W0 (x^2 + y^2), with
    tr(grad u grad u) = W0^2 (2 x^2 y^2 - 2 (1 - x^2)(1 - y^2)) from the surface gradients
    grad u_x = -W0 (y^ - y r), grad u_y = W0 (x^ - x r).  Checked with the leaf rule; with one
    level for the reference rule (its doubled metric is exact only at panel centres, so it is
    compared with the leaf rule there)."""
    W = sb.workloads
    L = 3 if rule == 1 else 1
    o = orc.OracleFMM(L, 8, 8, 0.7, 0)
    x, y, z, _ = o.spec_points()
    W0, Om = 3e-6, 7.292e-5
    vel = [-W0 * y, W0 * x, 0 * z]
    vor, div, h = 2 * W0 * z, 0 * z, 1 + 0.1 * z
    lap = 0.3 * x * y
    vt, dt_, ht = o.swe_tendency(x.ravel(), y.ravel(), z.ravel(), vel, vor, div, h, lap, Om, rule)
    gc = W0 ** 2 * (2 * x * x * y * y - 2 * (1 - x * x) * (1 - y * y))
    r2 = x * x + y * y
    expect = 4 * Om * W0 * z * z - gc - lap - W0 ** 2 * r2 - 2 * Om * W0 * r2
    assert np.abs(vt).max() == 0 and np.abs(ht).max() == 0
    if rule == 1:
        assert np.abs(dt_ - expect).max() < 1e-8 * np.abs(expect).max()
    else:
        centre = (np.abs(np.abs(x) - 1) < 1e-14) | (np.abs(np.abs(y) - 1) < 1e-14)
        assert np.abs(dt_ - expect)[centre].max() < 1e-12 * np.abs(expect).max()


@pytest.mark.parametrize("rule", [0, 1])
def test_cone_topography(orc, sb, rule):
    W = sb.workloads
    o = orc.OracleFMM(3, 6, 6, 0.7, 0)
    x, y, z, _ = o.spec_points()
    h = 1 + 0.01 * x
    rng = np.random.default_rng(3)
    dp = [1e-2 * rng.normal(size=x.shape) for _ in range(3)]
    dx = [1e-2 * rng.normal(size=(o.P, o.n)) for _ in range(3)]
    eff = o.apply_topography(x.ravel(), y.ravel(), z.ravel(), h, 1, rule, dp, dx)
    # the displacement each particle reads: its own (leaf rule) or row i of the [P][n] array
    # whose leaf rows sit at i + offset (reference rule)
    if rule == 1:
        d = [v.copy() for v in dp]
    else:
        off = o.P - o.Pl
        d = [np.concatenate([dx[f][:off], dp[f]])[: o.Pl] for f in range(3)]
    p = np.stack([x + d[0], y + d[1], z + d[2]])
    p /= np.linalg.norm(p, axis=0)
    lat = np.arcsin(p[2])
    lon = np.mod(np.arctan2(p[1], p[0]), 2 * np.pi)
    r2 = (lon - np.pi / 2) ** 2 + (lat - np.pi / 6) ** 2
    assert np.abs(eff - (h + 2000.0 / 6371000.0 * np.exp(-16 * r2))).max() < 1e-14
    assert np.array_equal(o.apply_topography(x.ravel(), y.ravel(), z.ravel(), h, 0, rule, dp, dx), h)


def test_laplacian_kernel_integral_and_eigenvalues(orc, sb):
    W = sb.workloads
    o = orc.OracleFMM(4, 8, 8, 0.7, 0)
    x, y, z, a = o.spec_points()

    def apply(f):
        qp, qx = o.upward_pass(a, f)
        pp, px = o.laplacian_interactions(qp, qx)
        return o.downward_pass([px], [pp])[0]

    one = apply(np.ones_like(x))
    assert abs(np.mean(one) / 80.0 - 1) < 1e-3 and np.std(one) < 0.01 * 80
    for l, m, tol in ((1, 0, 0.01), (2, 1, 0.03), (3, -2, 0.05)):
        f = sh(W, l, m, x, y, z)
        est = apply(f) - 80.0 * f
        lam = np.sum(est * f * a) / np.sum(f * f * a)
        assert abs(lam / (-l * (l + 1)) - 1) < tol


def test_laplacian_fmm_matches_direct_sum(orc, sb):
    """FMM of the kernel over displaced sources against the O(N^2) sum of the same kernel
    (sources P(x + disp), targets at the particles)."""
    from scipy.special import erf
    W = sb.workloads
    o = orc.OracleFMM(3, 8, 8, 0.7, 0)
    x, y, z, a = o.spec_points()
    f = sh(W, 2, 1, x, y, z) + 0.3
    vel = [1e-6 * (y * z + 0.5), 2e-6 * np.sin(3 * x), -2e-6 * x * y]
    _, dp, dx = o.bve_displacement(x, y, z, vel, 0 * x, 900.0, 0.0)
    qp, qx = o.upward_pass(a, f)
    pp, px = o.laplacian_interactions(qp, qx, dp, dx)
    fmm = o.downward_pass([px], [pp])[0].ravel()
    s = np.stack([x + dp[0], y + dp[1], z + dp[2]]).reshape(3, -1)
    s /= np.linalg.norm(s, axis=0)
    t = np.stack([x, y, z]).reshape(3, -1)
    eps = 0.1
    dot = t.T @ s
    d2 = dot * dot
    h1 = np.exp(-1 / eps) * np.pi ** -1.5 * (2 * dot * (-1 + 3 * eps + d2))
    h2 = (1 + erf(dot / eps)) * np.exp((d2 - 1) / eps) / (np.sqrt(eps) * np.pi) * (
        2 * eps * eps + 2 * d2 * (d2 - 1) + eps * (7 * d2 - 1))
    direct = ((h1 + h2) / eps ** 2.5) @ (a * f).ravel()
    assert W.rel_l2(fmm, direct) < 1e-6


def test_height_correction_at_rest_is_the_node_value(orc, sb):
    W = sb.workloads
    o = orc.OracleFMM(3, 6, 6, 0.7, 0)
    x, y, z, _ = o.spec_points()
    h, ht = 1 + 0.2 * x * y, 0.01 * z
    lap = 0.5 * z
    got = o.swe_height_correction(x.ravel(), y.ravel(), z.ravel(), [0 * x] * 3, h, ht, 300.0, lap)
    assert np.abs(got - (lap - 80.0 * (h + 300.0 * ht))).max() < 1e-11


@pytest.mark.parametrize("rule", [0, 1])
def test_step_at_rest_keeps_the_state(orc, sb, rule):
    W = sb.workloads
    o = orc.OracleFMM(2, 6, 6, 0.7, 0)
    x, y, z, a = o.spec_points()
    H = 5960.0 / 6371000.0
    v, d, h, vel, lap, _ = o.swe_back_rk4_step(x.ravel(), y.ravel(), z.ravel(), a, 0 * x, 0 * x, H + 0 * x, 900.0,
                                               7.292e-5, 1e-4, 0, 0, rule)
    assert np.abs(v).max() == 0 and np.abs(d).max() == 0
    assert np.abs(h / H - 1).max() < 1e-14
    assert max(np.abs(u).max() for u in vel) == 0 and np.abs(lap).max() == 0


@pytest.mark.parametrize("rule", [0, 1])
def test_gaussian_topography(orc, sb, rule):
    """h_s = H exp(-(r/w)^2) against the great-circle distance taken as arccos(p.c)."""
    o = orc.OracleFMM(3, 6, 6, 0.7, 0)
    x, y, z, _ = o.spec_points()
    mtn = (1500.0, 0.2, 0.4, -1.1)
    o.set_swe_physics(1.0, mtn, 0)
    h = 5000.0 + 10.0 * x
    rng = np.random.default_rng(4)
    dp = [1e-2 * rng.normal(size=x.shape) for _ in range(3)]
    dx = [1e-2 * rng.normal(size=(o.P, o.n)) for _ in range(3)]
    eff = o.apply_topography(x.ravel(), y.ravel(), z.ravel(), h, 2, rule, dp, dx)
    if rule == 1:
        d = [v.copy() for v in dp]
    else:
        off = o.P - o.Pl
        d = [np.concatenate([dx[f][:off], dp[f]])[: o.Pl] for f in range(3)]
    p = np.stack([x + d[0], y + d[1], z + d[2]])
    p /= np.linalg.norm(p, axis=0)
    c = np.array([np.cos(mtn[2]) * np.cos(mtn[3]), np.cos(mtn[2]) * np.sin(mtn[3]), np.sin(mtn[2])])
    r = np.arccos(np.clip(np.tensordot(c, p, 1), -1, 1))
    assert np.abs(eff - (h + 1500.0 * np.exp(-(r / 0.2) ** 2))).max() < 1e-8
    assert (eff - h).max() > 1000.0  # some particles sit near the top


def test_gravity_scales_the_surface_laplacian_term(orc, sb):
    o = orc.OracleFMM(2, 6, 6, 0.7, 0)
    x, y, z, _ = o.spec_points()
    vel = [1e-6 * (-y), 1e-6 * x, 0 * z]
    vor, div, h, lap = 2e-6 * z, 1e-7 * x * y, 5000 + 100 * z, 3.0 + x
    t1 = o.swe_tendency(x, y, z, vel, vor, div, h, lap, 7.292e-5, 1)
    o.set_swe_physics(2.5e-13)
    tg = o.swe_tendency(x, y, z, vel, vor, div, h, lap, 7.292e-5, 1)
    assert np.array_equal(t1[0], tg[0]) and np.array_equal(t1[2], tg[2])
    assert np.abs((tg[1] - t1[1]) - (1 - 2.5e-13) * lap).max() < 1e-12 * np.abs(lap).max()


def test_regularised_stage1_laplacian_and_surface_correction(orc, sb):
    """Stage-1 surface Laplacian of the regularised form at rest.  A flat free surface over a
    mountain (h = c - h_s) has zero Laplacian with the surface correction; without it the
    correction leaves 8/0.1 h_s.  A degree-2 harmonic height has eigenvalue near -6."""
    W = sb.workloads
    o = orc.OracleFMM(4, 6, 6, 0.7, 1)
    x, y, z, a = o.spec_points()
    mtn = (1500.0, 0.25, 0.3, 2.0)
    hs = W.mountain_height(x, y, z, mtn)
    h = 5000.0 - hs
    args = (x.ravel(), y.ravel(), z.ravel(), a, 0 * x, 0 * x)
    lap = {}
    for form in (1, 3):
        o.set_swe_physics(1.0, mtn, form)
        lap[form] = o.swe_back_rk4_step(*args, h, 900.0, 7.292e-5, 1e-4, 0, 2, 1)[4]
    # K * 1 = 8/0.1 holds to the quadrature's ~1% at this size (K * c - 80 c leaves 3.2% of
    # 80 max h_s at c = 5000), against the full 80 h_s the height-only correction leaves
    scale = 80.0 * hs.max()
    assert np.abs(lap[3]).max() < 0.05 * scale
    assert np.abs(lap[1] - 80.0 * hs).max() < 0.05 * scale and np.abs(lap[1]).max() > 0.9 * scale
    o.set_swe_physics(1.0, mtn, 3)
    f = sh(W, 2, 1, x, y, z)
    got = o.swe_back_rk4_step(*args, f, 900.0, 7.292e-5, 1e-4, 0, 0, 1)[4]
    lam = np.sum(got * f * a) / np.sum(f * f * a)
    assert abs(lam / -6.0 - 1) < 0.03
