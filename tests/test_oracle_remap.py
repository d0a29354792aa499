"""Pins the oracle's BVE displacement, departure points, remap and step
(bve_time_step.cpp:15-419) against exact special cases."""
import numpy as np
import pytest

OMEGA = 7.292e-5
DT = 3600.0


def _rot(x, y, z, axis, ang):
    """Rodrigues rotation of unit vectors by `ang` about unit `axis`."""
    k = np.asarray(axis, float) / np.linalg.norm(axis)
    v = np.stack([x, y, z], -1)
    c, s = np.cos(ang), np.sin(ang)
    kv = np.cross(np.broadcast_to(k, v.shape), v)
    return v * c + kv * s + np.outer((v @ k).ravel(), k).reshape(v.shape) * (1 - c)


@pytest.fixture(scope="module")
def setup(orc):
    o = orc.OracleFMM(3, 8, 7, 0.7, 1)
    x, y, z, a = o.spec_points()
    return o, x, y, z, a


def test_zero_velocity_is_identity_remap(orc, setup):
    o, x, y, z, a = setup
    zero = [np.zeros_like(x) for _ in range(3)]
    for combine in (0, 1):
        dep = o.bve_find_departure_points(x, y, z, [zero] * 4, DT, combine)
        np.testing.assert_allclose(dep[0], x, atol=2e-16)
        vor = np.sin(3 * x) * np.cos(2 * z) + y
        tr = np.stack([x * y, z], -1).copy()
        nv, ntr = o.bve_departure_to_target(z, dep, vor, OMEGA, tr)
        np.testing.assert_allclose(nv, vor, rtol=0, atol=1e-15)
        np.testing.assert_allclose(ntr, tr, rtol=0, atol=1e-15)


@pytest.mark.parametrize("combine", [0, 1])
def test_solid_body_rotation_departure_points(orc, setup, combine):
    """v = w x X (steady): textbook RK4 gives the exact backward rotation to
    O((w dt)^5); the reference combine (x - dt/6 k1, then + dt/3 k2, + dt/3 k3,
    + dt/6 k4, projecting after each) is evaluated independently here."""
    o, x, y, z, a = setup
    w = np.array([0.3, -0.2, 1.0]) * 2e-5
    vel = lambda X: np.cross(np.broadcast_to(w, X.shape), X)
    X = np.stack([x, y, z], -1)
    V = vel(X)
    stage = [V[..., 0].copy(), V[..., 1].copy(), V[..., 2].copy()]
    dep = o.bve_find_departure_points(x, y, z, [stage] * 4, DT, combine)
    dep = np.stack(dep, -1)
    P = lambda v: v / np.linalg.norm(v, axis=-1, keepdims=True)
    k1 = V
    k2 = vel(P(X - 0.5 * DT * k1))
    k3 = vel(P(X - 0.5 * DT * k2))
    k4 = vel(P(X - DT * k3))
    if combine == 1:
        expect = P(X - DT / 6 * (k1 + 2 * k2 + 2 * k3 + k4))
        exact = _rot(x, y, z, w, -np.linalg.norm(w) * DT)
        assert np.max(np.abs(dep - exact)) < 1e-8
    else:
        e = P(X - DT / 6 * k1)
        e = P(e + DT / 3 * k2)
        e = P(e + DT / 3 * k3)
        expect = P(e + DT / 6 * k4)
    # only the BLI error of the interpolated stage velocities separates the two
    assert np.max(np.abs(dep - expect)) < 1e-9


def test_absolute_vorticity_is_conserved(orc, setup):
    o, x, y, z, a = setup
    c = 3e-5
    vor = c - 2 * OMEGA * z  # absolute vorticity == c everywhere
    rng = np.random.default_rng(7)
    dep = np.stack([x, y, z], -1) + 0.01 * rng.normal(size=x.shape + (3,))
    dep /= np.linalg.norm(dep, axis=-1, keepdims=True)
    dep = [dep[..., i].copy() for i in range(3)]
    nv, _ = o.bve_departure_to_target(z, dep, vor, OMEGA)
    np.testing.assert_allclose(nv + 2 * OMEGA * z, c, rtol=0, atol=1e-12)


def test_displacement_update_and_proxy_interpolation(orc, setup):
    o, x, y, z, a = setup
    tau = 1800.0
    vel = [1e-6 * (y * z + 1), 1e-6 * x, -2e-6 * x * y]
    vor = np.cos(x) * z
    nv, dp, dx = o.bve_displacement(x, y, z, vel, vor, tau, OMEGA)
    X = np.stack([x, y, z], -1) + tau * np.stack(vel, -1)
    zp = X[..., 2] / np.linalg.norm(X, axis=-1)
    np.testing.assert_allclose(nv, vor + 2 * OMEGA * z - 2 * OMEGA * zp, rtol=0, atol=1e-18)
    for f in range(3):
        np.testing.assert_allclose(dp[f], tau * vel[f], rtol=0, atol=0)
    # non-leaf proxies: the displacement field evaluated at the proxy node (BLI accuracy)
    px, py, pz = o.proxy_points()
    field = [tau * 1e-6 * (py * pz + 1), tau * 1e-6 * px, -2 * tau * 1e-6 * px * py]
    for f in range(3):
        np.testing.assert_allclose(dx[f][: o.offset], field[f][: o.offset], rtol=0, atol=1e-8 * tau * 1e-6)


@pytest.mark.parametrize("combine", [0, 1])
def test_step_keeps_zonal_l1_field(orc, setup, combine):
    """omega = c z induces solid rotation about z with rate c/2 (psi = c z/2,
    u = -x cross grad psi); the field is invariant under it, so one step with
    Omega = 0 returns it to discretisation accuracy, and the stage-0 velocity is
    (c/2) z_hat cross x."""
    o, x, y, z, a = setup
    c = 2e-5
    vor = c * z
    nv, vel, _ = o.bve_back_rk4_step(x, y, z, a, vor, DT, 0.0, 1e-10, combine)
    expect = [-(c / 2) * y, (c / 2) * x, np.zeros_like(z)]
    err = np.sqrt(sum(np.sum((v - e) ** 2) for v, e in zip(vel, expect)) / sum(np.sum(e ** 2) for e in expect[:2]))
    assert err < 2e-2
    assert np.max(np.abs(nv - vor)) < 2e-2 * c


def test_swe_remap_identity_and_rotation(orc, setup):
    """swe_departure_to_target (swe_time_step.cpp:19-74): identity at the nodes,
    and under a small solid-body rotation each field is interpolated to the
    accuracy of the degree-(m-1) grid (smooth fields, no Coriolis terms)."""
    o, x, y, z, a = setup
    f = [np.sin(3 * x) * np.cos(2 * z) + y, x * z - 0.3, np.cos(x + 2 * y)]
    tr = np.stack([x * y, z], -1).copy()
    out, ntr = o.swe_departure_to_target([x, y, z], f, tr)
    for g, h in zip(out, f):
        np.testing.assert_allclose(g, h, rtol=0, atol=1e-15)
    np.testing.assert_allclose(ntr, tr, rtol=0, atol=1e-15)
    d = _rot(x, y, z, [0.2, -0.4, 0.9], 0.01)
    dep = [d[..., i].copy() for i in range(3)]
    out, _ = o.swe_departure_to_target(dep, f)
    exact = [np.sin(3 * dep[0]) * np.cos(2 * dep[2]) + dep[1], dep[0] * dep[2] - 0.3, np.cos(dep[0] + 2 * dep[1])]
    for g, h in zip(out, exact):
        assert np.max(np.abs(g - h)) < 1e-5
    # with no rotation rate the BVE remap of each field is the same interpolation
    for g, h in zip(out, f):
        nv, _ = o.bve_departure_to_target(z, dep, h, 0.0)
        np.testing.assert_allclose(g, nv, rtol=0, atol=1e-15)


@pytest.mark.parametrize("combine", [0, 1])
def test_combine_first_order_limit(orc, setup, combine):
    """Pins the RK4 combine of bve_find_departure_points (bve_time_step.cpp:288-303)
    without restating it.  The four stages read four different solid rotations
    v_s = w_s x X (k1 from stage 3, k2 from stage 2 at the first half point, k3 from
    stage 1, k4 from stage 0: bve_time_step.cpp:212-286).  As dt -> 0 every evaluation
    point tends to X and every projection removes only O(dt^2), so
        (dep - X) / dt -> (c3 w3 + c2 w2 + c1 w1 + c0 w0) x X,
    with (c3, c2, c1, c0) = (-1/6, +1/3, +1/3, +1/6) for the reference's signs
    (x - dt/6 k1, then x -= -dt/3 k2, -= -dt/3 k3, -= -dt/6 k4) and
    -(1/6, 1/3, 1/3, 1/6) for the textbook combine.  The rotations are not
    parallel, so flipping any one sign, or swapping which stage feeds which k,
    moves the limit by O(1); the remainder shrinks like dt."""
    o, x, y, z, a = setup
    ws = [np.array(v) * 1e-5 for v in ([0.3, -0.2, 1.0], [-0.7, 0.5, 0.1], [0.2, 0.9, -0.4], [1.1, 0.3, 0.6])]
    X = np.stack([x, y, z], -1)
    stages = []
    for w in ws:  # stage s of the caller's list holds v_s
        V = np.cross(np.broadcast_to(w, X.shape), X)
        stages.append([V[..., 0].copy(), V[..., 1].copy(), V[..., 2].copy()])
    coef = (1 / 6, 1 / 3, 1 / 3, -1 / 6) if combine == 0 else (-1 / 6, -1 / 3, -1 / 3, -1 / 6)  # c0, c1, c2, c3
    wl = sum(c * w for c, w in zip(coef, ws))
    limit = np.cross(np.broadcast_to(wl, X.shape), X)
    scale = np.max(np.abs(limit))
    errs = []
    for dt in (3600.0, 360.0, 36.0):
        dep = np.stack(o.bve_find_departure_points(x, y, z, stages, dt, combine), -1)
        errs.append(np.max(np.abs((dep - X) / dt - limit)) / scale)
    # first order: each factor 10 in dt divides the remainder by about 10
    assert errs[2] < 2e-3, errs
    assert errs[1] / errs[2] > 7 and errs[0] / errs[1] > 7, errs
    # the other combine's limit is O(1) away: the pin tells the two apart
    other = (-1 / 6, -1 / 3, -1 / 3, -1 / 6) if combine == 0 else (1 / 6, 1 / 3, 1 / 3, -1 / 6)
    wo = sum(c * w for c, w in zip(other, ws))
    lo = np.cross(np.broadcast_to(wo, X.shape), X)
    assert np.max(np.abs(lo - limit)) / scale > 0.5
