"""Pins the oracle's FMM evaluation at arbitrary targets (the `_ll` pass
structure of fmm_funcs.cpp:394-419: far field by BLI from the containing leaf,
child_panel_interp_3_ll, fmm_funcs_impl.hpp:318-364; near field over the
leaf's type-0/1 interactions, swe_vel_impl.hpp:188-255).

Checks that do not restate the function:
  * at interior leaf particles it must reproduce the particle solve, which is
    itself pinned against direct sums and closed forms;
  * at random and special points (face centres, cube corners, edge and
    panel mid-lines) it must agree with an O(N^2) direct sum to FMM accuracy;
  * the sh43-type field has psi = f/20 at every point of the sphere."""
import numpy as np
import pytest

CASES = {
    "ref-L3-d8": (3, 8, 8, 0.7, 0),
    "A'-L3-d8-m8": (3, 8, 7, 0.7, 1),
    "A'-L3-d10-m6": (3, 10, 5, 0.6, 1),
}
EPS = 1e-4


def interior(o):
    m = o.md + 1
    a, c = np.divmod(np.arange(m * m), m)
    return (a > 0) & (a < m - 1) & (c > 0) & (c < m - 1)


def special_points():
    pts = [np.eye(3)[i] * s for i in range(3) for s in (1, -1)]                     # face centres
    pts += [np.array([sx, sy, sz]) / np.sqrt(3) for sx in (1, -1) for sy in (1, -1) for sz in (1, -1)]
    pts += [np.array(v, dtype=float) / np.sqrt(2) for v in ([1, 1, 0], [0, -1, 1], [-1, 0, -1])]
    for xi, eta in ((0.0, 0.3), (0.2, 0.0), (np.pi / 8, -0.1), (-np.pi / 16, np.pi / 16)):
        X, Y = np.tan(xi), np.tan(eta)
        v = np.array([1.0, X, Y])
        pts.append(v / np.linalg.norm(v))                                           # panel mid-lines, face 0
    return np.array(pts)


def bs_direct(tx, ty, tz, x, y, z, q, eps):
    """Biot-Savart direct sum at separate targets: sum_j -q_j/(4 pi) (t x y_j)/(1 - t.y_j + eps)."""
    t = np.stack([tx, ty, tz], 1)
    s = np.stack([x, y, z], 1)
    g = -q[None, :] / (4 * np.pi) / (1 - t @ s.T + eps)
    w = g @ s
    return list(np.cross(t, w).T)


@pytest.mark.parametrize("name", list(CASES))
def test_targets_at_interior_particles_reproduce_solve(orc, sb, name):
    W = sb.workloads
    L, d, md, th, rule = CASES[name]
    o = orc.OracleFMM(L, d, md, th, rule)
    x, y, z, a = o.spec_points()
    f = W.gaussian_vortices(x, y, z, a, seed=21)
    mk = interior(o)
    tx, ty, tz = [v[:, mk].ravel().copy() for v in (x, y, z)]
    psi = o.poisson_solve(a, f)
    got = o.eval_targets("poisson", a, f, tx, ty, tz)[0]
    assert W.rel_l2(got, psi[:, mk].ravel()) < 1e-12
    vel = o.bve_velocity(a, f, EPS)
    gv = o.eval_targets("bve", a, f, tx, ty, tz, EPS)
    assert W.vec_rel_l2(gv, [v[:, mk].ravel() for v in vel]) < 1e-12


@pytest.mark.parametrize("name", ["A'-L3-d8-m8", "A'-L3-d10-m6"])
def test_targets_match_direct_sum(orc, sb, name):
    W = sb.workloads
    L, d, md, th, rule = CASES[name]
    o = orc.OracleFMM(L, d, md, th, rule)
    x, y, z, a = o.spec_points()
    f = W.random_sh_forcing(x, y, z, lmax=12, seed=4)[0]
    T = np.concatenate([np.stack(W.random_unit_vectors(400, seed=8), 1), special_points()])
    tx, ty, tz = [T[:, i].copy() for i in range(3)]
    flat = [v.ravel().copy() for v in (x, y, z, a, f)]
    psi = o.eval_targets("poisson", a, f, tx, ty, tz)[0]
    ref = orc.direct_sum_inv_lap_targets(tx, ty, tz, *flat)
    assert W.rel_l2(psi, ref) < 1e-6
    vel = o.eval_targets("bve", a, f, tx, ty, tz, EPS)
    refv = bs_direct(tx, ty, tz, flat[0], flat[1], flat[2], flat[3] * flat[4], EPS)
    assert W.vec_rel_l2(vel, refv) < 1e-6
    # a transposed BLI basis or a missing far field would show at every target, not on average only
    assert W.rel_linf(psi, ref) < 1e-5


def test_targets_closed_form(orc, sb):
    """sh43-type field: psi = f/20 everywhere (initial_conditions.cpp:52-74)."""
    W = sb.workloads
    o = orc.OracleFMM(4, 8, 7, 0.7, 1)
    x, y, z, a = o.spec_points()
    sh = lambda x, y, z: 1.770130769779931 * x * (x * x - 3.0 * y * y) * z
    T = np.concatenate([np.stack(W.random_unit_vectors(500, seed=9), 1), special_points()])
    tx, ty, tz = [T[:, i].copy() for i in range(3)]
    psi = o.eval_targets("poisson", a, sh(x, y, z), tx, ty, tz)[0]
    assert W.rel_l2(psi, sh(tx, ty, tz) / 20.0) < 0.02


def test_targets_nan_is_a_geometry_error(orc):
    o = orc.OracleFMM(2, 4, 4, 0.7, 0)
    x, y, z, a = o.spec_points()
    t = np.array([np.nan, 0.0])
    with pytest.raises(RuntimeError):
        o.eval_targets("poisson", a, np.ones_like(a), t, np.array([0.0, 1.0]), np.array([1.0, 0.0]))


def test_bs_direct_targets_is_rotated_gradient_of_log_sum(orc, sb):
    """With eps = 0 the Biot-Savart sum at a target t is -t x grad_t of the log sum
    psi_h(t) = sum_j -1/(4 pi) log(1 - t.y_j) q_j (ambient gradient; its radial part
    drops out of the cross product).  The gradient is taken by central differences of
    the log direct sum at targets moved off the sphere, so a wrong sign, a swapped
    cross-product order or a dropped weight in the Biot-Savart sum fails here."""
    W = sb.workloads
    o = orc.OracleFMM(2, 6, 5, 0.7, 1)
    x, y, z, a = [v.ravel().copy() for v in o.spec_points()]
    f = W.random_sh_forcing(x, y, z, lmax=6, seed=3)[0]
    T = np.stack(W.random_unit_vectors(64, seed=5), 1)
    tx, ty, tz = [T[:, i].copy() for i in range(3)]
    vel = orc.direct_sum_bve_vel_targets(tx, ty, tz, x, y, z, a, f, 0.0)
    h = 1e-6
    psi_at = lambda P: orc.direct_sum_inv_lap_targets(P[:, 0].copy(), P[:, 1].copy(), P[:, 2].copy(), x, y, z, a, f)
    grad = np.zeros_like(T)
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        # fourth-order central difference
        grad[:, k] = (8 * (psi_at(T + e) - psi_at(T - e)) - (psi_at(T + 2 * e) - psi_at(T - 2 * e))) / (12 * h)
    expect = list((-np.cross(T, grad)).T)
    assert W.vec_rel_l2(vel, expect) < 1e-8
    # the opposite orientation is far off
    assert W.vec_rel_l2(vel, [-v for v in expect]) > 1.0


def test_bs_direct_targets_at_particles_equals_particle_sum(orc, sb):
    """Targets placed on particles lb..ub reproduce the particle direct sum there (the
    self pair and coincident edge nodes contribute t x t = 0)."""
    W = sb.workloads
    o = orc.OracleFMM(2, 6, 5, 0.7, 1)
    x, y, z, a = [v.ravel().copy() for v in o.spec_points()]
    f = W.gaussian_vortices(x, y, z, a, seed=4)
    lb, ub = 100, 400
    got = orc.direct_sum_bve_vel_targets(x[lb:ub].copy(), y[lb:ub].copy(), z[lb:ub].copy(), x, y, z, a, f, EPS)
    ref = orc.direct_sum_bve_vel(x, y, z, a, f, EPS, None, lb, ub)
    assert W.vec_rel_l2(got, [r[lb:ub] for r in ref]) < 1e-14
