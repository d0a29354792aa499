"""Pins the oracle's SWE pair sums (fmm_interactions/swe_vel.cpp:9-67,
swe_vel_impl.hpp:12-147) without restating them:
  * with zero divergence both variants reduce to the Biot-Savart sums of
    bve_vel_impl.hpp (already pinned) -- same sources, same kernel;
  * the divergent part of a spherical harmonic has the closed form
    u = -/+ grad_T(delta) / (l(l+1)) (variant 1 / variant 2), converging O(h^2);
  * FMM against an O(N^2) direct sum of the same kernel, with displaced sources
    (variant 2 reads the divergence sources undisplaced)."""
import numpy as np
import pytest

EPS = 1e-4


def sh43(x, y, z):
    return 1.770130769779931 * x * (x * x - 3.0 * y * y) * z  # degree-4 harmonic, initial_conditions.cpp:71


def grad_t_sh43(x, y, z):
    c = 1.770130769779931
    g = [c * (3 * x * x - 3 * y * y) * z, c * (-6 * x * y) * z, c * x * (x * x - 3 * y * y)]
    f = sh43(x, y, z)
    return [g[0] - 4 * f * x, g[1] - 4 * f * y, g[2] - 4 * f * z]  # x.grad f = 4 f (homogeneous)


@pytest.mark.parametrize("variant", [1, 2])
@pytest.mark.parametrize("md", [8, 7])
def test_swe_without_divergence_is_biot_savart(orc, sb, variant, md):
    W = sb.workloads
    o = orc.OracleFMM(3, 8, md, 0.7, 1)
    x, y, z, a = o.spec_points()
    vor = W.gaussian_vortices(x, y, z, a, seed=3)
    rng = np.random.default_rng(4)
    dp = [1e-3 * rng.normal(size=(o.Pl, o.m2)) for _ in range(3)]
    dx = [1e-3 * rng.normal(size=(o.P, o.n)) for _ in range(3)]
    qp, qx = o.upward_pass(a, vor)
    zp, zx = np.zeros_like(qp), np.zeros_like(qx)
    sp, sx = o.swe_vel_interactions(variant, qp, qx, zp, zx, EPS, dp, dx)
    bp, bx = o.bve_vel_interactions(qp, qx, EPS, dp, dx)
    assert W.vec_rel_l2(sp, bp) < 1e-14
    assert W.vec_rel_l2([v[: o.offset] for v in sx], [v[: o.offset] for v in bx]) < 1e-14


@pytest.mark.parametrize("variant,sign", [(1, -1.0), (2, 1.0)])
def test_swe_divergent_closed_form_converges(orc, sb, variant, sign):
    W = sb.workloads
    errs = []
    for L in (2, 3, 4):
        o = orc.OracleFMM(L, 8, 8, 0.7, 0)
        x, y, z, a = o.spec_points()
        vel = o.swe_velocity(variant, a, np.zeros_like(a), sh43(x, y, z), 1e-10)
        expect = [sign * g / 20.0 for g in grad_t_sh43(x, y, z)]
        errs.append(W.vec_rel_l2(vel, expect))
    assert errs[-1] < 0.02
    assert errs[0] / errs[1] > 2.5 and errs[1] / errs[2] > 2.5


@pytest.mark.parametrize("variant", [1, 2])
def test_swe_fmm_matches_direct_sum(orc, sb, variant):
    W = sb.workloads
    o = orc.OracleFMM(3, 8, 7, 0.7, 1)
    x, y, z, a = o.spec_points()
    vor = W.gaussian_vortices(x, y, z, a, seed=5)
    div = W.gaussian_vortices(x, y, z, a, seed=6)
    # a smooth displacement (rotation about a tilted axis); proxy displacements
    # interpolated from the particles, as the stepper does
    ax = np.array([0.3, -0.5, 0.81]) * 0.02
    dpart = [ax[1] * z - ax[2] * y, ax[2] * x - ax[0] * z, ax[0] * y - ax[1] * x]
    _, _, dprox = o.bve_displacement(x, y, z, dpart, np.zeros_like(x), 1.0, 0.0)
    vel = o.swe_velocity(variant, a, vor, div, EPS, dpart, dprox)
    X = np.stack([v.ravel() for v in (x, y, z)], 1)
    Y0 = X.copy()
    Y = X + np.stack([d.ravel() for d in dpart], 1)
    Y /= np.linalg.norm(Y, axis=1, keepdims=True)
    qz, qd = (a * vor).ravel(), (a * div).ravel()
    dpv = X @ Y.T
    g = 1.0 / (4 * np.pi) / (1 - dpv + EPS)
    rot = -(g * qz[None, :]) @ Y
    rot = np.cross(X, rot)
    if variant == 1:
        gd = -g * qd[None, :]
        dvel = gd @ Y - X * (gd * dpv).sum(1)[:, None]
    else:
        dp2 = X @ Y0.T
        gd = 1.0 / (4 * np.pi) / (1 - dp2 + EPS) * qd[None, :]
        dvel = gd @ Y0 - X * (gd * dp2).sum(1)[:, None]
    ds = rot + dvel
    err = W.vec_rel_l2([v.ravel() for v in vel], [ds[:, i] for i in range(3)])
    assert err < 5e-7
