"""The lat-lon path's _ll operators on the irregular particle tree, on the GPU,
against the oracle's sequential restatement of the reference functors
(fmm_funcs_impl.hpp:190-364, swe_vel_impl.hpp:149-288) and against the
O(N^2) direct sums of the same kernels."""
import numpy as np
import pytest

from conftest import to_gpu, to_np

pytestmark = pytest.mark.gpu

EPS = 1e-4


def points(n, seed):
    x = np.random.default_rng(seed).normal(size=(3, n))
    return x / np.linalg.norm(x, axis=0)


@pytest.fixture(scope="module", params=[(4000, 40, 6), (3000, 60, 8)])
def setup(request, sb, orc, torch_gpu):
    n, leaf, deg = request.param
    print(f"\n[case irregular N={n} leaf={leaf} d={deg}]", flush=True)
    X = points(n, deg)
    ctx = sb.Context(sb.RunConfig(levels=2, interp_degree=deg, kernel_eps=EPS))
    irr = ctx.irregular(*X, leaf)
    o = orc.OracleIrreg(*X, leaf, deg, 0.7)
    rng = np.random.default_rng(1)
    area = np.full(n, 4 * np.pi / n) * (1 + 0.1 * rng.random(n))
    vor = X[2] + 0.3 * X[0] * X[1]
    div = 0.5 * X[0] - 0.2 * X[1] * X[2]
    return ctx, irr, o, X, area, vor, div


def test_upward_pass_matches_oracle(setup, torch_gpu):
    ctx, irr, o, X, area, vor, div = setup
    t = torch_gpu
    got = to_np(irr.upward_pass(to_gpu(t, area), to_gpu(t, vor)))
    ctx.synchronize()
    ref = o.upward_pass(area, vor)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13 * np.abs(ref).max())


def test_interactions_and_downward_match_oracle(setup, torch_gpu):
    ctx, irr, o, X, area, vor, div = setup
    t = torch_gpu
    qv, qd = o.upward_pass(area, vor), o.upward_pass(area, div)
    tp, vel = irr.swe_vel_interactions(to_gpu(t, qv), to_gpu(t, qd), to_gpu(t, vor), to_gpu(t, div), to_gpu(t, area))
    ctx.synchronize()
    rtp, rvel = o.swe_vel_interactions(qv, qd, vor, div, area, EPS)
    scale = max(np.abs(v).max() for v in rvel)
    for g, r in zip(tp, rtp):
        np.testing.assert_allclose(to_np(g), r, rtol=0, atol=1e-12 * scale)
    for g, r in zip(vel, rvel):
        np.testing.assert_allclose(to_np(g), r, rtol=0, atol=1e-12 * scale)
    sols = irr.downward_pass_3([tt.clone() for tt in tp], [v.clone() for v in vel])
    ctx.synchronize()
    _, rsol = o.downward_pass_3(rtp, rvel)
    for g, r in zip(sols, rsol):
        np.testing.assert_allclose(to_np(g), r, rtol=0, atol=1e-12 * scale)


def test_swe_velocity_matches_oracle_and_direct_sum(setup, sb, torch_gpu):
    ctx, irr, o, X, area, vor, div = setup
    t = torch_gpu
    got = np.stack([to_np(v) for v in irr.swe_velocity(to_gpu(t, area), to_gpu(t, vor), to_gpu(t, div))], 1)
    ctx.synchronize()
    ref = np.stack(o.swe_velocity(area, vor, div, EPS), 1)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    # direct sum of the same kernel over all points (swe_vel_impl.hpp:265-276)
    x = X.T
    dp = x @ x.T
    g = 0.07957747154594767 / (1 - dp + EPS)
    qv, qd = vor * area, div * area
    direct = -np.cross(x, (g * qv[None, :]) @ x) + (g * qd[None, :]) @ x - x * np.sum(g * qd[None, :] * dp, 1)[:, None]
    err = np.linalg.norm(got - direct) / np.linalg.norm(direct)
    print(f"  irregular FMM vs direct sum rel-L2 {err:.2e}")
    assert err < 1e-4


def test_bve_velocity_equals_divergence_free_swe(setup, torch_gpu):
    ctx, irr, o, X, area, vor, div = setup
    t = torch_gpu
    ga, gv = to_gpu(t, area), to_gpu(t, vor)
    bve = np.stack([to_np(v) for v in irr.bve_velocity(ga, gv)], 1)
    swe = np.stack([to_np(v) for v in irr.swe_velocity(ga, gv, t.zeros_like(gv))], 1)
    ctx.synchronize()
    # same pairs, different summation order (the shallow-water kernel splits each tile's
    # source sets over four waves): equal to rounding
    np.testing.assert_allclose(bve, swe, rtol=0, atol=1e-13 * np.abs(swe).max())
