"""GPU parity of the shallow-water pair sums (fmm_interactions/swe_vel.cpp:9-67,
swe_vel_impl.hpp:12-147) and of the fused SWE velocity evaluation
(swe_time_step.cpp:640-653) against the FP64 oracle.

FP64 tolerance: relative 1e-12 (summation order and the regrouping
x cross sum(g y) + sum(g y') - x sum(g x.y') differ from the reference)."""
import numpy as np
import pytest

from conftest import to_gpu, to_np

pytestmark = pytest.mark.gpu

RULE = {"reference": 0, "clenshaw_curtis": 1}
CASES = {
    "ref-layout-L3-d8": dict(levels=3, interp_degree=8, leaf_degree=8, fmm_theta=0.7, weights="reference"),
    "A'-L4-d8-m8": dict(levels=4, interp_degree=8, leaf_degree=7, fmm_theta=0.7, weights="clenshaw_curtis"),
}
EPS = 1e-4


@pytest.fixture(scope="module", params=list(CASES))
def case(request, sb, orc, torch_gpu):
    name = request.param
    print(f"\n[case {name}]", flush=True)
    kw = CASES[name]
    ctx = sb.Context(sb.RunConfig(kernel_eps=EPS, **kw))
    o = orc.OracleFMM(kw["levels"], kw["interp_degree"], kw["leaf_degree"], kw["fmm_theta"], RULE[kw["weights"]])
    x, y, z, a = o.spec_points()
    W = sb.workloads
    vor = W.gaussian_vortices(x, y, z, a, seed=31)
    div = 0.3 * W.gaussian_vortices(x, y, z, a, seed=32)
    ax = np.array([0.2, 0.4, -0.7]) * 0.01
    dpart = [ax[1] * z - ax[2] * y, ax[2] * x - ax[0] * z, ax[0] * y - ax[1] * x]
    _, _, dprox = o.bve_displacement(x, y, z, dpart, np.zeros_like(x), 1.0, 0.0)
    ctx_mixed = sb.Context(sb.RunConfig(kernel_eps=EPS, precision="mixed", **kw))
    return dict(ctx=ctx, ctx_mixed=ctx_mixed, o=o, x=x, y=y, z=z, a=a, vor=vor, div=div, dpart=dpart, dprox=dprox,
                t=torch_gpu, sb=sb)


@pytest.mark.parametrize("variant", [1, 2])
def test_swe_vel_interactions(case, variant):
    ctx, o, t = case["ctx"], case["o"], case["t"]
    qzp, qzx = o.upward_pass(case["a"], case["vor"])
    qdp, qdx = o.upward_pass(case["a"], case["div"])
    ep, ex = o.swe_vel_interactions(variant, qzp, qzx, qdp, qdx, EPS, case["dpart"], case["dprox"])
    gp, gx = ctx.swe_vel_interactions(*[to_gpu(t, v) for v in (qzp, qzx, qdp, qdx)],
                                      disp_part=[to_gpu(t, v) for v in case["dpart"]],
                                      disp_proxy=[to_gpu(t, v) for v in case["dprox"]], variant=variant)
    ctx.synchronize()
    W = case["sb"].workloads
    e1 = W.vec_rel_l2([to_np(g) for g in gp], ep)
    e2 = W.vec_rel_l2([to_np(g)[: o.offset] for g in gx], [v[: o.offset] for v in ex])
    print(f"  variant {variant}: particles {e1:.2e}, proxies {e2:.2e}")
    assert e1 < 1e-12 and e2 < 1e-12


@pytest.mark.parametrize("variant", [1, 2])
def test_swe_velocity(case, variant):
    ctx, o, t = case["ctx"], case["o"], case["t"]
    expect = o.swe_velocity(variant, case["a"], case["vor"], case["div"], EPS, case["dpart"], case["dprox"])
    got = ctx.swe_velocity(*[to_gpu(t, case[k]) for k in ("a", "vor", "div")],
                           disp_part=[to_gpu(t, v) for v in case["dpart"]],
                           disp_proxy=[to_gpu(t, v) for v in case["dprox"]], variant=variant)
    ctx.synchronize()
    err = case["sb"].workloads.vec_rel_l2([to_np(g) for g in got], expect)
    print(f"  variant {variant}: velocity rel-L2 {err:.2e}")
    assert err < 1e-12


def test_swe_displacement_and_remap(case):
    """swe_displacement_upward_pass (swe_time_step.cpp:265-276), swe_find_departure_points
    (:422-556) and swe_departure_to_target (:19-74) against the oracle."""
    ctx, o, t = case["ctx"], case["o"], case["t"]
    W = case["sb"].workloads
    x, y, z = case["x"], case["y"], case["z"]
    vel = W.rh4_velocity(x, y, z)
    gvel = [to_gpu(t, v) for v in vel]
    gdp, gdx = ctx.swe_displacement(gvel, 1800.0)
    ctx.synchronize()
    _, dp, dx = o.bve_displacement(x, y, z, vel, np.zeros_like(x), 1800.0, 0.0)
    for f in range(3):
        np.testing.assert_array_equal(to_np(gdp[f]), dp[f])
        np.testing.assert_allclose(to_np(gdx[f])[: o.offset], dx[f][: o.offset], rtol=0, atol=1e-16)
    stages = [vel, [1.1 * v for v in vel], [0.9 * v for v in vel], [1.05 * v for v in vel]]
    gst = [[to_gpu(t, v) for v in s] for s in stages]
    gx, gy, gz = [to_gpu(t, v) for v in (x, y, z)]
    gdep = ctx.swe_find_departure_points(gx, gy, gz, gst, 3600.0)
    ctx.synchronize()
    dep = o.bve_find_departure_points(x, y, z, stages, 3600.0, 0)
    for f in range(3):
        np.testing.assert_allclose(to_np(gdep[f]), dep[f], rtol=0, atol=1e-14)
    fields = [case["vor"], case["div"], 1.0 + 0.1 * z]
    tr = np.stack([x * y, z], -1).copy()
    got = ctx.swe_departure_to_target(gdep, *[to_gpu(t, v) for v in fields], tracers=to_gpu(t, tr))
    ctx.synchronize()
    exp, etr = o.swe_departure_to_target([to_np(g) for g in gdep], fields, tr)
    for g, e in zip(got[:3], exp):
        assert W.rel_l2(to_np(g), e) < 1e-13
    assert W.rel_l2(to_np(got[3]), etr) < 1e-13


@pytest.mark.parametrize("variant", [1, 2])
def test_swe_velocity_mixed_context(case, variant):
    """On a mixed-precision context variants 1 and 3 run FP32 pairs (FP64 accumulation per
    source set); variant 2 stays FP64.  Bar: 1e-6 relative (FP32 pair arithmetic in
    panel-centred coordinates, as for the Biot-Savart mixed kernels)."""
    sb, o, t = case["sb"], case["o"], case["t"]
    ctx = case["ctx_mixed"]
    expect = o.swe_velocity(variant, case["a"], case["vor"], case["div"], EPS, case["dpart"], case["dprox"])
    got = ctx.swe_velocity(*[to_gpu(t, case[k]) for k in ("a", "vor", "div")],
                           disp_part=[to_gpu(t, v) for v in case["dpart"]],
                           disp_proxy=[to_gpu(t, v) for v in case["dprox"]], variant=variant)
    ctx.synchronize()
    e = sb.workloads.vec_rel_l2([to_np(g) for g in got], expect)
    print(f"  mixed context, variant {variant}: {e:.2e}")
    assert e < (1e-6 if variant == 1 else 1e-12)
