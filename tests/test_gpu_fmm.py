"""GPU parity of the FMM passes and pair sums against the FP64 oracle
(fmm_funcs.cpp:341-392, bve_vel.cpp:9-37, inverse_poisson.cpp:9-33).

FP64 tolerance: relative 1e-12 (summation order and the x cross sum(g y)
rearrangement differ from the reference; both are O(1e-15) effects).
Mixed precision: relative 1e-6 against the FP64 result (the metric's bound)."""
import numpy as np
import pytest

from conftest import to_gpu, to_np

pytestmark = pytest.mark.gpu

RULE = {"reference": 0, "clenshaw_curtis": 1, "equal_area": 2}
CASES = {
    "ref-layout-L3-d8": dict(levels=3, interp_degree=8, leaf_degree=8, fmm_theta=0.7, weights="reference"),
    "A'-L4-d8-m8": dict(levels=4, interp_degree=8, leaf_degree=7, fmm_theta=0.7, weights="clenshaw_curtis"),
    "A'-L3-d10-m6": dict(levels=3, interp_degree=10, leaf_degree=5, fmm_theta=0.6, weights="clenshaw_curtis"),
}
EPS = 1e-4


def rel(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module", params=list(CASES))
def case(request, sb, orc, torch_gpu):
    name = request.param
    print(f"\n[case {name}]", flush=True)
    kw = CASES[name]
    cfg = sb.RunConfig(kernel_eps=EPS, **kw)
    ctx = sb.Context(cfg)
    o = orc.OracleFMM(kw["levels"], kw["interp_degree"], kw["leaf_degree"], kw["fmm_theta"], RULE[kw["weights"]])
    x, y, z, a = o.spec_points()
    vor = sb.workloads.gaussian_vortices(x, y, z, a, seed=11)
    return dict(name=name, ctx=ctx, o=o, x=x, y=y, z=z, a=a, vor=vor, torch=torch_gpu, sb=sb)


def test_spec_points_match(case):
    ctx, o = case["ctx"], case["o"]
    for g, h in zip(ctx.spec_points(), o.spec_points()):
        np.testing.assert_allclose(g, h, rtol=0, atol=1e-17)
    np.testing.assert_array_equal(ctx.interaction_list(), o.interactions())


def test_upward_pass(case):
    ctx, o, t = case["ctx"], case["o"], case["torch"]
    qp, qx = ctx.upward_pass(to_gpu(t, case["a"]), to_gpu(t, case["vor"]))
    ctx.synchronize()
    rp, rx = o.upward_pass(case["a"], case["vor"])
    np.testing.assert_array_equal(to_np(qp), rp)
    assert rel(to_np(qx), rx) < 1e-13


@pytest.mark.parametrize("displaced", [False, True])
def test_bve_vel_interactions(case, displaced):
    ctx, o, t = case["ctx"], case["o"], case["torch"]
    rp, rx = o.upward_pass(case["a"], case["vor"])
    dp = dx = None
    if displaced:
        rng = np.random.default_rng(3)
        dp = [1e-3 * rng.normal(size=(o.Pl, o.m2)) for _ in range(3)]
        dx = [1e-3 * rng.normal(size=(o.P, o.n)) for _ in range(3)]
    ep, ex = o.bve_vel_interactions(rp, rx, EPS, dp, dx)
    gp, gx = ctx.bve_vel_interactions(to_gpu(t, rp), to_gpu(t, rx),
                                      None if dp is None else [to_gpu(t, v) for v in dp],
                                      None if dx is None else [to_gpu(t, v) for v in dx])
    ctx.synchronize()
    for f in range(3):
        assert rel(to_np(gp[f]), ep[f]) < 1e-12
        nonleaf = slice(0, o.offset)
        assert rel(to_np(gx[f])[nonleaf], ex[f][nonleaf]) < 1e-12


@pytest.mark.parametrize("mode", ["mixed", "mixed_valu"])
def test_bve_vel_interactions_mixed(case, mode):
    """Mixed-precision pair sums (displaced sources) against the FP64 oracle,
    per output class (leaf particles, non-leaf proxies)."""
    ctx, o, t = case["ctx"], case["o"], case["torch"]
    rp, rx = o.upward_pass(case["a"], case["vor"])
    rng = np.random.default_rng(13)
    dp = [1e-3 * rng.normal(size=(o.Pl, o.m2)) for _ in range(3)]
    dx = [1e-3 * rng.normal(size=(o.P, o.n)) for _ in range(3)]
    ep, ex = o.bve_vel_interactions(rp, rx, EPS, dp, dx)
    ctx.set_precision(mode)
    try:
        gp, gx = ctx.bve_vel_interactions(to_gpu(t, rp), to_gpu(t, rx), [to_gpu(t, v) for v in dp],
                                          [to_gpu(t, v) for v in dx])
        ctx.synchronize()
    finally:
        ctx.set_precision("fp64")
    W = case["sb"].workloads
    e1 = W.vec_rel_l2([to_np(g) for g in gp], ep)
    e2 = W.vec_rel_l2([to_np(g)[: o.offset] for g in gx], [v[: o.offset] for v in ex])
    print(f"  {mode}: particles {e1:.2e}, proxies {e2:.2e}")
    assert e1 < 1e-6 and e2 < 1e-6


def test_poisson_fmm_interactions(case):
    ctx, o, t = case["ctx"], case["o"], case["torch"]
    rp, rx = o.upward_pass(case["a"], case["vor"])
    ep, ex = o.poisson_fmm_interactions(rp, rx)
    gp, gx = ctx.poisson_fmm_interactions(to_gpu(t, rp), to_gpu(t, rx))
    ctx.synchronize()
    assert rel(to_np(gp), ep) < 1e-12
    assert rel(to_np(gx)[: o.offset], ex[: o.offset]) < 1e-12


def test_downward_pass_3(case):
    ctx, o, t = case["ctx"], case["o"], case["torch"]
    rng = np.random.default_rng(5)
    phi = [rng.normal(size=(o.P, o.n)) for _ in range(3)]
    part = [rng.normal(size=(o.Pl, o.m2)) for _ in range(3)]
    expect = o.downward_pass([p.copy() for p in phi], part)
    gphi = [to_gpu(t, p) for p in phi]
    got = ctx.downward_pass_3(gphi, [to_gpu(t, p) for p in part])
    ctx.synchronize()
    for f in range(3):
        assert rel(to_np(got[f]), expect[f]) < 1e-13
    single = ctx.downward_pass(to_gpu(t, phi[0]))
    ctx.synchronize()
    assert rel(to_np(single), o.downward_pass([phi[0].copy()])[0]) < 1e-13


def test_bve_velocity_fp64_and_mixed(case):
    ctx, o, t = case["ctx"], case["o"], case["torch"]
    expect = o.bve_velocity(case["a"], case["vor"], EPS)
    ga, gv = to_gpu(t, case["a"]), to_gpu(t, case["vor"])
    ctx.set_precision("fp64")
    v64 = [to_np(v) for v in ctx.bve_velocity(ga, gv)]
    ctx.synchronize()
    assert case["sb"].workloads.vec_rel_l2(v64, expect) < 1e-12
    for mode in ("mixed", "mixed_valu"):
        ctx.set_precision(mode)
        vmx = [to_np(v) for v in ctx.bve_velocity(ga, gv)]
        ctx.synchronize()
        ctx.set_precision("fp64")
        err = case["sb"].workloads.vec_rel_l2(vmx, v64)
        print(f"  {mode} vs fp64 velocity rel-L2 {err:.2e}")
        assert err < 1e-6


def test_poisson_solve_fp64_and_mixed(case):
    ctx, o, t = case["ctx"], case["o"], case["torch"]
    expect = o.poisson_solve(case["a"], case["vor"])
    ga, gv = to_gpu(t, case["a"]), to_gpu(t, case["vor"])
    p64 = to_np(ctx.poisson_solve(ga, gv))
    ctx.synchronize()
    assert rel(p64, expect) < 1e-12
    ctx.set_precision("mixed")
    pmx = to_np(ctx.poisson_solve(ga, gv))
    ctx.synchronize()
    ctx.set_precision("fp64")
    err = rel(pmx, p64)
    print(f"  mixed vs fp64 psi rel-L2 {err:.2e}")
    assert err < 1e-6


def test_rh4_closed_form_on_gpu(sb, torch_gpu):
    """Same closed-form bound the oracle meets (no oracle involved)."""
    cfg = sb.RunConfig(levels=4, interp_degree=8, leaf_degree=8, kernel_eps=1e-10)
    ctx = sb.Context(cfg)
    x, y, z, a = ctx.spec_points()
    W = sb.workloads
    vel = ctx.bve_velocity(to_gpu(torch_gpu, a), to_gpu(torch_gpu, W.rh4_vorticity(x, y, z)))
    ctx.synchronize()
    assert W.vec_rel_l2([to_np(v) for v in vel], W.rh4_velocity(x, y, z)) < 0.01


@pytest.mark.parametrize("levels,deg,leaf", [(6, 8, 7), (5, 5, 5), (5, 6, 4)])
def test_fused_passes_match_level_by_level(sb, torch_gpu, levels, deg, leaf):
    """bve_velocity and poisson_solve run the fused passes: one upward launch for the P2M and
    the first M2M level, and one downward launch for every L2L level and the L2P, with each
    ancestor's expansion evaluated at the leaves' parent through a composite matrix.  The
    C-ABI's upward_pass, interactions and downward_pass run the reference's passes level by
    level (pinned against the oracle above).  With the same pair kernels the two may differ
    only by the passes' rounding: relative 1e-14.  (8, 7) and (5, 5) take the register-tiled
    kernels, (6, 4) the generic ones."""
    t = torch_gpu
    print(f"\n[case fused passes L={levels} d={deg} m={leaf + 1}]", flush=True)
    ctx = sb.Context(sb.RunConfig(levels=levels, interp_degree=deg, leaf_degree=leaf, fmm_theta=0.7,
                                  weights="clenshaw_curtis", kernel_eps=EPS))
    x, y, z, a = ctx.spec_points()
    vor = sb.workloads.gaussian_vortices(x, y, z, a, seed=3)
    ga, gv = to_gpu(t, a), to_gpu(t, vor)
    fused = [to_np(v) for v in ctx.bve_velocity(ga, gv)]
    qp, qx = ctx.upward_pass(ga, gv)
    pp, px = ctx.bve_vel_interactions(qp, qx)
    by_level = [to_np(v) for v in ctx.downward_pass_3(px, pp)]
    ctx.synchronize()
    err = sb.workloads.vec_rel_l2(fused, by_level)
    psi_f = to_np(ctx.poisson_solve(ga, gv))
    qp, qx = ctx.upward_pass(ga, gv)
    sp, sx = ctx.poisson_fmm_interactions(qp, qx)
    psi_l = to_np(ctx.downward_pass(sx, sp))
    ctx.synchronize()
    err_psi = rel(psi_f, psi_l)
    print(f"  velocity {err:.2e}, psi {err_psi:.2e}")
    assert err < 1e-14 and err_psi < 1e-14
    assert np.abs(fused[0]).max() > 0 and np.abs(psi_f).max() > 0
