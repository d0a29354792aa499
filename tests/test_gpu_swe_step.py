"""GPU parity of the shallow-water step (swe_back_rk4_step, swe_time_step.cpp:606-811)
and of its pieces against the FP64 oracle: the panel Laplacian (deriv_funcs.cpp:27-37),
the tendencies with both indexing rules (swe_time_step.cpp:76-161), the cone
topography (topography.cpp:6-56), the regularised-Laplacian pair sums
(laplacian_impl.hpp:11-69), the height correction (swe_time_step.cpp:558-604), and
the full step with and without topography and tracers, in the reference layout and
in layout A' (m != d+1).

FP64 tolerances: 1e-12 relative for sums and interpolations (summation order differs),
1e-10 for quantities that go through the spectral derivatives (they amplify the
rounding of the velocities by the inverse node spacing), 1e-9 for the whole step."""
import numpy as np
import pytest

from conftest import to_gpu, to_np

pytestmark = pytest.mark.gpu

OMEGA, DT, EPS = 7.292e-5, 900.0, 1e-4
H0 = 5960.0 / 6371000.0
CASES = {
    "ref-layout-L3-d6": dict(levels=3, interp_degree=6, leaf_degree=6, fmm_theta=0.7, weights="reference"),
    "A'-L3-d8-m6": dict(levels=3, interp_degree=8, leaf_degree=5, fmm_theta=0.7, weights="clenshaw_curtis"),
}
WRULE = {"reference": 0, "clenshaw_curtis": 1}


@pytest.fixture(scope="module", params=list(CASES))
def case(request, sb, orc, torch_gpu):
    name = request.param
    print(f"\n[case {name}]", flush=True)
    kw = CASES[name]
    o = orc.OracleFMM(kw["levels"], kw["interp_degree"], kw["leaf_degree"], kw["fmm_theta"], WRULE[kw["weights"]])
    x, y, z, a = o.spec_points()
    W = sb.workloads
    vor = W.gaussian_vortices(x, y, z, a, seed=41)
    div = 0.2 * W.gaussian_vortices(x, y, z, a, seed=42)
    h = H0 * (1 + 1e-3 * W.gaussian_vortices(x, y, z, a, seed=43) / 3e-5)
    ctxs = {}
    for rule in ("reference", "leaf"):
        for topo in ("none", "cone"):
            ctxs[rule, topo] = sb.Context(sb.RunConfig(kernel_eps=EPS, omega=OMEGA, delta_t=DT, swe_rule=rule,
                                                       topography=topo, tracer_count=2, **kw))
    return dict(o=o, ctxs=ctxs, ref_layout=kw["leaf_degree"] == kw["interp_degree"], x=x, y=y, z=z, a=a, vor=vor,
                div=div, h=h, t=torch_gpu, W=W, sb=sb)


def g(case, *arrs):
    return [to_gpu(case["t"], v) for v in arrs]


def test_panel_laplacian(case):
    ctx, o, W = case["ctxs"]["leaf", "none"], case["o"], case["W"]
    expect = o.panel_laplacian(case["h"])
    got = to_np(ctx.laplacian(*g(case, case["h"])))
    ctx.synchronize()
    e = W.rel_l2(got, expect)
    print(f"  panel Laplacian {e:.2e}")
    assert e < 1e-12


@pytest.mark.parametrize("rule", ["reference", "leaf"])
def test_tendency(case, rule):
    ctx, o, W = case["ctxs"][rule, "none"], case["o"], case["W"]
    x, y, z = case["x"], case["y"], case["z"]
    vel = o.swe_velocity(1, case["a"], case["vor"], case["div"], EPS)
    lap = o.panel_laplacian(case["h"])
    expect = o.swe_tendency(x.ravel(), y.ravel(), z.ravel(), vel, case["vor"], case["div"], case["h"], lap, OMEGA,
                            0 if rule == "reference" else 1)
    got = ctx.swe_tendency(*g(case, x, y, z), g(case, *vel), *g(case, case["vor"], case["div"], case["h"], lap))
    ctx.synchronize()
    errs = [W.rel_l2(to_np(gv), ev) for gv, ev in zip(got, expect)]
    print(f"  rule {rule}: vor_t {errs[0]:.2e} div_t {errs[1]:.2e} h_t {errs[2]:.2e}")
    assert max(errs) < 1e-10


@pytest.mark.parametrize("rule", ["reference", "leaf"])
def test_topography(case, rule):
    ctx, o, W = case["ctxs"][rule, "cone"], case["o"], case["W"]
    x, y, z = case["x"], case["y"], case["z"]
    rng = np.random.default_rng(5)
    dp = [1e-2 * rng.normal(size=x.shape) for _ in range(3)]
    dx = [1e-2 * rng.normal(size=(o.P, o.n)) for _ in range(3)]
    if rule == "reference" and not case["ref_layout"]:
        with pytest.raises(case["sb"].SbslpmError):
            ctx.apply_topography(*g(case, x, y, z, case["h"]), g(case, *dp), g(case, *dx))
        return
    expect = o.apply_topography(x.ravel(), y.ravel(), z.ravel(), case["h"], 1, 0 if rule == "reference" else 1, dp, dx)
    got = to_np(ctx.apply_topography(*g(case, x, y, z, case["h"]), g(case, *dp), g(case, *dx)))
    ctx.synchronize()
    e = np.abs(got - expect).max() / np.abs(expect).max()
    print(f"  topography rule {rule}: {e:.2e}")
    assert e < 1e-14


def test_laplacian_interactions(case):
    ctx, o, W = case["ctxs"]["leaf", "none"], case["o"], case["W"]
    x, y, z = case["x"], case["y"], case["z"]
    vel = [1e-6 * (y * z + 0.5), 2e-6 * np.sin(3 * x), -2e-6 * x * y]
    _, dp, dx = o.bve_displacement(x, y, z, vel, 0 * x, 900.0, 0.0)
    qp, qx = o.upward_pass(case["a"], case["h"])
    ep, ex = o.laplacian_interactions(qp, qx, dp, dx)
    gp, gx = ctx.laplacian_interactions(*g(case, qp, qx), g(case, *dp), g(case, *dx))
    ctx.synchronize()
    e1 = W.rel_l2(to_np(gp), ep)
    e2 = W.rel_l2(to_np(gx)[: o.offset], ex[: o.offset])
    print(f"  laplacian interactions: particles {e1:.2e}, proxies {e2:.2e}")
    assert e1 < 1e-12 and e2 < 1e-12


def test_height_correction(case):
    ctx, o, W = case["ctxs"]["leaf", "none"], case["o"], case["W"]
    x, y, z = case["x"], case["y"], case["z"]
    vel = o.swe_velocity(1, case["a"], case["vor"], case["div"], EPS)
    ht = -case["h"] * case["div"]
    lap = 0.1 * case["h"]
    expect = o.swe_height_correction(x.ravel(), y.ravel(), z.ravel(), vel, case["h"], ht, 0.5 * DT, lap)
    got = g(case, lap)[0]
    ctx.swe_height_correction(*g(case, x, y, z), g(case, *vel), *g(case, case["h"], ht), 0.5 * DT, got)
    ctx.synchronize()
    e = W.rel_l2(to_np(got), expect)
    print(f"  height correction {e:.2e}")
    assert e < 1e-12


@pytest.mark.parametrize("rule", ["reference", "leaf"])
@pytest.mark.parametrize("topo", ["none", "cone"])
def test_step(case, rule, topo):
    if topo == "cone" and rule == "reference" and not case["ref_layout"]:
        pytest.skip("the reference topography rule reads [P][n] displacement rows at particle indices; "
                    "it is defined for the reference layout only (the library rejects it elsewhere)")
    ctx, o, W = case["ctxs"][rule, topo], case["o"], case["W"]
    x, y, z, a = case["x"], case["y"], case["z"], case["a"]
    rng = np.random.default_rng(7)
    tr = rng.normal(size=x.shape + (2,))
    ev, ed, eh, evel, elap, etr = o.swe_back_rk4_step(x.ravel(), y.ravel(), z.ravel(), a, case["vor"], case["div"],
                                                      case["h"], DT, OMEGA, EPS, 0, 1 if topo == "cone" else 0,
                                                      0 if rule == "reference" else 1, tr)
    gv, gd, gh, gtr = g(case, case["vor"], case["div"], case["h"], tr)
    vel, lap = ctx.swe_back_rk4_step(*g(case, x, y, z, a), gv, gd, gh, tracers=gtr)
    ctx.synchronize()
    errs = dict(vor=W.rel_l2(to_np(gv), ev), div=W.rel_l2(to_np(gd), ed), h=W.rel_l2(to_np(gh), eh),
                vel=W.vec_rel_l2([to_np(v) for v in vel], evel), lap=W.rel_l2(to_np(lap), elap),
                tracers=W.rel_l2(to_np(gtr), etr))
    print(f"  step rule {rule} topo {topo}: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    assert errs["vel"] < 1e-12 and errs["lap"] < 1e-12
    # The remap and the height correction locate moved points in leaves; particles whose
    # points land within rounding of a leaf edge may be interpolated from either neighbour,
    # so the step amplifies FP64 rounding there.  The bar is the oracle's own spread under
    # 1e-13 relative perturbations of the input state (three draws), times 10.  With the
    # reference's cone the stage-2..4 divergence is dominated by the Laplacian of the
    # mountain, which the GPU's and the host's exp/atan2 round differently at grid scale, so
    # the height is perturbed too in that case.
    spread = dict(vor=0.0, div=0.0, h=0.0, tracers=0.0)
    for _ in range(3):
        pv = case["vor"] * (1 + 1e-13 * rng.normal(size=x.shape))
        pd = case["div"] * (1 + 1e-13 * rng.normal(size=x.shape))
        ph = case["h"] * (1 + 1e-13 * rng.normal(size=x.shape)) if topo == "cone" else case["h"]
        r = o.swe_back_rk4_step(x.ravel(), y.ravel(), z.ravel(), a, pv, pd, ph, DT, OMEGA, EPS, 0,
                                1 if topo == "cone" else 0, 0 if rule == "reference" else 1, tr)
        for k, name in ((0, "vor"), (1, "div"), (2, "h"), (5, "tracers")):
            spread[name] = max(spread[name], W.rel_l2(r[k], (ev, ed, eh, None, None, etr)[k]))
    print("    oracle spread: " + ", ".join(f"{k} {v:.2e}" for k, v in spread.items()))
    for name in spread:
        assert errs[name] <= 10 * spread[name] + 1e-10, name


def test_step_at_rest(case):
    W = case["W"]
    x, y, z, a = case["x"], case["y"], case["z"], case["a"]
    for rule in ("reference", "leaf"):
        ctx = case["ctxs"][rule, "none"]
        gv, gd, gh = g(case, 0 * x, 0 * x, H0 + 0 * x)
        tr = g(case, np.ones(x.shape + (2,)))[0]
        vel, lap = ctx.swe_back_rk4_step(*g(case, x, y, z, a), gv, gd, gh, tracers=tr)
        ctx.synchronize()
        assert np.abs(to_np(gv)).max() == 0 and np.abs(to_np(gd)).max() == 0
        assert np.abs(to_np(gh) / H0 - 1).max() < 1e-14
        assert np.abs(to_np(tr) - 1).max() < 1e-14


def test_mixed_context_laplacian_and_step(case, sb):
    """A mixed-precision context runs the regularised-Laplacian and shallow-water pairs in FP32
    (FP64 accumulation per source set): the Laplacian pair sums and the step's stage-1 velocity
    agree with the FP64 oracle to 1e-6; the spectral stage-1 Laplacian stays FP64 (1e-12)."""
    o, W = case["o"], case["W"]
    x, y, z, a = case["x"], case["y"], case["z"], case["a"]
    kw = CASES[[k for k in CASES if (CASES[k]["leaf_degree"] == o.md and CASES[k]["interp_degree"] == o.d)][0]]
    ctx = sb.Context(sb.RunConfig(kernel_eps=EPS, omega=OMEGA, delta_t=DT, swe_rule="leaf", precision="mixed", **kw))
    vel = [1e-6 * (y * z + 0.5), 2e-6 * np.sin(3 * x), -2e-6 * x * y]
    _, dp, dx = o.bve_displacement(x, y, z, vel, 0 * x, 900.0, 0.0)
    qp, qx = o.upward_pass(a, case["h"])
    ep, ex = o.laplacian_interactions(qp, qx, dp, dx)
    gp, gx = ctx.laplacian_interactions(*g(case, qp, qx), g(case, *dp), g(case, *dx))
    ctx.synchronize()
    e1 = W.rel_l2(to_np(gp), ep)
    e2 = W.rel_l2(to_np(gx)[: o.offset], ex[: o.offset])
    ev, ed, eh, evel, elap, _ = o.swe_back_rk4_step(x.ravel(), y.ravel(), z.ravel(), a, case["vor"], case["div"],
                                                    case["h"], DT, OMEGA, EPS, 0, 0, 1)
    gv, gd, gh = g(case, case["vor"], case["div"], case["h"])
    gvel, glap = ctx.swe_back_rk4_step(*g(case, x, y, z, a), gv, gd, gh)
    ctx.synchronize()
    e3 = W.vec_rel_l2([to_np(v) for v in gvel], evel)
    e4 = W.rel_l2(to_np(glap), elap)
    finite = all(np.isfinite(to_np(v)).all() for v in (gv, gd, gh))
    print(f"  mixed: laplacian pairs {e1:.2e} / {e2:.2e}, stage-1 velocity {e3:.2e}, stage-1 Laplacian {e4:.2e}")
    assert e1 < 1e-6 and e2 < 1e-6 and e3 < 1e-6 and e4 < 1e-12 and finite


def test_step_rejects_aliased_outputs(case):
    ctx, W = case["ctxs"]["leaf", "none"], case["W"]
    x, y, z, a = case["x"], case["y"], case["z"], case["a"]
    gv, gd, gh = g(case, case["vor"], case["div"], case["h"])
    with pytest.raises(case["sb"].SbslpmError):
        ctx.swe_back_rk4_step(*g(case, x, y, z, a), gv, gd, gh, height_lap=gh)
    with pytest.raises(case["sb"].SbslpmError):
        ctx.swe_back_rk4_step(*g(case, x, y, z, a), gv, gv, gh)


def test_step_graph_replay_matches_eager(case, sb, torch_gpu):
    """With use_graph on a non-default stream the step is captured once and replayed; three
    replays, each from the same state with eager work between them, equal the eager step to
    the bit (the step's kernels are deterministic: no atomics)."""
    t, o, W = torch_gpu, case["o"], case["W"]
    x, y, z, a = case["x"], case["y"], case["z"], case["a"]
    kw = CASES[[k for k in CASES if (CASES[k]["leaf_degree"] == o.md and CASES[k]["interp_degree"] == o.d)][0]]
    stream = t.cuda.Stream()
    with t.cuda.stream(stream):
        cg = sb.Context(sb.RunConfig(kernel_eps=EPS, omega=OMEGA, delta_t=DT, swe_rule="leaf", topography="cone",
                                     tracer_count=2, use_graph=True, **kw))
    ce = case["ctxs"]["leaf", "cone"]
    rng = np.random.default_rng(11)
    tr0 = rng.normal(size=x.shape + (2,))
    geo = g(case, x, y, z, a)
    ev, ed, eh, etr = g(case, case["vor"], case["div"], case["h"], tr0)
    evel, elap = ce.swe_back_rk4_step(*geo, ev, ed, eh, tracers=etr)
    ce.synchronize()
    expect = [to_np(v) for v in (ev, ed, eh, etr, *evel, elap)]
    state0 = g(case, case["vor"], case["div"], case["h"], tr0)
    st = [v.clone() for v in state0]
    vel, lap = cg.zeros_part(3), cg.zeros_part()
    for rep in range(3):
        with t.cuda.stream(stream):
            for v, v0 in zip(st, state0):
                v.copy_(v0)
            cg.swe_back_rk4_step(*geo, *st[:3], vel=vel, tracers=st[3], height_lap=lap)
        stream.synchronize()
        cg.synchronize()
        got = [to_np(v) for v in (*st, *vel, lap)]
        same = all(np.array_equal(gv, e) for gv, e in zip(got, expect))
        print(f"  replay {rep}: bit-identical to the eager step: {same}")
        assert same
        _ = (st[0] * 2).sum().item()  # eager work between replays


def test_step_graph_after_set_precision(case, sb, torch_gpu):
    """A captured step holds the kernels of the precision it was captured in.  After
    set_precision the graph-mode step must run the new precision: it equals, to the
    bit, the eager step of a fresh context of that precision (before and after the
    switch)."""
    t, o = torch_gpu, case["o"]
    x, y, z, a = case["x"], case["y"], case["z"], case["a"]
    kw = CASES[[k for k in CASES if (CASES[k]["leaf_degree"] == o.md and CASES[k]["interp_degree"] == o.d)][0]]
    common = dict(kernel_eps=EPS, omega=OMEGA, delta_t=DT, swe_rule="leaf", topography="cone", **kw)
    stream = t.cuda.Stream()
    with t.cuda.stream(stream):
        cg = sb.Context(sb.RunConfig(precision="fp64", use_graph=True, **common))
        eager = {p: sb.Context(sb.RunConfig(precision=p, **common)) for p in ("fp64", "mixed")}
    geo = g(case, x, y, z, a)
    state0 = g(case, case["vor"], case["div"], case["h"])
    # the graph context keeps one set of buffers, so its graph key (the pointers)
    # is the same before and after every switch
    st_g = [v.clone() for v in state0]
    vel_g, lap_g = cg.zeros_part(3), cg.zeros_part()
    for prec in ("fp64", "fp64", "mixed", "mixed", "fp64"):
        if prec != cg.cfg.precision:
            cg.set_precision(prec)
        print(f"  graph step in {prec}", flush=True)
        with t.cuda.stream(stream):
            for v, v0 in zip(st_g, state0):
                v.copy_(v0)
            cg.swe_back_rk4_step(*geo, *st_g, vel=vel_g, height_lap=lap_g)
            st = [v.clone() for v in state0]
            vel, lap = eager[prec].swe_back_rk4_step(*geo, *st)
        stream.synchronize()
        cg.synchronize()
        eager[prec].synchronize()
        for gv, e in zip((*st_g, *vel_g, lap_g), (*st, *vel, lap)):
            np.testing.assert_array_equal(to_np(gv), to_np(e))


@pytest.mark.parametrize("levels,m", [(2, 8), (3, 5)])
def test_unify_boundary_vals_matches_oracle(sb, orc, torch_gpu, levels, m):
    """unify_boundary_vals (initialize_cubed_sphere.cpp:364-370) on the GPU against the oracle
    (FP64 atomics: the at most four copies may be summed in another order)."""
    t = torch_gpu
    print(f"\n[case unify_boundary_vals levels={levels} m={m}]", flush=True)
    ctx = sb.Context(sb.RunConfig(levels=levels, interp_degree=8, leaf_degree=m - 1))
    mp, cnt, *_ = sb.HostTree(sb.RunConfig(levels=levels, interp_degree=8, leaf_degree=m - 1)).dedupe()
    vals = np.random.default_rng(7).standard_normal(mp.shape)
    want = orc.unify_boundary_vals(mp, cnt, vals)
    gm, gc = [t.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).cuda() for a in (mp, cnt)]
    gv = to_gpu(t, vals)
    for _ in range(2):  # the second call reuses the scratch and must change nothing
        ctx.unify_boundary_vals(gm, gc, gv)
        ctx.synchronize()
        np.testing.assert_allclose(to_np(gv), want, rtol=1e-15, atol=1e-15)
    with pytest.raises(sb.SbslpmError):
        ctx.unify_boundary_vals(gm, gc[:0], gv)
