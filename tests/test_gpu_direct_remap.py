"""GPU parity of the direct sums (direct_sum.cpp:8-15 and the Biot-Savart
analogue), the displacement / departure / remap kernels and the full
bve_back_rk4_step (bve_time_step.cpp:15-419) against the FP64 oracle."""
import numpy as np
import pytest

from conftest import to_gpu, to_np

pytestmark = pytest.mark.gpu

OMEGA = 7.292e-5
DT = 3600.0
EPS = 1e-4


def rel(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def c1(sb, orc, torch_gpu):
    """6 x 64^2 particles (leaf 8x8 grid, L=4, d=8), the CPU-runnable configuration."""
    print("\n[case C1 6x64^2 L4 d8 m8]", flush=True)
    cfg = sb.RunConfig(levels=4, interp_degree=8, leaf_degree=7, weights="equal_area", kernel_eps=EPS,
                       delta_t=DT, omega=OMEGA, tracer_count=2)
    ctx = sb.Context(cfg)
    o = orc.OracleFMM(4, 8, 7, 0.7, 2)
    x, y, z, a = o.spec_points()
    return dict(ctx=ctx, o=o, x=x, y=y, z=z, a=a, t=torch_gpu, sb=sb, orc=orc)


def test_direct_sum_inv_lap(c1):
    t, ctx, orc = c1["t"], c1["ctx"], c1["orc"]
    rng = np.random.default_rng(1)
    n = 6000
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    x, y, z = [v[:, i].copy() for i in range(3)]
    area, pots = rng.uniform(0.5, 1.5, n) * 4 * np.pi / n, rng.normal(size=n)
    expect = orc.direct_sum_inv_lap(x, y, z, area, pots, 100, 5100)
    g = ctx.direct_sum_inv_lap(*[to_gpu(t, q) for q in (x, y, z, area, pots)], lb=100, ub=5100)
    ctx.synchronize()
    got = to_np(g)
    assert np.all(got[:100] == 0) and np.all(got[5100:] == 0)
    # random-sign charges cancel; bound the error by the sum of |contributions|
    scale = np.abs(orc.direct_sum_inv_lap(x, y, z, area, np.abs(pots), 100, 5100))
    assert np.max(np.abs(got - expect)) < 1e-13 * np.max(scale)


def test_direct_sum_targets_and_bve(c1):
    t, ctx, orc, W = c1["t"], c1["ctx"], c1["orc"], c1["sb"].workloads
    flat = [q.ravel().copy() for q in (c1["x"], c1["y"], c1["z"], c1["a"])]
    vor = W.gaussian_vortices(*flat, seed=2)
    tx, ty, tz = W.random_unit_vectors(777, seed=3)
    expect = orc.direct_sum_inv_lap_targets(tx, ty, tz, *flat, vor)
    g = ctx.direct_sum_inv_lap_targets(*[to_gpu(t, q) for q in (tx, ty, tz)], *[to_gpu(t, q) for q in flat],
                                       to_gpu(t, vor))
    ctx.synchronize()
    assert rel(to_np(g), expect) < 1e-12
    rng = np.random.default_rng(4)
    disp = [1e-3 * rng.normal(size=flat[0].size) for _ in range(3)]
    ev = orc.direct_sum_bve_vel(*flat, vor, EPS, disp)
    gv = ctx.direct_sum_bve_vel(*[to_gpu(t, q) for q in flat], to_gpu(t, vor), [to_gpu(t, d) for d in disp])
    ctx.synchronize()
    assert W.vec_rel_l2([to_np(q) for q in gv], ev) < 1e-12
    # Biot-Savart at separate targets: the oracle's sum, and on particles lb..ub the particle sum
    ev = orc.direct_sum_bve_vel_targets(tx, ty, tz, *flat, vor, EPS)
    gv = ctx.direct_sum_bve_vel_targets(*[to_gpu(t, q) for q in (tx, ty, tz)], *[to_gpu(t, q) for q in flat],
                                        to_gpu(t, vor))
    lb, ub = 300, 1100
    gs = ctx.direct_sum_bve_vel_targets(*[to_gpu(t, q[lb:ub].copy()) for q in flat[:3]],
                                        *[to_gpu(t, q) for q in flat], to_gpu(t, vor))
    ep = orc.direct_sum_bve_vel(*flat, vor, EPS, None, lb, ub)
    ctx.synchronize()
    assert W.vec_rel_l2([to_np(q) for q in gv], ev) < 1e-12
    assert W.vec_rel_l2([to_np(q) for q in gs], [e[lb:ub] for e in ep]) < 1e-12


def test_displacement(c1):
    t, ctx, o = c1["t"], c1["ctx"], c1["o"]
    x, y, z = c1["x"], c1["y"], c1["z"]
    vel = [3e-6 * (y * z + 0.5), 2e-6 * np.sin(3 * x), -2e-6 * x * y]
    vor = np.cos(2 * x) * z * 1e-5
    nv, dp, dx = o.bve_displacement(x, y, z, vel, vor, 1800.0, OMEGA)
    gnv, gdp, gdx = ctx.bve_displacement(*[to_gpu(t, q) for q in (x, y, z)], [to_gpu(t, v) for v in vel],
                                         to_gpu(t, vor), 1800.0)
    ctx.synchronize()
    assert rel(to_np(gnv), nv) < 1e-13
    for f in range(3):
        np.testing.assert_array_equal(to_np(gdp[f]), dp[f])
        assert rel(to_np(gdx[f])[: o.offset], dx[f][: o.offset]) < 1e-13


@pytest.mark.parametrize("combine", [0, 1])
def test_departure_points_and_remap(c1, combine):
    t, ctx, o, sb = c1["t"], c1["ctx"], c1["o"], c1["sb"]
    x, y, z = c1["x"], c1["y"], c1["z"]
    rng = np.random.default_rng(9)
    vels = []
    for s in range(4):
        w = np.array([0.3, -0.2, 1.0]) * 2e-5 + 2e-6 * rng.normal(size=3)
        X = np.stack([x, y, z], -1)
        V = np.cross(np.broadcast_to(w, X.shape), X) + 1e-7 * np.sin(5 * X)
        vels.append([V[..., i].copy() for i in range(3)])
    gctx = sb.Context(sb.RunConfig(levels=4, interp_degree=8, leaf_degree=7, weights="equal_area",
                                   rk4_combine=["reference", "textbook"][combine]))
    dep = o.bve_find_departure_points(x, y, z, vels, DT, combine)
    gdep = gctx.bve_find_departure_points(*[to_gpu(t, q) for q in (x, y, z)],
                                          [[to_gpu(t, v) for v in st] for st in vels], DT)
    gctx.synchronize()
    for f in range(3):
        np.testing.assert_allclose(to_np(gdep[f]), dep[f], rtol=0, atol=1e-14)
    vor = sb.workloads.gaussian_vortices(x, y, z, c1["a"], seed=4)
    tr = np.stack([x * y, np.cos(z)], -1).copy()
    nv, ntr = o.bve_departure_to_target(z, dep, vor, OMEGA, tr)
    gnv, gtr = ctx.bve_departure_to_target(to_gpu(t, z), [to_gpu(t, d) for d in dep], to_gpu(t, vor),
                                           to_gpu(t, tr))
    ctx.synchronize()
    assert rel(to_np(gnv), nv) < 1e-13
    assert rel(to_np(gtr), ntr) < 1e-13


def test_geometry_error_is_reported(c1):
    t, ctx, sb = c1["t"], c1["ctx"], c1["sb"]
    x, y, z = c1["x"], c1["y"], c1["z"]
    bad = x.copy()
    bad[3, 5] = np.nan
    with pytest.raises(sb.SbslpmError) as e:
        ctx.bve_departure_to_target(to_gpu(t, z), [to_gpu(t, bad), to_gpu(t, y), to_gpu(t, z)],
                                    to_gpu(t, np.zeros_like(x)))
        ctx.synchronize()
    assert e.value.code == 2
    ctx.synchronize()  # the flag is cleared once reported


@pytest.mark.parametrize("layout,combine", [("ref", 0), ("A'", 0), ("A'", 1)])
def test_bve_back_rk4_step(sb, orc, torch_gpu, layout, combine):
    t = torch_gpu
    print(f"\n[case step layout={layout} combine={combine}]", flush=True)
    if layout == "ref":
        L, d, md, rule, w = 3, 8, 8, 0, "reference"
    else:
        L, d, md, rule, w = 3, 8, 7, 1, "clenshaw_curtis"
    cfg = sb.RunConfig(levels=L, interp_degree=d, leaf_degree=md, weights=w, kernel_eps=EPS, delta_t=DT,
                       omega=OMEGA, rk4_combine=["reference", "textbook"][combine], tracer_count=1)
    ctx = sb.Context(cfg)
    o = orc.OracleFMM(L, d, md, 0.7, rule)
    x, y, z, a = o.spec_points()
    vor = sb.workloads.gaussian_vortices(x, y, z, a, seed=6) * 3
    tr = (x * z)[..., None].copy()
    nv, vel, ntr = o.bve_back_rk4_step(x, y, z, a, vor, DT, OMEGA, EPS, combine, tr)
    gx, gy, gz, ga = [to_gpu(t, q) for q in (x, y, z, a)]
    gvor, gtr = to_gpu(t, vor), to_gpu(t, tr)
    gvel = ctx.zeros_part(3)
    ctx.bve_back_rk4_step(gx, gy, gz, ga, gvor, gvel, gtr)
    ctx.synchronize()
    W = sb.workloads
    assert W.vec_rel_l2([to_np(v) for v in gvel], vel) < 1e-12
    assert rel(to_np(gvor), nv) < 1e-11
    assert rel(to_np(gtr), ntr) < 1e-11


def test_graph_replays_match_eager(sb, torch_gpu):
    """Each step captured once in a hipGraph and replayed; every replay must
    match the eager result, with eager torch work and syncs in between."""
    t = torch_gpu
    print("\n[case graph replay L4 d8 m8]", flush=True)
    kw = dict(levels=4, interp_degree=8, leaf_degree=7, weights="clenshaw_curtis", kernel_eps=EPS, delta_t=DT,
              omega=OMEGA)
    eager = sb.Context(sb.RunConfig(**kw))
    s = t.cuda.Stream()
    with t.cuda.stream(s):
        graph = sb.Context(sb.RunConfig(use_graph=True, **kw))
    x, y, z, a = eager.spec_points()
    vor = sb.workloads.gaussian_vortices(x, y, z, a, seed=8)
    gx, gy, gz, ga = [to_gpu(t, q) for q in (x, y, z, a)]
    v_e, v_g = to_gpu(t, vor), to_gpu(t, vor)
    vel_e, vel_g = eager.zeros_part(3), graph.zeros_part(3)
    t.cuda.synchronize()
    for step in range(3):
        eager.bve_back_rk4_step(gx, gy, gz, ga, v_e, vel_e)
        eager.synchronize()
        with t.cuda.stream(s):
            graph.bve_back_rk4_step(gx, gy, gz, ga, v_g, vel_g)
        graph.synchronize()
        _ = (v_e * 2).sum().item()  # eager torch work between replays
        np.testing.assert_array_equal(to_np(v_g), to_np(v_e))
        for f in range(3):
            np.testing.assert_array_equal(to_np(vel_g[f]), to_np(vel_e[f]))
