"""GPU parity of the FMM evaluation at arbitrary targets (far field by BLI from
the containing leaf, child_panel_interp_3_ll, fmm_funcs_impl.hpp:318-364; near
field over the leaf's type-0/1 interactions, swe_vel_impl.hpp:188-255)
against the FP64 oracle, and against the direct sum at the 4,096-target size
of the Poisson workload.

FP64 tolerance: relative 1e-12 (summation order differs from the oracle)."""
import numpy as np
import pytest

from conftest import to_gpu, to_np
from test_oracle_targets import special_points

pytestmark = pytest.mark.gpu

RULE = {"reference": 0, "clenshaw_curtis": 1, "equal_area": 2}
CASES = {
    "ref-layout-L3-d8": dict(levels=3, interp_degree=8, leaf_degree=8, fmm_theta=0.7, weights="reference"),
    "A'-L4-d8-m8": dict(levels=4, interp_degree=8, leaf_degree=7, fmm_theta=0.7, weights="clenshaw_curtis"),
    "A'-L3-d10-m6": dict(levels=3, interp_degree=10, leaf_degree=5, fmm_theta=0.6, weights="clenshaw_curtis"),
    "one-level-L1-d6": dict(levels=1, interp_degree=6, leaf_degree=6, fmm_theta=0.7, weights="reference"),
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
    f = sb.workloads.gaussian_vortices(x, y, z, a, seed=17)
    T = np.concatenate([np.stack(sb.workloads.random_unit_vectors(1500, seed=23), 1), special_points()])
    # the particles themselves, edge (duplicated) nodes included
    T = np.concatenate([T, np.stack([x.ravel()[::7], y.ravel()[::7], z.ravel()[::7]], 1)])
    return dict(ctx=ctx, o=o, a=a, f=f, T=T, t=torch_gpu, sb=sb)


def test_poisson_targets_match_oracle(case):
    ctx, o, t, T = case["ctx"], case["o"], case["t"], case["T"]
    tx, ty, tz = [T[:, i].copy() for i in range(3)]
    expect = o.eval_targets("poisson", case["a"], case["f"], tx, ty, tz)[0]
    got = ctx.poisson_solve_targets(to_gpu(t, case["a"]), to_gpu(t, case["f"]), *[to_gpu(t, v) for v in (tx, ty, tz)])
    ctx.synchronize()
    err = case["sb"].workloads.rel_l2(to_np(got), expect)
    print(f"  psi rel-L2 vs oracle {err:.2e}")
    assert err < 1e-12


def test_bve_targets_match_oracle(case):
    ctx, o, t, T = case["ctx"], case["o"], case["t"], case["T"]
    tx, ty, tz = [T[:, i].copy() for i in range(3)]
    expect = o.eval_targets("bve", case["a"], case["f"], tx, ty, tz, EPS)
    got = ctx.bve_velocity_targets(to_gpu(t, case["a"]), to_gpu(t, case["f"]), *[to_gpu(t, v) for v in (tx, ty, tz)])
    ctx.synchronize()
    err = case["sb"].workloads.vec_rel_l2([to_np(g) for g in got], expect)
    print(f"  velocity rel-L2 vs oracle {err:.2e}")
    assert err < 1e-12


def test_targets_empty_and_nan(sb, torch_gpu):
    t = torch_gpu
    ctx = sb.Context(sb.RunConfig(levels=2, interp_degree=4, leaf_degree=4))
    x, y, z, a = ctx.spec_points()
    ga, gf = to_gpu(t, a), to_gpu(t, np.ones_like(a))
    e = t.zeros(0, dtype=t.float64, device="cuda")
    ctx.poisson_solve_targets(ga, gf, e, e, e, psi=e)
    ctx.synchronize()
    bad = [to_gpu(t, np.array(v)) for v in ([np.nan, 0.0], [0.0, 1.0], [1.0, 0.0])]
    ctx.poisson_solve_targets(ga, gf, *bad)
    with pytest.raises(sb.SbslpmError):
        ctx.synchronize()


def test_poisson_workload_4096_targets_vs_direct_sum(sb, orc, torch_gpu):
    """6 x 256^2 Poisson workload (L=6, m=8, d=10, theta=0.6), SH forcing l <= 32:
    FMM at 4,096 uniform targets against the FP64 direct sum at the same targets
    and against the closed form psi = sum a_lm Y_lm / (l(l+1))."""
    t, W = torch_gpu, sb.workloads
    ctx = sb.Context(sb.RunConfig(levels=6, interp_degree=10, leaf_degree=7, fmm_theta=0.6,
                                  weights="clenshaw_curtis"))
    x, y, z, a = ctx.spec_points()
    f = W.random_sh_forcing(x, y, z, lmax=32, seed=1)[0]
    tx, ty, tz = W.random_unit_vectors(4096, seed=2)
    exact = W.random_sh_forcing(tx, ty, tz, lmax=32, seed=1)[1]
    g = [to_gpu(t, v) for v in (tx, ty, tz)]
    fmm = to_np(ctx.poisson_solve_targets(to_gpu(t, a), to_gpu(t, f), *g))
    flat = [to_gpu(t, v.ravel().copy()) for v in (x, y, z, a, f)]
    ds = to_np(ctx.direct_sum_inv_lap_targets(*g, *flat))
    ctx.synchronize()
    e_ds, e_ex = W.rel_l2(fmm, ds), W.rel_l2(fmm, exact)
    print(f"  FMM vs direct sum {e_ds:.2e} (Linf {W.rel_linf(fmm, ds):.2e}), vs closed form {e_ex:.2e}")
    assert e_ds < 1e-9
    assert e_ex < 1e-3
