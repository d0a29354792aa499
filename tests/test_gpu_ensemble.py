"""Ensemble stepping: several BVE members share every kernel launch
(gridDim.y = members) and the tree, interaction list and work tiles.  Each
member must come out bit for bit as the single-member bve_back_rk4_step
(bve_time_step.cpp:311-419) leaves it, in every precision, eagerly and as a
replayed hipGraph.  With the multi-right-hand-side first stage (mixed
precision, members' charges as columns of one product) members agree with
the single-member steps to FP32 rounding."""
import numpy as np
import pytest

from conftest import to_gpu, to_np

pytestmark = pytest.mark.gpu

EPS, DT, OMEGA = 1e-4, 3600.0, 7.292e-5


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed_valu"])
def test_ensemble_members_match_single_steps(sb, torch_gpu, precision):
    t = torch_gpu
    print(f"\n[case ensemble precision={precision}]", flush=True)
    kw = dict(levels=3, interp_degree=8, leaf_degree=7, weights="clenshaw_curtis", kernel_eps=EPS, delta_t=DT,
              omega=OMEGA, precision=precision)
    ctx = sb.Context(sb.RunConfig(**kw))
    x, y, z, a = ctx.spec_points()
    gx, gy, gz, ga = [to_gpu(t, q) for q in (x, y, z, a)]
    members = [sb.workloads.gaussian_vortices(x, y, z, a, seed=s) * (1 + 0.5 * s) for s in (1, 2, 3)]
    single_vor, single_vel = [], []
    for vor in members:
        gv = to_gpu(t, vor)
        vel = ctx.zeros_part(3)
        ctx.bve_back_rk4_step(gx, gy, gz, ga, gv, vel)
        ctx.synchronize()
        single_vor.append(to_np(gv))
        single_vel.append([to_np(v) for v in vel])
    ens = to_gpu(t, np.stack(members))
    evel = [t.zeros_like(ens) for _ in range(3)]
    for _ in range(2):  # the second call reuses the workspace
        e2 = to_gpu(t, np.stack(members))
        ctx.bve_back_rk4_step_ensemble(gx, gy, gz, ga, e2, evel)
        ctx.synchronize()
    for j in range(len(members)):
        np.testing.assert_array_equal(to_np(e2[j]), single_vor[j])
        for f in range(3):
            np.testing.assert_array_equal(to_np(evel[f][j]), single_vel[j][f])


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("nmem", [3, 6, 9])
def test_ensemble_multi_rhs_first_stage(sb, torch_gpu, nmem):
    """nmem = 3: one chunk of four members (one padded); 6: two chunks in one
    wave (two padded); 9: two waves of two chunks (seven padded).  Stage-0
    velocities and stepped vorticities against the single-member mixed and FP64
    steps; the mixed-precision bar is rel-L2 1e-6 against FP64."""
    t = torch_gpu
    print(f"\n[case ensemble multi_rhs nmem={nmem}]", flush=True)
    kw = dict(levels=3, interp_degree=8, leaf_degree=7, weights="clenshaw_curtis", kernel_eps=EPS, delta_t=DT,
              omega=OMEGA, precision="mixed")
    ctx = sb.Context(sb.RunConfig(multi_rhs=True, **kw))
    cmx = sb.Context(sb.RunConfig(**kw))
    c64 = sb.Context(sb.RunConfig(**dict(kw, precision="fp64")))
    x, y, z, a = ctx.spec_points()
    gx, gy, gz, ga = [to_gpu(t, q) for q in (x, y, z, a)]
    members = [sb.workloads.gaussian_vortices(x, y, z, a, seed=s) * (1 + 0.5 * s) for s in range(1, nmem + 1)]
    single = {}
    for name, c in (("mixed", cmx), ("fp64", c64)):
        single[name] = []
        for vor in members:
            gv = to_gpu(t, vor)
            vel = c.zeros_part(3)
            c.bve_back_rk4_step(gx, gy, gz, ga, gv, vel)
            c.synchronize()
            single[name].append((to_np(gv), np.stack([to_np(v) for v in vel])))
    ctx.enable_timing(True)
    for _ in range(2):  # the second call reuses the workspace and the charge columns
        ens = to_gpu(t, np.stack(members))
        evel = [t.zeros_like(ens) for _ in range(3)]
        ctx.bve_back_rk4_step_ensemble(gx, gy, gz, ga, ens, evel)
        ctx.synchronize()
        for j in range(nmem):
            vel = np.stack([to_np(evel[f][j]) for f in range(3)])
            vor = to_np(ens[j])
            errs = (rel_l2(vel, single["mixed"][j][1]), rel_l2(vel, single["fp64"][j][1]),
                    rel_l2(vor, single["mixed"][j][0]), rel_l2(vor, single["fp64"][j][0]))
            base = rel_l2(single["mixed"][j][1], single["fp64"][j][1])
            print(f"member {j}: vel vs mixed {errs[0]:.2e} vs fp64 {errs[1]:.2e} (single mixed vs fp64 "
                  f"{base:.2e}); vor vs mixed {errs[2]:.2e} vs fp64 {errs[3]:.2e}", flush=True)
            # two FP32 summation orders, each at the FP32 rounding level of the eps = 1e-4 near field
            assert errs[0] < 1e-6 and errs[1] < 1e-6, j
            assert errs[2] < 1e-6 and errs[3] < 1e-6, j
    _, launches = ctx.kernel_time("bve_mrhs", reset=True)
    assert launches == 2  # one first-stage launch per ensemble step
    _, launches = ctx.kernel_time("bve_mixed", reset=True)
    assert launches == 6  # stages 1..3 on the per-member kernel


def test_ensemble_multi_rhs_autotune_keeps_results(sb, torch_gpu):
    """The multi-RHS launch shape (waves, XCD mapping, chunks per wave) changes no bit."""
    t = torch_gpu
    print("\n[case ensemble multi_rhs autotune]", flush=True)
    kw = dict(levels=3, interp_degree=8, leaf_degree=7, weights="clenshaw_curtis", kernel_eps=EPS, delta_t=DT,
              omega=OMEGA, precision="mixed", multi_rhs=True)
    plain, tuned = sb.Context(sb.RunConfig(**kw)), sb.Context(sb.RunConfig(**kw))
    x, y, z, a = plain.spec_points()
    gx, gy, gz, ga = [to_gpu(t, q) for q in (x, y, z, a)]
    vors = np.stack([sb.workloads.gaussian_vortices(x, y, z, a, seed=s) for s in range(6)])
    rows, best = tuned.autotune_ensemble(ga, to_gpu(t, vors))
    assert len(rows) == 18 and all(r[3] > 0 for r in rows)
    assert best[2] in (1, 2) and best[0] in (1, 2, 4)
    out = []
    for c in (plain, tuned):
        v = to_gpu(t, vors)
        vel = [t.zeros_like(v) for _ in range(3)]
        c.bve_back_rk4_step_ensemble(gx, gy, gz, ga, v, vel)
        c.synchronize()
        out.append([to_np(v)] + [to_np(q) for q in vel])
    for p, q in zip(*out):
        np.testing.assert_array_equal(p, q)
    c64 = sb.Context(sb.RunConfig(**dict(kw, precision="fp64")))
    with pytest.raises(sb.SbslpmError):
        c64.autotune_ensemble(ga, to_gpu(t, vors))


@pytest.mark.parametrize("multi_rhs", [False, True])
def test_ensemble_graph_replays_match_eager(sb, torch_gpu, multi_rhs):
    t = torch_gpu
    print(f"\n[case ensemble graph replay multi_rhs={multi_rhs}]", flush=True)
    kw = dict(levels=4, interp_degree=8, leaf_degree=7, weights="clenshaw_curtis", kernel_eps=EPS, delta_t=DT,
              omega=OMEGA, multi_rhs=multi_rhs, precision="mixed" if multi_rhs else "fp64")
    eager = sb.Context(sb.RunConfig(**kw))
    s = t.cuda.Stream()
    with t.cuda.stream(s):
        graph = sb.Context(sb.RunConfig(use_graph=True, **kw))
    x, y, z, a = eager.spec_points()
    gx, gy, gz, ga = [to_gpu(t, q) for q in (x, y, z, a)]
    vors = np.stack([sb.workloads.gaussian_vortices(x, y, z, a, seed=s) for s in (4, 5)])
    v_e, v_g = to_gpu(t, vors), to_gpu(t, vors)
    vel_e = [t.zeros_like(v_e) for _ in range(3)]
    vel_g = [t.zeros_like(v_g) for _ in range(3)]
    t.cuda.synchronize()
    for step in range(3):
        eager.bve_back_rk4_step_ensemble(gx, gy, gz, ga, v_e, vel_e)
        eager.synchronize()
        with t.cuda.stream(s):
            graph.bve_back_rk4_step_ensemble(gx, gy, gz, ga, v_g, vel_g)
        graph.synchronize()
        _ = (v_e * 2).sum().item()
        np.testing.assert_array_equal(to_np(v_g), to_np(v_e))
        for f in range(3):
            np.testing.assert_array_equal(to_np(vel_g[f]), to_np(vel_e[f]))


def test_ensemble_rejects_bad_arguments(sb, torch_gpu):
    t = torch_gpu
    ctx = sb.Context(sb.RunConfig(levels=1, interp_degree=4, leaf_degree=4))
    x, y, z, a = ctx.spec_points()
    g = [to_gpu(t, q) for q in (x, y, z, a)]
    vors = to_gpu(t, np.stack([np.zeros_like(x)] * 2))
    with pytest.raises(sb.SbslpmError):
        ctx.bve_back_rk4_step_ensemble(*g, vors, [t.zeros_like(vors) for _ in range(3)])


@pytest.mark.parametrize("first,second", [("fp64", "mixed"), ("mixed", "fp64"), ("mixed", "mixed_valu")])
def test_ensemble_after_set_precision(sb, torch_gpu, first, second):
    """An ensemble workspace holds the source records of the precision it was
    built for; after set_precision the next ensemble step must rebuild it.
    The step on the switched context equals the step of a fresh context of the
    second precision, bit for bit, eagerly and as a replayed graph."""
    t = torch_gpu
    print(f"\n[case ensemble set_precision {first}->{second}]", flush=True)
    kw = dict(levels=3, interp_degree=8, leaf_degree=7, weights="clenshaw_curtis", kernel_eps=EPS, delta_t=DT,
              omega=OMEGA, multi_rhs=True)
    stream = t.cuda.Stream()
    with t.cuda.stream(stream):
        sw = sb.Context(sb.RunConfig(precision=first, use_graph=True, **kw))
        fresh = sb.Context(sb.RunConfig(precision=second, use_graph=True, **kw))
    x, y, z, a = sw.spec_points()
    gx, gy, gz, ga = [to_gpu(t, q) for q in (x, y, z, a)]
    members = np.stack([sb.workloads.gaussian_vortices(x, y, z, a, seed=s) for s in (1, 2, 3, 4, 5)])
    with t.cuda.stream(stream):
        e = to_gpu(t, members)
        sw.bve_back_rk4_step_ensemble(gx, gy, gz, ga, e, [t.zeros_like(e) for _ in range(3)])
    sw.synchronize()
    sw.set_precision(second)
    out = {}
    for name, c in (("switched", sw), ("fresh", fresh)):
        with t.cuda.stream(stream):
            e = to_gpu(t, members)
            vel = [t.zeros_like(e) for _ in range(3)]
            for _ in range(2):  # capture, then replay
                e.copy_(to_gpu(t, members))
                c.bve_back_rk4_step_ensemble(gx, gy, gz, ga, e, vel)
        c.synchronize()
        t.cuda.synchronize()
        out[name] = (to_np(e), [to_np(v) for v in vel])
    np.testing.assert_array_equal(out["switched"][0], out["fresh"][0])
    for f in range(3):
        np.testing.assert_array_equal(out["switched"][1][f], out["fresh"][1][f])
