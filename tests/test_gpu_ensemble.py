"""Ensemble stepping: several BVE members share every kernel launch
(gridDim.y = members) and the tree, interaction list and work tiles.  Each
member must come out bit for bit as the single-member bve_back_rk4_step
(bve_time_step.cpp:311-419) leaves it, in every precision, eagerly and as a
replayed hipGraph."""
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


def test_ensemble_graph_replays_match_eager(sb, torch_gpu):
    t = torch_gpu
    print("\n[case ensemble graph replay]", flush=True)
    kw = dict(levels=4, interp_degree=8, leaf_degree=7, weights="clenshaw_curtis", kernel_eps=EPS, delta_t=DT,
              omega=OMEGA)
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
