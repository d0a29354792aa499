"""The pair-kernel autotuner: it times every candidate launch shape on the
context's own interaction list, keeps the fastest, and the results of the
evaluation and of the (graph-captured) step do not change, since each target
is still summed by one lane in the same order whatever the shape."""
import numpy as np
import pytest

from conftest import to_gpu, to_np

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed_valu"])
def test_autotune_picks_fastest_and_keeps_results(sb, torch_gpu, precision):
    t = torch_gpu
    print(f"\n[case autotune {precision}]", flush=True)
    kw = dict(levels=5, interp_degree=8, leaf_degree=7, weights="clenshaw_curtis", kernel_eps=1e-4,
              precision=precision)
    ctx = sb.Context(sb.RunConfig(**kw))
    x, y, z, a = ctx.spec_points()
    g = [to_gpu(t, q) for q in (x, y, z, a)]
    vor = to_gpu(t, sb.workloads.gaussian_vortices(x, y, z, a, seed=5))
    before = [to_np(v) for v in ctx.bve_velocity(g[3], vor)]
    table, shape = ctx.autotune(g[3], vor)
    ctx.synchronize()
    print("  ", [(int(w), int(c), round(ms, 4)) for w, c, ms in table], "->", shape)
    assert len(table) == 9 and all(ms > 0 for _, _, ms in table)
    best = min(table, key=lambda r: r[2])
    assert shape == (int(best[0]), int(best[1]))
    after = [to_np(v) for v in ctx.bve_velocity(g[3], vor)]
    for b, c in zip(before, after):
        np.testing.assert_array_equal(b, c)


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed_valu"])
def test_every_xcd_mapping_covers_every_tile(sb, torch_gpu, precision):
    """Mappings 0 (round-robin), 1 (equal contiguous ranges) and 2 (cost-balanced contiguous
    ranges) at 1, 2 and 4 waves per workgroup give the same velocity to the bit.  Each run
    follows one on a different field, so a tile that a mapping skipped would keep the other
    field's values."""
    t = torch_gpu
    print(f"\n[case xcd mappings {precision}]", flush=True)
    kw = dict(levels=5, interp_degree=8, leaf_degree=7, weights="clenshaw_curtis", kernel_eps=1e-4,
              precision=precision)
    ctx = sb.Context(sb.RunConfig(**kw))
    x, y, z, a = ctx.spec_points()
    g = [to_gpu(t, q) for q in (x, y, z, a)]
    va = to_gpu(t, sb.workloads.gaussian_vortices(x, y, z, a, seed=5))
    vb = to_gpu(t, sb.workloads.gaussian_vortices(x, y, z, a, seed=6))
    ctx.set_pair_shape(4, 0)
    ref = [to_np(v) for v in ctx.bve_velocity(g[3], vb)]
    for waves in (1, 2, 4):
        for xcd in (0, 1, 2):
            ctx.set_pair_shape(4, 0)
            ctx.bve_velocity(g[3], va)
            ctx.set_pair_shape(waves, xcd)
            assert ctx.pair_shape() == (waves, xcd)
            got = [to_np(v) for v in ctx.bve_velocity(g[3], vb)]
            for r, q in zip(ref, got):
                np.testing.assert_array_equal(r, q)


def test_autotuned_graph_step_matches_eager(sb, torch_gpu):
    t = torch_gpu
    print("\n[case autotune + graph]", flush=True)
    kw = dict(levels=4, interp_degree=8, leaf_degree=7, weights="clenshaw_curtis", kernel_eps=1e-4, precision="mixed")
    s = t.cuda.Stream()
    with t.cuda.stream(s):
        gctx = sb.Context(sb.RunConfig(use_graph=True, **kw))
    ectx = sb.Context(sb.RunConfig(**kw))
    x, y, z, a = ectx.spec_points()
    g = [to_gpu(t, q) for q in (x, y, z, a)]
    v0 = sb.workloads.gaussian_vortices(x, y, z, a, seed=6)
    ve, vg = to_gpu(t, v0), to_gpu(t, v0)
    vel_e, vel_g = ectx.zeros_part(3), gctx.zeros_part(3)
    t.cuda.synchronize()
    with t.cuda.stream(s):
        gctx.bve_back_rk4_step(*g, vg, vel_g)  # captured with the default shape
        gctx.autotune(g[3], vg)               # drops the graph
    gctx.synchronize()
    ectx.bve_back_rk4_step(*g, ve, vel_e)
    ectx.synchronize()
    for _ in range(2):
        with t.cuda.stream(s):
            gctx.bve_back_rk4_step(*g, vg, vel_g)
        gctx.synchronize()
        ectx.bve_back_rk4_step(*g, ve, vel_e)
        ectx.synchronize()
        np.testing.assert_array_equal(to_np(vg), to_np(ve))
