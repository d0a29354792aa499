"""The opt-in bf16 M2L (SBSLPM_FLAG_BF16_M2L): on a mixed context the kernel product of
proxy-to-proxy interactions takes bf16 operands.  Its velocity stays close to FP64 but its
error is far above the FP32 form's (which meets 1e-6); leaf and near-field interactions keep
FP32, so a tree without proxy-to-proxy sets is unchanged to the bit."""
import numpy as np
import pytest

from conftest import to_gpu, to_np

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_bf16_m2l_error(sb, torch_gpu):
    t = torch_gpu
    print("\n[case bf16 M2L L5]", flush=True)
    kw = dict(levels=5, interp_degree=8, leaf_degree=7, weights="clenshaw_curtis", kernel_eps=1e-4)
    c64, c32, cbf = (sb.Context(sb.RunConfig(precision=p, bf16_m2l=b, **kw))
                     for p, b in (("fp64", False), ("mixed", False), ("mixed", True)))
    x, y, z, a = c64.spec_points()
    g = [to_gpu(t, q) for q in (x, y, z, a)]
    vor = to_gpu(t, sb.workloads.gaussian_vortices(x, y, z, a, seed=5))
    u64, u32, ubf = [np.stack([to_np(v) for v in c.bve_velocity(g[3], vor)]) for c in (c64, c32, cbf)]
    e32, ebf = rel(u32, u64), rel(ubf, u64)
    print(f"  velocity rel-L2 vs FP64: FP32 {e32:.2e}, bf16 M2L {ebf:.2e}")
    assert e32 < 1e-6
    assert e32 < ebf < 1e-2


def test_bf16_m2l_without_m2l_sets_is_unchanged(sb, torch_gpu):
    """One level (the six faces are the leaves): no panel pair is proxy to proxy, so the
    flag changes nothing."""
    t = torch_gpu
    print("\n[case bf16 M2L L1]", flush=True)
    kw = dict(levels=1, interp_degree=8, leaf_degree=7, weights="clenshaw_curtis", kernel_eps=1e-4,
              precision="mixed")
    c32, cbf = sb.Context(sb.RunConfig(**kw)), sb.Context(sb.RunConfig(bf16_m2l=True, **kw))
    x, y, z, a = c32.spec_points()
    g = [to_gpu(t, q) for q in (x, y, z, a)]
    vor = to_gpu(t, sb.workloads.gaussian_vortices(x, y, z, a, seed=5))
    for p, q in zip(c32.bve_velocity(g[3], vor), cbf.bve_velocity(g[3], vor)):
        np.testing.assert_array_equal(to_np(p), to_np(q))
