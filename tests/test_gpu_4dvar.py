"""Vorticity 4D-Var on the BVE through the hand-written adjoints: the
observation operator (BLI at arbitrary points) against its transpose by a dot
product, and the gradient of L = 1/2 |H(vor_N) - obs|^2 over several
checkpointed steps against central finite differences of the library's FP64
forward run."""
import importlib

import numpy as np
import pytest

from conftest import to_gpu, to_np

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(sb, torch_gpu):
    t = torch_gpu
    print("\n[case 4dvar L4 m6]", flush=True)
    ctx = sb.Context(sb.RunConfig(levels=4, interp_degree=8, leaf_degree=5, weights="clenshaw_curtis",
                                  kernel_eps=1e-4, omega=7.292e-5, delta_t=1800.0))
    x, y, z, a = ctx.spec_points()
    rng = np.random.default_rng(40)
    obs = rng.normal(size=(3, 512))
    obs /= np.linalg.norm(obs, axis=0)
    return ctx, [to_gpu(t, q) for q in (x, y, z, a)], [to_gpu(t, o) for o in obs], (x, y, z, a)


def test_observation_operator_transpose(setup, torch_gpu):
    t = torch_gpu
    ctx, g, obs, host = setup
    rng = np.random.default_rng(41)
    f, v = rng.normal(size=host[0].shape), rng.normal(size=obs[0].numel())
    hf = to_np(ctx.interp_points(*obs, [to_gpu(t, f)])[0])
    htv = to_np(ctx.interp_points_adjoint(*obs, [to_gpu(t, v)])[0])
    ctx.synchronize()
    lhs, rhs = float(np.dot(hf, v)), float(np.sum(f * htv))
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))


def test_4dvar_gradient_matches_finite_differences(setup, sb, torch_gpu):
    t = torch_gpu
    ag = importlib.import_module("spectral-bslpm_amd.autograd")
    ctx, g, obs, host = setup
    x, y, z, a = host
    truth = to_gpu(t, sb.workloads.gaussian_vortices(x, y, z, a, seed=42))
    vn = ag.bve_steps(ctx, *g, truth, 6, 6)
    obs_val = ag.interp_points(ctx, *obs, vn).detach()
    first = to_gpu(t, sb.workloads.gaussian_vortices(x, y, z, a, seed=43)) * 0.3 + truth
    L, grad = ag.bve_4dvar(ctx, *g, first, 6, 2, obs, obs_val)

    def loss(w):
        with t.no_grad():
            hx = ag.interp_points(ctx, *obs, ag.bve_steps(ctx, *g, w, 6, 6))
            return float(0.5 * ((hx - obs_val) ** 2).sum())

    assert abs(loss(first) - float(L)) <= 1e-12 * float(L)
    rng = np.random.default_rng(44)
    for k in range(3):
        d = to_gpu(t, rng.normal(size=x.shape))
        h = 1e-4 * float(first.norm()) / float(d.norm())
        fd = (loss(first + h * d) - loss(first - h * d)) / (2 * h)
        an = float((grad * d).sum())
        rel = abs(fd - an) / abs(fd)
        print(f"  direction {k}: fd {fd:.10e} adjoint {an:.10e} rel {rel:.2e}")
        assert rel < 1e-6
