"""Gradient of the full BVE backward RK4 step (bve_time_step.cpp:311-419) and
of several steps with checkpointing, from the hand-written reverse pass
(sbslpm_bve_back_rk4_step_adjoint), against central finite differences of the
library's own FP64 forward step along random directions.  The step is
nonlinear in the vorticity through the displaced sources and the departure
trajectories, so these checks cover the position derivatives of the FMM pair
sums, the transposed passes and proxy-displacement interpolation, the
departure search (BLI differentiated in the point, re-projections, both
combine rules) and the remap with its Coriolis terms."""
import importlib

import numpy as np
import pytest

from conftest import to_gpu, to_np

pytestmark = pytest.mark.gpu

EPS, DT, OMEGA = 1e-4, 3600.0, 7.292e-5


def make(sb, combine, levels=3):
    ctx = sb.Context(sb.RunConfig(levels=levels, interp_degree=8, leaf_degree=7, weights="clenshaw_curtis",
                                  kernel_eps=EPS, omega=OMEGA, delta_t=DT, rk4_combine=combine))
    x, y, z, a = ctx.spec_points()
    return ctx, x, y, z, a


def directional_checks(t, loss, grad, w0, rng, ndir, rel_h=1e-5):
    worst = 0.0
    for k in range(ndir):
        d = to_gpu(t, rng.normal(size=w0.shape))
        h = rel_h * float(w0.norm()) / float(d.norm())
        fd = (loss(w0 + h * d) - loss(w0 - h * d)) / (2 * h)
        an = float((grad * d).sum())
        rel = abs(fd - an) / abs(fd)
        print(f"  direction {k}: fd {fd:.12e} adjoint {an:.12e} rel {rel:.2e}", flush=True)
        worst = max(worst, rel)
    return worst


@pytest.mark.parametrize("combine", ["reference", "textbook"])
def test_step_adjoint_matches_finite_differences(sb, torch_gpu, combine):
    t = torch_gpu
    print(f"\n[case step adjoint L3 m8 d8 combine={combine}]", flush=True)
    ctx, x, y, z, a = make(sb, combine)
    g = [to_gpu(t, q) for q in (x, y, z, a)]
    rng = np.random.default_rng(11)
    w0 = to_gpu(t, sb.workloads.gaussian_vortices(x, y, z, a, seed=12))
    r = to_gpu(t, rng.normal(size=x.shape))
    vel = ctx.zeros_part(3)

    def loss(w):
        v = w.clone()
        ctx.bve_back_rk4_step(*g, v, vel)
        return float((r * v).sum())

    grad = ctx.bve_back_rk4_step_adjoint(*g, w0, r)
    ctx.synchronize()
    assert np.all(np.isfinite(to_np(grad)))
    assert directional_checks(t, loss, grad, w0, rng, 4) < 1e-6


def test_step_adjoint_at_rest_is_remap_transpose(sb, torch_gpu):
    """At zero vorticity and Omega = 0 nothing moves: the departure points are the
    particles themselves, the gradient of every velocity-dependent path vanishes
    (the field has zero gradient), and the step's Jacobian is the remap
    interpolation at those points.  Particles on a leaf edge coincide with a
    node of the neighbouring leaf, where point location may take them, so the
    Jacobian gathers from that leaf's node: the gradient is B(dep)^T r, not r."""
    t = torch_gpu
    print("\n[case step adjoint at rest]", flush=True)
    ctx = sb.Context(sb.RunConfig(levels=3, interp_degree=8, leaf_degree=7, weights="clenshaw_curtis",
                                  kernel_eps=EPS, omega=0.0, delta_t=DT))
    x, y, z, a = ctx.spec_points()
    g = [to_gpu(t, q) for q in (x, y, z, a)]
    r = to_gpu(t, np.random.default_rng(2).normal(size=x.shape))
    zero = to_gpu(t, np.zeros_like(x))
    grad = ctx.bve_back_rk4_step_adjoint(*g, zero, r)
    dep = ctx.bve_find_departure_points(*g[:3], [[zero, zero, zero]] * 4, DT)
    ref = ctx.bve_departure_to_target_adjoint(dep, r)
    ctx.synchronize()
    np.testing.assert_allclose(to_np(grad), to_np(ref), rtol=0, atol=1e-12)
    # and the remap itself at rest is the identity in value for a continuous field
    f = to_gpu(t, x * y + z)
    moved, _ = ctx.bve_departure_to_target(g[2], dep, f)
    np.testing.assert_allclose(to_np(moved), to_np(f), rtol=0, atol=1e-13)


def test_checkpointed_steps_gradient(sb, torch_gpu):
    t = torch_gpu
    print("\n[case 4 steps, checkpoints every 1, 2, 4]", flush=True)
    ag = importlib.import_module("spectral-bslpm_amd.autograd")
    ctx, x, y, z, a = make(sb, "reference")
    g = [to_gpu(t, q) for q in (x, y, z, a)]
    rng = np.random.default_rng(21)
    w0 = to_gpu(t, sb.workloads.gaussian_vortices(x, y, z, a, seed=22))
    r = to_gpu(t, rng.normal(size=x.shape))
    grads = []
    for every in (1, 2, 4):
        w = w0.clone().requires_grad_(True)
        out = ag.bve_steps(ctx, *g, w, 4, every)
        (r * out).sum().backward()
        grads.append(w.grad.detach().clone())
    for gr in grads[1:]:
        np.testing.assert_allclose(to_np(gr), to_np(grads[0]), rtol=1e-12, atol=1e-12 * float(grads[0].abs().max()))

    def loss(w):
        with t.no_grad():
            return float((r * ag.bve_steps(ctx, *g, w, 4, 4)).sum())

    assert directional_checks(t, loss, grads[1], w0, rng, 3) < 1e-6


def test_step_adjoint_rejects_aliasing(sb, torch_gpu):
    t = torch_gpu
    ctx, x, y, z, a = make(sb, "reference", levels=2)
    g = [to_gpu(t, q) for q in (x, y, z, a)]
    w = to_gpu(t, np.zeros_like(x))
    with pytest.raises(sb.SbslpmError):
        ctx.bve_back_rk4_step_adjoint(*g, w, w.clone(), w)


def test_step_gradient_against_direct_sum_reference(sb, torch_gpu):
    """Against torch.autograd through the FP64 direct-sum step (tests/torch_ref_bve.py, CPU):
    the two differ by the FMM's approximation of the displaced-source sums."""
    t = torch_gpu
    from torch_ref_bve import StepModel
    print("\n[case step gradient vs direct-sum autograd, L3]", flush=True)
    ctx, x, y, z, a = make(sb, "reference")
    g = [to_gpu(t, q) for q in (x, y, z, a)]
    w = sb.workloads.gaussian_vortices(x, y, z, a, seed=12)
    r = np.random.default_rng(13).normal(size=x.shape)
    got = to_np(ctx.bve_back_rk4_step_adjoint(*g, to_gpu(t, w), to_gpu(t, r))).ravel()
    ctx.synchronize()
    M = StepModel(x, y, z, a, 3, EPS, OMEGA, DT)
    wt = t.tensor(w.ravel(), dtype=t.float64, requires_grad=True)
    (t.tensor(r.ravel(), dtype=t.float64) * M.step(wt)).sum().backward()
    ref = wt.grad.numpy()
    err = sb.workloads.rel_l2(got, ref)
    print(f"  gradient rel-L2 vs direct sum {err:.2e}")
    assert err < 1e-4


def test_mixed_precision_gradient_matches_fp64(sb, torch_gpu):
    """A mixed-precision context runs the transposed pair sums in FP32 on its own (FP32)
    trajectory; the gradient stays within the 1e-6 relative bar of the FP64 one."""
    t = torch_gpu
    print("\n[case step adjoint mixed vs fp64]", flush=True)
    kw = dict(levels=4, interp_degree=8, leaf_degree=7, weights="clenshaw_curtis", kernel_eps=EPS, omega=OMEGA,
              delta_t=DT)
    c64 = sb.Context(sb.RunConfig(precision="fp64", **kw))
    cmx = sb.Context(sb.RunConfig(precision="mixed", **kw))
    x, y, z, a = c64.spec_points()
    g = [to_gpu(t, q) for q in (x, y, z, a)]
    w = to_gpu(t, sb.workloads.gaussian_vortices(x, y, z, a, seed=31))
    r = to_gpu(t, np.random.default_rng(32).normal(size=x.shape))
    ub = [r, 0.5 * r, -r]
    g64, gmx = c64.bve_back_rk4_step_adjoint(*g, w, r), cmx.bve_back_rk4_step_adjoint(*g, w, r)
    v64, vmx = c64.bve_velocity_adjoint(g[3], ub), cmx.bve_velocity_adjoint(g[3], ub)
    c64.synchronize()
    cmx.synchronize()
    W = sb.workloads
    e_step = W.rel_l2(to_np(gmx), to_np(g64))
    e_vel = W.rel_l2(to_np(vmx), to_np(v64))
    print(f"  mixed vs fp64: velocity adjoint {e_vel:.2e}, step gradient {e_step:.2e}")
    assert e_vel < 1e-6 and e_step < 1e-6


def test_recorded_step_and_stage_adjoint(sb, torch_gpu):
    """The recording step equals bve_back_rk4_step bit for bit, and the reverse pass on the
    recorded stages equals the step adjoint (which recomputes them) to rounding."""
    t = torch_gpu
    print("\n[case recorded step]", flush=True)
    ctx, x, y, z, a = make(sb, "reference")
    g = [to_gpu(t, q) for q in (x, y, z, a)]
    w = to_gpu(t, sb.workloads.gaussian_vortices(x, y, z, a, seed=51))
    r = to_gpu(t, np.random.default_rng(52).normal(size=x.shape))
    ref = w.clone()
    ctx.bve_back_rk4_step(*g, ref, ctx.zeros_part(3))
    out, stages = ctx.bve_back_rk4_step_record(*g, w)
    g1 = ctx.bve_back_rk4_step_adjoint(*g, w, r)
    g2 = ctx.bve_back_rk4_step_adjoint_stages(*g, w, stages, r)
    ctx.synchronize()
    np.testing.assert_array_equal(to_np(out), to_np(ref))
    np.testing.assert_allclose(to_np(g2), to_np(g1), rtol=0, atol=1e-13 * float(g1.abs().max()))
