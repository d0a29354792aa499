"""The shallow-water step in the physical form of the 4D-Var configuration, on the GPU
against the FP64 oracle: heights in metres with the free-surface Laplacian scaled by
g/a^2 (g = 9.80616, a = 6.37122e6 m), the Gaussian mountain, the regularised stage-1
Laplacian and the free-surface height correction (sbslpm.h, SBSLPM_SWE_FORM_*), from
balanced zonal flow u = 20 cos(lat) m/s over the mountain.  Layout of the reference
(m = d + 1 = 6), both indexing rules.

Bars as in the step's own parity tests: 1e-12 for the stage-1 velocity and Laplacian,
10x the oracle's spread under 1e-13 input perturbations for the remapped state."""
import numpy as np
import pytest

from conftest import to_gpu, to_np

pytestmark = pytest.mark.gpu

OMEGA, DT, EPS = 7.292e-5, 900.0, 1e-4


@pytest.fixture(scope="module")
def c5small(sb, orc, torch_gpu):
    print("\n[case C5 form L3 d5 m6]", flush=True)
    W = sb.workloads
    o = orc.OracleFMM(3, 5, 5, 0.7, 1)
    x, y, z, a = o.spec_points()
    mtn = W.gaussian_mountain(seed=5)
    G = W.GRAVITY / W.EARTH_RADIUS ** 2
    vor, div, h = W.swe_zonal_flow(x, y, z, mtn)
    return dict(o=o, x=x, y=y, z=z, a=a, mtn=mtn, G=G, vor=vor, div=div, h=h, W=W, sb=sb, t=torch_gpu)


def _ctx(c, rule, form, precision="fp64"):
    sb = c["sb"]
    return sb.Context(sb.RunConfig(levels=3, interp_degree=5, leaf_degree=5, fmm_theta=0.7, weights="clenshaw_curtis",
                                   kernel_eps=EPS, omega=OMEGA, delta_t=DT, swe_rule=rule, topography="gaussian",
                                   gravity=c["G"], mountain=c["mtn"], swe_form=form, precision=precision))


FORMS = {"reference": (), "reg": ("reg_laplacian",), "reg+surface": ("reg_laplacian", "surface_correction")}


@pytest.mark.parametrize("rule", ["reference", "leaf"])
@pytest.mark.parametrize("form", list(FORMS))
def test_physical_step_matches_oracle(c5small, rule, form):
    c = c5small
    o, W, t = c["o"], c["W"], c["t"]
    x, y, z, a = c["x"], c["y"], c["z"], c["a"]
    bits = sum({"reg_laplacian": 1, "surface_correction": 2}[f] for f in FORMS[form])
    o.set_swe_physics(c["G"], c["mtn"], bits)
    r = 0 if rule == "reference" else 1
    ev, ed, eh, evel, elap, _ = o.swe_back_rk4_step(x.ravel(), y.ravel(), z.ravel(), a, c["vor"], c["div"], c["h"], DT,
                                                    OMEGA, EPS, 0, 2, r)
    ctx = _ctx(c, rule, FORMS[form])
    gv, gd, gh = [to_gpu(t, q) for q in (c["vor"], c["div"], c["h"])]
    vel, lap = ctx.swe_back_rk4_step(*[to_gpu(t, q) for q in (x, y, z, a)], gv, gd, gh)
    ctx.synchronize()
    errs = dict(vor=W.rel_l2(to_np(gv), ev), div=W.rel_l2(to_np(gd), ed), h=W.rel_l2(to_np(gh), eh),
                vel=W.vec_rel_l2([to_np(v) for v in vel], evel), lap=W.rel_l2(to_np(lap), elap))
    print(f"  {form} rule {rule}: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    assert errs["vel"] < 1e-12 and errs["lap"] < 1e-12
    rng = np.random.default_rng(11)
    spread = dict(vor=0.0, div=0.0, h=0.0)
    for _ in range(3):
        pv = c["vor"] * (1 + 1e-13 * rng.normal(size=x.shape))
        pd = 1e-13 * np.abs(c["vor"]).max() * rng.normal(size=x.shape)
        ph = c["h"] * (1 + 1e-13 * rng.normal(size=x.shape))
        q = o.swe_back_rk4_step(x.ravel(), y.ravel(), z.ravel(), a, pv, pd, ph, DT, OMEGA, EPS, 0, 2, r)
        for k, name in ((0, "vor"), (1, "div"), (2, "h")):
            spread[name] = max(spread[name], W.rel_l2(q[k], (ev, ed, eh)[k]))
    print("    oracle spread: " + ", ".join(f"{k} {v:.2e}" for k, v in spread.items()))
    for name in spread:
        assert errs[name] <= 10 * spread[name] + 1e-10, name


def test_physical_form_rejects_bad_parameters(c5small):
    c = c5small
    sb = c["sb"]
    with pytest.raises(sb.SbslpmError):
        sb.Context(sb.RunConfig(levels=2, interp_degree=5, topography="gaussian", mountain=(1000.0, 0.0, 0.0, 0.0)))
    with pytest.raises(sb.SbslpmError):
        sb.Context(sb.RunConfig(levels=2, interp_degree=5, gravity=-1.0))


# ---- the step's hand-written adjoint ------------------------------------------------

SCALES = (1e-6, 1e-7, 10.0)  # typical sizes of vorticity, divergence, height changes


def _perturbed_state(c):
    """The zonal flow with smooth seeded perturbations, so that no moved point sits exactly on
    a leaf edge: the balanced flow carries the equator nodes along the equator, which is a leaf
    edge, and a point on an edge is interpolated from either leaf by the tie rule, a jump that
    finite differences would see."""
    W = c["W"]
    x, y, z, a = c["x"], c["y"], c["z"], c["a"]
    return (c["vor"] + 0.3 * W.gaussian_vortices(x, y, z, a, seed=31),
            0.05 * W.gaussian_vortices(x, y, z, a, seed=32),
            c["h"] + 2e6 * W.gaussian_vortices(x, y, z, a, seed=33))


def _fd_check(c, ctx, rng, ndir=4, h=1e-4):
    """Directional derivatives of L = sum_q r_q . S_q(after one step) from the adjoint against
    central differences of the library's own step, along random directions.

    The step is piecewise smooth: a moved point is interpolated from the leaf that contains
    it, and the copies of a leaf-edge node that neighbouring leaves hold differ (each leaf
    differentiates its own values), so the interpolant jumps where a point crosses a leaf
    edge.  A difference quotient whose perturbation moves some point across an edge sees that
    jump.  Each direction is therefore differenced at h and h/8; a direction whose two
    quotients disagree (a jump inside the stencil) is reported and replaced, and the adjoint,
    the derivative of the smooth pieces, is compared on the consistent ones."""
    t, W = c["t"], c["W"]
    x, y, z, a = [to_gpu(t, q) for q in (c["x"], c["y"], c["z"], c["a"])]
    s0 = [to_gpu(t, q) for q in _perturbed_state(c)]
    r = [to_gpu(t, rng.normal(size=c["x"].shape) / sc) for sc in SCALES]
    gb = ctx.swe_back_rk4_step_adjoint(x, y, z, a, *s0, r)

    def loss(st):
        out = ctx.swe_back_rk4_step_record(x, y, z, a, *st)
        return sum(float((rq * oq).sum()) for rq, oq in zip(r, out[:3]))

    def quotient(d, step):
        return (loss([s + step * dq for s, dq in zip(s0, d)]) - loss([s - step * dq for s, dq in zip(s0, d)])) / (
            2 * step)

    worst, used, tried = 0.0, 0, 0
    while used < ndir and tried < 2 * ndir:
        tried += 1
        d = [to_gpu(t, sc * rng.normal(size=c["x"].shape)) for sc in SCALES]
        fd, fd8 = quotient(d, h), quotient(d, h / 8)
        ad = sum(float((g * dq).sum()) for g, dq in zip(gb, d))
        if abs(fd - fd8) > 1e-6 * abs(fd):
            print(f"    jump inside the stencil: fd(h) {fd:.6e} fd(h/8) {fd8:.6e} adjoint {ad:.6e}", flush=True)
            continue
        used += 1
        worst = max(worst, abs(fd - ad) / max(abs(fd), 1e-300))
        print(f"    fd {fd:.10e} adjoint {ad:.10e} rel {abs(fd - ad) / abs(fd):.2e}", flush=True)
    ctx.synchronize()
    assert used == ndir, "too many directions crossed a leaf edge"
    return worst


@pytest.mark.parametrize("rule", ["reference", "leaf"])
@pytest.mark.parametrize("form", list(FORMS))
def test_step_adjoint_vs_finite_differences(c5small, rule, form):
    c = c5small
    ctx = _ctx(c, rule, FORMS[form])
    print(f"  form {form} rule {rule}", flush=True)
    worst = _fd_check(c, ctx, np.random.default_rng(21))
    assert worst < 1e-5


def test_mixed_precision_gradient_matches_fp64(c5small):
    """FP32 pair sums (forward and transposed) on a mixed context: the gradient stays within
    FP32 rounding of the FP64 one."""
    c = c5small
    t, W = c["t"], c["W"]
    g = [to_gpu(t, q) for q in (c["x"], c["y"], c["z"], c["a"])]
    s0 = [to_gpu(t, q) for q in _perturbed_state(c)]
    rng = np.random.default_rng(23)
    r = [to_gpu(t, rng.normal(size=c["x"].shape) / sc) for sc in SCALES]
    form = FORMS["reg+surface"]
    g64 = _ctx(c, "leaf", form).swe_back_rk4_step_adjoint(*g, *s0, r)
    gmx = _ctx(c, "leaf", form, "mixed").swe_back_rk4_step_adjoint(*g, *s0, r)
    t.cuda.synchronize()
    errs = [W.rel_l2(to_np(a), to_np(b)) for a, b in zip(gmx, g64)]
    print("  mixed vs fp64 gradient (vor, div, h): " + ", ".join(f"{e:.2e}" for e in errs))
    assert max(errs) < 1e-5


def test_gradient_vs_direct_sum_reference(c5small):
    """The library's gradient (FMM, hand-written adjoint) against torch autograd through the
    FP64 direct-sum restatement of the same step (tests/torch_ref_swe.py).  They differ by the
    FMM's truncation and its interpolated proxy displacements, not by the adjoint."""
    import torch
    from torch_ref_swe import SweModel
    c = c5small
    t, W = c["t"], c["W"]
    x, y, z, a = c["x"], c["y"], c["z"], c["a"]
    st = _perturbed_state(c)
    rng = np.random.default_rng(24)
    r = [rng.normal(size=x.shape) / sc for sc in SCALES]
    ctx = _ctx(c, "leaf", FORMS["reg+surface"])
    gl = ctx.swe_back_rk4_step_adjoint(*[to_gpu(t, q) for q in (x, y, z, a)], *[to_gpu(t, q) for q in st],
                                       [to_gpu(t, q) for q in r])
    ctx.synchronize()
    M = SweModel(x, y, z, a, 3, EPS, OMEGA, DT, c["G"], c["mtn"], 0)
    s_t = [torch.tensor(np.ravel(q), requires_grad=True) for q in st]
    out = M.step(*s_t)
    loss = sum((torch.tensor(np.ravel(rq)) * oq).sum() for rq, oq in zip(r, out))
    gt = torch.autograd.grad(loss, s_t)
    errs = [W.rel_l2(to_np(a_).ravel(), b.numpy()) for a_, b in zip(gl, gt)]
    print("  library vs direct-sum autograd gradient (vor, div, h): " + ", ".join(f"{e:.2e}" for e in errs))
    assert max(errs) < 1e-3
