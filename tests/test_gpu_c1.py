"""The plain-method workload itself on the GPU: one BVE RK4 step on the
6 x 64^2 = 24,576-particle cubed sphere (L=4, 8x8 particle grid per leaf,
proxy degree 8, theta 0.7, equal-area weights 4 pi / N) from the Rossby-Haurwitz
wave 4 (omega0 = K = 7.848e-6), dt = 3600 s, Omega = 7.292e-5, eps = 1e-4.

Checked against the FP64 oracle (step, velocity), the FP64 direct sum on all
24,576 particles, and the closed form of the wave: its velocity is the l=1
plus l=5 stream-function field, and the wave pattern travels east at
nu = (n(3+n) w - 2 Omega) / ((1+n)(2+n)), n = 4, w = omega0 + K (Haurwitz 1940),
so one textbook step must carry the field that far.  The reference combine
(bve_time_step.cpp:288-303) displaces by +2/3 dt v to first order instead of
-dt v, so its step moves the pattern by -2/3 of that angle."""
import numpy as np
import pytest

from conftest import to_gpu, to_np

pytestmark = pytest.mark.gpu

EPS, DT, OMEGA, K = 1e-4, 3600.0, 7.292e-5, 7.848e-6
NU = (4 * 7 * (2 * K) - 2 * OMEGA) / (5 * 6)


def _rotated_rh4(W, x, y, z, angle):
    """RH4 vorticity with its pattern moved east by `angle` (longitude - angle)."""
    c, s = np.cos(angle), np.sin(angle)
    return W.rh4_vorticity(c * x + s * y, -s * x + c * y, z)


@pytest.fixture(scope="module")
def c1(sb, orc, torch_gpu):
    print("\n[case C1 6x64^2 L4 d8 m8 equal-area RH4]", flush=True)
    o = orc.OracleFMM(4, 8, 7, 0.7, 2)
    x, y, z, a = o.spec_points()
    assert x.size == 24576
    np.testing.assert_allclose(a, 4 * np.pi / 24576, rtol=1e-14)
    vor = sb.workloads.rh4_vorticity(x, y, z)
    return o, x, y, z, a, vor


def _ctx(sb, combine="reference", precision="fp64"):
    return sb.Context(sb.RunConfig(levels=4, interp_degree=8, leaf_degree=7, fmm_theta=0.7, weights="equal_area",
                                   kernel_eps=EPS, delta_t=DT, omega=OMEGA, rk4_combine=combine,
                                   precision=precision))


def test_c1_velocity_vs_oracle_direct_sum_and_closed_form(c1, sb, orc, torch_gpu):
    t = torch_gpu
    W = sb.workloads
    o, x, y, z, a, vor = c1
    ctx = _ctx(sb)
    gx, gy, gz, ga, gv = [to_gpu(t, q) for q in (x, y, z, a, vor)]
    vel = [to_np(v) for v in ctx.bve_velocity(ga, gv)]
    flat = [q.reshape(-1) for q in (gx, gy, gz, ga, gv)]
    ds = [to_np(d).reshape(x.shape) for d in ctx.direct_sum_bve_vel(*flat)]
    mx = _ctx(sb, precision="mixed")
    velm = [to_np(v) for v in mx.bve_velocity(ga, gv)]
    ctx.synchronize()
    mx.synchronize()
    ref = o.bve_velocity(a, vor, EPS)
    ods = [d.reshape(x.shape) for d in orc.direct_sum_bve_vel(*[q.ravel() for q in (x, y, z, a, vor)], EPS)]
    e_orc = W.vec_rel_l2(vel, ref)
    e_ods = W.vec_rel_l2(ds, ods)
    e_ds, e_ds_inf = W.vec_rel_l2(vel, ds), W.vec_rel_linf(vel, ds)
    e_mx = W.vec_rel_l2(velm, ds)
    e_cf = W.vec_rel_l2(vel, W.rh4_velocity(x, y, z))
    e_cf_ds = W.vec_rel_l2(ds, W.rh4_velocity(x, y, z))
    print(f"  FMM vs oracle {e_orc:.2e}; direct sum vs oracle direct sum {e_ods:.2e}; FMM vs direct sum "
          f"rel-L2 {e_ds:.2e} Linf {e_ds_inf:.2e}; mixed vs direct sum {e_mx:.2e}; "
          f"FMM vs closed form {e_cf:.2e} (direct sum {e_cf_ds:.2e})")
    assert e_orc < 1e-12
    assert e_ods < 1e-12
    # the FMM's truncation at d = 8, theta = 0.7 (9.8e-8 measured by the oracle)
    assert e_ds < 3e-7 and e_ds_inf < 3e-6
    assert e_mx < 1e-6
    # equal weights on the per-leaf Chebyshev grid are a first-order quadrature: the
    # direct sum itself sits 7.5e-2 from the closed form, and the FMM with it
    assert e_cf < 0.1 and abs(e_cf - e_cf_ds) < 1e-6


@pytest.mark.parametrize("combine,shift", [("reference", -2.0 / 3.0), ("textbook", 1.0)])
def test_c1_step_vs_oracle_and_wave_speed(c1, sb, torch_gpu, combine, shift):
    t = torch_gpu
    W = sb.workloads
    o, x, y, z, a, vor = c1
    ctx = _ctx(sb, combine)
    gx, gy, gz, ga, gv = [to_gpu(t, q) for q in (x, y, z, a, vor)]
    gvel = ctx.zeros_part(3)
    ctx.bve_back_rk4_step(gx, gy, gz, ga, gv, gvel)
    ctx.synchronize()
    nv, vel, _ = o.bve_back_rk4_step(x, y, z, a, vor, DT, OMEGA, EPS, 0 if combine == "reference" else 1)
    out = to_np(gv)
    e_vel = W.vec_rel_l2([to_np(v) for v in gvel], vel)
    e_vor = W.rel_l2(out, nv)
    moved = _rotated_rh4(W, x, y, z, shift * NU * DT)
    e_wave = W.rel_l2(out, moved)
    e_still = W.rel_l2(out, vor)
    e_other = W.rel_l2(out, _rotated_rh4(W, x, y, z, -shift * NU * DT))
    print(f"  {combine}: stage-0 velocity vs oracle {e_vel:.2e}, vorticity vs oracle {e_vor:.2e}; vs the wave "
          f"moved by {shift:+.3f} nu dt {e_wave:.2e} (unmoved {e_still:.2e}, moved the other way {e_other:.2e})")
    assert e_vel < 1e-12
    assert e_vor < 1e-11
    if combine == "textbook":
        # 1.4e-2 at these weights (1.3e-3 with Clenshaw-Curtis weights), against 0.12
        # between the initial field and the moved wave
        assert e_wave < 0.03 and e_still > 0.1
    else:
        assert e_wave < 0.05 and e_other > 0.15
