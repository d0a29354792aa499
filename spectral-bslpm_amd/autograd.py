"""torch.autograd.Function wrappers whose backward passes are the hand-written
adjoint kernels of the library (sbslpm_bve_velocity_adjoint,
sbslpm_bve_departure_to_target_adjoint).

* ``bve_velocity(ctx, area, vor)`` -- the Biot-Savart velocity of the
  vorticity with undisplaced sources (the first RK4 stage of
  bve_back_rk4_step, bve_time_step.cpp:321-344); differentiable in ``vor``.
* ``remap(ctx, z, dep, vor)`` -- bve_departure_to_target
  (bve_time_step.cpp:38-66) at given departure points; differentiable in
  ``vor`` (the departure points are constants of this function).

Both are linear in the differentiated argument, so their Jacobian-vector and
vector-Jacobian products are exact up to the FMM truncation (velocity) or
rounding (remap).
"""
import torch


class _BVEVelocity(torch.autograd.Function):
    @staticmethod
    def forward(fctx, sbctx, area, vor):
        vel = sbctx.bve_velocity(area, vor)
        fctx.sbctx = sbctx
        fctx.save_for_backward(area)
        return tuple(vel)

    @staticmethod
    def backward(fctx, gx, gy, gz):
        (area,) = fctx.saved_tensors
        g = [torch.zeros_like(area) if v is None else v.contiguous() for v in (gx, gy, gz)]
        vor_bar = fctx.sbctx.bve_velocity_adjoint(area, g)
        return None, None, vor_bar


class _Remap(torch.autograd.Function):
    @staticmethod
    def forward(fctx, sbctx, z, dep_x, dep_y, dep_z, vor):
        new_vor, _ = sbctx.bve_departure_to_target(z, [dep_x, dep_y, dep_z], vor)
        fctx.sbctx = sbctx
        fctx.save_for_backward(dep_x, dep_y, dep_z)
        return new_vor

    @staticmethod
    def backward(fctx, g):
        dep = fctx.saved_tensors
        vor_bar = fctx.sbctx.bve_departure_to_target_adjoint(list(dep), g.contiguous())
        return None, None, None, None, None, vor_bar


def bve_velocity(sbctx, area, vor):
    """(vel_x, vel_y, vel_z) = K(vor); backward: vor_bar = K^T (vel_bar)."""
    return _BVEVelocity.apply(sbctx, area, vor)


def remap(sbctx, z, dep, vor):
    """new_vor = B(dep) vor + 2 Omega (dep_z - z); backward: vor_bar = B(dep)^T new_vor_bar."""
    return _Remap.apply(sbctx, z, dep[0], dep[1], dep[2], vor)
