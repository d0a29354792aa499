"""torch.autograd.Function wrappers whose backward passes are the hand-written
adjoint kernels of the library (sbslpm_bve_velocity_adjoint,
sbslpm_bve_departure_to_target_adjoint, sbslpm_bve_back_rk4_step_adjoint).

* ``bve_velocity(ctx, area, vor)`` -- the Biot-Savart velocity of the
  vorticity with undisplaced sources (the first RK4 stage of
  bve_back_rk4_step, bve_time_step.cpp:321-344); differentiable in ``vor``.
* ``remap(ctx, z, dep, vor)`` -- bve_departure_to_target
  (bve_time_step.cpp:38-66) at given departure points; differentiable in
  ``vor`` (the departure points are constants of this function).

* ``bve_steps(ctx, x, y, z, area, vor0, nsteps, every)`` -- ``nsteps``
  bve_back_rk4_step calls (bve_time_step.cpp:311-419) from ``vor0``;
  differentiable in ``vor0`` through the full nonlinear step (trajectories
  included), checkpointed every ``every`` steps.
* ``swe_steps(ctx, x, y, z, area, vor0, div0, h0, nsteps, every)`` -- the
  same for swe_back_rk4_step (swe_time_step.cpp:606-811), differentiable in
  the three initial fields (sbslpm_swe_back_rk4_step_adjoint_stages).

The first two are linear in the differentiated argument; their backward
passes are the exact transposes of the library's FP64 operators.
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


class _BVESteps(torch.autograd.Function):
    """Forward: nsteps steps, keeping the state at every `every`-th step (the
    checkpoints) and, for the last segment, each step's stage velocities.
    Backward: segment by segment from the last, the segment's steps are rerun
    from its checkpoint with their stage velocities recorded
    (sbslpm_bve_back_rk4_step_record), then the reverse passes run in reverse
    order on the recorded stages (sbslpm_bve_back_rk4_step_adjoint_stages).
    Memory: nsteps/every states + every x (state + 12 stage fields); work: the
    forward run, one recorded rerun of every segment but the last, and the
    reverse passes."""

    @staticmethod
    def forward(fctx, sbctx, x, y, z, area, vor0, nsteps, every):
        nseg = (nsteps + every - 1) // every
        last = (nseg - 1) * every
        v = vor0.detach().clone()
        checkpoints, tail = [v.clone()], []
        vel = None
        for n in range(nsteps):
            if n >= last:
                nv, stages = sbctx.bve_back_rk4_step_record(x, y, z, area, v)
                tail.append((v, stages))
                v = nv
            else:
                vel = sbctx.zeros_part(3) if vel is None else vel
                sbctx.bve_back_rk4_step(x, y, z, area, v, vel)
            if (n + 1) % every == 0 and n + 1 < nsteps:
                checkpoints.append(v.clone())
        fctx.sbctx, fctx.nsteps, fctx.every = sbctx, nsteps, every
        fctx.geom = (x, y, z, area)
        fctx.checkpoints, fctx.tail = checkpoints, tail
        return v

    @staticmethod
    def backward(fctx, g):
        sbctx, nsteps, every = fctx.sbctx, fctx.nsteps, fctx.every
        x, y, z, area = fctx.geom
        gbar = g.contiguous()
        for seg in reversed(range(len(fctx.checkpoints))):
            start, end = seg * every, min((seg + 1) * every, nsteps)
            if seg == len(fctx.checkpoints) - 1:
                rec = fctx.tail
            else:
                rec, v = [], fctx.checkpoints[seg]
                for _ in range(start, end):
                    nv, stages = sbctx.bve_back_rk4_step_record(x, y, z, area, v)
                    rec.append((v, stages))
                    v = nv
            for state, stages in reversed(rec):
                gbar = sbctx.bve_back_rk4_step_adjoint_stages(x, y, z, area, state, stages, gbar)
        fctx.tail = None
        return None, None, None, None, None, gbar, None, None


def bve_steps(sbctx, x, y, z, area, vor0, nsteps, every=1):
    """Vorticity after nsteps backward RK4 steps; backward: the hand-written step
    adjoints with checkpoints every `every` steps."""
    if nsteps < 1 or every < 1:
        raise ValueError("nsteps and every must be positive")
    return _BVESteps.apply(sbctx, x, y, z, area, vor0, nsteps, every)


class _SWESteps(torch.autograd.Function):
    """nsteps shallow-water steps (swe_back_rk4_step, swe_time_step.cpp:606-811) from the state
    (vor, div, h), checkpointed like _BVESteps: the forward keeps the state every `every` steps
    and, for the last segment, each step's record (sbslpm_swe_back_rk4_step_record: the stage
    velocities and tendencies); the backward reruns each earlier segment from its checkpoint
    with records and runs the reverse passes (sbslpm_swe_back_rk4_step_adjoint_stages) in
    reverse order."""

    @staticmethod
    def forward(fctx, sbctx, x, y, z, area, vor0, div0, h0, nsteps, every):
        nseg = (nsteps + every - 1) // every
        last = (nseg - 1) * every
        st = tuple(t.detach().clone() for t in (vor0, div0, h0))
        checkpoints, tail = [tuple(t.clone() for t in st)], []
        vel = lap = None
        for n in range(nsteps):
            if n >= last:
                nv, nd, nh, rec = sbctx.swe_back_rk4_step_record(x, y, z, area, *st)
                tail.append((st, rec))
                st = (nv, nd, nh)
            else:
                if vel is None:
                    vel, lap = sbctx.zeros_part(3), sbctx.zeros_part()
                sbctx.swe_back_rk4_step(x, y, z, area, *st, vel=vel, height_lap=lap)
            if (n + 1) % every == 0 and n + 1 < nsteps:
                checkpoints.append(tuple(t.clone() for t in st))
        fctx.sbctx, fctx.nsteps, fctx.every = sbctx, nsteps, every
        fctx.geom = (x, y, z, area)
        fctx.checkpoints, fctx.tail = checkpoints, tail
        return st

    @staticmethod
    def backward(fctx, gv, gd, gh):
        sbctx, nsteps, every = fctx.sbctx, fctx.nsteps, fctx.every
        x, y, z, area = fctx.geom
        ref = fctx.checkpoints[0][0]
        gbar = [torch.zeros_like(ref) if g is None else g.contiguous() for g in (gv, gd, gh)]
        for seg in reversed(range(len(fctx.checkpoints))):
            start, end = seg * every, min((seg + 1) * every, nsteps)
            if seg == len(fctx.checkpoints) - 1:
                rec = fctx.tail
            else:
                rec, st = [], fctx.checkpoints[seg]
                for _ in range(start, end):
                    nv, nd, nh, stages = sbctx.swe_back_rk4_step_record(x, y, z, area, *st)
                    rec.append((st, stages))
                    st = (nv, nd, nh)
            for state, stages in reversed(rec):
                gbar = list(sbctx.swe_back_rk4_step_adjoint(x, y, z, area, *state, gbar, stages=stages))
        fctx.tail = None
        return None, None, None, None, None, gbar[0], gbar[1], gbar[2], None, None


def swe_steps(sbctx, x, y, z, area, vor0, div0, h0, nsteps, every=1):
    """(vor, div, h) after nsteps shallow-water steps; backward: the hand-written step adjoints
    with checkpoints every `every` steps.  The context needs tracer_count == 0."""
    if nsteps < 1 or every < 1:
        raise ValueError("nsteps and every must be positive")
    return _SWESteps.apply(sbctx, x, y, z, area, vor0, div0, h0, nsteps, every)


def swe_4dvar(sbctx, x, y, z, area, state0, nsteps, every, obs_xyz, obs_vor, obs_h, sigma_vor=1.0, sigma_h=1.0):
    """Loss and gradient of a shallow-water 4D-Var problem: the initial state (vor, div, h) is
    advanced nsteps steps and its vorticity and height are compared with observations at the
    points obs_xyz (three 1-D tensors) at the final time,
      L = 1/2 sum ((H(vor_N) - obs_vor) / sigma_vor)^2 + 1/2 sum ((H(h_N) - obs_h) / sigma_h)^2,
    H the BLI at the points.  The gradient comes from the hand-written step adjoints
    (checkpoints every `every` steps) and the transposed observation operator.  Returns
    (L, (dL/dvor0, dL/ddiv0, dL/dh0))."""
    w = [t.detach().clone().requires_grad_(True) for t in state0]
    vn, dn, hn = swe_steps(sbctx, x, y, z, area, w[0], w[1], w[2], nsteps, every)
    hv = interp_points(sbctx, obs_xyz[0], obs_xyz[1], obs_xyz[2], vn)
    hh = interp_points(sbctx, obs_xyz[0], obs_xyz[1], obs_xyz[2], hn)
    loss = 0.5 * (((hv - obs_vor) / sigma_vor) ** 2).sum() + 0.5 * (((hh - obs_h) / sigma_h) ** 2).sum()
    loss.backward()
    return loss.detach(), tuple(t.grad.detach() for t in w)


class _InterpPoints(torch.autograd.Function):
    @staticmethod
    def forward(fctx, sbctx, x, y, z, field):
        out = sbctx.interp_points(x, y, z, [field.contiguous()])[0]
        fctx.sbctx = sbctx
        fctx.save_for_backward(x, y, z)
        return out

    @staticmethod
    def backward(fctx, g):
        x, y, z = fctx.saved_tensors
        fb = fctx.sbctx.interp_points_adjoint(x, y, z, [g.contiguous()])[0]
        return None, None, None, None, fb


def interp_points(sbctx, x, y, z, field):
    """BLI of a leaf field [Pl, m2] at points (1-D x, y, z): latlon_interp_host's functor
    (interp_funcs.hpp:22-60); backward: the transposed interpolation."""
    return _InterpPoints.apply(sbctx, x, y, z, field)


def bve_4dvar(sbctx, x, y, z, area, vor0, nsteps, every, obs_xyz, obs_values):
    """Loss and gradient of a vorticity 4D-Var problem on the BVE: the initial field vor0 is
    advanced nsteps steps and compared with observations at the points obs_xyz (three 1-D
    tensors) at the final time, L = 1/2 sum (H(vor_N) - obs)^2, H the BLI at the points.
    The gradient comes from the hand-written step adjoints (checkpoints every `every` steps)
    and the transposed observation operator.  Returns (L, dL/dvor0)."""
    w = vor0.detach().clone().requires_grad_(True)
    vn = bve_steps(sbctx, x, y, z, area, w, nsteps, every)
    hx = interp_points(sbctx, obs_xyz[0], obs_xyz[1], obs_xyz[2], vn)
    loss = 0.5 * ((hx - obs_values) ** 2).sum()
    loss.backward()
    return loss.detach(), w.grad.detach()
