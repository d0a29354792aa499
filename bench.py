"""Benchmark: BVE RK4 steps/s at 6 x 512^2 = 1,572,864 particles on one MI355X,
mixed precision vs plain FP64, with the accuracy of mixed vs FP64.

Workload: uniform cubed-sphere quadtree with 7 levels (24,576 leaves), an 8 x 8
Chebyshev particle grid per leaf, proxy degree 8, MAC theta 0.7,
kernel_eps 1e-4, dt 3600 s, Omega 7.292e-5; initial vorticity = 12 seeded
Gaussian vortices (synthetic).  One step = bve_back_rk4_step: four FMM
velocity evaluations with displaced sources, backward RK4 departure points and
the semi-Lagrangian remap, captured once in a hipGraph and replayed.

Prints one JSON line (see README of the driver contract in the task):
value = mixed-precision steps/s; the plain FP64 rate, the speed-up, the
rel-L2 / Linf velocity and vorticity errors of mixed vs FP64, the FP64 FMM
error vs the FP64 direct sum on 4,096 targets, the pair-kernel roofline
measured with HIP events on the context stream, and the CPU baseline (the
oracle restatement of the reference algorithm timed on a bounded sample on
this host's cores).
"""
import argparse
import ctypes
import importlib
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "BVE RK4 steps/sec at 1.57M particles, mixed vs FP64 plain, rel-L2 err ≤1e-6"
PEAK_FP64 = 78.6   # TFLOP/s, MI355X dense FP64 (vector and matrix)
PEAK_FP32 = 157.3  # TFLOP/s, MI355X FP32
FLOP_BS = 27       # flop per Biot-Savart pair (difference form; div/rcp/sqrt counted as 1)
FLOP_LOG = 13      # flop per log (Poisson) pair, same convention
FLOP_SWE = 40      # shallow-water pair (variant 1): the Biot-Savart 27 + q_d (1) + y - x(x.y) (6) + accumulation (6)
FLOP_LAP = 30      # regularised-Laplacian pair: dot (5), t^2 (1), erf/exp terms and the polynomial (20), acc (4)

CONFIGS = {
    # levels, proxy degree, particle-grid degree, theta
    "c3": dict(levels=7, interp_degree=8, leaf_degree=7, fmm_theta=0.7),
    "c1": dict(levels=4, interp_degree=8, leaf_degree=7, fmm_theta=0.7),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pair_count(il, P, Pl, n, m2):
    t, s, typ = il[:, 0], il[:, 1], il[:, 2]
    nt = np.where(typ <= 1, m2, n).astype(np.int64)
    ns = np.where((typ == 0) | (typ == 2), m2, n).astype(np.int64)
    return int(np.sum(nt * ns))


def time_steps(ctx, torch, stream, x, y, z, a, vor0, steps, warmup):
    vor = vor0.clone()
    vel = ctx.zeros_part(3)
    with torch.cuda.stream(stream):
        for _ in range(warmup):
            ctx.bve_back_rk4_step(x, y, z, a, vor, vel)
    ctx.synchronize()
    torch.cuda.synchronize()
    vor.copy_(vor0)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        e0.record(stream)
        for _ in range(steps):
            ctx.bve_back_rk4_step(x, y, z, a, vor, vel)
        e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ctx.synchronize()
    return wall, e0.elapsed_time(e1) / 1e3, vor, vel


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cfg, x, y, z, a, vor, budget_s):
    """Oracle (FP64 restatement of the reference loops) on this host's cores: one full
    bve_back_rk4_step, measured, when a probe of its interaction phase puts it within
    4 x budget_s; otherwise composed from a bounded sample (one upward pass, one downward pass,
    one displacement, one departure + remap at full size, and the interaction phase on every
    `stride`-th target panel; one step = 4 x (upward + interactions + downward) + 3
    displacements + departure + remap)."""
    from oracle import oracle as O
    threads = len(os.sched_getaffinity(0))
    O.set_threads(threads)
    o = O.OracleFMM(cfg["levels"], cfg["interp_degree"], cfg["leaf_degree"], cfg["fmm_theta"], 1)
    qp, qx = o.upward_pass(a, vor)
    t0 = time.perf_counter()
    o.bve_vel_interactions_sample(qp, qx, 1e-4, 256)
    t_probe = time.perf_counter() - t0
    est = 4 * t_probe / max(np.mean(np.arange(o.P) % 256 == 0), 1e-9)
    if est < 4 * budget_s:
        t0 = time.perf_counter()
        o.bve_back_rk4_step(x, y, z, a, vor.copy(), 3600.0, 7.292e-5, 1e-4, 0)
        sec = time.perf_counter() - t0
        return dict(value=round(1.0 / sec, 6), unit="steps/s", cores=threads, cpu_model=cpu_model(), kind="port",
                    sample=f"one full oracle bve_back_rk4_step at this size, measured ({sec:.1f} s)",
                    seconds_per_step=round(sec, 2))
    t0 = time.perf_counter()
    qp, qx = o.upward_pass(a, vor)
    t_up = time.perf_counter() - t0
    il = o.interactions()
    # pick the sample stride from a small probe so the sample takes ~budget/2
    stride = 256
    t0 = time.perf_counter()
    o.bve_vel_interactions_sample(qp, qx, 1e-4, stride)
    t_probe = time.perf_counter() - t0
    frac_probe = np.mean(np.arange(o.P) % stride == 0)
    est_full = t_probe / max(frac_probe, 1e-9)
    want = max(0.5 * budget_s, t_probe)
    stride2 = int(max(1, min(256, round(est_full / want))))
    t0 = time.perf_counter()
    pp, px = o.bve_vel_interactions_sample(qp, qx, 1e-4, stride2)
    t_int_sample = time.perf_counter() - t0
    sampled = np.isin(il[:, 0], np.arange(0, o.P, stride2))
    pairs_all = pair_count(il, o.P, o.Pl, o.n, o.m2)
    pairs_sample = pair_count(il[sampled], o.P, o.Pl, o.n, o.m2)
    t_int = t_int_sample * pairs_all / max(pairs_sample, 1)
    t0 = time.perf_counter()
    vel = o.downward_pass(px, pp)
    t_down = time.perf_counter() - t0
    t0 = time.perf_counter()
    vfake = [1e-6 * np.ones_like(x), 1e-6 * y, -1e-6 * x]
    nv, dpart, dprox = o.bve_displacement(x, y, z, vfake, vor, 1800.0, 7.292e-5)
    t_disp = time.perf_counter() - t0
    t0 = time.perf_counter()
    dep = o.bve_find_departure_points(x, y, z, [vfake] * 4, 3600.0, 0)
    t_dep = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.bve_departure_to_target(z, dep, vor, 7.292e-5)
    t_remap = time.perf_counter() - t0
    step = 4 * (t_up + t_int + t_down) + 3 * t_disp + t_dep + t_remap
    measured = t_up + t_probe + t_int_sample + t_down + t_disp + t_dep + t_remap
    sample = (f"full-size upward/downward pass, displacement, departure+remap; interactions of every "
              f"{stride2}-th target panel ({pairs_sample:.3e} of {pairs_all:.3e} pairs, scaled by pair count); "
              f"step = 4x(up+interactions+down)+3x displacement+departure+remap; {measured:.1f} s measured")
    return dict(value=1.0 / step, unit="steps/s", cores=threads, cpu_model=cpu_model(), kind="port", sample=sample,
                seconds_per_step_extrapolated=step)


ENSEMBLE = [  # (particles per face side, levels, particle-grid side m); proxy degree 8, theta 0.7
    (96, 5, 6), (128, 5, 8), (192, 6, 6)]


def bench_ensemble(args, torch, sb):
    """Ensemble of 48 BVE members, 16 at each of 6x96^2, 6x128^2 and 6x192^2:
    one context (tree, interaction list, tiles) per resolution, each step of
    its 16 members in shared kernel launches, captured in a hipGraph.  Member
    fields: the 12-vortex field plus seeded spherical-harmonic noise (l <= 24,
    1e-3 of the field RMS).  Reports member-steps/s over all 48 members and
    the speed-up over stepping the members one at a time."""
    W = sb.workloads
    nmem = 16
    stream = torch.cuda.Stream()
    dev = lambda v: torch.from_numpy(np.ascontiguousarray(v)).cuda()
    groups, mshapes = [], {}
    for side, levels, m in ENSEMBLE:
        kw = dict(levels=levels, interp_degree=8, leaf_degree=m - 1, fmm_theta=0.7, weights="clenshaw_curtis",
                  kernel_eps=1e-4, delta_t=3600.0, omega=7.292e-5, rk4_combine="reference",
                  use_graph=not args.no_graph, precision="mixed", multi_rhs=not args.no_multi_rhs)
        with torch.cuda.stream(stream):
            ctx = sb.Context(sb.RunConfig(**kw))
        x, y, z, a = ctx.spec_points()
        base = W.gaussian_vortices(x, y, z, a, count=12, seed=0)
        rms = float(np.sqrt(np.mean(base ** 2)))
        mem = []
        for j in range(nmem):
            noise = W.random_sh_forcing(x, y, z, lmax=24, seed=100 + j)[0]
            mem.append(base + 1e-3 * rms * noise / float(np.sqrt(np.mean(noise ** 2))))
        vors = dev(np.stack(mem))
        if not args.no_tune:
            with torch.cuda.stream(stream):
                _, shape = ctx.autotune(dev(a), vors[0].contiguous())
                if not args.no_multi_rhs:
                    _, mshape = ctx.autotune_ensemble(dev(a), vors)
            torch.cuda.synchronize()
            log(f"6x{side}^2: pair-kernel launch shape {shape}")
            if not args.no_multi_rhs:
                log(f"6x{side}^2: multi-RHS launch shape (waves, xcd, chunks per wave) {mshape}")
                mshapes[f"6x{side}^2"] = list(mshape)
        g = dict(side=side, ctx=ctx, kw=kw, xyz=[dev(v) for v in (x, y, z, a)], vors0=vors,
                 vors=vors.clone(), vel=[torch.zeros_like(vors) for _ in range(3)])
        groups.append(g)
    torch.cuda.synchronize()

    def run(steps):
        with torch.cuda.stream(stream):
            for _ in range(steps):
                for g in groups:
                    g["ctx"].bve_back_rk4_step_ensemble(*g["xyz"], g["vors"], g["vel"])
    run(args.warmup)
    torch.cuda.synchronize()
    for g in groups:
        g["vors"].copy_(g["vors0"])
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(stream):
        e0.record(stream)
    run(args.steps)
    with torch.cuda.stream(stream):
        e1.record(stream)
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) / 1e3
    total_members = nmem * len(groups)
    value = total_members * args.steps / sec

    # the same members stepped one at a time (same contexts, single-member entry point)
    one = []
    for g in groups:
        v1 = g["vors0"][0].clone()
        vel1 = g["ctx"].zeros_part(3)
        with torch.cuda.stream(stream):
            g["ctx"].bve_back_rk4_step(*g["xyz"], v1, vel1)
        torch.cuda.synchronize()
        f0, f1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(stream):
            f0.record(stream)
            for _ in range(args.steps):
                g["ctx"].bve_back_rk4_step(*g["xyz"], v1, vel1)
            f1.record(stream)
        torch.cuda.synchronize()
        one.append(f0.elapsed_time(f1) / 1e3 / args.steps)
    seq = nmem * sum(one)  # seconds per step of all 48 members, one member at a time
    # the three resolutions as concurrent branches of one captured graph per step
    # (spectral-bslpm_amd/ensemble.py), on eager contexts with the same tuned shapes
    one_graph = None
    if not args.no_graph:
        UE = importlib.import_module("spectral-bslpm_amd.ensemble").UnevenEnsemble
        eg = []
        for g in groups:
            ce = sb.Context(sb.RunConfig(**dict(g["kw"], use_graph=False)))
            ce.set_pair_shape(*g["ctx"].pair_shape())
            if not args.no_multi_rhs and not args.no_tune:
                ce.autotune_ensemble(g["xyz"][3], g["vors0"])
            eg.append((ce, g["xyz"], g["vors0"].clone()))
        ue = UE(eg)
        for _ in range(max(args.warmup, 2)):
            ue.step()
        torch.cuda.synchronize()
        h0, h1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        h0.record()
        for _ in range(args.steps):
            ue.step()
        h1.record()
        torch.cuda.synchronize()
        sec1 = h0.elapsed_time(h1) / 1e3
        one_graph = dict(member_steps_per_s=round(total_members * args.steps / sec1, 3),
                         ms_per_step=round(1e3 * sec1 / args.steps, 3),
                         gain_vs_three_graphs=round(sec / sec1, 3))
        log(f"c4 one graph over the three resolutions: {1e3 * sec1 / args.steps:.3f} ms/step "
            f"(three graphs back to back: {1e3 * sec / args.steps:.3f})")
        del ue, eg
    # accuracy: member 0 of each resolution after the timed steps against an FP64 single-member run
    acc = {}
    for g in groups:
        kw = dict(g["kw"], precision="fp64", use_graph=False)
        c64 = sb.Context(sb.RunConfig(**kw))
        v64 = g["vors0"][0].clone()
        vel64 = c64.zeros_part(3)
        for _ in range(args.steps):
            c64.bve_back_rk4_step(*g["xyz"], v64, vel64)
        c64.synchronize()
        acc[f"6x{g['side']}^2"] = float(f"{W.rel_l2(g['vors'][0].cpu().numpy(), v64.cpu().numpy()):.3e}")
        del c64
    # bf16 M2L (opt-in): the same step with bf16 kernel-product operands for proxy-to-proxy sets;
    # member 0 against the FP64 run and the members' own spread (each member's distance from
    # member 0 after the same steps)
    m2l = {}
    if not args.no_graph:
        for g in groups:
            kwb = dict(g["kw"], bf16_m2l=True)
            with torch.cuda.stream(stream):
                cb_ = sb.Context(sb.RunConfig(**kwb))
            cb_.set_pair_shape(*g["ctx"].pair_shape())
            if not args.no_multi_rhs and not args.no_tune:
                cb_.autotune_ensemble(g["xyz"][3], g["vors0"])
            vb = g["vors0"].clone()
            velb = [torch.zeros_like(vb) for _ in range(3)]
            with torch.cuda.stream(stream):
                cb_.bve_back_rk4_step_ensemble(*g["xyz"], vb, velb)  # capture
            torch.cuda.synchronize()
            vb.copy_(g["vors0"])
            torch.cuda.synchronize()
            b0, b1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            with torch.cuda.stream(stream):
                b0.record(stream)
                for _ in range(args.steps):
                    cb_.bve_back_rk4_step_ensemble(*g["xyz"], vb, velb)
                b1.record(stream)
            torch.cuda.synchronize()
            kw64 = dict(g["kw"], precision="fp64", use_graph=False)
            c64 = sb.Context(sb.RunConfig(**kw64))
            v64 = g["vors0"][0].clone()
            vel64 = c64.zeros_part(3)
            for _ in range(args.steps):
                c64.bve_back_rk4_step(*g["xyz"], v64, vel64)
            c64.synchronize()
            vh = vb.cpu().numpy()
            spread = float(np.median([W.rel_l2(vh[j], vh[0]) for j in range(1, nmem)]))
            m2l[f"6x{g['side']}^2"] = {"ms_per_step": round(b0.elapsed_time(b1) / args.steps, 3),
                                       "member0_vor_rel_l2_vs_fp64": float(f"{W.rel_l2(vh[0], v64.cpu().numpy()):.3e}"),
                                       "member_spread_rel_l2": float(f"{spread:.3e}")}
            del cb_, c64
        log("bf16 M2L", m2l)
    # pair-kernel rooflines per resolution, from HIP events on an eager context (graph replays
    # carry no event timers): algorithmic flop = pairs per evaluation x members x 27
    roofs = {}
    for g in groups:
        ce = sb.Context(sb.RunConfig(**dict(g["kw"], use_graph=False)))
        ce.set_pair_shape(*g["ctx"].pair_shape())
        if not args.no_multi_rhs and not args.no_tune:
            ce.autotune_ensemble(g["xyz"][3], g["vors0"])
        s_ = ce.sizes
        pairs = pair_count(ce.interaction_list(), s_.P, s_.Pl, s_.n, s_.m2)
        v, vel = g["vors0"].clone(), [torch.zeros_like(g["vors0"]) for _ in range(3)]
        ce.enable_timing(True)
        ce.bve_back_rk4_step_ensemble(*g["xyz"], v, vel)
        ce.synchronize()
        r = {}
        for kern, name in (("bve_mrhs", "first_stage_multi_rhs"), ("bve_mixed", "per_member_kernel")):
            ms, nl = ce.kernel_time(kern)
            if nl == 0:
                continue
            achieved = pairs * nmem * FLOP_BS / (ms / nl / 1e3) / 1e12
            r[name] = dict(bound="mfma", achieved=round(achieved, 3), peak=PEAK_FP32, unit="TFLOP/s",
                           frac=round(achieved / PEAK_FP32, 4), avg_launch_ms=round(ms / nl, 4), launches=nl,
                           members_per_launch=nmem, pairs_per_member=pairs, flop_per_pair=FLOP_BS)
            if kern == "bve_mrhs":
                # executed work: per pair and member the 16x16x4 products against four charge
                # columns (8 flop); the kernel matrix (16x16x4 product, 8 flop, and the
                # reciprocal) once per pair for the 4 * cb members of a wave
                cb = mshapes.get(f"6x{g['side']}^2", [0, 0, 2])[2]
                exe = 8.0 + 9.0 / (4 * cb)
                ach = pairs * nmem * exe / (ms / nl / 1e3) / 1e12
                r[name].update(executed_flop_per_pair_member=exe, achieved_executed=round(ach, 3),
                               frac_executed=round(ach / PEAK_FP32, 4))
        roofs[f"6x{g['side']}^2"] = r
        del ce
    log("c4 rooflines", roofs)
    line = {"metric": "BVE ensemble member-steps/sec, 48 members at 6x96^2, 6x128^2, 6x192^2",
            "value": round(value, 3), "unit": "member-steps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(1e3 * sec / args.steps, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32 pair kernels, f64 accumulation/trajectories",
            "data": "synthetic (12 seeded Gaussian vortices + seeded SH noise per member)",
            "config": {"workload": "BVE ensemble (c4): 3 x 16 members", "members": total_members,
                       "resolutions": [f"6x{s}^2 (L={l}, m={m}, d=8)" for s, l, m in ENSEMBLE],
                       "graph": not args.no_graph, "multi_rhs_first_stage": not args.no_multi_rhs},
            "one_graph_over_resolutions": one_graph,
            "one_member_at_a_time_ms_per_step": round(1e3 * seq, 3),
            "speedup_vs_one_member_at_a_time": round(seq / (sec / args.steps), 3),
            "member0_vor_rel_l2_vs_fp64": acc,
            "bf16_m2l": m2l,
            "pair_kernel_launch_shapes": {f"6x{g['side']}^2": list(g["ctx"].pair_shape()) for g in groups},
            "multi_rhs_launch_shapes": mshapes, "pair_kernel_rooflines": roofs}
    print(json.dumps(line), flush=True)


def bench_adjoint(args, torch, sb):
    """Gradient of one BVE step at 6x512^2 from the hand-written reverse pass
    (sbslpm_bve_back_rk4_step_adjoint: the step's forward recomputed inside,
    then the FP64 reverse pass), against the forward step itself (hipGraph
    replay, FP64 and mixed), with one central finite-difference check of the
    gradient along a random direction at this size."""
    W = sb.workloads
    cfgd = CONFIGS["c3"]
    stream = torch.cuda.Stream()
    common = dict(weights="clenshaw_curtis", kernel_eps=1e-4, delta_t=3600.0, omega=7.292e-5,
                  rk4_combine="reference", **cfgd)
    with torch.cuda.stream(stream):
        ctx64 = sb.Context(sb.RunConfig(precision="fp64", use_graph=not args.no_graph, **common))
        ctxmx = sb.Context(sb.RunConfig(precision="mixed", use_graph=not args.no_graph, **common))
    x, y, z, a = ctx64.spec_points()
    dev = lambda v: torch.from_numpy(np.ascontiguousarray(v)).cuda()
    vor_h = W.gaussian_vortices(x, y, z, a, count=12, seed=0)
    rng = np.random.default_rng(1)
    gx, gy, gz, ga, gvor0, r = [dev(v) for v in (x, y, z, a, vor_h, rng.normal(size=x.shape))]
    for ctx in (ctx64, ctxmx):
        if args.shape:
            ctx.set_pair_shape(*[int(v) for v in args.shape.split(",")])
        elif not args.no_tune:
            with torch.cuda.stream(stream):
                ctx.autotune(ga, gvor0)
    torch.cuda.synchronize()
    _, ev64, _, _ = time_steps(ctx64, torch, stream, gx, gy, gz, ga, gvor0, args.steps, args.warmup)
    _, evmx, _, _ = time_steps(ctxmx, torch, stream, gx, gy, gz, ga, gvor0, args.steps, args.warmup)
    fwd64, fwdmx = ev64 / args.steps, evmx / args.steps
    log(f"forward step: fp64 {1e3 * fwd64:.2f} ms, mixed {1e3 * fwdmx:.2f} ms")

    def time_adjoint(ctx):
        out = ctx.zeros_part()
        with torch.cuda.stream(stream):
            for _ in range(max(args.warmup, 1)):
                ctx.bve_back_rk4_step_adjoint(gx, gy, gz, ga, gvor0, r, out)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(stream):
            e0.record(stream)
            for _ in range(args.steps):
                ctx.bve_back_rk4_step_adjoint(gx, gy, gz, ga, gvor0, r, out)
            e1.record(stream)
        torch.cuda.synchronize()
        ctx.synchronize()
        return e0.elapsed_time(e1) / 1e3 / args.steps, out

    adj64, grad = time_adjoint(ctx64)
    adjmx, gradmx = time_adjoint(ctxmx)
    log(f"step adjoint: fp64 forward {1e3 * adj64:.2f} ms, mixed forward {1e3 * adjmx:.2f} ms")

    # central finite difference of L(w) = sum r * step(w) along one random direction
    d = dev(rng.normal(size=x.shape))
    h = 1e-5 * float(gvor0.norm()) / float(d.norm())
    vel = ctx64.zeros_part(3)

    def loss(w):
        v = w.clone()
        with torch.cuda.stream(stream):
            ctx64.bve_back_rk4_step(gx, gy, gz, ga, v, vel)
        torch.cuda.synchronize()
        return float((r * v).sum())

    fd = (loss(gvor0 + h * d) - loss(gvor0 - h * d)) / (2 * h)
    an = float((grad * d).sum())
    fd_rel = abs(fd - an) / abs(fd)
    mixed_vs_fp64 = W.rel_l2(gradmx.cpu().numpy(), grad.cpu().numpy())
    log(f"finite difference {fd:.10e} adjoint {an:.10e} rel {fd_rel:.2e}; gradient mixed vs fp64 {mixed_vs_fp64:.2e}")
    s = ctx64.sizes
    line = {"metric": "BVE step gradient (hand-written adjoint) cost / forward step cost at 6x512^2",
            "value": round(adj64 / fwd64, 4), "unit": "x forward step (fp64)", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(1e3 * adj64, 3), "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64 reverse pass",
            "data": "synthetic (12 seeded Gaussian vortices, seeded cotangent)",
            "config": {"workload": "BVE RK4 step gradient: 6x512^2 particles", "particles": s.N,
                       "levels": cfgd["levels"], "proxy_degree": cfgd["interp_degree"],
                       "leaf_grid": cfgd["leaf_degree"] + 1, "theta": cfgd["fmm_theta"]},
            "forward_fp64_ms": round(1e3 * fwd64, 3), "forward_mixed_ms": round(1e3 * fwdmx, 3),
            "adjoint_fp64_forward_ms": round(1e3 * adj64, 3), "adjoint_mixed_forward_ms": round(1e3 * adjmx, 3),
            "adjoint_over_mixed_forward": round(adjmx / fwdmx, 4),
            "fd_check": {"rel_err": float(f"{fd_rel:.3e}"), "relative_step": 1e-5},
            "gradient_mixed_vs_fp64_rel_l2": float(f"{mixed_vs_fp64:.3e}")}
    print(json.dumps(line), flush=True)


def bench_4dvar(args, torch, sb):
    """Vorticity 4D-Var on the BVE at 6x192^2 = 221,184 particles (L=6, 6x6 grid per leaf,
    proxy degree 8): 24 RK4 steps of dt = 900 s, the loss 1/2 |H(vor_24) - obs|^2 at 2,048
    seeded observation points (BLI), the gradient from the hand-written step adjoints with
    checkpoints every 6 steps and the transposed observation operator.  Reports the gradient
    cost against the 24-step forward run and central finite differences along 16 random
    directions (FP64)."""
    ag = importlib.import_module("spectral-bslpm_amd.autograd")
    W = sb.workloads
    nsteps, every, nobs, ndir = 24, 6, 2048, 16
    kw = dict(levels=6, interp_degree=8, leaf_degree=5, fmm_theta=0.7, weights="clenshaw_curtis", kernel_eps=1e-4,
              delta_t=900.0, omega=7.292e-5, rk4_combine="reference")
    ctx = sb.Context(sb.RunConfig(precision="fp64", **kw))
    x, y, z, a = ctx.spec_points()
    dev = lambda v: torch.from_numpy(np.ascontiguousarray(v)).cuda()
    g = [dev(v) for v in (x, y, z, a)]
    truth = dev(W.gaussian_vortices(x, y, z, a, count=12, seed=0))
    first = truth + 0.3 * dev(W.gaussian_vortices(x, y, z, a, count=12, seed=1))
    rng = np.random.default_rng(2)
    o = rng.normal(size=(3, nobs))
    o /= np.linalg.norm(o, axis=0)
    obs = [dev(v) for v in o]
    with torch.no_grad():
        obs_val = ag.interp_points(ctx, *obs, ag.bve_steps(ctx, *g, truth, nsteps, nsteps)).detach()
    torch.cuda.synchronize()

    def timed(f, reps):
        f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            out = f()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, out

    t_fwd, _ = timed(lambda: ag.bve_steps(ctx, *g, first, nsteps, nsteps), args.steps)
    t_grad, (L, grad) = timed(lambda: ag.bve_4dvar(ctx, *g, first, nsteps, every, obs, obs_val), args.steps)
    log(f"forward {nsteps} steps {1e3 * t_fwd:.1f} ms, gradient {1e3 * t_grad:.1f} ms")

    def loss(w):
        with torch.no_grad():
            hx = ag.interp_points(ctx, *obs, ag.bve_steps(ctx, *g, w, nsteps, nsteps))
            return float(0.5 * ((hx - obs_val) ** 2).sum())

    rels = []
    for k in range(ndir):
        d = dev(rng.normal(size=x.shape))
        h = 1e-4 * float(first.norm()) / float(d.norm())
        fd = (loss(first + h * d) - loss(first - h * d)) / (2 * h)
        an = float((grad * d).sum())
        rels.append(abs(fd - an) / abs(fd))
    log(f"finite differences: max rel {max(rels):.2e}, median {float(np.median(rels)):.2e}")
    # the same on a mixed-precision context (FP32 pair sums forward and transposed)
    ctxm = sb.Context(sb.RunConfig(precision="mixed", **kw))
    m_fwd, _ = timed(lambda: ag.bve_steps(ctxm, *g, first, nsteps, nsteps), args.steps)
    m_grad, (Lm, gradm) = timed(lambda: ag.bve_4dvar(ctxm, *g, first, nsteps, every, obs, obs_val), args.steps)
    m_err = float((gradm - grad).norm() / grad.norm())
    log(f"mixed: forward {1e3 * m_fwd:.1f} ms, gradient {1e3 * m_grad:.1f} ms ({m_grad / m_fwd:.2f}x), "
        f"gradient rel-L2 vs fp64 {m_err:.2e}")
    line = {"metric": "BVE 4D-Var gradient cost / forward cost, 24 steps at 6x192^2",
            "value": round(t_grad / t_fwd, 4), "unit": "x forward run", "n_gpus": 1, "steps": args.steps,
            "warmup": 1, "ms_per_step": round(1e3 * t_grad / nsteps, 3), "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (12 seeded Gaussian vortices; first guess + 0.3 x a second field; 2,048 seeded points)",
            "config": {"workload": "BVE vorticity 4D-Var: 6x192^2 particles, 24 steps dt=900 s, checkpoints every 6",
                       "particles": int(x.size), "observations": nobs},
            "forward_ms": round(1e3 * t_fwd, 3), "gradient_ms": round(1e3 * t_grad, 3), "loss": float(L),
            "fd_check": {"directions": ndir, "max_rel_err": float(f"{max(rels):.3e}"),
                         "median_rel_err": float(f"{float(np.median(rels)):.3e}"), "relative_step": 1e-4},
            "mixed": {"forward_ms": round(1e3 * m_fwd, 3), "gradient_ms": round(1e3 * m_grad, 3),
                      "ratio": round(m_grad / m_fwd, 4), "loss": float(Lm),
                      "gradient_rel_l2_vs_fp64": float(f"{m_err:.3e}")}}
    print(json.dumps(line), flush=True)


def bench_c5(args, torch, sb):
    """Shallow-water 4D-Var at 6x192^2 = 221,184 particles (L=6, the reference layout with a 6x6
    grid per leaf, proxy degree 5, theta 0.7): 24 RK4 steps of dt = 900 s from balanced zonal
    flow u = 20 cos(lat) m/s over a seeded Gaussian mountain (height U(1000, 2500) m, width
    U(0.1, 0.3) rad), g = 9.80616, mean depth 5960 m, a = 6.37122e6 m; the step in its bounded
    physical form (free-surface Laplacian times g/a^2, regularised stage-1 Laplacian, height
    correction on the free surface, leaf indexing rule).  The loss is the squared misfit of
    vorticity (sigma 1e-6 1/s) and height (sigma 10 m) at 2,048 seeded observation points at the
    final time against the truth run; the first guess adds seeded vortices and a height
    perturbation.  The gradient comes from the hand-written step adjoints with checkpoints every
    6 steps and the transposed observation operator.  Headline: mixed precision (FP32 pair sums
    forward and transposed, FP64 trajectories); the FP64 run is timed beside it.  Reports the
    gradient cost against the 24-step forward run, central finite differences of the FP64 run
    along 16 random directions, the mixed-vs-FP64 gradient, and one oracle forward step on the
    host cores as the CPU baseline."""
    ag = importlib.import_module("spectral-bslpm_amd.autograd")
    W = sb.workloads
    nsteps, every, nobs, ndir = 24, 6, 2048, args.fd_directions
    sig_v, sig_h = 1e-6, 10.0
    mtn = W.gaussian_mountain(seed=7)
    G = W.GRAVITY / W.EARTH_RADIUS ** 2
    kw = dict(levels=6, interp_degree=5, leaf_degree=5, fmm_theta=0.7, weights="clenshaw_curtis", kernel_eps=1e-4,
              delta_t=900.0, omega=7.292e-5, swe_rule="leaf", topography="gaussian", mountain=mtn, gravity=G,
              swe_form=("reg_laplacian", "surface_correction"))
    ctxs = {p: sb.Context(sb.RunConfig(precision=p, **kw)) for p in ("mixed", "fp64")}
    x, y, z, a = ctxs["fp64"].spec_points()
    dev = lambda v: torch.from_numpy(np.ascontiguousarray(v)).cuda()
    g = [dev(v) for v in (x, y, z, a)]
    truth = W.swe_zonal_flow(x, y, z, mtn)
    first = [dev(truth[0] + 0.2 * W.gaussian_vortices(x, y, z, a, seed=71)), dev(truth[1]),
             dev(truth[2] + 1e6 * W.gaussian_vortices(x, y, z, a, seed=72))]
    ox, oy, oz = W.random_unit_vectors(nobs, seed=73)
    obs = [dev(v) for v in (ox, oy, oz)]
    with torch.no_grad():
        tr = ag.swe_steps(ctxs["fp64"], *g, *[dev(v) for v in truth], nsteps, nsteps)
        ov, oh = [ctxs["fp64"].interp_points(*obs, [f])[0] for f in (tr[0], tr[2])]
    torch.cuda.synchronize()

    def forward(ctx, state):
        st = [v.clone() for v in state]
        vel, lap = ctx.zeros_part(3), ctx.zeros_part()
        for _ in range(nsteps):
            ctx.swe_back_rk4_step(*g, *st, vel=vel, height_lap=lap)
        return st

    def loss(ctx, state):
        st = forward(ctx, state)
        hv, hh = [ctx.interp_points(*obs, [f])[0] for f in (st[0], st[2])]
        return 0.5 * float((((hv - ov) / sig_v) ** 2).sum()) + 0.5 * float((((hh - oh) / sig_h) ** 2).sum())

    def timed(f, reps):
        f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            out = f()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, out

    res = {}
    for p, ctx in ctxs.items():
        t_fwd, st = timed(lambda: forward(ctx, first), args.steps)
        t_grad, (L, grad) = timed(lambda: ag.swe_4dvar(ctx, *g, first, nsteps, every, obs, ov, oh, sig_v, sig_h),
                                  args.steps)
        # the same gradient with every step's stage record kept in HBM by the forward run (one
        # segment: no rerun; 24 x 21 fields of 221,184 doubles = 0.89 GB)
        t_hbm, (L_hbm, g_hbm) = timed(lambda: ag.swe_4dvar(ctx, *g, first, nsteps, nsteps, obs, ov, oh, sig_v,
                                                          sig_h), args.steps)
        # (the reverse passes sum transposed interpolations with atomics, so gradients agree to
        # rounding, not bit for bit)
        same = max(W.rel_l2(v.cpu().numpy(), u.cpu().numpy()) for u, v in zip(grad, g_hbm))
        bounded = all(bool(torch.isfinite(v).all()) for v in st)
        res[p] = dict(fwd=t_fwd, grad=t_grad, L=float(L), grad_fields=grad, bounded=bounded,
                      hmin=float(st[2].min()), hmax=float(st[2].max()), vmax=float(st[0].abs().max()),
                      grad_hbm=t_hbm, hbm_same=same)
        log(f"c5 {p}: forward {nsteps} steps {1e3 * t_fwd:.1f} ms, gradient {1e3 * t_grad:.1f} ms "
            f"({t_grad / t_fwd:.2f}x), loss {float(L):.6e}, h in [{res[p]['hmin']:.0f}, {res[p]['hmax']:.0f}]; "
            f"records in HBM: {1e3 * t_hbm:.1f} ms ({t_hbm / t_fwd:.2f}x, gradient rel-L2 {same:.1e} from the "
            f"checkpointed one)")
    gerr = [W.rel_l2(u.cpu().numpy(), v.cpu().numpy()) for u, v in zip(res["mixed"]["grad_fields"],
                                                                        res["fp64"]["grad_fields"])]
    # finite differences of the FP64 run (relative step 1.25e-3 of the direction; leaf-edge
    # crossings inside the stencil set a floor near 1e-5, see tests/test_gpu_swe_4dvar.py)
    rng = np.random.default_rng(74)
    rels, hstep = [], 1.25e-3
    c64, g64 = ctxs["fp64"], res["fp64"]["grad_fields"]
    for _ in range(ndir):
        d = [dev(s * rng.normal(size=x.shape)) for s in (1e-7, 1e-8, 1.0)]
        q = (loss(c64, [b + hstep * e for b, e in zip(first, d)]) -
             loss(c64, [b - hstep * e for b, e in zip(first, d)])) / (2 * hstep)
        an = sum(float((gg * e).sum()) for gg, e in zip(g64, d))
        rels.append(abs(q - an) / abs(q))
    if rels:
        log(f"c5 finite differences: {len(rels)} directions, max rel {max(rels):.2e}, median "
            f"{float(np.median(rels)):.2e}; mixed vs fp64 gradient {gerr}")

    cpu = None
    if args.cpu_budget > 0:
        from oracle import oracle as O
        threads = len(os.sched_getaffinity(0))
        O.set_threads(threads)
        o = O.OracleFMM(kw["levels"], kw["interp_degree"], kw["leaf_degree"], kw["fmm_theta"], 1)
        o.set_swe_physics(G, mtn, 3)
        s0 = [v.cpu().numpy() for v in first]
        t0 = time.perf_counter()
        o.swe_back_rk4_step(x.ravel(), y.ravel(), z.ravel(), a, *s0, kw["delta_t"], kw["omega"], kw["kernel_eps"],
                            0, 2, 1)
        osec = time.perf_counter() - t0
        cpu = dict(value=round(1.0 / osec, 5), unit="forward steps/s", cores=threads, kind="port",
                   sample=f"one full oracle shallow-water step (physical form) at 6x192^2 ({osec:.1f} s); "
                          "the oracle has no adjoint")
        log(f"cpu baseline: {cpu}")

    m, f = res["mixed"], res["fp64"]
    line = {"metric": "SWE 4D-Var gradient cost / forward cost, 24 steps at 6x192^2",
            "value": round(m["grad"] / m["fwd"], 4), "unit": "x forward run", "n_gpus": 1, "steps": args.steps,
            "warmup": 1, "ms_per_step": round(1e3 * m["grad"] / nsteps, 3), "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32 pair sums, f64 trajectories",
            "data": "synthetic (balanced zonal flow over a seeded Gaussian mountain; first guess with seeded "
                    "vortices and height perturbation; 2,048 seeded observation points)",
            "config": {"workload": "Shallow-water 4D-Var: 6x192^2 particles, 24 steps dt=900 s, checkpoints every 6",
                       "particles": int(x.size), "observations": nobs, "mountain": [round(v, 5) for v in mtn],
                       "gravity_over_a2": G, "swe_form": "reg_laplacian+surface_correction", "swe_rule": "leaf"},
            "forward_ms": round(1e3 * m["fwd"], 3), "gradient_ms": round(1e3 * m["grad"], 3), "loss": m["L"],
            "fp64": {"forward_ms": round(1e3 * f["fwd"], 3), "gradient_ms": round(1e3 * f["grad"], 3),
                     "ratio": round(f["grad"] / f["fwd"], 4), "loss": f["L"]},
            "records_in_hbm": {"what": "every step's stage record kept by the forward run (one segment, no "
                                       "rerun; 0.89 GB) instead of checkpoints every 6 steps",
                               "mixed_gradient_ms": round(1e3 * m["grad_hbm"], 3),
                               "mixed_ratio": round(m["grad_hbm"] / m["fwd"], 4),
                               "fp64_gradient_ms": round(1e3 * f["grad_hbm"], 3),
                               "fp64_ratio": round(f["grad_hbm"] / f["fwd"], 4),
                               # the 24 reverse steps alone against the 24 forward steps: the records-in-HBM
                               # gradient less its forward run
                               "mixed_reverse_vs_forward": round((m["grad_hbm"] - m["fwd"]) / m["fwd"], 4),
                               "fp64_reverse_vs_forward": round((f["grad_hbm"] - f["fwd"]) / f["fwd"], 4),
                               "gradient_rel_l2_vs_checkpointed": float(f"{max(m['hbm_same'], f['hbm_same']):.3e}")},
            "mixed_vs_fp64_gradient_rel_l2": [float(f"{e:.3e}") for e in gerr],
            "forward_bounded": m["bounded"] and f["bounded"],
            "final_height_range_m": [round(m["hmin"], 1), round(m["hmax"], 1)],
            "fd_check": {"directions": len(rels), "max_rel_err": float(f"{max(rels):.3e}"),
                         "median_rel_err": float(f"{float(np.median(rels)):.3e}"), "relative_step": hstep}
            if rels else None,
            "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)


PROBE_MIXES = {
    0: "16 v_fma_f32 (independent)",
    1: "16 v_rcp_f32 (independent)",
    2: "4 v_mfma_f32_16x16x4_f32 (independent accumulators)",
    3: "4 v_mfma_f32_4x4x1_16b_f32 (independent accumulators)",
    4: "k_bve_mfma 256-pair step: 16x16x4 + 4 add + 4 rcp + 4 x 4x4x1",
    5: "the same with the 4 reciprocals batched: 16x16x4 + 4 add + 9 mul + 1 rcp + 4 x 4x4x1",
    6: "half the waves mix 2, half mix 0 (vector/matrix co-issue)",
    7: "4 v_mfma_f32_16x16x32_bf16 (independent accumulators)",
    8: "step with a 16x16x32 bf16 first product, alpha as its accumulator, batched reciprocals",
    9: "step with a 16x16x32 bf16 first product, alpha as its accumulator, 4 reciprocals",
}


def bench_probe(args, torch):
    """Register-only issue-rate probes (spectral-bslpm_amd/csrc/probe/issue_rates.hip) behind the
    floor table of the mixed pair kernel: shader-clock cycles per iteration of each instruction
    mix, per wave and per SIMD, at 1 and 4 waves per SIMD (256 compute units x 4 SIMDs)."""
    lib_path = os.path.join(REPO, "spectral-bslpm_amd", "libsbslpm_probe.so")
    if not os.path.exists(lib_path):
        raise SystemExit("libsbslpm_probe.so is not built: run __graft_entry__.build()")
    torch.cuda.init()
    lib = ctypes.CDLL(lib_path)
    lib.sbslpm_probe_issue.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    lib.sbslpm_probe_issue.restype = ctypes.c_int
    iters = 20000
    out = {}
    for mix, what in PROBE_MIXES.items():
        row = {"mix": what}
        for wps in (1, 4):
            blocks, wpb = 256 * wps, 4  # one block of 4 waves per compute unit and wps blocks per unit
            cyc, ms = ctypes.c_double(), ctypes.c_double()
            lib.sbslpm_probe_issue(mix, blocks, wpb, iters // 10, ctypes.byref(cyc), ctypes.byref(ms))  # warm-up
            rc = lib.sbslpm_probe_issue(mix, blocks, wpb, iters, ctypes.byref(cyc), ctypes.byref(ms))
            if rc:
                raise SystemExit(f"probe mix {mix} failed with {rc}")
            row[f"{wps}_wave_per_simd"] = {"cycles_per_iter_per_wave": round(cyc.value, 2),
                                           "cycles_per_iter_per_simd": round(cyc.value / wps, 2),
                                           "launch_ms": round(ms.value, 3)}
        out[str(mix)] = row
        log(f"probe {mix}: {what}: {row}")
    print(json.dumps({"metric": "register-only issue probes", "unit": "shader cycles per iteration",
                      "iterations": iters, "probes": out}), flush=True)


C2 = dict(levels=6, interp_degree=10, leaf_degree=7, fmm_theta=0.6)


def copy_probe(args, torch):
    """A plain device copy of --copy-probe MiB, three times: the scale reference for the
    HBM byte counters read in the same profiled run."""
    if args.copy_probe <= 0:
        return
    nbytes = args.copy_probe << 20
    src = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"copy probe: 3 x {nbytes} bytes read + written")


def tangent_gradient(fun, tx, ty, tz, h=1e-4):
    """Surface gradient of a function of unit vectors by fourth-order central differences
    in arc length along the great circles through each point in two tangent directions."""
    T = np.stack([tx, ty, tz], 1)
    ref = np.where(np.abs(T[:, 2:3]) < 0.9, np.array([[0.0, 0.0, 1.0]]), np.array([[1.0, 0.0, 0.0]]))
    e1 = np.cross(T, ref)
    e1 /= np.linalg.norm(e1, axis=1, keepdims=True)
    e2 = np.cross(T, e1)
    grad = np.zeros_like(T)
    for e in (e1, e2):
        def at(s):
            P = np.cos(s) * T + np.sin(s) * e
            return fun(P[:, 0].copy(), P[:, 1].copy(), P[:, 2].copy())
        d = (8 * (at(h) - at(-h)) - (at(2 * h) - at(-2 * h))) / (12 * h)
        grad += d[:, None] * e
    return grad


def bench_c2(args, torch, sb):
    """Sphere Poisson FMM solve at 6x256^2 = 393,216 particles: L=6 (6,144 leaves), an
    8x8 Clenshaw-Curtis particle grid per leaf, proxy degree 10, theta 0.6.  Forcing
    f = sum a_lm Y_lm, 1 <= l <= 32, a_lm ~ N(0, (1+l)^-2) (seeded; zero mean).  Times
    poisson_solve (upward pass, log pair sums, downward pass) in FP64 -- the value -- and
    in mixed precision; checks the stream function and the Biot-Savart velocity (eps 0)
    at 4,096 seeded targets against the FP64 direct sums at those targets and against the
    closed forms psi = sum a_lm Y_lm / (l(l+1)), u = -x cross grad psi; log pair-kernel
    roofline from HIP events on the context stream; the oracle's full solve as the CPU
    baseline."""
    W = sb.workloads
    stream = torch.cuda.Stream()
    common = dict(weights="clenshaw_curtis", kernel_eps=0.0, **C2)
    with torch.cuda.stream(stream):
        ctx64 = sb.Context(sb.RunConfig(precision="fp64", **common))
        ctxmx = sb.Context(sb.RunConfig(precision="mixed", **common))
    s = ctx64.sizes
    x, y, z, a = ctx64.spec_points()
    f_h = W.random_sh_forcing(x, y, z, lmax=32, seed=1)[0]
    dev = lambda v: torch.from_numpy(np.ascontiguousarray(v)).cuda()
    ga, gf = dev(a), dev(f_h)
    torch.cuda.synchronize()

    def time_solves(ctx):
        psi = ctx.zeros_part()
        with torch.cuda.stream(stream):
            for _ in range(max(args.warmup, 1)):
                ctx.poisson_solve(ga, gf, psi)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(stream):
            e0.record(stream)
            for _ in range(args.steps):
                ctx.poisson_solve(ga, gf, psi)
            e1.record(stream)
        torch.cuda.synchronize()
        ctx.synchronize()
        return e0.elapsed_time(e1) / 1e3 / args.steps, psi

    t64, psi64 = time_solves(ctx64)
    tmx, psimx = time_solves(ctxmx)
    log(f"poisson solve: fp64 {1e3 * t64:.3f} ms, mixed {1e3 * tmx:.3f} ms")

    # accuracy at 4,096 seeded targets: FMM vs the FP64 direct sums there and the closed forms
    tx, ty, tz = W.random_unit_vectors(4096, seed=2)
    gt = [dev(v) for v in (tx, ty, tz)]
    flat = [dev(v.ravel().copy()) for v in (x, y, z, a, f_h)]
    with torch.cuda.stream(stream):
        p_fmm = ctx64.poisson_solve_targets(ga, gf, *gt)
        p_ds = ctx64.direct_sum_inv_lap_targets(*gt, *flat)
        u_fmm = ctx64.bve_velocity_targets(ga, gf, *gt)
        u_ds = ctx64.direct_sum_bve_vel_targets(*gt, *flat)
        p_fmm_mx = ctxmx.poisson_solve_targets(ga, gf, *gt)
    torch.cuda.synchronize()
    ctx64.synchronize()
    ctxmx.synchronize()
    h = lambda v: v.cpu().numpy()
    p_fmm, p_ds, p_fmm_mx = h(p_fmm), h(p_ds), h(p_fmm_mx)
    u_fmm, u_ds = [h(v) for v in u_fmm], [h(v) for v in u_ds]
    p_exact = W.random_sh_forcing(tx, ty, tz, lmax=32, seed=1)[1]
    grad = tangent_gradient(lambda px, py, pz: W.random_sh_forcing(px, py, pz, lmax=32, seed=1)[1], tx, ty, tz)
    u_exact = list((-np.cross(np.stack([tx, ty, tz], 1), grad)).T)
    acc = dict(psi_fmm_vs_direct_rel_l2=W.rel_l2(p_fmm, p_ds), psi_fmm_vs_direct_rel_linf=W.rel_linf(p_fmm, p_ds),
               vel_fmm_vs_direct_rel_l2=W.vec_rel_l2(u_fmm, u_ds), vel_fmm_vs_direct_rel_linf=W.vec_rel_linf(u_fmm, u_ds),
               psi_direct_vs_closed_form_rel_l2=W.rel_l2(p_ds, p_exact),
               vel_direct_vs_closed_form_rel_l2=W.vec_rel_l2(u_ds, u_exact),
               psi_mixed_targets_vs_direct_rel_l2=W.rel_l2(p_fmm_mx, p_ds),
               psi_mixed_vs_fp64_particles_rel_l2=W.rel_l2(h(psimx), h(psi64)))
    log("accuracy", acc)

    # log pair-kernel roofline: HIP events around each pair-kernel launch on the context stream
    pairs = pair_count(ctx64.interaction_list(), s.P, s.Pl, s.n, s.m2)
    roof = {}
    for ctx, name, peak in ((ctx64, "log_fp64", PEAK_FP64), (ctxmx, "log_mixed", PEAK_FP32)):
        ctx.enable_timing(True)
        with torch.cuda.stream(stream):
            for _ in range(5):
                ctx.poisson_solve(ga, gf)
        torch.cuda.synchronize()
        ms, nl = ctx.kernel_time(name)
        ctx.enable_timing(False)
        avg = ms / nl / 1e3
        achieved = pairs * FLOP_LOG / avg / 1e12
        # both log kernels run on the vector unit (FP64 FMA; FP32 unpacked, with the hardware log)
        roof[name] = dict(bound="valu", achieved=round(achieved, 3), peak=peak, unit="TFLOP/s",
                          frac=round(achieved / peak, 4), traffic=None, avg_launch_ms=round(ms / nl, 4),
                          pairs_per_launch=pairs, flop_per_pair=FLOP_LOG)
    tfile = os.path.join(REPO, "profiles", "poisson_c2_hbm_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            tr = json.load(fh)["kernels"]
        for name, kern in (("log_fp64", "k_log64"), ("log_mixed", "k_log_mixed")):
            if tr.get(kern):
                roof[name]["traffic"] = tr[kern]["hbm_bytes"]
                roof[name]["traffic_kernel"] = kern
    log("roofline", roof)
    copy_probe(args, torch)

    cpu = None
    if args.cpu_budget > 0:
        from oracle import oracle as O
        threads = len(os.sched_getaffinity(0))
        O.set_threads(threads)
        o = O.OracleFMM(C2["levels"], C2["interp_degree"], C2["leaf_degree"], C2["fmm_theta"], 1)
        t0 = time.perf_counter()
        o.poisson_solve(a, f_h)
        sec = time.perf_counter() - t0
        cpu = dict(value=round(1.0 / sec, 5), unit="solves/s", cores=threads, kind="port",
                   sample=f"one full oracle Poisson solve at 6x256^2 (upward pass, all {pairs:.3e} log pairs, "
                          f"downward pass), {sec:.1f} s")
        log(f"cpu baseline: {cpu}")

    line = {"metric": "Sphere Poisson FMM solves/sec at 6x256^2 = 393,216 particles (FP64)",
            "value": round(1.0 / t64, 4), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
            "warmup": max(args.warmup, 1), "ms_per_step": round(1e3 * t64, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded spherical-harmonic forcing, l <= 32; 4,096 seeded targets)",
            "config": {"workload": "Sphere Poisson solve (c2): 6x256^2 particles", "particles": s.N,
                       "levels": C2["levels"], "proxy_degree": C2["interp_degree"],
                       "leaf_grid": C2["leaf_degree"] + 1, "theta": C2["fmm_theta"], "interactions": s.I,
                       "pairs_per_solve": pairs, "targets": 4096},
            "mixed_solves_per_s": round(1.0 / tmx, 4), "mixed_ms_per_solve": round(1e3 * tmx, 4),
            "speedup_mixed_vs_fp64": round(t64 / tmx, 3),
            "accuracy": {k: float(f"{v:.3e}") for k, v in acc.items()},
            "roofline": roof["log_fp64"], "roofline_mixed": roof["log_mixed"], "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)


def bench_swe(args, torch, sb):
    """Shallow-water RK4 step (swe_back_rk4_step) at 6x192^2 = 221,184 particles: L=6 (6,144
    leaves), a 6x6 particle grid per leaf that is also the proxy grid (degree 5, the reference
    layout), theta 0.7, eps 1e-4, dt 900 s, Omega 7.292e-5, the reference's cone mountain
    (2000 m, e^{-16 r^2}), the reference indexing rule.  Initial state: the reference's test
    case 5 (initial_conditions.cpp:472-484; lengths in Earth radii, g = 9.80616): vor = 2 u0 z,
    div = 0, h = 5960/a - (a/g)(u0 pi^2/86400^2 + u0^2/2) z^2, u0 = 20 m/s / a.  The reference's
    step does not keep this state bounded (its tendency takes the Laplacian of h without the
    g/a^2 factor, so the divergence grows by orders of magnitude per step), so every timed step
    starts from the initial state again (a device copy of three fields, inside the timing).
    One step = 4 shallow-water FMM velocities, 2 regularised-Laplacian FMM evaluations, the
    panel Laplacian, 3 tendency passes, departure search and remap, captured in a hipGraph and
    replayed.  Reports FP64 and mixed-precision steps/s, the two pair kernels' rooflines (HIP
    events around eager launches on contexts without the graph), the mixed-vs-FP64 stage-1
    velocity, and one full oracle step on the host cores as the CPU baseline."""
    W = sb.workloads
    kw = dict(levels=6, interp_degree=5, leaf_degree=5, fmm_theta=0.7, weights="reference", kernel_eps=1e-4,
              delta_t=900.0, omega=7.292e-5, topography="cone", swe_rule="reference")
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        probe = sb.Context(sb.RunConfig(**kw))
    s = probe.sizes
    x, y, z, a = probe.spec_points()
    pairs = pair_count(probe.interaction_list(), s.P, s.Pl, s.n, s.m2)
    del probe
    A, Om, g0 = 6371000.0, 7.292e-5, 9.80616
    u0 = 20.0 / A
    vor_h = 2 * u0 * z
    div_h = np.zeros_like(z)
    h_h = 5960.0 / A - A / g0 * (u0 * np.pi ** 2 / 86400.0 ** 2 + 0.5 * u0 * u0) * z * z
    dev = lambda v: torch.from_numpy(np.ascontiguousarray(v)).cuda()
    gx, gy, gz, ga = [dev(v) for v in (x, y, z, a)]
    state0 = [dev(v) for v in (vor_h, div_h, h_h)]

    def leg(precision):
        """Steps/s on a graph-captured context; rooflines and the stage-1 fields from an eager one."""
        with torch.cuda.stream(stream):
            cg = sb.Context(sb.RunConfig(precision=precision, use_graph=not args.no_graph, **kw))
            ce = sb.Context(sb.RunConfig(precision=precision, **kw))
        st = [v.clone() for v in state0]
        vel, lap = cg.zeros_part(3), cg.zeros_part()

        def run(ctx, n):
            with torch.cuda.stream(stream):
                for _ in range(n):
                    for v, v0 in zip(st, state0):
                        v.copy_(v0)
                    ctx.swe_back_rk4_step(gx, gy, gz, ga, *st, vel=vel, height_lap=lap)

        run(cg, max(args.warmup, 1))
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(stream):
            e0.record(stream)
        run(cg, args.steps)
        with torch.cuda.stream(stream):
            e1.record(stream)
        torch.cuda.synchronize()
        cg.synchronize()
        sec = e0.elapsed_time(e1) / 1e3 / args.steps
        finite = all(bool(torch.isfinite(v).all()) for v in st)
        ce.enable_timing(True)
        run(ce, 2)
        torch.cuda.synchronize()
        ce.synchronize()
        roof, peak = {}, PEAK_FP64 if precision == "fp64" else PEAK_FP32
        tag = "fp64" if precision == "fp64" else "mixed"
        for kern, flop in (("swe", FLOP_SWE), ("lap", FLOP_LAP)):
            ms, nl = ce.kernel_time(f"{kern}_{tag}")
            avg = ms / nl / 1e3
            achieved = pairs * flop / avg / 1e12
            # every shallow-water and Laplacian pair kernel runs on the vector unit
            roof[kern] = dict(bound="valu", achieved=round(achieved, 3), peak=peak, unit="TFLOP/s",
                              frac=round(achieved / peak, 4), traffic=None, avg_launch_ms=round(ms / nl, 4),
                              launches_per_step=nl // 2, pairs_per_launch=pairs, flop_per_pair=flop)
        ce.enable_timing(False)
        log(f"swe step {precision}: {1e3 * sec:.3f} ms", roof)
        return sec, roof, [v.cpu().numpy() for v in vel], lap.cpu().numpy(), finite

    sec, roof, vel64, lap64, finite = leg("fp64")
    secm, roofm, velm, lapm, finitem = leg("mixed")
    vel_err = W.vec_rel_l2(velm, vel64)
    log(f"mixed vs fp64: stage-1 velocity {vel_err:.2e}, Laplacian {W.rel_l2(lapm, lap64):.2e}")

    cpu = None
    if args.cpu_budget > 0:
        from oracle import oracle as O
        threads = len(os.sched_getaffinity(0))
        O.set_threads(threads)
        o = O.OracleFMM(kw["levels"], kw["interp_degree"], kw["leaf_degree"], kw["fmm_theta"], 0)
        t0 = time.perf_counter()
        o.swe_back_rk4_step(x.ravel(), y.ravel(), z.ravel(), a, vor_h, div_h, h_h, kw["delta_t"], Om,
                            kw["kernel_eps"], 0, 1, 0)
        osec = time.perf_counter() - t0
        cpu = dict(value=round(1.0 / osec, 5), unit="steps/s", cores=threads, kind="port",
                   sample=f"one full oracle shallow-water step at 6x192^2 ({osec:.1f} s)")
        log(f"cpu baseline: {cpu}")

    line = {"metric": "SWE RK4 steps/sec at 6x192^2 = 221,184 particles (FP64)",
            "value": round(1.0 / sec, 4), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": max(args.warmup, 1), "ms_per_step": round(1e3 * sec, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (the reference's test case 5 state with its cone mountain; every step from that state)",
            "config": {"workload": "Shallow-water RK4 step: 6x192^2 particles", "particles": s.N,
                       "levels": kw["levels"], "proxy_degree": kw["interp_degree"], "leaf_grid": kw["leaf_degree"] + 1,
                       "theta": kw["fmm_theta"], "interactions": s.I, "pairs_per_evaluation": pairs,
                       "dt": kw["delta_t"], "topography": "cone", "swe_rule": "reference",
                       "graph": not args.no_graph},
            "state_finite_after_steps": finite and finitem,
            "mixed_steps_per_s": round(1.0 / secm, 4), "mixed_ms_per_step": round(1e3 * secm, 4),
            "speedup_mixed_vs_fp64": round(sec / secm, 3),
            "mixed_vs_fp64_stage1_velocity_rel_l2": float(f"{vel_err:.3e}"),
            "roofline": roof["swe"], "roofline_laplacian": roof["lap"],
            "roofline_mixed": roofm["swe"], "roofline_laplacian_mixed": roofm["lap"], "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)


def bench_ll(args, torch, sb):
    """Lat-lon shallow-water step (swe_back_rk4_step_ll) on the reference's grid: 1 degree,
    180 x 360 = 64,800 points (swe_time_evo_ll.cpp:48-80), interp_degree 10, leaf size
    4 (d+1)^2 = 484, theta 0.5, kernel_eps 1e-4, delta_t 50 s (namelist.txt), the lat-lon
    test case 5 with the cone mountain (initial_conditions.cpp:486-506, topography.cpp:111-131).
    One step = the irregular-tree FMM velocity (two upward passes, the FP64 pair kernel, the
    downward pass), the tendencies, the pole means, the arrival values, the divergence
    smoothing and the polar filter; timed steps advance the state.  The CPU baseline is one
    oracle step from the same state on the host cores, also checked against the GPU's."""
    L = sb.latlon
    dx, deg, leaf, dt = 1.0, 10, 484, 50.0
    lats, lons, x, y, z, area = L.cell_grid(dx)
    topo = L.cone_topography(lats, lons)
    vor, div, h = L.initial_state("tc5", lats, lons, topo)
    nlat, nlon = x.shape
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        ctx = sb.Context(sb.RunConfig(levels=2, interp_degree=deg, kernel_eps=1e-4, fmm_theta=0.5))
        irr = ctx.irregular(x, y, z, leaf)
    dev = lambda v: torch.from_numpy(np.ascontiguousarray(v)).cuda()
    g = [dev(v) for v in (x, y, z, area, topo)]
    st0 = [dev(v) for v in (vor, div, h)]
    st = [v.clone() for v in st0]
    vel = [torch.zeros(nlat * nlon, dtype=torch.float64, device="cuda") for _ in range(3)]
    grid = (nlat, nlon, dx)

    def run(n):
        with torch.cuda.stream(stream):
            for _ in range(n):
                irr.swe_back_rk4_step_ll(grid, dt, g[0], g[1], g[2], g[3], *st, g[4], vel=vel)

    run(max(args.warmup, 1))
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(stream):
        e0.record(stream)
    run(args.steps)
    with torch.cuda.stream(stream):
        e1.record(stream)
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) / 1e3 / args.steps
    finite = all(bool(torch.isfinite(v).all()) for v in st)
    log(f"lat-lon step: {1e3 * sec:.3f} ms ({1 / sec:.1f} steps/s), {irr.P} panels, {irr.I} interactions")

    cpu, parity = None, None
    if args.cpu_budget > 0:
        from oracle import oracle as O
        threads = len(os.sched_getaffinity(0))
        O.set_threads(threads)
        o = O.OracleIrreg(x, y, z, leaf, deg, 0.5)
        one = [v.clone() for v in st0]
        with torch.cuda.stream(stream):
            gvel = irr.swe_back_rk4_step_ll(grid, dt, g[0], g[1], g[2], g[3], *one, g[4])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        *ref, rvel = o.swe_back_rk4_step_ll(nlat, nlon, dx, dt, L.OMEGA, area.ravel(), vor.ravel(), div.ravel(),
                                            h.ravel(), topo.ravel(), 1e-4)
        osec = time.perf_counter() - t0
        cpu = dict(value=round(1.0 / osec, 4), unit="steps/s", cores=threads, kind="port",
                   sample=f"one full oracle lat-lon step at 1 degree ({osec:.2f} s)")
        # The step interpolates along the row where |v| < 1e-16 (swe_time_step_ll.cpp:446).  In
        # this zonal state v is rounding noise (|v| < 1e-16 at every point in the oracle), so the
        # two sides of that test are picked by rounding at some points; those points are counted
        # and left out of the comparison of vorticity and height (divergence and velocity do not
        # depend on the branch).
        _, v_gpu = O.xyz_vel_to_uv_vel(x, y, z, *[t.cpu().numpy() for t in gvel])
        _, v_orc = O.xyz_vel_to_uv_vel(x, y, z, *rvel)
        flip = (np.abs(v_gpu) < 1e-16) != (np.abs(v_orc) < 1e-16)
        errs = []
        vscale = max(np.abs(v).max() for v in rvel)  # one scale for the three velocity components
        for i, (a, b) in enumerate(zip(one + gvel, ref + list(rvel))):
            keep = ~flip if i in (0, 2) else np.ones_like(flip)
            a = a.cpu().numpy().ravel()
            scale = vscale if i >= 3 else np.abs(b).max()
            errs.append(float(np.abs(a - b)[keep].max() / max(scale, 1e-300)))
        parity = dict(max_rel=float(f"{max(errs):.3e}"), branch_flips=int(flip.sum()),
                      oracle_points_on_row_branch=int((np.abs(v_orc) < 1e-16).sum()))
        log(f"cpu baseline: {cpu}; GPU vs oracle step: {parity}")

    line = {"metric": "lat-lon SWE steps/sec at 1 degree = 64,800 points (FP64)",
            "value": round(1.0 / sec, 3), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": max(args.warmup, 1), "ms_per_step": round(1e3 * sec, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (the reference's lat-lon test case 5 with its cone mountain)",
            "config": {"workload": "Lat-lon shallow-water step: 180 x 360 points", "points": nlat * nlon,
                       "interp_degree": deg, "leaf_size": leaf, "theta": 0.5, "panels": irr.P,
                       "interactions": irr.I, "dt": dt},
            "state_finite_after_steps": finite, "gpu_vs_oracle_step": parity, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS) + ["c2", "c4", "c5", "adjoint", "4dvar", "swe", "ll", "probe"],
                    help="c3: the 6x512^2 step (the metric), c1: 6x64^2, c2: the 6x256^2 Poisson solve, "
                         "swe: the shallow-water step at 6x192^2, ll: the lat-lon shallow-water step at 1 degree, "
                         "c4: the 48-member ensemble, "
                         "adjoint: cost of the step gradient at 6x512^2, 4dvar: BVE 4D-Var gradient at 6x192^2, "
                         "c5: shallow-water 4D-Var gradient at 6x192^2")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of oracle CPU work (0 = skip)")
    ap.add_argument("--fd-directions", type=int, default=16, help="c5: finite-difference directions (0 = skip)")
    ap.add_argument("--long-steps", type=int, default=100,
                    help="c1/c3: steps of the mixed-vs-FP64 accuracy run (the configuration's 100; 0 = skip)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-multi-rhs", action="store_true",
                    help="c4: first RK4 stage on the per-member pair kernel instead of the multi-RHS product")
    ap.add_argument("--no-tune", action="store_true", help="skip the pair-kernel launch-shape autotuner")
    ap.add_argument("--shape", default=None, metavar="WAVES,XCD",
                    help="fix the pair-kernel launch shape instead of autotuning (e.g. 4,0)")
    ap.add_argument("--valu", action="store_true", help="also time the mixed-precision vector-unit pair kernel")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy of MIB MiB three times (scale reference for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(more than one GPU is out of scope: members and panels would shard with no change to the kernels)", file=sys.stderr)
        sys.exit(2)

    import torch
    sb = importlib.import_module("spectral-bslpm_amd")
    if args.config == "c4":
        bench_ensemble(args, torch, sb)
        return
    if args.config == "adjoint":
        bench_adjoint(args, torch, sb)
        return
    if args.config == "4dvar":
        bench_4dvar(args, torch, sb)
        return
    if args.config == "c2":
        bench_c2(args, torch, sb)
        return
    if args.config == "swe":
        bench_swe(args, torch, sb)
        return
    if args.config == "c5":
        bench_c5(args, torch, sb)
        return
    if args.config == "ll":
        bench_ll(args, torch, sb)
        return
    if args.config == "probe":
        bench_probe(args, torch)
        return
    W = sb.workloads
    cfgd = CONFIGS[args.config]
    stream = torch.cuda.Stream()
    common = dict(weights="clenshaw_curtis", kernel_eps=1e-4, delta_t=3600.0, omega=7.292e-5,
                  rk4_combine="reference", use_graph=not args.no_graph, **cfgd)
    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        ctx64 = sb.Context(sb.RunConfig(precision="fp64", **common))
        ctxmx = sb.Context(sb.RunConfig(precision="mixed", **common))
    log(f"setup {time.perf_counter() - t0:.1f} s: N={ctx64.sizes.N} P={ctx64.sizes.P} I={ctx64.sizes.I}")
    x, y, z, a = ctx64.spec_points()
    vor_h = W.gaussian_vortices(x, y, z, a, count=12, seed=0)
    dev = lambda v: torch.from_numpy(np.ascontiguousarray(v)).cuda()
    gx, gy, gz, ga, gvor0 = [dev(v) for v in (x, y, z, a, vor_h)]
    torch.cuda.synchronize()
    tuned = {}
    for name, ctx in (("fp64", ctx64), ("mixed", ctxmx)):
        if args.shape:
            w_, x_ = [int(v) for v in args.shape.split(",")]
            ctx.set_pair_shape(w_, x_)
            tuned[name] = {"waves_per_workgroup": w_, "xcd_mapping": x_, "fixed": True}
        elif not args.no_tune:
            with torch.cuda.stream(stream):
                table, shape = ctx.autotune(ga, gvor0)
            # xcd mapping 0: round-robin over the XCDs, 1: equal contiguous ranges, 2: cost-balanced ranges
            tuned[name] = {"waves_per_workgroup": shape[0], "xcd_mapping": shape[1],
                           "candidates_ms": [[int(w), int(c), round(ms, 4)] for w, c, ms in table]}
    torch.cuda.synchronize()
    log("launch shape", tuned)

    wall64, ev64, vor64, vel64 = time_steps(ctx64, torch, stream, gx, gy, gz, ga, gvor0, args.steps, args.warmup)
    log(f"fp64 : {args.steps / ev64:.3f} steps/s ({1e3 * ev64 / args.steps:.2f} ms/step)")
    wallmx, evmx, vormx, velmx = time_steps(ctxmx, torch, stream, gx, gy, gz, ga, gvor0, args.steps, args.warmup)
    log(f"mixed: {args.steps / evmx:.3f} steps/s ({1e3 * evmx / args.steps:.2f} ms/step)")
    valu = None
    if args.valu:
        with torch.cuda.stream(stream):
            ctxva = sb.Context(sb.RunConfig(precision="mixed_valu", **common))
        _, evva, vorva, _ = time_steps(ctxva, torch, stream, gx, gy, gz, ga, gvor0, args.steps, args.warmup)
        valu = dict(steps_per_s=round(args.steps / evva, 4),
                    vor_rel_l2_after_steps=float(f"{W.rel_l2(vorva.cpu().numpy(), vor64.cpu().numpy()):.3e}"))
        log(f"mixed (vector unit): {args.steps / evva:.3f} steps/s")
        del ctxva

    # accuracy of mixed vs FP64: one velocity evaluation from the initial field,
    # and the vorticity after the timed steps (both runs start from the same field)
    with torch.cuda.stream(stream):
        v64 = ctx64.bve_velocity(ga, gvor0)
        vmx = ctxmx.bve_velocity(ga, gvor0)
    torch.cuda.synchronize()
    v64h = [v.cpu().numpy() for v in v64]
    vmxh = [v.cpu().numpy() for v in vmx]
    acc = dict(
        vel_rel_l2=W.vec_rel_l2(vmxh, v64h), vel_rel_linf=W.vec_rel_linf(vmxh, v64h),
        vor_rel_l2_after_steps=W.rel_l2(vormx.cpu().numpy(), vor64.cpu().numpy()),
        vor_rel_linf_after_steps=W.rel_linf(vormx.cpu().numpy(), vor64.cpu().numpy()))
    # FP64 FMM vs FP64 direct sum at 4,096 targets (a contiguous block of 64 leaves)
    rng = np.random.default_rng(0)
    lb = int(rng.integers(0, ctx64.sizes.Pl - 64)) * ctx64.sizes.m2
    ub = lb + 4096
    flat = [t.reshape(-1) for t in (gx, gy, gz, ga, gvor0)]
    with torch.cuda.stream(stream):
        ds = ctx64.direct_sum_bve_vel(*flat, lb=lb, ub=ub)
    torch.cuda.synchronize()
    dsh = [d.cpu().numpy()[lb:ub] for d in ds]
    fmm64 = [v.reshape(-1)[lb:ub] for v in v64h]
    fmmmx = [v.reshape(-1)[lb:ub] for v in vmxh]
    acc["fp64_fmm_vs_direct_rel_l2"] = W.vec_rel_l2(fmm64, dsh)
    acc["mixed_fmm_vs_direct_rel_l2"] = W.vec_rel_l2(fmmmx, dsh)
    # the configuration's 100 steps: mixed against FP64 from the same field
    if args.long_steps > 0:
        w64, wmx = gvor0.clone(), gvor0.clone()
        vel = ctx64.zeros_part(3)
        with torch.cuda.stream(stream):
            for _ in range(args.long_steps):
                ctx64.bve_back_rk4_step(gx, gy, gz, ga, w64, vel)
                ctxmx.bve_back_rk4_step(gx, gy, gz, ga, wmx, vel)
        torch.cuda.synchronize()
        a64, amx = w64.cpu().numpy(), wmx.cpu().numpy()
        acc[f"vor_rel_l2_after_{args.long_steps}_steps"] = W.rel_l2(amx, a64)
        acc[f"vor_rel_linf_after_{args.long_steps}_steps"] = W.rel_linf(amx, a64)
    log("accuracy", acc)

    # pair-kernel roofline: HIP events around each pair-kernel launch on the context stream
    il = ctx64.interaction_list()
    s = ctx64.sizes
    pairs = pair_count(il, s.P, s.Pl, s.n, s.m2)
    roof = {}
    # executed work per pair besides the 27-flop convention (which credits a per-pair cross
    # product both kernels do once per target and source set):
    #   FP64: x.y (5), (1 - x.y) + eps (2), rcp + one Newton step (5), g q (1), 3 fma (6) = 19;
    #   mixed: the 16x16x4 denominator product (8 per pair), four 4x4x1 accumulation products
    #   (8 per pair) on the matrix cores, the add of 1/2|p|^2 and the reciprocal on the vector
    #   unit (2) = 18.
    executed = {"bve_fp64": 19, "bve_mixed": 18}
    limiter = {"bve_fp64": "FP64 vector FMA and reciprocal issue (78.6 TF/s vector peak)",
               "bve_mixed": "MFMA products (64 cycles per 256 pairs) and vector work with four v_rcp_f32 "
                            "(about 90 cycles), serialized on each SIMD (register-only probe: no vector/matrix "
                            "co-issue; profiles/bve_mixed_issue_rates.md)"}
    for ctx, name, peak in ((ctx64, "bve_fp64", PEAK_FP64), (ctxmx, "bve_mixed", PEAK_FP32)):
        ctx.enable_timing(True)
        with torch.cuda.stream(stream):
            for _ in range(5):
                ctx.bve_velocity(ga, gvor0)
        torch.cuda.synchronize()
        ms, nl = ctx.kernel_time(name)
        ctx.enable_timing(False)
        avg = ms / nl / 1e3
        achieved = pairs * FLOP_BS / avg / 1e12
        roof[name] = dict(bound="mfma" if name == "bve_mixed" else "valu", achieved=round(achieved, 3), peak=peak,
                          unit="TFLOP/s", frac=round(achieved / peak, 4), traffic=None,
                          avg_launch_ms=round(ms / nl, 4), pairs_per_launch=pairs, flop_per_pair=FLOP_BS,
                          executed_flop_per_pair=executed[name],
                          executed_tflops=round(pairs * executed[name] / avg / 1e12, 3), limiter=limiter[name])
    # HBM bytes per launch of the pair kernels, from the committed FETCH_SIZE / WRITE_SIZE
    # runs of this workload (scaled by the copy-kernel factors recorded with them)
    tfile = os.path.join(REPO, "profiles", "bve_step_c3_hbm_traffic.json")
    if args.config == "c3" and os.path.exists(tfile):
        with open(tfile) as fh:
            tr = json.load(fh)["kernels"]
        for name, kern in (("bve_fp64", "k_bve64"), ("bve_mixed", "k_bve_mfma")):
            if tr.get(kern):
                roof[name]["traffic"] = tr[kern]["hbm_bytes"]
                roof[name]["traffic_kernel"] = kern
    log("roofline", roof)

    copy_probe(args, torch)

    cpu = None
    if args.cpu_budget > 0:
        t0 = time.perf_counter()
        cpu = cpu_baseline(cfgd, x, y, z, a, vor_h, args.cpu_budget)
        log(f"cpu baseline {time.perf_counter() - t0:.1f} s: {cpu}")

    value = args.steps / evmx
    line = {
        "metric": METRIC, "value": round(value, 4), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * evmx / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 pair kernels, f64 accumulation/trajectories",
        "data": "synthetic (12 seeded Gaussian vortices)",
        "config": {"workload": f"BVE RK4 step ({args.config}): 6x512^2 particles" if args.config == "c3" else
                   f"BVE RK4 step ({args.config})", "particles": s.N, "levels": cfgd["levels"],
                   "proxy_degree": cfgd["interp_degree"], "leaf_grid": cfgd["leaf_degree"] + 1,
                   "theta": cfgd["fmm_theta"], "interactions": s.I, "pairs_per_evaluation": pairs,
                   "kernel_eps": 1e-4, "dt": 3600.0, "graph": not args.no_graph},
        "fp64_steps_per_s": round(args.steps / ev64, 4), "fp64_ms_per_step": round(1e3 * ev64 / args.steps, 3),
        "speedup_mixed_vs_fp64": round(ev64 / evmx, 3),
        "accuracy": {k: float(f"{v:.3e}") for k, v in acc.items()},
        "wall_s": {"fp64": round(wall64, 4), "mixed": round(wallmx, 4)},
        "roofline": roof["bve_mixed"], "roofline_fp64": roof["bve_fp64"],
        "pair_kernel_launch_shape": tuned,
        "cpu_baseline": cpu,
    }
    if valu:
        line["mixed_valu"] = valu
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
