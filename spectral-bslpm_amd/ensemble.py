"""Ensembles of uneven resolutions advanced as one graph launch per step.

Each resolution is a Context (tree, interaction list, tiles, workspace) whose
members share every kernel launch (sbslpm_bve_back_rk4_step_ensemble).  The
groups are independent, so one step of all of them is captured once as a
single torch CUDA graph in which each group is a branch on its own stream
(forked from and joined back to the capturing stream); a replay is one launch
for the whole ensemble, and the small groups' kernels fill the GPU beside the
large group's instead of waiting behind it.  No reference counterpart: the
reference steps one field per bve_back_rk4_step call (bve_time_step.cpp:311-419).
"""
import torch


class UnevenEnsemble:
    """groups: list of (ctx, (x, y, z, area), vors [members, Pl, m2]) with contexts created
    without SBSLPM_FLAG_GRAPH (the whole step is captured here).  vors are updated in place by
    every step; vel[i] receives group i's stage-0 velocities."""

    def __init__(self, groups):
        self.groups = list(groups)
        self.vel = [[torch.zeros_like(v) for _ in range(3)] for _, _, v in self.groups]
        self.streams = [torch.cuda.Stream() for _ in self.groups]  # default priority
        self.main = torch.cuda.Stream()
        self.graph = None
        for (ctx, _, _), s in zip(self.groups, self.streams):
            ctx.set_stream(s)

    def _step_branches(self):
        fork = torch.cuda.Event()
        fork.record(self.main)
        joins = []
        for (ctx, geom, vors), vel, s in zip(self.groups, self.vel, self.streams):
            s.wait_event(fork)
            with torch.cuda.stream(s):
                ctx.bve_back_rk4_step_ensemble(*geom, vors, vel)
            e = torch.cuda.Event()
            e.record(s)
            joins.append(e)
        for e in joins:
            self.main.wait_event(e)

    def step_eager(self):
        """One step of every group, launched on the groups' streams without a graph."""
        self.main.wait_stream(torch.cuda.current_stream())
        self._step_branches()
        torch.cuda.current_stream().wait_stream(self.main)

    def capture(self):
        """Captures one step of every group as one graph.  The workspaces must exist (one eager
        step first), since nothing may be allocated under capture."""
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=self.main):
            self._step_branches()
        torch.cuda.synchronize()

    def step(self):
        """One step of every group: a replay of the captured graph.  The first call runs the
        step eagerly (allocating the workspaces) and captures it for the later calls."""
        if self.graph is None:
            self.step_eager()
            self.capture()
            return
        self.graph.replay()
