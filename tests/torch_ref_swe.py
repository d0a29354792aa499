"""FP64 direct-sum restatement of one shallow-water step in torch (CPU),
differentiable by torch.autograd: the reference gradient for the library's
hand-written shallow-water step adjoint.

It follows swe_back_rk4_step (swe_time_step.cpp:606-811) in the physical form
of the 4D-Var configuration: heights in metres with the free-surface Laplacian
scaled by g/a^2, the Gaussian mountain, the stage-1 surface Laplacian by the
regularised kernel, the height correction on the free surface, each leaf's own
face and displacement (the leaf rule).  Every FMM evaluation is an O(N^2) sum
over all particles with sources displaced by their own displacement (no
proxies): the shallow-water velocity of swe_vel_impl.hpp:12-76 (variant 1)
and the regularised-Laplacian kernel of laplacian_impl.hpp:46-66.  The
velocity gradient is the surface gradient f_xi grad(xi) + f_eta grad(eta),
with f_xi, f_eta from the barycentric differentiation matrix of each leaf's
Chebyshev-Lobatto grid (Berrut & Trefethen 2004) and grad(xi), grad(eta) the
gradients of the label coordinates taken by torch.autograd.  Departure search
and remap as tests/torch_ref_bve.py.  Test infrastructure only."""
import math

import torch

from torch_ref_bve import StepModel, face_xieta, project

LAP_EPS = 0.1


def lap_kernel(t):
    """K(t) of laplacian_impl.hpp:46-66, eps = 0.1."""
    e = LAP_EPS
    t2 = t * t
    h1 = math.exp(-1.0 / e) * math.pi ** -1.5 * (2.0 * t * (-1.0 + 3.0 * e + t2))
    h2 = (1.0 + torch.erf(t / e)) * torch.exp((t2 - 1.0) / e) / (math.sqrt(e) * math.pi) * (
        2.0 * e * e + 2.0 * t2 * (t2 - 1.0) + e * (7.0 * t2 - 1.0))
    return (h1 + h2) / e ** 2.5


class SweModel(StepModel):
    def __init__(self, x, y, z, area, levels, eps, omega, dt, gravity, mountain, combine=0):
        super().__init__(x, y, z, area, levels, eps, omega, dt)
        self.gravity, self.combine = gravity, combine
        h0, w, lat, lon = mountain
        self.mtn_h, self.mtn_w = h0, w
        self.mc = torch.tensor([math.cos(lat) * math.cos(lon), math.cos(lat) * math.sin(lon), math.sin(lat)],
                               dtype=torch.float64)
        m = self.m
        # differentiation matrix on [-1, 1] Chebyshev-Lobatto nodes cos(pi i / (m-1)), scaled to a
        # leaf of label width (pi/2) / 2^(L-1)
        nodes = self.cheb
        D = torch.zeros((m, m), dtype=torch.float64)
        for i in range(m):
            for k in range(m):
                if i != k:
                    D[i, k] = (self.bw[k] / self.bw[i]) / (nodes[i] - nodes[k])
            D[i, i] = -D[i].sum()
        self.D = D / (0.5 * (math.pi / 2) / self.n1)
        # gradients of the particles' label coordinates on their own face
        p = self.X.clone().requires_grad_(True)
        _, xi, eta = face_xieta(p)
        self.gxi = torch.autograd.grad(xi.sum(), p, retain_graph=True)[0].detach()
        self.geta = torch.autograd.grad(eta.sum(), p)[0].detach()

    def mountain(self, p):
        c = self.mc
        cr = torch.cross(p, c.expand_as(p), dim=1)
        r = torch.atan2(cr.norm(dim=1), p @ c)
        return self.mtn_h * torch.exp(-(r / self.mtn_w) ** 2)

    def swe_velocity(self, qz, qd, disp=None):
        """u_t = sum_s r (Qz (x cross y) + Qd (y - x (x.y))), r = 1/(1 - x.y + eps),
        Q = -q/(4 pi), y = P(x + disp)."""
        X = self.X
        y = X if disp is None else project(X + disp)
        dot = X @ y.T
        r = 1.0 / ((1.0 - dot) + self.eps)
        Qz, Qd = -qz / (4 * math.pi), -qd / (4 * math.pi)
        w = r @ (Qz.unsqueeze(1) * y)
        d = r @ (Qd.unsqueeze(1) * y) - X * ((r * dot) @ Qd).unsqueeze(1)
        return torch.cross(X, w, dim=1) + d

    def lap(self, q, disp=None):
        y = self.X if disp is None else project(self.X + disp)
        return lap_kernel(self.X @ y.T) @ q

    def gradient(self, f):
        """Surface gradient of a leaf field [Pl*m2] at the particles -> [N][3]."""
        m = self.m
        F = f.view(self.Pl, m, m)
        fx = torch.einsum("ak,lkb->lab", self.D, F).reshape(-1)
        fe = torch.einsum("bk,lak->lab", self.D, F).reshape(-1)
        return fx.unsqueeze(1) * self.gxi + fe.unsqueeze(1) * self.geta

    def tendency(self, u, vor, div, h, lap):
        """swe_tendency_computation (swe_time_step.cpp:76-161) with the surface gradient."""
        X, om = self.X, self.omega
        G = [self.gradient(u[:, c]) for c in range(3)]  # G[a][:, b] = d_b u_a
        gg = sum(G[a][:, b] * G[b][:, a] for a in range(3) for b in range(3))
        z = X[:, 2]
        vor_t = -div * (vor + 2 * om * z) - 2 * om * u[:, 2]
        cor = torch.where(X[:, 0] ** 2 + X[:, 1] ** 2 > 1e-16, -X[:, 1] * u[:, 0] + X[:, 0] * u[:, 1],
                          torch.zeros_like(z))
        div_t = 2 * om * z * vor - gg - self.gravity * lap - (u * u).sum(1) - 2 * om * cor
        return vor_t, div_t, -h * div

    def step(self, vor, div, h):
        dt, X, A = self.dt, self.X, self.area
        S0 = (vor, div, h)
        e1 = h + self.mountain(X)
        lap = self.lap(A * e1) - 8.0 / 0.1 * e1
        V = [self.swe_velocity(A * vor, A * div)]
        T = [self.tendency(V[0], vor, div, h, lap)]
        taus = (0.5 * dt, 0.5 * dt, dt)
        for s in (1, 2, 3):
            tau = taus[s - 1]
            disp = tau * V[s - 1]
            S = [S0[q] + tau * T[s - 1][q] for q in range(3)]
            As = A * (1.0 + tau * div)
            V.append(self.swe_velocity(As * S[0], As * S[1], disp))
            if s == 3:
                break
            p = project(X + disp)
            e = S[2] + self.mountain(p)
            corr = self.interp([h + tau * T[s - 1][2]], p)[0] + self.mountain(p)
            lap = self.lap(As * e, disp) - 8.0 / 0.1 * corr
            T.append(self.tendency(V[s], vor, div, h, lap))
        R = [S0[q] + dt * T[0][q] for q in range(3)]
        k1 = V[3]
        p2 = project(X - 0.5 * dt * k1)
        k2 = self.interp(V[2].T, p2).T
        p3 = project(X - 0.5 * dt * k2)
        k3 = self.interp(V[1].T, p3).T
        p4 = project(X - dt * k3)
        k4 = self.interp(V[0].T, p4).T
        if self.combine == 0:
            a = project(X - dt / 6.0 * k1)
            a = project(a + dt / 3.0 * k2)
            a = project(a + dt / 3.0 * k3)
            dep = project(a + dt / 6.0 * k4)
        else:
            dep = project(X - dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4))
        return tuple(self.interp(R, dep))
