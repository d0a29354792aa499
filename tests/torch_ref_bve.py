"""FP64 direct-sum restatement of one BVE backward RK4 step in torch (CPU),
differentiable by torch.autograd: the reference gradient for the library's
hand-written step adjoint.

It follows bve_back_rk4_step (bve_time_step.cpp:311-419) with every velocity
evaluation done as the O(N^2) Biot-Savart sum of bve_vel_impl.hpp:54-66 over
all particles (sources displaced by their own displacement, no proxies), the
displacement update of :91-104, the backward RK4 departure search with the
reference's combine sequence (:206-303) and the remap of :38-66.  Point
location uses the uniform leaf grid of each face (the quadtree descent of
interp_funcs_impl.hpp:114-142 agrees with it away from panel edges); the BLI
basis is the second barycentric form on the leaf's Chebyshev-Lobatto grid
(interp_funcs_impl.hpp:11-112) without the node snap.  Test infrastructure
only."""
import math

import numpy as np
import torch


def face_xieta(p):
    """Face (cubed_sphere_transforms_impl.hpp:8-32) and label coordinates (:100-123)."""
    x, y, z = p[..., 0], p[..., 1], p[..., 2]
    ax, ay, az = x.abs(), y.abs(), z.abs()
    fx = (ax >= ay) & (ax >= az)
    fy = ~fx & (ay >= ax) & (ay >= az)
    face = torch.where(fx, torch.where(x >= 0, 0, 2),
                       torch.where(fy, torch.where(y >= 0, 1, 3), torch.where(z >= 0, 4, 5)))
    n1 = torch.stack([y, -x, y, -x, y, -y], -1)
    n2 = torch.stack([z, z, -z, -z, -x, -x], -1)
    d = torch.stack([x, y, x, y, z, z], -1)
    idx = face.unsqueeze(-1)
    xi = torch.atan(n1.gather(-1, idx).squeeze(-1) / d.gather(-1, idx).squeeze(-1))
    eta = torch.atan(n2.gather(-1, idx).squeeze(-1) / d.gather(-1, idx).squeeze(-1))
    return face, xi, eta


def project(p):
    return p / p.norm(dim=-1, keepdim=True)


class StepModel:
    """Geometry of a uniform cubed-sphere particle set [Pl][m*m] (leaf-major,
    node index a*m + c with a along xi) at levels L."""

    def __init__(self, x, y, z, area, levels, eps, omega, dt):
        self.X = torch.tensor(np.stack([x.ravel(), y.ravel(), z.ravel()], 1), dtype=torch.float64)
        self.area = torch.tensor(area.ravel(), dtype=torch.float64)
        self.Pl, m2 = x.shape
        self.m = int(round(math.sqrt(m2)))
        self.n1 = 1 << (levels - 1)
        self.eps, self.omega, self.dt = eps, omega, dt
        m = self.m
        self.cheb = torch.tensor([math.cos(math.pi * i / (m - 1)) for i in range(m)], dtype=torch.float64)
        bw = torch.tensor([(0.5 if i in (0, m - 1) else 1.0) * (-1) ** i for i in range(m)], dtype=torch.float64)
        self.bw = bw
        # leaf of each (face, i, j): from the mean label coordinates of its particles
        lookup = -np.ones((6, self.n1, self.n1), dtype=np.int64)
        f, xi, eta = face_xieta(self.X.view(self.Pl, m2, 3).mean(1))
        h = (math.pi / 2) / self.n1
        i = torch.floor((xi + math.pi / 4) / h).long().numpy()
        j = torch.floor((eta + math.pi / 4) / h).long().numpy()
        for leaf in range(self.Pl):
            lookup[f[leaf], i[leaf], j[leaf]] = leaf
        assert (lookup >= 0).all()
        self.lookup = torch.tensor(lookup)

    def velocity(self, vor_eff, disp=None):
        y = self.X if disp is None else project(self.X + disp)
        q = -(self.area * vor_eff) / (4 * math.pi)
        g = 1.0 / ((1.0 - self.X @ y.T) + self.eps)
        s = g @ (q.unsqueeze(1) * y)
        return torch.cross(self.X, s, dim=1)

    def interp(self, fields, p):
        """BLI of leaf fields [k][Pl*m2] at points p [N][3] -> [k][N]."""
        face, xi, eta = face_xieta(p)
        h = (math.pi / 2) / self.n1
        i = torch.clamp(torch.floor((xi + math.pi / 4) / h).long(), 0, self.n1 - 1)
        j = torch.clamp(torch.floor((eta + math.pi / 4) / h).long(), 0, self.n1 - 1)
        leaf = self.lookup[face, i, j]

        def weights(t, k):
            lo = -math.pi / 4 + k.double() * h
            node = self.cheb.unsqueeze(0) * (0.5 * h) + (lo + 0.5 * h).unsqueeze(1)
            r = self.bw.unsqueeze(0) / (t.unsqueeze(1) - node)
            return r / r.sum(1, keepdim=True)

        wx, we = weights(xi, i), weights(eta, j)
        m = self.m
        out = []
        for f in fields:
            F = f.view(self.Pl, m, m)[leaf]
            out.append(torch.einsum("na,nac,nc->n", wx, F, we))
        return torch.stack(out)

    def step(self, vor):
        dt, om = self.dt, self.omega
        X = self.X
        v = [self.velocity(vor)]
        for tau in (0.5 * dt, 0.5 * dt, dt):
            p = project(X + tau * v[-1])
            new_vor = vor + 2 * om * X[:, 2] - 2 * om * p[:, 2]
            v.append(self.velocity(new_vor, tau * v[-1]))
        k1 = v[3]
        p2 = project(X - 0.5 * dt * k1)
        k2 = self.interp(v[2].T, p2).T
        p3 = project(X - 0.5 * dt * k2)
        k3 = self.interp(v[1].T, p3).T
        p4 = project(X - dt * k3)
        k4 = self.interp(v[0].T, p4).T
        a = project(X - dt / 6.0 * k1)
        a = project(a + dt / 3.0 * k2)
        a = project(a + dt / 3.0 * k3)
        dep = project(a + dt / 6.0 * k4)
        return self.interp([vor], dep)[0] + 2 * om * dep[:, 2] - 2 * om * X[:, 2]
