// remap_adjoint.hip -- reverse-mode kernels of the BVE step's remap side
// (bve_time_step.cpp:15-309), FP64:
//   * k_remap_adj       -- transpose of the remap interpolation in the field;
//   * k_departure_adj   -- the backward RK4 departure search and the remap,
//                          differentiated in the stage velocities, the field
//                          and the departure points (recomputed per thread);
//   * k_disp_interp_adj -- transpose of the proxy-displacement interpolation;
//   * k_stage_adj       -- reverse of the per-stage displacement update.
// Point location is piecewise constant and contributes nothing; the BLI
// interpolant is differentiated in its label coordinates (bli_wd) and those
// in the point (xieta_gradient), the re-projections through
// project_adjoint.  Transposed interpolations are summed per wave and leaf
// in LDS (wave_scatter_leaf) and added to the leaf nodes with FP64 atomics
// (summation order not fixed).  Threads past the end take the last point's
// data and scatter nothing, so every lane reaches the wave-wide scatters.
#include "bli.hpp"

namespace sbslpm {
namespace {

using namespace bli;

struct CPtr3 {
  const double* p[3];
};
struct Ptr3 {
  double* p[3];
};
struct CVel12 {
  const double* p[12];
};
struct Vel12 {
  double* p[12];
};

// Transpose of the BLI interpolation of a leaf field at n points:
// wbar[leaf nodes] += B newbar (the remap at the departure points, n = Pl*m2,
// or an observation operator at any points).
template <int M>
__global__ __launch_bounds__(128) void k_remap_adj(DevTree t, int64_t n, const double* __restrict__ dx,
                                                   const double* __restrict__ dy, const double* __restrict__ dz,
                                                   const double* __restrict__ newbar, double* __restrict__ wbar) {
  __shared__ double lds[2][wave_scatter_lds<M, 1>()];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  const int64_t u = valid ? i : n - 1;
  double xe[2], wx[M], we[M];
  int leaf = locate_leaf(t, dx[u], dy[u], dz[u], xe);
  if (leaf < 0) {
    if (valid) flag(t, kErrLocate);
  } else {
    const DevPanel& P = t.panels[leaf];
    bool ok = bli_w<M>(xe[0], P.min_xi, P.max_xi, t.cheb_m, wx);
    ok = bli_w<M>(xe[1], P.min_eta, P.max_eta, t.cheb_m, we) && ok;
    if (!ok && valid) flag(t, kErrBliBounds);
  }
  if (!valid) leaf = -1;
  const double g = newbar[u];
  double* dst[1] = {wbar};
  wave_scatter_leaf<M, 1>(lds[threadIdx.x >> 6], leaf, wx, we, &g, dst, t.offset);
}

// Forward value of three leaf fields at a point (as vel_at in the departure search).
template <int M>
__device__ inline void interp3_at(const DevTree& t, double x, double y, double z, const double* const* f, double* k) {
  double xe[2];
  const int leaf = locate_leaf(t, x, y, z, xe);
  if (leaf < 0) {
    k[0] = k[1] = k[2] = 0;
    return;
  }
  const DevPanel& P = t.panels[leaf];
  double wx[M], we[M];
  bli_w<M>(xe[0], P.min_xi, P.max_xi, t.cheb_m, wx);
  bli_w<M>(xe[1], P.min_eta, P.max_eta, t.cheb_m, we);
  const int64_t base = (int64_t)(leaf - t.offset) * (M * M);
#pragma unroll
  for (int q = 0; q < 3; ++q) k[q] = tensor_gather<M>(f[q] + base, wx, we);
}

struct CF3 {
  const double* p[3];
};
struct F3 {
  double* p[3];
};

// Reverse of k_departure followed by the remap of NF fields for one arrival
// particle: k_dep_to_target (NF = 1, the Omega terms included) or the
// three-field remap of the shallow-water step (NF = 3, omega = 0).  Inputs:
// the four stage velocities v (v0..v3, as in the departure search), the
// fields before the remap and the cotangents of the fields after it.
// fbar and vbar[0..8] (v0..v2) are accumulated, vbar[9..11] (v3, read only at
// the arrival node) assigned.
template <int M, int NF>
__global__ __launch_bounds__(128) void k_departure_adj(DevTree t, const double* __restrict__ x,
                                                       const double* __restrict__ y, const double* __restrict__ z,
                                                       CVel12 v, CF3 fld, CF3 nbar, double dt, int combine,
                                                       double omega, Vel12 vbar, F3 fbar) {
  __shared__ double lds[2][wave_scatter_lds<M, 3>()];
  double* wl = lds[threadIdx.x >> 6];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < t.npart;
  const int64_t u = valid ? i : t.npart - 1;
  const double X = x[u], Y = y[u], Z = z[u];
  // forward, the arithmetic of k_departure
  double k1[3] = {v.p[9][u], v.p[10][u], v.p[11][u]}, k2[3], k3[3], k4[3];
  double w2[3] = {X - 0.5 * dt * k1[0], Y - 0.5 * dt * k1[1], Z - 0.5 * dt * k1[2]};
  double a = w2[0], b = w2[1], c = w2[2];
  project_rcp(a, b, c);
  const double p2[3] = {a, b, c};
  interp3_at<M>(t, a, b, c, v.p + 6, k2);
  double w3[3] = {X - 0.5 * dt * k2[0], Y - 0.5 * dt * k2[1], Z - 0.5 * dt * k2[2]};
  a = w3[0]; b = w3[1]; c = w3[2];
  project_rcp(a, b, c);
  const double p3[3] = {a, b, c};
  interp3_at<M>(t, a, b, c, v.p + 3, k3);
  double w4[3] = {X - dt * k3[0], Y - dt * k3[1], Z - dt * k3[2]};
  a = w4[0]; b = w4[1]; c = w4[2];
  project_rcp(a, b, c);
  const double p4[3] = {a, b, c};
  interp3_at<M>(t, a, b, c, v.p + 0, k4);
  double c1[3], c2[3], c3[3], c4[3];  // pre-projection points of the combine
  if (combine == 0) {
    a = X - 1.0 / 6.0 * dt * k1[0]; b = Y - 1.0 / 6.0 * dt * k1[1]; c = Z - 1.0 / 6.0 * dt * k1[2];
    c1[0] = a; c1[1] = b; c1[2] = c;
    project_rcp(a, b, c);
    a -= -1.0 / 3.0 * dt * k2[0]; b -= -1.0 / 3.0 * dt * k2[1]; c -= -1.0 / 3.0 * dt * k2[2];
    c2[0] = a; c2[1] = b; c2[2] = c;
    project_rcp(a, b, c);
    a -= -1.0 / 3.0 * dt * k3[0]; b -= -1.0 / 3.0 * dt * k3[1]; c -= -1.0 / 3.0 * dt * k3[2];
    c3[0] = a; c3[1] = b; c3[2] = c;
    project_rcp(a, b, c);
    a -= -1.0 / 6.0 * dt * k4[0]; b -= -1.0 / 6.0 * dt * k4[1]; c -= -1.0 / 6.0 * dt * k4[2];
    c4[0] = a; c4[1] = b; c4[2] = c;
    project_rcp(a, b, c);
  } else {
    a = X - dt / 6.0 * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0]);
    b = Y - dt / 6.0 * (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1]);
    c = Z - dt / 6.0 * (k1[2] + 2 * k2[2] + 2 * k3[2] + k4[2]);
    c4[0] = a; c4[1] = b; c4[2] = c;
    project_rcp(a, b, c);
  }
  double wx[M], we[M];
  // remap: new_q = BLI[f_q](dep) (+ 2 Omega dep_z - 2 Omega z for the vorticity, NF = 1)
  double g[NF];
#pragma unroll
  for (int q = 0; q < NF; ++q) g[q] = nbar.p[q][u];
  double db[3] = {0, 0, 2 * omega * g[0]};
  {
    int leaf = interp_grad<M, NF>(t, a, b, c, fld.p, g, db, wx, we, valid);
    wave_scatter_leaf<M, NF>(wl, valid ? leaf : -1, wx, we, g, fbar.p, t.offset);
  }
  // combine
  double kb1[3], kb2[3], kb3[3], kb4[3];
  if (combine == 0) {
    project_adjoint(c4[0], c4[1], c4[2], db);
    for (int k = 0; k < 3; ++k) kb4[k] = dt / 6.0 * db[k];
    project_adjoint(c3[0], c3[1], c3[2], db);
    for (int k = 0; k < 3; ++k) kb3[k] = dt / 3.0 * db[k];
    project_adjoint(c2[0], c2[1], c2[2], db);
    for (int k = 0; k < 3; ++k) kb2[k] = dt / 3.0 * db[k];
    project_adjoint(c1[0], c1[1], c1[2], db);
    for (int k = 0; k < 3; ++k) kb1[k] = -dt / 6.0 * db[k];
  } else {
    project_adjoint(c4[0], c4[1], c4[2], db);
    for (int k = 0; k < 3; ++k) {
      kb1[k] = -dt / 6.0 * db[k];
      kb2[k] = -dt / 3.0 * db[k];
      kb3[k] = -dt / 3.0 * db[k];
      kb4[k] = -dt / 6.0 * db[k];
    }
  }
  // k4 = BLI[v0](P(X - dt k3))
  double pb[3] = {0, 0, 0};
  int leaf = interp_grad<M, 3>(t, p4[0], p4[1], p4[2], v.p + 0, kb4, pb, wx, we, valid);
  wave_scatter_leaf<M, 3>(wl, valid ? leaf : -1, wx, we, kb4, vbar.p + 0, t.offset);
  project_adjoint(w4[0], w4[1], w4[2], pb);
  for (int k = 0; k < 3; ++k) kb3[k] -= dt * pb[k];
  // k3 = BLI[v1](P(X - dt/2 k2))
  pb[0] = pb[1] = pb[2] = 0;
  leaf = interp_grad<M, 3>(t, p3[0], p3[1], p3[2], v.p + 3, kb3, pb, wx, we, valid);
  wave_scatter_leaf<M, 3>(wl, valid ? leaf : -1, wx, we, kb3, vbar.p + 3, t.offset);
  project_adjoint(w3[0], w3[1], w3[2], pb);
  for (int k = 0; k < 3; ++k) kb2[k] -= 0.5 * dt * pb[k];
  // k2 = BLI[v2](P(X - dt/2 k1))
  pb[0] = pb[1] = pb[2] = 0;
  leaf = interp_grad<M, 3>(t, p2[0], p2[1], p2[2], v.p + 6, kb2, pb, wx, we, valid);
  wave_scatter_leaf<M, 3>(wl, valid ? leaf : -1, wx, we, kb2, vbar.p + 6, t.offset);
  project_adjoint(w2[0], w2[1], w2[2], pb);
  for (int k = 0; k < 3; ++k) kb1[k] -= 0.5 * dt * pb[k];
  // k1 = v3 at the arrival node
  if (valid) {
    vbar.p[9][u] = kb1[0];
    vbar.p[10][u] = kb1[1];
    vbar.p[11][u] = kb1[2];
  }
}

// Transpose of k_disp_interp: dbar_part[leaf nodes] += B dbar_prox (non-leaf
// proxies, taken in leaf order so that a wave mostly shares one leaf).
template <int M>
__global__ __launch_bounds__(256) void k_disp_interp_adj(DevTree t, CPtr3 dprox, Ptr3 dpart) {
  __shared__ double lds[4][wave_scatter_lds<M, 3>()];
  const int64_t cnt = (int64_t)t.offset * t.n;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < cnt;
  const int64_t v = t.proxy_order[valid ? i : cnt - 1];
  int leaf = valid ? t.proxy_leaf[v] : -1;  // a failed location was flagged by the forward interpolation
  double wx[M], we[M];
  const double g[3] = {dprox.p[0][v], dprox.p[1][v], dprox.p[2][v]};
  if (leaf >= 0) {
    const DevPanel& P = t.panels[leaf];
    bli_w<M>(t.proxy_xieta[2 * v], P.min_xi, P.max_xi, t.cheb_m, wx);
    bli_w<M>(t.proxy_xieta[2 * v + 1], P.min_eta, P.max_eta, t.cheb_m, we);
  }
  wave_scatter_leaf<M, 3>(lds[threadIdx.x >> 6], leaf, wx, we, g, dpart.p, t.offset);
}

// Reverse of bve_disp_update (bve_time_step.cpp:91-104) for one stage, given
// qbar = dL/dq_part (q_part = area * new_vor) and dbar = dL/d disp_part:
//   vor_bar += area qbar,
//   vbar += tau (dbar - 2 Omega area qbar dP_z(x + tau v)/dw).
// Without vbar only the charge part is applied (the undisplaced stage).
__global__ void k_stage_adj(DevTree t, const double* __restrict__ x, const double* __restrict__ y,
                            const double* __restrict__ z, CPtr3 vel, double tau, double omega,
                            const double* __restrict__ area, const double* __restrict__ qbar, CPtr3 dbar,
                            double* __restrict__ vor_bar, Ptr3 vbar) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= t.npart) return;
  const double nb = area[u] * qbar[u];
  vor_bar[u] += nb;
  if (!vbar.p[0]) return;
  double g[3] = {0, 0, -2 * omega * nb};
  project_adjoint(x[u] + vel.p[0][u] * tau, y[u] + vel.p[1][u] * tau, z[u] + vel.p[2][u] * tau, g);
  for (int f = 0; f < 3; ++f) vbar.p[f][u] += tau * (dbar.p[f][u] + g[f]);
}

template <int M>
struct RemapAdj {
  static void run(const DevTree& t, int64_t n, const double* dx, const double* dy, const double* dz, const double* nb,
                  double* wb, hipStream_t s) {
    hipLaunchKernelGGL(k_remap_adj<M>, dim3(blocks_for(n, 128)), dim3(128), 0, s, t, n, dx, dy, dz, nb, wb);
  }
};

template <int M>
struct DepartureAdj {
  static void run(const DevTree& t, const double* x, const double* y, const double* z, CVel12 v, int nf, CF3 f,
                  CF3 nb, double dt, int combine, double omega, Vel12 vb, F3 fb, hipStream_t s) {
    if (nf == 1)
      hipLaunchKernelGGL((k_departure_adj<M, 1>), dim3(blocks_for(t.npart, 128)), dim3(128), 0, s, t, x, y, z, v, f,
                         nb, dt, combine, omega, vb, fb);
    else
      hipLaunchKernelGGL((k_departure_adj<M, 3>), dim3(blocks_for(t.npart, 128)), dim3(128), 0, s, t, x, y, z, v, f,
                         nb, dt, combine, omega, vb, fb);
  }
};

template <int M>
struct DispInterpAdj {
  static void run(const DevTree& t, CPtr3 dx, Ptr3 dp, hipStream_t s) {
    const int64_t cnt = (int64_t)t.offset * t.n;
    if (cnt > 0) hipLaunchKernelGGL(k_disp_interp_adj<M>, dim3(blocks_for(cnt, 256)), dim3(256), 0, s, t, dx, dp);
  }
};

}  // namespace

void launch_remap_adjoint(const DevTree& t, const double* dx, const double* dy, const double* dz, const double* newbar,
                          double* wbar, hipStream_t s) {
  dispatch_m<RemapAdj>(t.m, t, t.npart, dx, dy, dz, newbar, wbar, s);
}

void launch_interp_points_adjoint(const DevTree& t, int64_t n, const double* x, const double* y, const double* z,
                                  const double* out_bar, double* field_bar, hipStream_t s) {
  if (n > 0) dispatch_m<RemapAdj>(t.m, t, n, x, y, z, out_bar, field_bar, s);
}

void launch_departure_adjoint(const DevTree& t, const double* x, const double* y, const double* z,
                              const double* const* vels, const double* vor, const double* new_vor_bar, double dt,
                              int combine, double omega, double* const* vel_bar, double* vor_bar, hipStream_t s) {
  CVel12 v{};
  Vel12 vb{};
  for (int i = 0; i < 12; ++i) {
    v.p[i] = vels[i];
    vb.p[i] = vel_bar[i];
  }
  CF3 f{{vor, nullptr, nullptr}}, nb{{new_vor_bar, nullptr, nullptr}};
  F3 fb{{vor_bar, nullptr, nullptr}};
  dispatch_m<DepartureAdj>(t.m, t, x, y, z, v, 1, f, nb, dt, combine, omega, vb, fb, s);
}

void launch_departure_adjoint3(const DevTree& t, const double* x, const double* y, const double* z,
                               const double* const* vels, const double* const* fields, const double* const* new_bar,
                               double dt, int combine, double* const* vel_bar, double* const* field_bar,
                               hipStream_t s) {
  CVel12 v{};
  Vel12 vb{};
  for (int i = 0; i < 12; ++i) {
    v.p[i] = vels[i];
    vb.p[i] = vel_bar[i];
  }
  CF3 f{}, nb{};
  F3 fb{};
  for (int q = 0; q < 3; ++q) {
    f.p[q] = fields[q];
    nb.p[q] = new_bar[q];
    fb.p[q] = field_bar[q];
  }
  dispatch_m<DepartureAdj>(t.m, t, x, y, z, v, 3, f, nb, dt, combine, 0.0, vb, fb, s);
}

void launch_disp_interp_adjoint(const DevTree& t, const double* const* dbar_proxy, double* const* dbar_part,
                                hipStream_t s) {
  CPtr3 dx{{dbar_proxy[0], dbar_proxy[1], dbar_proxy[2]}};
  Ptr3 dp{{dbar_part[0], dbar_part[1], dbar_part[2]}};
  dispatch_m<DispInterpAdj>(t.m, t, dx, dp, s);
}

void launch_stage_adjoint(const DevTree& t, const double* x, const double* y, const double* z,
                          const double* const* vel, double tau, double omega, const double* area, const double* qbar,
                          const double* const* dbar, double* vor_bar, double* const* vbar, hipStream_t s) {
  CPtr3 v{}, d{};
  Ptr3 vb{};
  if (vbar) {
    for (int f = 0; f < 3; ++f) {
      v.p[f] = vel[f];
      d.p[f] = dbar[f];
      vb.p[f] = vbar[f];
    }
  }
  hipLaunchKernelGGL(k_stage_adj, dim3(blocks_for(t.npart, 256)), dim3(256), 0, s, t, x, y, z, v, tau, omega, area,
                     qbar, d, vor_bar, vb);
}

}  // namespace sbslpm
