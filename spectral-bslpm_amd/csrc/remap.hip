// remap.hip -- BVE source displacement, backward RK4 departure points and the
// semi-Lagrangian remap (bve_time_step.cpp:15-309), all in FP64.
//
// One thread per particle (or proxy node).  Point location descends the
// quadtree from the face panel (interp_funcs_impl.hpp:114-142); the
// barycentric basis is evaluated as two 1-D weight vectors (the reference's
// tensor basis is their outer product, interp_funcs_impl.hpp:105-111).  The
// kernels are templated on the particle-grid side M so the weight vectors
// and the m x m gathers stay in registers.  They are gather-bound: per
// particle and stage, one leaf tile of three velocity components
// (3 * M^2 * 8 bytes) is read, mostly from L2 because neighbouring particles
// share leaves.
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
struct Vel12 {
  const double* p[12];  // stage velocities v0x v0y v0z v1x .. v3z
};

// Interpolate NF leaf fields [Pl][M*M] at label point (xi, eta) of leaf `leaf`.
template <int M, int NF>
__device__ inline bool interp_leaf(const DevTree& t, int leaf, double xi, double eta, const double* const* f,
                                   double* out) {
  const DevPanel& P = t.panels[leaf];
  double wx[M], we[M];
  bool ok = bli_w<M>(xi, P.min_xi, P.max_xi, t.cheb_m, wx);
  ok = bli_w<M>(eta, P.min_eta, P.max_eta, t.cheb_m, we) && ok;
  const int64_t base = (int64_t)(leaf - t.offset) * (M * M);
#pragma unroll
  for (int k = 0; k < NF; ++k) out[k] = tensor_gather<M>(f[k] + base, wx, we);
  return ok;
}

// bve_disp_update, bve_time_step.cpp:91-104.
__global__ void k_disp_update(DevTree t, const double* __restrict__ x, const double* __restrict__ y,
                              const double* __restrict__ z, CPtr3 vel, const double* __restrict__ vor, double tau,
                              double omega, double* __restrict__ new_vor, Ptr3 disp) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= t.npart) return;
  {  // member slice (x, y, z are shared by the members)
    const int64_t mp = member_part(t);
    for (int f = 0; f < 3; ++f) {
      vel.p[f] += mp;
      disp.p[f] += mp;
    }
    if (new_vor) {
      vor += mp;
      new_vor += mp;
    }
  }
  const double vx = vel.p[0][u], vy = vel.p[1][u], vz = vel.p[2][u];
  if (new_vor) {  // null for the shallow-water displacement (swe_disp_update, swe_time_step.cpp:194-210)
    double a = x[u] + vx * tau, b = y[u] + vy * tau, c = z[u] + vz * tau;
    project_to_sphere(a, b, c);
    new_vor[u] = vor[u] + 2 * omega * z[u] - 2 * omega * c;
  }
  disp.p[0][u] = tau * vx;
  disp.p[1][u] = tau * vy;
  disp.p[2][u] = tau * vz;
}

// bve_disp_interp, bve_time_step.cpp:131-159: displacement of each non-leaf
// proxy node interpolated from the leaf that contains it (leaf and label
// coordinates precomputed on the host, the tree being fixed in label space).
// Threads take the nodes in leaf order, so neighbouring lanes gather the same
// leaf tile.
template <int M>
__global__ __launch_bounds__(256) void k_disp_interp(DevTree t, CPtr3 dpart, Ptr3 dprox) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)t.offset * t.n) return;
  for (int f = 0; f < 3; ++f) {
    dpart.p[f] += member_part(t);
    dprox.p[f] += member_proxy(t);
  }
  const int64_t v = t.proxy_order[i];
  const int leaf = t.proxy_leaf[v];
  double out[3] = {0, 0, 0};
  if (leaf < 0) {
    flag(t, kErrLocate);
  } else if (!interp_leaf<M, 3>(t, leaf, t.proxy_xieta[2 * v], t.proxy_xieta[2 * v + 1], dpart.p, out)) {
    flag(t, kErrBliBounds);
  }
  dprox.p[0][v] = out[0];
  dprox.p[1][v] = out[1];
  dprox.p[2][v] = out[2];
}

template <int M>
__device__ inline void vel_at(const DevTree& t, double x, double y, double z, const double* const* f, double* k) {
  double xe[2];
  const int leaf = locate_leaf(t, x, y, z, xe);
  if (leaf < 0) {
    flag(t, kErrLocate);
    k[0] = k[1] = k[2] = 0;
    return;
  }
  if (!interp_leaf<M, 3>(t, leaf, xe[0], xe[1], f, k)) flag(t, kErrBliBounds);
}

// bve_find_departure_points, bve_time_step.cpp:206-308.
template <int M>
__global__ __launch_bounds__(128) void k_departure(DevTree t, const double* __restrict__ x, const double* __restrict__ y,
                            const double* __restrict__ z, Vel12 v, double dt, int combine, double* __restrict__ dx,
                            double* __restrict__ dy, double* __restrict__ dz) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= t.npart) return;
  {  // member slice (x, y, z are shared by the members)
    const int64_t mp = member_part(t);
    for (int i = 0; i < 12; ++i) v.p[i] += mp;
    dx += mp;
    dy += mp;
    dz += mp;
  }
  const double X = x[u], Y = y[u], Z = z[u];
  double k1[3] = {v.p[9][u], v.p[10][u], v.p[11][u]}, k2[3], k3[3], k4[3];
  double a = X - 0.5 * dt * k1[0], b = Y - 0.5 * dt * k1[1], c = Z - 0.5 * dt * k1[2];
  project_rcp(a, b, c);
  vel_at<M>(t, a, b, c, v.p + 6, k2);
  a = X - 0.5 * dt * k2[0]; b = Y - 0.5 * dt * k2[1]; c = Z - 0.5 * dt * k2[2];
  project_rcp(a, b, c);
  vel_at<M>(t, a, b, c, v.p + 3, k3);
  a = X - dt * k3[0]; b = Y - dt * k3[1]; c = Z - dt * k3[2];
  project_rcp(a, b, c);
  vel_at<M>(t, a, b, c, v.p + 0, k4);
  if (combine == 0) {
    // the reference's sequence (:288-303)
    a = X - 1.0 / 6.0 * dt * k1[0]; b = Y - 1.0 / 6.0 * dt * k1[1]; c = Z - 1.0 / 6.0 * dt * k1[2];
    project_rcp(a, b, c);
    a -= -1.0 / 3.0 * dt * k2[0]; b -= -1.0 / 3.0 * dt * k2[1]; c -= -1.0 / 3.0 * dt * k2[2];
    project_rcp(a, b, c);
    a -= -1.0 / 3.0 * dt * k3[0]; b -= -1.0 / 3.0 * dt * k3[1]; c -= -1.0 / 3.0 * dt * k3[2];
    project_rcp(a, b, c);
    a -= -1.0 / 6.0 * dt * k4[0]; b -= -1.0 / 6.0 * dt * k4[1]; c -= -1.0 / 6.0 * dt * k4[2];
    project_rcp(a, b, c);
  } else {
    a = X - dt / 6.0 * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0]);
    b = Y - dt / 6.0 * (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1]);
    c = Z - dt / 6.0 * (k1[2] + 2 * k2[2] + 2 * k3[2] + k4[2]);
    project_rcp(a, b, c);
  }
  dx[u] = a; dy[u] = b; dz[u] = c;
}

// bve_departure_to_target (bve_time_step.cpp:38-66; NF = 1, Omega terms) and
// swe_departure_to_target (swe_time_step.cpp:47-74; NF = 3 fields vor, div, h,
// no Omega terms): fields and tracers [Pl][m2][T] interpolated at the
// departure points.
struct Fields3 {
  const double* in[3];
  double* out[3];
};

template <int M, int NF>
__global__ __launch_bounds__(128) void k_dep_to_target(DevTree t, const double* __restrict__ z, const double* __restrict__ dx,
                                const double* __restrict__ dy, const double* __restrict__ dz, Fields3 f,
                                const double* __restrict__ tracers, int ntr, double omega, int coriolis,
                                double* __restrict__ new_tracers) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= t.npart) return;
  {  // member slice (z is shared by the members)
    const int64_t mp = member_part(t);
    dx += mp;
    dy += mp;
    dz += mp;
    for (int q = 0; q < NF; ++q) {
      f.in[q] += mp;
      f.out[q] += mp;
    }
    if (ntr > 0) {
      tracers += mp * ntr;
      new_tracers += mp * ntr;
    }
  }
  double xe[2];
  const double X = dx[u], Y = dy[u], Z = dz[u];
  const int leaf = locate_leaf(t, X, Y, Z, xe);
  if (leaf < 0) {
    flag(t, kErrLocate);
#pragma unroll
    for (int q = 0; q < NF; ++q) f.out[q][u] = 0;
    for (int k = 0; k < ntr; ++k) new_tracers[u * ntr + k] = 0;
    return;
  }
  const DevPanel& P = t.panels[leaf];
  double wx[M], we[M];
  bool ok = bli_w<M>(xe[0], P.min_xi, P.max_xi, t.cheb_m, wx);
  ok = bli_w<M>(xe[1], P.min_eta, P.max_eta, t.cheb_m, we) && ok;
  if (!ok) flag(t, kErrBliBounds);
  const int64_t base = (int64_t)(leaf - t.offset) * (M * M);
#pragma unroll
  for (int q = 0; q < NF; ++q) {
    double v = tensor_gather<M>(f.in[q] + base, wx, we);
    if (coriolis) {
      v += 2 * omega * Z;
      v -= 2 * omega * z[u];
    }
    f.out[q][u] = v;
  }
  for (int k = 0; k < ntr; ++k) {
    double s = 0;
#pragma unroll
    for (int a = 0; a < M; ++a) {
      double r = 0;
#pragma unroll
      for (int c = 0; c < M; ++c) r = fma(tracers[(base + a * M + c) * ntr + k], we[c], r);
      s = fma(wx[a], r, s);
    }
    new_tracers[u * ntr + k] = s;
  }
}

template <int M>
struct DispInterp {
  static void run(const DevTree& t, CPtr3 dp, Ptr3 dx, hipStream_t s, int nmem) {
    int64_t cnt = (int64_t)t.offset * t.n;
    if (cnt > 0) hipLaunchKernelGGL(k_disp_interp<M>, dim3(blocks_for(cnt, 256), nmem), dim3(256), 0, s, t, dp, dx);
  }
};

template <int M>
struct Departure {
  static void run(const DevTree& t, const double* x, const double* y, const double* z, Vel12 v, double dt,
                  int combine, double* dx, double* dy, double* dz, hipStream_t s, int nmem) {
    hipLaunchKernelGGL(k_departure<M>, dim3(blocks_for(t.npart, 128), nmem), dim3(128), 0, s, t, x, y, z, v, dt, combine,
                       dx, dy, dz);
  }
};

template <int M>
struct DepToTarget {
  static void run(const DevTree& t, int nf, const double* z, const double* dx, const double* dy, const double* dz,
                  Fields3 f, const double* tr, int ntr, double omega, int coriolis, double* ntrc, hipStream_t s,
                  int nmem) {
    if (nf == 1)
      hipLaunchKernelGGL((k_dep_to_target<M, 1>), dim3(blocks_for(t.npart, 128), nmem), dim3(128), 0, s, t, z, dx, dy, dz, f,
                         tr, ntr, omega, coriolis, ntrc);
    else
      hipLaunchKernelGGL((k_dep_to_target<M, 3>), dim3(blocks_for(t.npart, 128), nmem), dim3(128), 0, s, t, z, dx, dy, dz, f,
                         tr, ntr, omega, coriolis, ntrc);
  }
};

}  // namespace

void launch_bve_displacement(const DevTree& t, const double* x, const double* y, const double* z,
                             const double* const* vel, const double* vor, double tau, double omega,
                             double* new_vor, double* const* disp_part, double* const* disp_proxy, hipStream_t s,
                             int nmem) {
  CPtr3 v{{vel[0], vel[1], vel[2]}};
  Ptr3 dp{{disp_part[0], disp_part[1], disp_part[2]}};
  hipLaunchKernelGGL(k_disp_update, dim3(blocks_for(t.npart, 256), nmem), dim3(256), 0, s, t, x, y, z, v, vor, tau, omega,
                     new_vor, dp);
  CPtr3 cdp{{disp_part[0], disp_part[1], disp_part[2]}};
  Ptr3 dx{{disp_proxy[0], disp_proxy[1], disp_proxy[2]}};
  dispatch_m<DispInterp>(t.m, t, cdp, dx, s, nmem);
}

void launch_bve_departure(const DevTree& t, const double* x, const double* y, const double* z,
                          const double* const* vels, double dt, int combine, double* dx, double* dy, double* dz,
                          hipStream_t s, int nmem) {
  Vel12 v{};
  for (int i = 0; i < 12; ++i) v.p[i] = vels[i];
  dispatch_m<Departure>(t.m, t, x, y, z, v, dt, combine, dx, dy, dz, s, nmem);
}

void launch_bve_departure_to_target(const DevTree& t, const double* z, const double* dx, const double* dy,
                                    const double* dz, const double* vor, const double* tracers, int ntr,
                                    double omega, double* new_vor, double* new_tracers, hipStream_t s, int nmem) {
  Fields3 f{{vor, nullptr, nullptr}, {new_vor, nullptr, nullptr}};
  dispatch_m<DepToTarget>(t.m, t, 1, z, dx, dy, dz, f, tracers, ntr, omega, 1, new_tracers, s, nmem);
}

void launch_swe_departure_to_target(const DevTree& t, const double* dx, const double* dy, const double* dz,
                                    const double* const* fields, const double* tracers, int ntr,
                                    double* const* new_fields, double* new_tracers, hipStream_t s) {
  Fields3 f{{fields[0], fields[1], fields[2]}, {new_fields[0], new_fields[1], new_fields[2]}};
  dispatch_m<DepToTarget>(t.m, t, 3, nullptr, dx, dy, dz, f, tracers, ntr, 0.0, 0, new_tracers, s, 1);
}

}  // namespace sbslpm
