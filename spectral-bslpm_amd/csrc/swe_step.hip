// swe_step.hip -- per-leaf and per-particle kernels of the shallow-water RK4
// step (swe_back_rk4_step, swe_time_step.cpp:606-811):
//   * k_topography: effective height with the cone mountain (topography.cpp:6-56);
//   * k_panel_lap<M>: spectral surface Laplacian on each leaf's particle grid
//     (panel_laplacian, deriv_funcs_impl.hpp:338-389);
//   * k_swe_tendency<M>: vorticity / divergence / height tendencies with the
//     velocity gradient of each leaf (swe_tendency_computation,
//     swe_time_step.cpp:76-161; single_panel_grad, deriv_funcs_impl.hpp:272-305);
//   * k_apply_tend: state + tau * tendency and the area update
//     (swe_apply_tendencies, swe_time_step.cpp:163-192);
//   * k_height_corr<M>: the surface-Laplacian correction at moved points
//     (swe_height_correction, swe_time_step.cpp:558-604).
// Per-leaf kernels run one workgroup per leaf with the leaf's m x m values in
// LDS, one thread per particle; node coordinates are computed once per
// workgroup.  All arithmetic is FP64.
#include "bli.hpp"

namespace sbslpm {
namespace {

using namespace bli;

constexpr double kPiD = 3.14159265358979323846;

struct CP3 {
  const double* p[3];
};
struct P3 {
  double* p[3];
};

// Barycentric weight of Chebyshev node i of degree D (deriv_funcs_impl.hpp:24-30).
__device__ inline double bary_w(int i, int D) {
  if (i == 0) return 0.5;
  if (i == D) return (D & 1) ? -0.5 : 0.5;
  return (i & 1) ? -1.0 : 1.0;
}

// Derivative at node (a, b) of the leaf tile s[M*M] along the first index
// (dir 0, xi) or the second (dir 1, eta), in the reference's difference form
// (bli_deriv_xi / bli_deriv_eta, deriv_funcs_impl.hpp:8-92).
template <int M>
__device__ inline double tile_deriv(const double* s, int a, int b, int dir, const double* nodes) {
  constexpr int D = M - 1;
  const int i = dir == 0 ? a : b;
  const double xi = nodes[i], fi = s[a * M + b];
  double acc = 0;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    if (k == i) continue;
    const double fk = dir == 0 ? s[k * M + b] : s[a * M + k];
    acc += (fk - fi) * bary_w(k, D) / (xi - nodes[k]);
  }
  return acc / bary_w(i, D);
}

// Node coordinates (radians) of a leaf: cos(pi i / D) * half + mid
// (deriv_funcs_impl.hpp:14-15, 31).
template <int M>
__device__ inline void leaf_nodes(const DevPanel& P, double* xn, double* en) {
  constexpr int D = M - 1;
  const int o = threadIdx.x;
  if (o < M) {
    xn[o] = cos(kPiD * o / D) * (0.5 * (P.max_xi - P.min_xi)) + 0.5 * (P.max_xi + P.min_xi);
    en[o] = cos(kPiD * o / D) * (0.5 * (P.max_eta - P.min_eta)) + 0.5 * (P.max_eta + P.min_eta);
  }
}

template <int M>
__global__ __launch_bounds__(128) void k_panel_lap(DevTree t, const double* __restrict__ f, double* __restrict__ lap) {
  __shared__ double sf[M * M], sx[M * M], se[M * M], xn[M], en[M];
  const int leaf = blockIdx.x, o = threadIdx.x, a = o / M, b = o % M;
  const DevPanel& P = t.panels[leaf + t.offset];
  leaf_nodes<M>(P, xn, en);
  if (o < M * M) sf[o] = f[(int64_t)leaf * (M * M) + o];
  __syncthreads();
  if (o < M * M) {
    sx[o] = tile_deriv<M>(sf, a, b, 0, xn);
    se[o] = tile_deriv<M>(sf, a, b, 1, en);
  }
  __syncthreads();
  if (o < M * M) {
    const double fxx = tile_deriv<M>(sx, a, b, 0, xn), fxe = tile_deriv<M>(se, a, b, 0, xn),
                 fee = tile_deriv<M>(se, a, b, 1, en);
    const double X = tan(xn[a]), Y = tan(en[b]);
    const double C2 = 1 + X * X, D2 = 1 + Y * Y, delta = 1 + X * X + Y * Y;
    lap[(int64_t)leaf * (M * M) + o] = delta * (fxx / C2 + fee / D2 + 2.0 * X * Y / (C2 * D2) * fxe);
  }
}

// xyzvec_from_xietavec (cubed_sphere_transforms_impl.hpp:153-239): metric
// combination, the 2x2 (colatitude, longitude) system solved by Cramer's rule,
// rotation to Cartesian components.  At the pole nodes of faces 4/5 the
// reference returns (g_xi, g_eta, 0); with pole_fix the surface gradient.
__device__ inline void xyzvec_from_xietavec(double xd, double ed, int face, double xi, double eta, bool pole_fix,
                                            double* out) {
  const double X = tan(xi), Y = tan(eta), X2 = X * X, Y2 = Y * Y;
  const double C = sqrt(1 + X2), D = sqrt(1 + Y2);
  const double gx = D * xd + X * Y / D * ed, ge = X * Y / C * xd + C * ed;
  const double isqd = 1.0 / sqrt(1 + X2 + Y2);
  double xyz[3];
  xyz_from_xieta(xi, eta, face, xyz);
  const double x = xyz[0], y = xyz[1], z = xyz[2];
  const double sqc = sqrt(x * x + y * y);
  double m00, m01, m10, m11;
  if (face <= 3) {
    m00 = 0; m01 = C * D * isqd; m10 = -1.0; m11 = X * Y * isqd;
  } else {
    if (fabs(xi) < 1e-16 && fabs(eta) < 1e-16) {
      if (!pole_fix) { out[0] = gx; out[1] = ge; }
      else if (face == 4) { out[0] = -ge; out[1] = gx; }
      else { out[0] = ge; out[1] = gx; }
      out[2] = 0;
      return;
    }
    const double isqc2 = 1.0 / sqrt(X2 + Y2), sg = face == 4 ? 1.0 : -1.0;
    m00 = sg * isqc2 * D * X; m01 = -sg * isqc2 * D * Y * isqd;
    m10 = sg * isqc2 * C * Y; m11 = sg * isqc2 * C * X * isqd;
  }
  const double det = m00 * m11 - m01 * m10;
  const double colat = (gx * m11 - m01 * ge) / det, lon = (m00 * ge - m10 * gx) / det;
  out[0] = x * z / sqc * colat - y / sqc * lon;
  out[1] = y * z / sqc * colat + x / sqc * lon;
  out[2] = -sqc * colat;
}

// One workgroup per leaf.  rule 0 (the reference): vectors converted with
// panels[leaf].face (leaf-local index into the global panel array, as
// swe_time_step.cpp:109-111 passes it) and the metric applied twice
// (deriv_funcs_impl.hpp:291-292, then again inside the conversion); rule 1: the
// leaf's own face, the raw derivatives, the pole nodes mapped correctly.
template <int M>
__global__ __launch_bounds__(128) void k_swe_tendency(DevTree t, const double* __restrict__ px,
                                                      const double* __restrict__ py, const double* __restrict__ pz,
                                                      CP3 vel, const double* __restrict__ vor,
                                                      const double* __restrict__ div, const double* __restrict__ h,
                                                      const double* __restrict__ lap, double omega, int rule,
                                                      P3 out) {
  __shared__ double sv[3][M * M], xn[M], en[M];
  const int leaf = blockIdx.x, o = threadIdx.x, a = o / M, b = o % M;
  const DevPanel& P = t.panels[leaf + t.offset];
  leaf_nodes<M>(P, xn, en);
  const int64_t k = (int64_t)leaf * (M * M) + o;
  if (o < M * M)
    for (int c = 0; c < 3; ++c) sv[c][o] = vel.p[c][k];
  __syncthreads();
  if (o >= M * M) return;
  const int face = rule == 0 ? t.panels[leaf].face : P.face;
  const double xi = xn[a], eta = en[b];
  const double X = tan(xi), Y = tan(eta), C = sqrt(1 + X * X), D = sqrt(1 + Y * Y);
  double G[3][3];
  for (int c = 0; c < 3; ++c) {
    const double fx = tile_deriv<M>(sv[c], a, b, 0, xn), fe = tile_deriv<M>(sv[c], a, b, 1, en);
    const double xd = rule == 0 ? D * fx + X * Y / D * fe : fx;
    const double ed = rule == 0 ? X * Y / C * fx + C * fe : fe;
    xyzvec_from_xietavec(xd, ed, face, xi, eta, rule != 0, G[c]);
  }
  const double x = px[k], y = py[k], z = pz[k];
  const double ux = sv[0][o], uy = sv[1][o], uz = sv[2][o];
  const double vk = vor[k], dk = div[k];
  out.p[0][k] = -dk * (vk + 2.0 * omega * z) - 2.0 * omega * uz;
  double d = 2.0 * omega * z * vk;
  d -= G[0][0] * G[0][0] + G[0][1] * G[1][0] + G[0][2] * G[2][0] + G[1][0] * G[0][1] + G[1][1] * G[1][1] +
       G[1][2] * G[2][1] + G[2][0] * G[0][2] + G[2][1] * G[1][2] + G[2][2] * G[2][2];
  d -= lap[k];
  d -= ux * ux + uy * uy + uz * uz;
  if (x * x + y * y > 1e-16) d -= 2.0 * omega * (-y * ux + x * uy);
  out.p[1][k] = d;
  out.p[2][k] = -h[k] * dk;
}

// cone_mountain (topography.cpp:6-45): eff = h + h0 exp(-16 r^2) at P(x + disp).
// The displacement of leaf-local i, node j: rule 0 reads row i of the full
// [P][n] displacement array (the reference's index; leaf rows live at
// i + offset there), rule 1 the leaf's own row.
__global__ void k_topography(DevTree t, const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ z, CP3 dpart, CP3 dprox, const double* __restrict__ h,
                             int rule, double* __restrict__ eff) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= t.npart) return;
  const int i = (int)(k / t.m2), j = (int)(k % t.m2);
  double d[3];
  for (int f = 0; f < 3; ++f) {
    if (rule != 0 || i >= t.offset) {
      const int64_t kk = rule != 0 ? k : (int64_t)(i - t.offset) * t.m2 + j;
      d[f] = dpart.p[f] ? dpart.p[f][kk] : 0.0;
    } else {
      d[f] = dprox.p[f] ? dprox.p[f][(int64_t)i * t.n + j] : 0.0;
    }
  }
  double a = x[k] + d[0], b = y[k] + d[1], c = z[k] + d[2];
  project_to_sphere(a, b, c);
  const double lat = kPiD / 2.0 - atan2(sqrt(a * a + b * b), c);
  double lon = atan2(b, a);
  lon += 2 * kPiD;
  lon = fmod(lon, 2 * kPiD);
  const double dl = lon - 0.5 * kPiD, dt = lat - kPiD / 6.0;
  eff[k] = h[k] + (2000.0 / 6371000.0) * exp(-16.0 * (dl * dl + dt * dt));
}

// out_q = f_q + tau ft_q (q = vor, div, h); new_area = area (1 + tau div) when area is given.
__global__ void k_apply_tend(int64_t n, CP3 f, CP3 ft, const double* __restrict__ area, double tau, P3 out,
                             double* __restrict__ new_area) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  for (int q = 0; q < 3; ++q) out.p[q][k] = f.p[q][k] + tau * ft.p[q][k];
  if (area) new_area[k] = area[k] * (1.0 + tau * f.p[1][k]);
}

// swe_height_correction: at P(x + tau vel) the BLI of h + tau h_t over the
// containing leaf; lap -= 8 new_h / 0.1.
template <int M>
__global__ __launch_bounds__(128) void k_height_corr(DevTree t, const double* __restrict__ x,
                                                     const double* __restrict__ y, const double* __restrict__ z,
                                                     CP3 vel, const double* __restrict__ h,
                                                     const double* __restrict__ ht, double tau,
                                                     double* __restrict__ lap) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= t.npart) return;
  double a = x[k] + tau * vel.p[0][k], b = y[k] + tau * vel.p[1][k], c = z[k] + tau * vel.p[2][k];
  project_to_sphere(a, b, c);
  double xe[2];
  const int leaf = locate_leaf(t, a, b, c, xe);
  if (leaf < 0) {
    flag(t, kErrLocate);
    return;
  }
  const DevPanel& P = t.panels[leaf];
  double wx[M], we[M];
  bool ok = bli_w<M>(xe[0], P.min_xi, P.max_xi, t.cheb_m, wx);
  ok = bli_w<M>(xe[1], P.min_eta, P.max_eta, t.cheb_m, we) && ok;
  if (!ok) flag(t, kErrBliBounds);
  const int64_t base = (int64_t)(leaf - t.offset) * (M * M);
  double v = 0;
#pragma unroll
  for (int i = 0; i < M; ++i) {
    double r = 0;
#pragma unroll
    for (int j = 0; j < M; ++j) r += (h[base + i * M + j] + tau * ht[base + i * M + j]) * we[j];
    v += wx[i] * r;
  }
  lap[k] -= 8.0 * v / 0.1;
}

template <int M>
struct PanelLap {
  static void run(const DevTree& t, const double* f, double* lap, hipStream_t s) {
    hipLaunchKernelGGL(k_panel_lap<M>, dim3(t.Pl), dim3(128), 0, s, t, f, lap);
  }
};

template <int M>
struct Tendency {
  static void run(const DevTree& t, const double* x, const double* y, const double* z, CP3 vel, const double* vor,
                  const double* div, const double* h, const double* lap, double omega, int rule, P3 out,
                  hipStream_t s) {
    hipLaunchKernelGGL(k_swe_tendency<M>, dim3(t.Pl), dim3(128), 0, s, t, x, y, z, vel, vor, div, h, lap, omega,
                       rule, out);
  }
};

template <int M>
struct HeightCorr {
  static void run(const DevTree& t, const double* x, const double* y, const double* z, CP3 vel, const double* h,
                  const double* ht, double tau, double* lap, hipStream_t s) {
    hipLaunchKernelGGL(k_height_corr<M>, dim3(blocks_for(t.npart, 128)), dim3(128), 0, s, t, x, y, z, vel, h, ht,
                       tau, lap);
  }
};

CP3 cp3(const double* const* p) {
  CP3 r{};
  for (int i = 0; i < 3; ++i) r.p[i] = p ? p[i] : nullptr;
  return r;
}
P3 p3(double* const* p) {
  P3 r{};
  for (int i = 0; i < 3; ++i) r.p[i] = p[i];
  return r;
}

}  // namespace

void launch_panel_laplacian(const DevTree& t, const double* f, double* lap, hipStream_t s) {
  if (t.Pl > 0) dispatch_m<PanelLap>(t.m, t, f, lap, s);
}

void launch_swe_tendency(const DevTree& t, const double* x, const double* y, const double* z,
                         const double* const* vel, const double* vor, const double* div, const double* h,
                         const double* lap, double omega, int rule, double* const* out, hipStream_t s) {
  if (t.Pl > 0) dispatch_m<Tendency>(t.m, t, x, y, z, cp3(vel), vor, div, h, lap, omega, rule, p3(out), s);
}

void launch_topography(const DevTree& t, const double* x, const double* y, const double* z,
                       const double* const* disp_part, const double* const* disp_proxy, const double* h, int rule,
                       double* eff, hipStream_t s) {
  if (t.npart > 0)
    hipLaunchKernelGGL(k_topography, dim3(blocks_for(t.npart, 256)), dim3(256), 0, s, t, x, y, z, cp3(disp_part),
                       cp3(disp_proxy), h, rule, eff);
}

void launch_apply_tendencies(int64_t n, const double* const* f, const double* const* ft, const double* area,
                             double tau, double* const* out, double* new_area, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_apply_tend, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, cp3(f), cp3(ft), area, tau, p3(out),
                       new_area);
}

void launch_height_correction(const DevTree& t, const double* x, const double* y, const double* z,
                              const double* const* vel, const double* h, const double* ht, double tau, double* lap,
                              hipStream_t s) {
  if (t.npart > 0) dispatch_m<HeightCorr>(t.m, t, x, y, z, cp3(vel), h, ht, tau, lap, s);
}

// unify_boundary_vals (initialize_cubed_sphere.cpp:328-370): every copy of a
// unique node (two_d_to_1d, cubed_sphere_2d_to_1d) takes the mean of its
// copies.  As in the reference's accumulate_vals the copies are summed with
// FP64 atomics, so the order of the (at most four) terms is not fixed.
// Entries whose index lies outside [0, n1d) are left as they are.
namespace {
__global__ void k_unify_sum(int64_t n, const int* __restrict__ map, const double* __restrict__ vals, int64_t n1d,
                            double* __restrict__ sum1d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = map[i];
  if (k >= 0 && k < n1d) atomicAdd(sum1d + k, vals[i]);
}
__global__ void k_unify_mean(int64_t n, const int* __restrict__ map, const int* __restrict__ counts, int64_t n1d,
                             const double* __restrict__ sum1d, double* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = map[i];
  if (k >= 0 && k < n1d) vals[i] = sum1d[k] / counts[k];
}
}  // namespace

void launch_unify_boundary_vals(int64_t n, const int* map, const int* counts, int64_t n1d, double* sum1d,
                                double* vals, hipStream_t s) {
  launch_fill_zero(sum1d, n1d, s);
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_unify_sum, dim3(blocks), dim3(256), 0, s, n, map, vals, n1d, sum1d);
  hipLaunchKernelGGL(k_unify_mean, dim3(blocks), dim3(256), 0, s, n, map, counts, n1d, sum1d, vals);
}

}  // namespace sbslpm
