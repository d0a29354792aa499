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
// (deriv_funcs_impl.hpp:14-15, 31); node o of each direction, o < M.
template <int M>
__device__ inline void leaf_nodes(const DevPanel& P, double* xn, double* en, int o) {
  constexpr int D = M - 1;
  if (o < M) {
    xn[o] = cos(kPiD * o / D) * (0.5 * (P.max_xi - P.min_xi)) + 0.5 * (P.max_xi + P.min_xi);
    en[o] = cos(kPiD * o / D) * (0.5 * (P.max_eta - P.min_eta)) + 0.5 * (P.max_eta + P.min_eta);
  }
}

// The leaf kernels below keep a leaf's m x m tile in LDS and give each of its
// nodes a thread; a 256-thread workgroup holds leaf_slots<M>() leaves (7 at
// m = 6, 4 at m = 8), one LDS slot each.  Threads past the last slot or the
// last leaf take part in the barriers only.
template <int M>
constexpr int leaf_slots() {
  return 256 / (M * M);
}
struct LeafSlot {
  int slot, leaf, o, a, b;
  bool live;
};
template <int M>
__device__ inline LeafSlot leaf_slot(const DevTree& t) {
  constexpr int M2 = M * M, S = leaf_slots<M>();
  static_assert(S >= 1, "leaf grid larger than a workgroup");
  LeafSlot L;
  L.slot = threadIdx.x / M2;
  L.o = threadIdx.x - L.slot * M2;
  L.leaf = blockIdx.x * S + L.slot;
  L.live = L.slot < S && L.leaf < t.Pl;
  if (!L.live) L.slot = 0;  // LDS slot 0 is only read through L.live-guarded code
  L.a = L.o / M;
  L.b = L.o % M;
  return L;
}
template <int M>
inline unsigned leaf_blocks(const DevTree& t) {
  return (unsigned)((t.Pl + leaf_slots<M>() - 1) / leaf_slots<M>());
}

template <int M>
__global__ __launch_bounds__(256) void k_panel_lap(DevTree t, const double* __restrict__ f, double* __restrict__ lap) {
  constexpr int M2 = M * M, S = leaf_slots<M>();
  __shared__ double sf_[S][M2], sx_[S][M2], se_[S][M2], xn_[S][M], en_[S][M];
  const LeafSlot L = leaf_slot<M>(t);
  const int a = L.a, b = L.b, o = L.o;
  double *sf = sf_[L.slot], *sx = sx_[L.slot], *se = se_[L.slot], *xn = xn_[L.slot], *en = en_[L.slot];
  const int64_t k = (int64_t)L.leaf * M2 + o;
  if (L.live) {
    leaf_nodes<M>(t.panels[L.leaf + t.offset], xn, en, o);
    sf[o] = f[k];
  }
  __syncthreads();
  if (L.live) {
    sx[o] = tile_deriv<M>(sf, a, b, 0, xn);
    se[o] = tile_deriv<M>(sf, a, b, 1, en);
  }
  __syncthreads();
  if (L.live) {
    const double fxx = tile_deriv<M>(sx, a, b, 0, xn), fxe = tile_deriv<M>(se, a, b, 0, xn),
                 fee = tile_deriv<M>(se, a, b, 1, en);
    const double X = tan(xn[a]), Y = tan(en[b]);
    const double C2 = 1 + X * X, D2 = 1 + Y * Y, delta = 1 + X * X + Y * Y;
    lap[k] = delta * (fxx / C2 + fee / D2 + 2.0 * X * Y / (C2 * D2) * fxe);
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
__global__ __launch_bounds__(256) void k_swe_tendency(DevTree t, const double* __restrict__ px,
                                                      const double* __restrict__ py, const double* __restrict__ pz,
                                                      CP3 vel, const double* __restrict__ vor,
                                                      const double* __restrict__ div, const double* __restrict__ h,
                                                      const double* __restrict__ lap, double omega, int rule,
                                                      double gravity, P3 out) {
  constexpr int M2 = M * M, S = leaf_slots<M>();
  __shared__ double sv_[S][3][M2], xn_[S][M], en_[S][M];
  const LeafSlot L = leaf_slot<M>(t);
  const int leaf = L.leaf, o = L.o, a = L.a, b = L.b;
  double(*sv)[M2] = sv_[L.slot];
  double *xn = xn_[L.slot], *en = en_[L.slot];
  const int64_t k = (int64_t)leaf * M2 + o;
  if (L.live) {
    leaf_nodes<M>(t.panels[leaf + t.offset], xn, en, o);
    for (int c = 0; c < 3; ++c) sv[c][o] = vel.p[c][k];
  }
  __syncthreads();
  if (!L.live) return;
  const DevPanel& P = t.panels[leaf + t.offset];
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
  d -= gravity * lap[k];
  d -= ux * ux + uy * uy + uz * uz;
  if (x * x + y * y > 1e-16) d -= 2.0 * omega * (-y * ux + x * uy);
  out.p[1][k] = d;
  out.p[2][k] = -h[k] * dk;
}

// apply_topography (topography.cpp:47-56): eff = h + h_s at P(x + disp), h_s the
// cone_mountain (:6-45) or the Gaussian mountain (topo_height).  The
// displacement of leaf-local i, node j: rule 0 reads row i of the full [P][n]
// displacement array (the reference's index; leaf rows live at i + offset
// there), rule 1 the leaf's own row.
__global__ void k_topography(DevTree t, const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ z, CP3 dpart, CP3 dprox, const double* __restrict__ h,
                             int rule, SwePhys ph, double* __restrict__ eff) {
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
  eff[k] = h[k] + topo_height(ph, a, b, c);
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
// containing leaf; lap -= 8 new_h / 0.1.  With the surface correction the
// topography at that point is added to new_h.  tau = 0 reads the nodes
// themselves (vel may then be null).
template <int M>
__global__ __launch_bounds__(128) void k_height_corr(DevTree t, const double* __restrict__ x,
                                                     const double* __restrict__ y, const double* __restrict__ z,
                                                     CP3 vel, const double* __restrict__ h,
                                                     const double* __restrict__ ht, double tau, SwePhys ph,
                                                     double* __restrict__ lap) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= t.npart) return;
  double a = x[k], b = y[k], c = z[k];
  if (tau != 0.0) {
    a += tau * vel.p[0][k]; b += tau * vel.p[1][k]; c += tau * vel.p[2][k];
  }
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
    for (int j = 0; j < M; ++j) r += (h[base + i * M + j] + (ht ? tau * ht[base + i * M + j] : 0.0)) * we[j];
    v += wx[i] * r;
  }
  if (ph.surface_correction) v += topo_height(ph, a, b, c);
  lap[k] -= 8.0 * v / 0.1;
}

// ---- reverse mode -------------------------------------------------------------

// Transpose of tile_deriv: the contribution of the derivative cotangents g
// (along dir, one per node) to node (a, b) of the differentiated tile.  The
// derivative at node i is sum_{k != i} (f_k - f_i) w_k / (w_i (x_i - x_k)).
template <int M>
__device__ inline double tile_deriv_t(const double* g, int a, int b, int dir, const double* nodes) {
  constexpr int D = M - 1;
  const int kk = dir == 0 ? a : b;
  const double xk = nodes[kk], wk = bary_w(kk, D);
  double acc = 0, self = 0;
#pragma unroll
  for (int i = 0; i < M; ++i) {
    if (i == kk) continue;
    const double gi = dir == 0 ? g[i * M + b] : g[a * M + i];
    const double wi = bary_w(i, D);
    acc += gi * wk / (wi * (nodes[i] - xk));
    self += wi / (xk - nodes[i]);
  }
  return acc - g[a * M + b] * self / wk;
}

// The velocity-gradient map of k_swe_tendency at one node is linear in the
// label derivatives (f_xi, f_eta): G = f_xi col0 + f_eta col1.
__device__ inline void grad_map(int rule, int face, double xi, double eta, double* col0, double* col1) {
  const double X = tan(xi), Y = tan(eta), C = sqrt(1 + X * X), D = sqrt(1 + Y * Y);
  const double xd0 = rule == 0 ? D : 1.0, ed0 = rule == 0 ? X * Y / C : 0.0;
  const double xd1 = rule == 0 ? X * Y / D : 0.0, ed1 = rule == 0 ? C : 1.0;
  xyzvec_from_xietavec(xd0, ed0, face, xi, eta, rule != 0, col0);
  xyzvec_from_xietavec(xd1, ed1, face, xi, eta, rule != 0, col1);
}

// Reverse of k_swe_tendency for the cotangents tbar = (vor_t, div_t, h_t)bar:
// vbar (velocity) and sbar (vor, div, h before the step) accumulated, lapbar
// assigned.  sum_ab G_ab G_ba has dG_ab = 2 G_ba; the label-derivative
// cotangents go back through the transposed leaf differentiation.
template <int M>
__global__ __launch_bounds__(256) void k_swe_tendency_adj(DevTree t, const double* __restrict__ px,
                                                          const double* __restrict__ py,
                                                          const double* __restrict__ pz, CP3 vel,
                                                          const double* __restrict__ vor,
                                                          const double* __restrict__ div,
                                                          const double* __restrict__ h, CP3 tbar, double omega,
                                                          int rule, double gravity, P3 vbar,
                                                          double* __restrict__ lapbar, P3 sbar) {
  constexpr int M2 = M * M, S = leaf_slots<M>();
  __shared__ double sv_[S][3][M2], gx_[S][3][M2], ge_[S][3][M2], xn_[S][M], en_[S][M];
  const LeafSlot L = leaf_slot<M>(t);
  const int leaf = L.leaf, o = L.o, a = L.a, b = L.b;
  double(*sv)[M2] = sv_[L.slot];
  double(*gx)[M2] = gx_[L.slot];
  double(*ge)[M2] = ge_[L.slot];
  double *xn = xn_[L.slot], *en = en_[L.slot];
  const int64_t k = (int64_t)leaf * M2 + o;
  if (L.live) {
    leaf_nodes<M>(t.panels[leaf + t.offset], xn, en, o);
    for (int c = 0; c < 3; ++c) sv[c][o] = vel.p[c][k];
  }
  __syncthreads();
  double vb[3] = {0, 0, 0};
  if (L.live) {
    const DevPanel& P = t.panels[leaf + t.offset];
    const int face = rule == 0 ? t.panels[leaf].face : P.face;
    const double xi = xn[a], eta = en[b];
    double col0[3], col1[3], G[3][3];
    grad_map(rule, face, xi, eta, col0, col1);
    for (int c = 0; c < 3; ++c) {
      const double fx = tile_deriv<M>(sv[c], a, b, 0, xn), fe = tile_deriv<M>(sv[c], a, b, 1, en);
      for (int i = 0; i < 3; ++i) G[c][i] = fx * col0[i] + fe * col1[i];
    }
    const double tv = tbar.p[0][k], td = tbar.p[1][k], th = tbar.p[2][k];
    for (int c = 0; c < 3; ++c) {
      double bx = 0, be = 0;
      for (int i = 0; i < 3; ++i) {
        const double gb = -2.0 * td * G[i][c];
        bx += gb * col0[i];
        be += gb * col1[i];
      }
      gx[c][o] = bx;
      ge[c][o] = be;
    }
    const double x = px[k], y = py[k], z = pz[k];
    const double ux = sv[0][o], uy = sv[1][o], uz = sv[2][o];
    const double vk = vor[k], dk = div[k];
    vb[0] = -2.0 * ux * td;
    vb[1] = -2.0 * uy * td;
    vb[2] = -2.0 * uz * td - 2.0 * omega * tv;
    if (x * x + y * y > 1e-16) {
      vb[0] += 2.0 * omega * y * td;
      vb[1] -= 2.0 * omega * x * td;
    }
    sbar.p[0][k] += -dk * tv + 2.0 * omega * z * td;
    sbar.p[1][k] += -(vk + 2.0 * omega * z) * tv - h[k] * th;
    sbar.p[2][k] += -dk * th;
    lapbar[k] = -gravity * td;
  }
  __syncthreads();
  if (L.live)
    for (int c = 0; c < 3; ++c)
      vbar.p[c][k] += vb[c] + tile_deriv_t<M>(gx[c], a, b, 0, xn) + tile_deriv_t<M>(ge[c], a, b, 1, en);
}

// Transpose of k_panel_lap: fbar += L^T lapbar.
template <int M>
__global__ __launch_bounds__(256) void k_panel_lap_adj(DevTree t, const double* __restrict__ lapbar,
                                                       double* __restrict__ fbar) {
  constexpr int M2 = M * M, S = leaf_slots<M>();
  __shared__ double sa_[S][M2], sb_[S][M2], sc_[S][M2], sx_[S][M2], se_[S][M2], xn_[S][M], en_[S][M];
  const LeafSlot L = leaf_slot<M>(t);
  const int o = L.o, a = L.a, b = L.b;
  double *sa = sa_[L.slot], *sb = sb_[L.slot], *sc = sc_[L.slot], *sx = sx_[L.slot], *se = se_[L.slot];
  double *xn = xn_[L.slot], *en = en_[L.slot];
  if (L.live) leaf_nodes<M>(t.panels[L.leaf + t.offset], xn, en, o);
  __syncthreads();
  const int64_t k = (int64_t)L.leaf * M2 + o;
  if (L.live) {
    const double X = tan(xn[a]), Y = tan(en[b]);
    const double C2 = 1 + X * X, D2 = 1 + Y * Y, delta = 1 + X * X + Y * Y, l = lapbar[k];
    sa[o] = delta / C2 * l;                       // d/d f_xixi
    sb[o] = delta / D2 * l;                       // d/d f_etaeta
    sc[o] = delta * 2.0 * X * Y / (C2 * D2) * l;  // d/d f_xieta
  }
  __syncthreads();
  if (L.live) {
    sx[o] = tile_deriv_t<M>(sa, a, b, 0, xn);
    se[o] = tile_deriv_t<M>(sc, a, b, 0, xn) + tile_deriv_t<M>(sb, a, b, 1, en);
  }
  __syncthreads();
  if (L.live) fbar[k] += tile_deriv_t<M>(sx, a, b, 0, xn) + tile_deriv_t<M>(se, a, b, 1, en);
}

// Reverse of k_height_corr for lapbar: with kb = -8/0.1 lapbar, the transposed
// interpolation kb B is scattered into hscat (h + tau h_t's cotangent; the
// caller splits it), and the moved point's cotangent (interpolant and, with the
// surface correction, topography gradients) goes back through the projection
// into vbar += tau pbar.  tau = 0: the nodes themselves, no velocity term.
template <int M>
__global__ __launch_bounds__(128) void k_height_corr_adj(DevTree t, const double* __restrict__ x,
                                                         const double* __restrict__ y,
                                                         const double* __restrict__ z, CP3 vel,
                                                         const double* __restrict__ hh, double tau, SwePhys ph,
                                                         const double* __restrict__ lapbar,
                                                         double* __restrict__ hscat, P3 vbar) {
  __shared__ double lds[2][wave_scatter_lds<M, 1>()];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < t.npart;
  const int64_t u = valid ? i : t.npart - 1;
  double w[3] = {x[u], y[u], z[u]};
  if (tau != 0.0) {
    w[0] += tau * vel.p[0][u]; w[1] += tau * vel.p[1][u]; w[2] += tau * vel.p[2][u];
  }
  double a = w[0], b = w[1], c = w[2];
  project_to_sphere(a, b, c);
  const double kb = -8.0 / 0.1 * lapbar[u];
  double pb[3] = {0, 0, 0}, wx[M], we[M];
  const double* f[1] = {hh};
  int leaf = interp_grad<M, 1>(t, a, b, c, f, &kb, pb, wx, we, valid);
  if (ph.surface_correction) {
    double g[3];
    topo_height(ph, a, b, c, g);
    for (int q = 0; q < 3; ++q) pb[q] += kb * g[q];
  }
  double* dst[1] = {hscat};
  wave_scatter_leaf<M, 1>(lds[threadIdx.x >> 6], valid ? leaf : -1, wx, we, &kb, dst, t.offset);
  if (valid && tau != 0.0 && vbar.p[0]) {
    project_adjoint(w[0], w[1], w[2], pb);
    for (int q = 0; q < 3; ++q) vbar.p[q][u] += tau * pb[q];
  }
}

// Reverse of one stage's state, area and (stages 2-3) effective-height
// updates, given the charge cotangents of its velocity evaluation (qz, qd:
// q = A_s S_vor, A_s S_div) and of its surface-Laplacian evaluation (qe:
// q = A_s eff, eff = S_h + h_s(P(x + d_row)); qe null at stage 4):
//   S_s = S0 + taua T,  A_s = area (1 + taua S0_div).
// sbar (vor, div, h of S0) and tbar accumulated; with displacements given,
// the topography's displacement cotangent is added to dbar at the row the
// forward read (rule 0: row i of the [P][n] array, rule 1: the particle).
__global__ void k_swe_stage_adj(DevTree t, const double* __restrict__ x, const double* __restrict__ y,
                                const double* __restrict__ z, CP3 s0, CP3 tprev, const double* __restrict__ area,
                                double taua, const double* __restrict__ qz, const double* __restrict__ qd,
                                const double* __restrict__ qe, CP3 dpart, CP3 dprox, int rule, SwePhys ph, P3 sbar,
                                P3 tbar, P3 dbar_part, P3 dbar_prox) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= t.npart) return;
  double st[3];
  for (int q = 0; q < 3; ++q) st[q] = s0.p[q][k] + (tprev.p[0] ? taua * tprev.p[q][k] : 0.0);
  const double as = area[k] * (1.0 + taua * s0.p[1][k]);
  const double bz = as * qz[k], bd = as * qd[k], be = qe ? as * qe[k] : 0.0;
  double ab = st[0] * qz[k] + st[1] * qd[k];
  if (qe) {
    double eff = st[2];
    if (ph.topo) {
      // the forward's displaced point
      const int i = (int)(k / t.m2), j = (int)(k % t.m2);
      const bool prox = rule == 0 && i < t.offset;
      const int64_t row = rule != 0 ? k : (prox ? (int64_t)i * t.n + j : (int64_t)(i - t.offset) * t.m2 + j);
      double w[3] = {x[k], y[k], z[k]};
      if (dpart.p[0])
        for (int q = 0; q < 3; ++q) w[q] += (prox ? dprox.p[q] : dpart.p[q])[row];
      double a = w[0], b = w[1], c = w[2];
      project_to_sphere(a, b, c);
      double g[3];
      eff += topo_height(ph, a, b, c, g);
      if (dpart.p[0]) {
        for (int q = 0; q < 3; ++q) g[q] *= be;
        project_adjoint(w[0], w[1], w[2], g);
        for (int q = 0; q < 3; ++q) (prox ? dbar_prox.p[q] : dbar_part.p[q])[row] += g[q];
      }
    }
    ab += eff * qe[k];
  }
  sbar.p[0][k] += bz;
  sbar.p[1][k] += bd + area[k] * taua * ab;
  sbar.p[2][k] += be;
  if (tbar.p[0]) {
    tbar.p[0][k] += taua * bz;
    tbar.p[1][k] += taua * bd;
    tbar.p[2][k] += taua * be;
  }
}

template <int M>
struct TendencyAdj {
  static void run(const DevTree& t, const double* x, const double* y, const double* z, CP3 vel, const double* vor,
                  const double* div, const double* h, CP3 tbar, double omega, int rule, double gravity, P3 vbar,
                  double* lapbar, P3 sbar, hipStream_t s) {
    hipLaunchKernelGGL(k_swe_tendency_adj<M>, dim3(leaf_blocks<M>(t)), dim3(256), 0, s, t, x, y, z, vel, vor, div, h, tbar, omega,
                       rule, gravity, vbar, lapbar, sbar);
  }
};

template <int M>
struct PanelLapAdj {
  static void run(const DevTree& t, const double* lapbar, double* fbar, hipStream_t s) {
    hipLaunchKernelGGL(k_panel_lap_adj<M>, dim3(leaf_blocks<M>(t)), dim3(256), 0, s, t, lapbar, fbar);
  }
};

template <int M>
struct HeightCorrAdj {
  static void run(const DevTree& t, const double* x, const double* y, const double* z, CP3 vel, const double* hh,
                  double tau, SwePhys ph, const double* lapbar, double* hscat, P3 vbar, hipStream_t s) {
    hipLaunchKernelGGL(k_height_corr_adj<M>, dim3(blocks_for(t.npart, 128)), dim3(128), 0, s, t, x, y, z, vel, hh,
                       tau, ph, lapbar, hscat, vbar);
  }
};

template <int M>
struct PanelLap {
  static void run(const DevTree& t, const double* f, double* lap, hipStream_t s) {
    hipLaunchKernelGGL(k_panel_lap<M>, dim3(leaf_blocks<M>(t)), dim3(256), 0, s, t, f, lap);
  }
};

template <int M>
struct Tendency {
  static void run(const DevTree& t, const double* x, const double* y, const double* z, CP3 vel, const double* vor,
                  const double* div, const double* h, const double* lap, double omega, int rule, double gravity,
                  P3 out, hipStream_t s) {
    hipLaunchKernelGGL(k_swe_tendency<M>, dim3(leaf_blocks<M>(t)), dim3(256), 0, s, t, x, y, z, vel, vor, div, h, lap, omega,
                       rule, gravity, out);
  }
};

template <int M>
struct HeightCorr {
  static void run(const DevTree& t, const double* x, const double* y, const double* z, CP3 vel, const double* h,
                  const double* ht, double tau, SwePhys ph, double* lap, hipStream_t s) {
    hipLaunchKernelGGL(k_height_corr<M>, dim3(blocks_for(t.npart, 128)), dim3(128), 0, s, t, x, y, z, vel, h, ht,
                       tau, ph, lap);
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
                         const double* lap, double omega, int rule, double gravity, double* const* out,
                         hipStream_t s) {
  if (t.Pl > 0) dispatch_m<Tendency>(t.m, t, x, y, z, cp3(vel), vor, div, h, lap, omega, rule, gravity, p3(out), s);
}

void launch_topography(const DevTree& t, const double* x, const double* y, const double* z,
                       const double* const* disp_part, const double* const* disp_proxy, const double* h, int rule,
                       const SwePhys& ph, double* eff, hipStream_t s) {
  if (t.npart > 0)
    hipLaunchKernelGGL(k_topography, dim3(blocks_for(t.npart, 256)), dim3(256), 0, s, t, x, y, z, cp3(disp_part),
                       cp3(disp_proxy), h, rule, ph, eff);
}

void launch_apply_tendencies(int64_t n, const double* const* f, const double* const* ft, const double* area,
                             double tau, double* const* out, double* new_area, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_apply_tend, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, cp3(f), cp3(ft), area, tau, p3(out),
                       new_area);
}

void launch_height_correction(const DevTree& t, const double* x, const double* y, const double* z,
                              const double* const* vel, const double* h, const double* ht, double tau,
                              const SwePhys& ph, double* lap, hipStream_t s) {
  if (t.npart > 0) dispatch_m<HeightCorr>(t.m, t, x, y, z, cp3(vel), h, ht, tau, ph, lap, s);
}

void launch_swe_tendency_adjoint(const DevTree& t, const double* x, const double* y, const double* z,
                                 const double* const* vel, const double* vor, const double* div, const double* h,
                                 const double* const* tbar, double omega, int rule, double gravity,
                                 double* const* vbar, double* lapbar, double* const* sbar, hipStream_t s) {
  if (t.Pl > 0)
    dispatch_m<TendencyAdj>(t.m, t, x, y, z, cp3(vel), vor, div, h, cp3(tbar), omega, rule, gravity, p3(vbar), lapbar,
                            p3(sbar), s);
}

void launch_panel_laplacian_adjoint(const DevTree& t, const double* lapbar, double* fbar, hipStream_t s) {
  if (t.Pl > 0) dispatch_m<PanelLapAdj>(t.m, t, lapbar, fbar, s);
}

void launch_height_correction_adjoint(const DevTree& t, const double* x, const double* y, const double* z,
                                      const double* const* vel, const double* hh, double tau, const SwePhys& ph,
                                      const double* lapbar, double* hscat, double* const* vbar, hipStream_t s) {
  P3 vb{};
  if (vbar)
    for (int i = 0; i < 3; ++i) vb.p[i] = vbar[i];
  if (t.npart > 0) dispatch_m<HeightCorrAdj>(t.m, t, x, y, z, cp3(vel), hh, tau, ph, lapbar, hscat, vb, s);
}

void launch_swe_stage_adjoint(const DevTree& t, const double* x, const double* y, const double* z,
                              const double* const* s0, const double* const* tprev, const double* area, double taua,
                              const double* qz, const double* qd, const double* qe, const double* const* dpart,
                              const double* const* dprox, int rule, const SwePhys& ph, double* const* sbar,
                              double* const* tbar, double* const* dbar_part, double* const* dbar_prox,
                              hipStream_t s) {
  P3 tb{}, dbp{}, dbx{};
  for (int i = 0; i < 3; ++i) {
    tb.p[i] = tbar ? tbar[i] : nullptr;
    dbp.p[i] = dbar_part ? dbar_part[i] : nullptr;
    dbx.p[i] = dbar_prox ? dbar_prox[i] : nullptr;
  }
  if (t.npart > 0)
    hipLaunchKernelGGL(k_swe_stage_adj, dim3(blocks_for(t.npart, 256)), dim3(256), 0, s, t, x, y, z, cp3(s0),
                       cp3(tprev), area, taua, qz, qd, qe, cp3(dpart), cp3(dprox), rule, ph, p3(sbar), tb, dbp, dbx);
}

namespace {
// The height correction at the nodes themselves (the regularised stage-1
// Laplacian): lap -= 8/0.1 (h + h_s(x)), h_s only with the surface correction.
__global__ void k_node_corr(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                            const double* __restrict__ z, const double* __restrict__ h, SwePhys ph,
                            double* __restrict__ lap) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double v = h[k];
  if (ph.surface_correction) v += topo_height(ph, x[k], y[k], z[k]);
  lap[k] -= 8.0 * v / 0.1;
}

// out = x + alpha y (out may alias x or y)
__global__ void k_lincomb(int64_t n, const double* __restrict__ x, double alpha, const double* __restrict__ y,
                          double* out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = x[k] + alpha * y[k];
}

// Zero fill of up to kFillArrays arrays in one launch: array blockIdx.y.
constexpr int kFillArrays = 6;
struct FillSet {
  double* p[kFillArrays];
  int64_t count[kFillArrays];
};
__global__ void k_fill_zero_sets(FillSet f) {
  const int y = blockIdx.y;
  double* p = f.p[0];
  int64_t count = f.count[0];
#pragma unroll
  for (int i = 1; i < kFillArrays; ++i)
    if (y == i) {
      p = f.p[i];
      count = f.count[i];
    }
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += (int64_t)gridDim.x * blockDim.x)
    p[k] = 0.0;
}

// out = x + alpha y, zero[k] = 0 (a scatter target cleared in the same pass).
__global__ void k_lincomb_zero(int64_t n, const double* __restrict__ x, double alpha, const double* __restrict__ y,
                               double* __restrict__ out, double* __restrict__ zero) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  out[k] = x[k] + alpha * y[k];
  zero[k] = 0.0;
}

// a += src, b += beta src (two accumulations of one cotangent).
__global__ void k_add2(int64_t n, const double* __restrict__ src, double* __restrict__ a, double beta,
                       double* __restrict__ b) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double v = src[k];
  a[k] += v;
  b[k] += beta * v;
}

// out1_q = f1_q + alpha1 r_q, out2_q = f2_q + alpha2 r_q for q = 0..2.
__global__ void k_apply_two(int64_t n, CP3 r, CP3 f1, double alpha1, P3 out1, CP3 f2, double alpha2, P3 out2) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  for (int q = 0; q < 3; ++q) {
    const double v = r.p[q][k];
    out1.p[q][k] = f1.p[q][k] + alpha1 * v;
    out2.p[q][k] = f2.p[q][k] + alpha2 * v;
  }
}
}  // namespace

void launch_fill_zero_sets(double* const* arrays, const int64_t* counts, int n, hipStream_t s) {
  for (int b = 0; b < n; b += kFillArrays) {
    FillSet f{};
    int64_t most = 0;
    const int m = n - b < kFillArrays ? n - b : kFillArrays;
    for (int i = 0; i < m; ++i) {
      f.p[i] = arrays[b + i];
      f.count[i] = counts[b + i];
      most = counts[b + i] > most ? counts[b + i] : most;
    }
    const unsigned gx = (unsigned)std::min<int64_t>(blocks_for(most, 256), 2048);
    if (most > 0) hipLaunchKernelGGL(k_fill_zero_sets, dim3(gx, m), dim3(256), 0, s, f);
  }
}

void launch_lincomb_zero(int64_t n, const double* x, double alpha, const double* y, double* out, double* zero,
                         hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_lincomb_zero, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, x, alpha, y, out, zero);
}

void launch_add2(int64_t n, const double* src, double* a, double beta, double* b, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_add2, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, src, a, beta, b);
}

void launch_apply_two(int64_t n, const double* const* r, const double* const* f1, double alpha1, double* const* out1,
                      const double* const* f2, double alpha2, double* const* out2, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_apply_two, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, cp3(r), cp3(f1), alpha1, p3(out1),
                       cp3(f2), alpha2, p3(out2));
}

void launch_node_correction(int64_t n, const double* x, const double* y, const double* z, const double* h,
                            const SwePhys& ph, double* lap, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_node_corr, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, x, y, z, h, ph, lap);
}

void launch_lincomb(int64_t n, const double* x, double alpha, const double* y, double* out, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_lincomb, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, x, alpha, y, out);
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
