// bli.hpp -- per-thread barycentric Lagrange interpolation on one leaf's
// m x m particle grid, shared by the remap kernels and their adjoints.
//
// The reference evaluates the tensor basis B_kl = l_k(xi) l_l(eta) node by
// node (interp_funcs_impl.hpp:11-112); here the two 1-D weight vectors are
// computed once (compile-time degree M-1, so they stay in registers) and
// applied one direction at a time.
#pragma once
#include "device.hpp"

namespace sbslpm {
namespace bli {

__device__ inline void flag(const DevTree& t, int bit) { atomicOr(t.err, bit); }

// Barycentric weight of Chebyshev-Lobatto node i of degree D (1/2 at the ends,
// alternating signs), interp_funcs.cpp:62-72.
__device__ inline double node_weight(int i, int D) {
  double bw = (i == 0 || i == D) ? 0.5 : 1.0;
  return (i & 1) ? -bw : bw;
}

// 1-D weights l_i(t) on [lo, hi] with nodes cheb[i]*(hi-lo)/2 + (hi+lo)/2 (see
// bli_weights_1d): second barycentric form, or the delta of a node within
// 1e-15 of t (interp_funcs_impl.hpp:57-79).  Returns false when t lies
// outside [lo, hi] by more than 1e-15.
template <int M>
__device__ inline bool bli_w(double t, double lo, double hi, const double* __restrict__ cheb, double* w) {
  constexpr int D = M - 1;
  const double range = 0.5 * (hi - lo), off = 0.5 * (hi + lo);
  const bool ok = (t >= lo - kSnapTol) && (t <= hi + kSnapTol);
  double node[M];
#pragma unroll
  for (int i = 0; i < M; ++i) node[i] = cheb[i] * range + off;
  int hit = -1;
#pragma unroll
  for (int i = M - 1; i >= 0; --i)
    if (fabs(t - node[i]) < kSnapTol) hit = i;
  if (hit >= 0) {
#pragma unroll
    for (int i = 0; i < M; ++i) w[i] = (i == hit) ? 1.0 : 0.0;
    return ok;
  }
  double den = 0;
#pragma unroll
  for (int i = 0; i < M; ++i) {
    w[i] = node_weight(i, D) * rcp64(t - node[i]);  // |t - node| >= 1e-15 here
    den += w[i];
  }
  const double inv = 1.0 / den;
#pragma unroll
  for (int i = 0; i < M; ++i) w[i] *= inv;
  return ok;
}

// Weights as bli_w and their derivatives dw[i] = l_i'(t).  Away from the
// nodes l_i' = l_i (S - 1/(t - t_i)), S = sum_j l_j/(t - t_j); the derivative
// of the nearest node's function is taken as minus the sum of the others
// (the l_i sum to one), which avoids the cancellation in its own formula.  At
// a node k (the snapped case) the derivatives are the differentiation-matrix
// row l_i'(t_k) = (w_i/w_k)/(t_k - t_i), i != k, l_k'(t_k) = -sum_{i != k}.
template <int M>
__device__ inline bool bli_wd(double t, double lo, double hi, const double* __restrict__ cheb, double* w, double* dw) {
  constexpr int D = M - 1;
  const double range = 0.5 * (hi - lo), off = 0.5 * (hi + lo);
  const bool ok = (t >= lo - kSnapTol) && (t <= hi + kSnapTol);
  double node[M];
#pragma unroll
  for (int i = 0; i < M; ++i) node[i] = cheb[i] * range + off;
  int hit = -1, near = 0;
  double best = fabs(t - node[0]);
#pragma unroll
  for (int i = M - 1; i >= 0; --i) {
    const double a = fabs(t - node[i]);
    if (a < kSnapTol) hit = i;
    if (a < best) { best = a; near = i; }
  }
  if (hit >= 0) {
    const double wk = node_weight(hit, D), tk = node[hit];
    double s = 0;
#pragma unroll
    for (int i = 0; i < M; ++i) {
      w[i] = (i == hit) ? 1.0 : 0.0;
      dw[i] = (i == hit) ? 0.0 : (node_weight(i, D) / wk) / (tk - node[i]);
      s += dw[i];
    }
#pragma unroll
    for (int i = 0; i < M; ++i)
      if (i == hit) dw[i] = -s;
    return ok;
  }
  double den = 0, inv_d[M];
#pragma unroll
  for (int i = 0; i < M; ++i) {
    inv_d[i] = rcp64(t - node[i]);
    w[i] = node_weight(i, D) * inv_d[i];
    den += w[i];
  }
  const double inv = 1.0 / den;
  double S = 0;
#pragma unroll
  for (int i = 0; i < M; ++i) {
    w[i] *= inv;
    S = fma(w[i], inv_d[i], S);
  }
  double s = 0;
#pragma unroll
  for (int i = 0; i < M; ++i) {
    dw[i] = (i == near) ? 0.0 : w[i] * (S - inv_d[i]);
    s += dw[i];
  }
#pragma unroll
  for (int i = 0; i < M; ++i)
    if (i == near) dw[i] = -s;
  return ok;
}

// sum_a wx[a] sum_c we[c] f[a*M + c]: the reference's tensor basis
// (bve_time_step.cpp:222-237 sums f * (wx_a * we_c)) applied one direction at a time.
template <int M>
__device__ inline double tensor_gather(const double* __restrict__ f, const double* wx, const double* we) {
  double out = 0;
#pragma unroll
  for (int a = 0; a < M; ++a) {
    double s = 0;
#pragma unroll
    for (int c = 0; c < M; ++c) s = fma(f[a * M + c], we[c], s);
    out = fma(wx[a], s, out);
  }
  return out;
}

// project_to_sphere (general_utils_impl.hpp:179-185) with one reciprocal of
// the norm instead of three divisions (each result within an ulp).
__device__ inline void project_rcp(double& x, double& y, double& z) {
  const double r = rcp64(sqrt(x * x + y * y + z * z));
  x *= r;
  y *= r;
  z *= r;
}

// Transpose of the projection's Jacobian applied to a cotangent:
// d P(w) = (I - y y^T) dw / |w|, y = P(w), which is symmetric.
__device__ inline void project_adjoint(double wx, double wy, double wz, double* bar) {
  const double r = sqrt(wx * wx + wy * wy + wz * wz), ir = 1.0 / r;
  const double yx = wx * ir, yy = wy * ir, yz = wz * ir;
  const double d = yx * bar[0] + yy * bar[1] + yz * bar[2];
  bar[0] = (bar[0] - yx * d) * ir;
  bar[1] = (bar[1] - yy * d) * ir;
  bar[2] = (bar[2] - yz * d) * ir;
}

// Gradients of the label coordinates (xi, eta) = xieta_from_xyz_face(p, face)
// with respect to p (cubed_sphere_transforms_impl.hpp:100-123): on every face
// xi = atan(N1/D), eta = atan(N2/D) with N1, N2, D signed coordinates, and
// d atan(N/D) = (D dN - N dD) / (N^2 + D^2).
__device__ inline void xieta_gradient(double x, double y, double z, int face, double* gxi, double* geta) {
  // (component index, sign) of N1, N2 and D per face
  int i1, i2, id;
  double s1, s2;
  switch (face) {
    case 0: i1 = 1; s1 = 1; i2 = 2; s2 = 1; id = 0; break;
    case 1: i1 = 0; s1 = -1; i2 = 2; s2 = 1; id = 1; break;
    case 2: i1 = 1; s1 = 1; i2 = 2; s2 = -1; id = 0; break;
    case 3: i1 = 0; s1 = -1; i2 = 2; s2 = -1; id = 1; break;
    case 4: i1 = 1; s1 = 1; i2 = 0; s2 = -1; id = 2; break;
    default: i1 = 1; s1 = -1; i2 = 0; s2 = -1; id = 2; break;
  }
  const double p[3] = {x, y, z};
  const double N1 = s1 * p[i1], N2 = s2 * p[i2], Dn = p[id];
  const double r1 = 1.0 / (N1 * N1 + Dn * Dn), r2 = 1.0 / (N2 * N2 + Dn * Dn);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    gxi[k] = 0;
    geta[k] = 0;
  }
  gxi[i1] += Dn * s1 * r1;
  gxi[id] -= N1 * r1;
  geta[i2] += Dn * s2 * r2;
  geta[id] -= N2 * r2;
}

// LDS doubles per wave used by wave_scatter_leaf<M, NF>.
template <int M, int NF>
constexpr int wave_scatter_lds() {
  return 64 * (2 * M + NF);
}

// Transposed leaf interpolation summed over the wave before it reaches memory:
// every lane holding a point (leaf >= 0) contributes g[q] wx[a] we[c] to node
// a*M + c of field q of its leaf.  The wave takes one distinct leaf at a time:
// the lanes on that leaf stage wx, we and g in the wave's LDS slice `lds`
// (wave_scatter_lds<M, NF>() doubles), then each lane sums the staged
// contributions to its nodes and issues one atomic add per node and field.
// Neighbouring particles mostly share a leaf, so a wave makes a few passes
// instead of 64 * M * M * NF contended atomics.  Every lane of the wave must
// call it (lanes without a point pass leaf = -1); the loop bounds are
// wave-uniform.
template <int M, int NF>
__device__ inline void wave_scatter_leaf(double* __restrict__ lds, int leaf, const double* wx, const double* we,
                                         const double* g, double* const* dst, int offset) {
  const int lane = threadIdx.x & 63;
  double* sx = lds;
  double* se = lds + 64 * M;
  double* sg = lds + 128 * M;
  uint64_t todo = __ballot(leaf >= 0);
  while (todo) {
    const int L = __shfl(leaf, __builtin_ctzll(todo));
    const bool mine = (leaf == L);
    const uint64_t mm = __ballot(mine);
    if (mine) {
#pragma unroll
      for (int i = 0; i < M; ++i) {
        sx[lane * M + i] = wx[i];
        se[lane * M + i] = we[i];
      }
#pragma unroll
      for (int q = 0; q < NF; ++q) sg[lane * NF + q] = g[q];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    const int64_t base = (int64_t)(L - offset) * (M * M);
    for (int j = lane; j < M * M; j += 64) {
      const int a = j / M, c = j % M;
      double acc[NF];
#pragma unroll
      for (int q = 0; q < NF; ++q) acc[q] = 0;
      uint64_t b = mm;
      while (b) {
        const int l = __builtin_ctzll(b);
        b &= b - 1;
        const double w = sx[l * M + a] * se[l * M + c];
#pragma unroll
        for (int q = 0; q < NF; ++q) acc[q] = fma(w, sg[l * NF + q], acc[q]);
      }
#pragma unroll
      for (int q = 0; q < NF; ++q) atomicAdd(dst[q] + base + j, acc[q]);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    todo &= ~mm;
  }
}

// Reverse of k = BLI[f](p) for NF leaf fields at point p, first part:
// pbar += sum_q kbar_q grad_p f_q(p); returns the leaf (-1 when location
// fails) and the 1-D weights for the transposed interpolation fbar += B kbar,
// which the caller applies with wave_scatter_leaf.
template <int M, int NF>
__device__ inline int interp_grad(const DevTree& t, double x, double y, double z, const double* const* f,
                                  const double* kbar, double* pbar, double* wx, double* we, bool report) {
  double xe[2];
  const int leaf = locate_leaf(t, x, y, z, xe);
  if (leaf < 0) {
    if (report) flag(t, kErrLocate);
    return -1;
  }
  const DevPanel& P = t.panels[leaf];
  double dwx[M], dwe[M];
  bool ok = bli_wd<M>(xe[0], P.min_xi, P.max_xi, t.cheb_m, wx, dwx);
  ok = bli_wd<M>(xe[1], P.min_eta, P.max_eta, t.cheb_m, we, dwe) && ok;
  if (!ok && report) flag(t, kErrBliBounds);
  double gxi[3], geta[3];
  xieta_gradient(x, y, z, P.face, gxi, geta);
  const int64_t base = (int64_t)(leaf - t.offset) * (M * M);
  double sxi = 0, seta = 0;  // sum_q kbar_q d f_q / d xi, d eta
#pragma unroll
  for (int q = 0; q < NF; ++q) {
    const double* __restrict__ F = f[q] + base;
    double dxi = 0, deta = 0;
#pragma unroll
    for (int a = 0; a < M; ++a) {
      double s = 0, sd = 0;
#pragma unroll
      for (int c = 0; c < M; ++c) {
        s = fma(F[a * M + c], we[c], s);
        sd = fma(F[a * M + c], dwe[c], sd);
      }
      dxi = fma(dwx[a], s, dxi);
      deta = fma(wx[a], sd, deta);
    }
    sxi = fma(kbar[q], dxi, sxi);
    seta = fma(kbar[q], deta, seta);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) pbar[k] += sxi * gxi[k] + seta * geta[k];
  return leaf;
}

inline unsigned blocks_for(int64_t count, int bs) { return (unsigned)((count + bs - 1) / bs); }

// Runs F<M>::run(args...) for the particle-grid side m (2..11).
template <template <int> class F, typename... A>
void dispatch_m(int m, A&&... a) {
  switch (m) {
    case 2: F<2>::run(a...); break;
    case 3: F<3>::run(a...); break;
    case 4: F<4>::run(a...); break;
    case 5: F<5>::run(a...); break;
    case 6: F<6>::run(a...); break;
    case 7: F<7>::run(a...); break;
    case 8: F<8>::run(a...); break;
    case 9: F<9>::run(a...); break;
    case 10: F<10>::run(a...); break;
    default: F<11>::run(a...); break;
  }
}

}  // namespace bli
}  // namespace sbslpm
