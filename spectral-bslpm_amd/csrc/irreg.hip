// irreg.hip -- passes of the irregular particle tree (the lat-lon path's _ll
// functors, fmm_funcs_impl.hpp:190-364) on leaf-sorted points:
//   * k_point_weights: each point's 1-D BLI basis on its leaf, once per tree;
//   * k_p2m_ll: base_pots_ll -- proxy charges of a leaf by BLI anterpolation
//     of its points (one workgroup per leaf, point weights staged in LDS);
//   * k_l2p_ll: child_panel_interp_3_ll -- proxy potentials of the point's
//     leaf interpolated at the point (one thread per point);
//   * gathers into and scatters out of the sorted order.
// M2M / L2L reuse the uniform tree's kernels over lists of non-leaf panels
// (children are exact quadrants, so the fixed transfer matrices apply), and
// the pair sums reuse the target-major kernels.
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

__device__ inline void xieta_auto(double x, double y, double z, double* xe) {
  xieta_from_xyz_face(x, y, z, face_from_xyz(x, y, z), xe);
}

// q[k] = area[perm[k]] * pots[perm[k]]
__global__ void k_gather_q(int64_t n, const int* __restrict__ perm, const double* __restrict__ area,
                           const double* __restrict__ pots, double* __restrict__ q) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = perm[k];
  q[k] = area ? area[i] * pots[i] : pots[i];
}

// out[f][perm[k]] (+)= in[f][k]
__global__ void k_scatter(int64_t n, int ncomp, const int* __restrict__ perm, CPtr3 in, Ptr3 out, int accumulate) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = perm[k];
  for (int f = 0; f < ncomp; ++f) out.p[f][i] = accumulate ? out.p[f][i] + in.p[f][k] : in.p[f][k];
}

// a[f][i] += b[f][i]
__global__ void k_add(int64_t n, int ncomp, Ptr3 a, CPtr3 b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int f = 0; f < ncomp; ++f) a.p[f][i] += b.p[f][i];
}

// The 1-D BLI basis of every sorted point on its leaf: w[k][0..d] at xi, then
// w[k][d+1..2d+1] at eta.  The points do not move, so P2M and L2P read these
// instead of recomputing them every pass (same values).
__global__ void k_point_weights(DevTree t, const int* __restrict__ point_leaf, double* __restrict__ w) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= t.npart) return;
  const DevPanel& P = t.panels[point_leaf[k]];
  const int dp1 = t.dp1;
  double xe[2];
  xieta_auto(t.px[k], t.py[k], t.pz[k], xe);
  double* wk = w + k * 2 * dp1;
  bool ok = bli_weights_1d(xe[0], P.min_xi, P.max_xi, dp1 - 1, t.cheb_d, wk);
  ok = bli_weights_1d(xe[1], P.min_eta, P.max_eta, dp1 - 1, t.cheb_d, wk + dp1) && ok;
  if (!ok) flag(t, kErrBliBounds);
}

// base_pots_ll (fmm_funcs_impl.hpp:207-223): proxy[leaf][a*(d+1)+b] =
// sum over the leaf's points of l_a(xi) l_b(eta) q (assigned), for nq = 1 or 2
// charge arrays at once (q + j*qstride -> proxy + j*pstride).  One workgroup of
// kP2mParts x 128 threads per leaf.  The leaf's weights and charges are staged
// in LDS in rounds of up to kP2mRound points, all threads loading (each point's
// weights are a new cache line, so the loads must be in flight together, not
// one point after another); thread (h, o) then sums output o over part h of
// the round, and the parts are added through LDS in a fixed order.
constexpr int kP2mParts = 4, kP2mRound = 256;

__global__ __launch_bounds__(128 * kP2mParts) void k_p2m_ll(DevTree t, const int* __restrict__ leaves,
                                                            const int* __restrict__ leaf_first,
                                                            const int* __restrict__ leaf_count,
                                                            const double* __restrict__ w, int nq,
                                                            const double* __restrict__ q, int64_t qstride,
                                                            double* __restrict__ proxy, int64_t pstride) {
  static_assert((kMaxDeg + 1) * (kMaxDeg + 1) <= 128, "k_p2m_ll: one output per thread of a part");
  __shared__ double sw[kP2mRound * 2 * (kMaxDeg + 1)];
  __shared__ double sq[2][kP2mRound];
  __shared__ double sacc[kP2mParts][2][128];
  const int leaf = leaves[blockIdx.x];
  const int first = leaf_first[leaf], cnt = leaf_count[leaf];
  const int dp1 = t.dp1, n = t.n, w2 = 2 * dp1;
  const int o = threadIdx.x & 127, h = threadIdx.x >> 7;
  const int a = o / dp1, b = o - a * dp1;
  double acc0 = 0, acc1 = 0;
  for (int c0 = 0; c0 < cnt; c0 += kP2mRound) {
    const int m = min(kP2mRound, cnt - c0);
    const double* __restrict__ wc = w + ((int64_t)first + c0) * w2;
    for (int i = threadIdx.x; i < m * w2; i += blockDim.x) sw[i] = wc[i];
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      const int64_t k = (int64_t)first + c0 + i;
      sq[0][i] = q[k];
      sq[1][i] = nq > 1 ? q[qstride + k] : 0.0;
    }
    __syncthreads();
    if (o < n) {
      const int part = (m + kP2mParts - 1) / kP2mParts, j0 = h * part, j1 = min(m, j0 + part);
      if (nq > 1) {
#pragma unroll 4
        for (int j = j0; j < j1; ++j) {
          const double ab = sw[j * w2 + a] * sw[j * w2 + dp1 + b];
          acc0 = fma(ab, sq[0][j], acc0);
          acc1 = fma(ab, sq[1][j], acc1);
        }
      } else {
#pragma unroll 4
        for (int j = j0; j < j1; ++j) acc0 = fma(sw[j * w2 + a] * sw[j * w2 + dp1 + b], sq[0][j], acc0);
      }
    }
    __syncthreads();
  }
  if (o < n) {
    sacc[h][0][o] = acc0;
    sacc[h][1][o] = acc1;
  }
  __syncthreads();
  if (h == 0 && o < n) {
    double v0 = sacc[0][0][o], v1 = sacc[0][1][o];
#pragma unroll
    for (int i = 1; i < kP2mParts; ++i) {
      v0 += sacc[i][0][o];
      v1 += sacc[i][1][o];
    }
    proxy[(int64_t)leaf * n + o] = v0;
    if (nq > 1) proxy[pstride + (int64_t)leaf * n + o] = v1;
  }
}

// child_panel_interp_3_ll (fmm_funcs_impl.hpp:341-362) for sorted point k:
// out[f][k] = sum_ab l_a(xi) l_b(eta) phi[f][leaf][a*(d+1)+b] (+ part[f][k]).
template <int DP1>
__global__ __launch_bounds__(256) void k_l2p_ll(DevTree t, const int* __restrict__ point_leaf,
                                                const double* __restrict__ w, int ncomp, CPtr3 phi, CPtr3 part,
                                                Ptr3 out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= t.npart) return;
  const int leaf = point_leaf[k];
  double wx[DP1], we[DP1];
  const double* wk = w + k * 2 * DP1;
#pragma unroll
  for (int a = 0; a < DP1; ++a) {
    wx[a] = wk[a];
    we[a] = wk[DP1 + a];
  }
  for (int f = 0; f < ncomp; ++f) {
    double v = tensor_gather<DP1>(phi.p[f] + (int64_t)leaf * (DP1 * DP1), wx, we);
    if (part.p[f]) v += part.p[f][k];
    out.p[f][k] = v;
  }
}

template <int DP1>
struct L2PLL {
  static void run(const DevTree& t, const int* point_leaf, const double* w, int ncomp, CPtr3 phi, CPtr3 part,
                  Ptr3 out, hipStream_t s) {
    hipLaunchKernelGGL(k_l2p_ll<DP1>, dim3(blocks_for(t.npart, 256)), dim3(256), 0, s, t, point_leaf, w, ncomp, phi,
                       part, out);
  }
};

}  // namespace

void launch_gather_charges(int64_t n, const int* perm, const double* area, const double* pots, double* q,
                           hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_gather_q, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, perm, area, pots, q);
}

void launch_scatter(int64_t n, int ncomp, const int* perm, const double* const* in, double* const* out,
                    bool accumulate, hipStream_t s) {
  CPtr3 a{};
  Ptr3 b{};
  for (int f = 0; f < ncomp; ++f) {
    a.p[f] = in[f];
    b.p[f] = out[f];
  }
  if (n > 0)
    hipLaunchKernelGGL(k_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, ncomp, perm, a, b, accumulate ? 1 : 0);
}

void launch_add(int64_t n, int ncomp, double* const* a, const double* const* b, hipStream_t s) {
  Ptr3 x{};
  CPtr3 y{};
  for (int f = 0; f < ncomp; ++f) {
    x.p[f] = a[f];
    y.p[f] = b[f];
  }
  if (n > 0) hipLaunchKernelGGL(k_add, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, ncomp, x, y);
}

void launch_point_weights(const DevTree& t, const int* point_leaf, double* w, hipStream_t s) {
  if (t.npart > 0)
    hipLaunchKernelGGL(k_point_weights, dim3(blocks_for(t.npart, 256)), dim3(256), 0, s, t, point_leaf, w);
}

void launch_p2m_ll(const DevTree& t, const int* leaves, int nleaves, const int* leaf_first, const int* leaf_count,
                   const double* w, int nq, const double* q, int64_t qstride, double* proxy, int64_t pstride,
                   hipStream_t s) {
  if (nleaves > 0)
    hipLaunchKernelGGL(k_p2m_ll, dim3(nleaves), dim3(128 * kP2mParts), 0, s, t, leaves, leaf_first, leaf_count, w, nq, q, qstride,
                       proxy, pstride);
}

void launch_l2p_ll(const DevTree& t, const int* point_leaf, const double* w, int ncomp, const double* const* phi,
                   const double* const* part, double* const* out, hipStream_t s) {
  CPtr3 a{}, b{};
  Ptr3 o{};
  for (int f = 0; f < ncomp; ++f) {
    a.p[f] = phi[f];
    b.p[f] = part ? part[f] : nullptr;
    o.p[f] = out[f];
  }
  if (t.npart > 0) dispatch_m<L2PLL>(t.dp1, t, point_leaf, w, ncomp, a, b, o, s);
}

}  // namespace sbslpm
