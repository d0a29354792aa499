// fmm_passes.hip -- upward (P2M, M2M) and downward (L2L, L2P) passes.
//
// The reference evaluates a tensor barycentric basis per child node per panel
// and scatters with atomics (fmm_funcs_impl.hpp:47-70, 138-171).  Children are
// exact quadrants, so the parent basis at child nodes is the tensor product of
// two fixed 1-D matrices (upper or lower half in each direction).  Each pass is
// a small dense contraction per panel: one workgroup per parent (or leaf),
// operands staged in LDS, each output owned by one thread (no atomics,
// deterministic order).  The passes move ~3 bytes per flop; they are bound by
// HBM and launch latency, not by arithmetic.
#include "device.hpp"

namespace sbslpm {
namespace {

constexpr int kMaxN = (kMaxDeg + 1) * (kMaxDeg + 1);

struct Ptr3 {
  double* p[3];
};
struct CPtr3 {
  const double* p[3];
};

// out(q, p) = sum_r X[q*R + r] * sum_s Y[p*R + s] * in[r*R + s]
__device__ inline double contract(const double* X, const double* Y, const double* in, int R, int q, int p) {
  double acc = 0;
  for (int r = 0; r < R; ++r) {
    double s = 0;
    for (int c = 0; c < R; ++c) s += Y[p * R + c] * in[r * R + c];
    acc += X[q * R + r] * s;
  }
  return acc;
}

// q_part = area * pots (base_pots, fmm_funcs_impl.hpp:29-33).
__global__ void k_base_pots(int64_t npart, const double* __restrict__ area, const double* __restrict__ pots,
                            double* __restrict__ q_part) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u < npart) q_part[u] = area[u] * pots[u];
}

// Leaf proxies from the particle grid: Q_leaf = T Q_part T^T, T[a][k] = l_a(x_k).
// For the reference layout (m == d+1) T is the identity and this is the
// base_pots copy into the leaf rows.
__global__ __launch_bounds__(128) void k_p2m(DevTree t, const double* __restrict__ q_part,
                                             double* __restrict__ q_proxy) {
  __shared__ double sin_[kMaxN];
  __shared__ double sT[kMaxN];
  const int leaf = blockIdx.x;
  const int m = t.m, m2 = t.m2, dp1 = t.dp1, n = t.n;
  for (int i = threadIdx.x; i < m2; i += blockDim.x) sin_[i] = q_part[(int64_t)leaf * m2 + i];
  for (int i = threadIdx.x; i < dp1 * m; i += blockDim.x) sT[i] = t.p2m[i];
  __syncthreads();
  for (int o = threadIdx.x; o < n; o += blockDim.x) {
    double v = contract(sT, sT, sin_, m, o / dp1, o % dp1);
    q_proxy[(int64_t)(t.offset + leaf) * n + o] = v;
  }
}

// M2M for all parents of one level (child_to_parent, fmm_funcs_impl.hpp:47-70),
// assigning the parent rows.
__global__ __launch_bounds__(128) void k_m2m(DevTree t, int pbegin, double* __restrict__ q) {
  __shared__ double sq[4 * kMaxN];
  __shared__ double shi[kMaxN], slo[kMaxN];
  const int par = pbegin + blockIdx.x;
  const int n = t.n, dp1 = t.dp1;
  const int c0 = t.panels[par].child0;
  for (int i = threadIdx.x; i < 4 * n; i += blockDim.x) sq[i] = q[(int64_t)c0 * n + i];
  for (int i = threadIdx.x; i < n; i += blockDim.x) { shi[i] = t.m2m_hi[i]; slo[i] = t.m2m_lo[i]; }
  __syncthreads();
  for (int o = threadIdx.x; o < n; o += blockDim.x) {
    const int a = o / dp1, b = o % dp1;
    double acc = 0;
    for (int c = 0; c < 4; ++c) {
      const double* X = (c == 0 || c == 3) ? shi : slo;  // xi: children 1,4 upper half
      const double* Y = (c <= 1) ? shi : slo;            // eta: children 1,2 upper half
      acc += contract(X, Y, sq + c * n, dp1, a, b);
    }
    q[(int64_t)par * n + o] = acc;
  }
}

// L2L for all parents of one level (parent_to_child_3, fmm_funcs_impl.hpp:138-171),
// accumulating into the four children.
__global__ __launch_bounds__(256) void k_l2l(DevTree t, int pbegin, int ncomp, Ptr3 phi) {
  __shared__ double sp[3 * kMaxN];
  __shared__ double shT[kMaxN], slT[kMaxN];
  const int par = pbegin + blockIdx.x;
  const int n = t.n, dp1 = t.dp1;
  const int c0 = t.panels[par].child0;
  for (int f = 0; f < ncomp; ++f)
    for (int i = threadIdx.x; i < n; i += blockDim.x) sp[f * n + i] = phi.p[f][(int64_t)par * n + i];
  for (int i = threadIdx.x; i < n; i += blockDim.x) { shT[i] = t.m2m_hiT[i]; slT[i] = t.m2m_loT[i]; }
  __syncthreads();
  for (int idx = threadIdx.x; idx < 4 * n; idx += blockDim.x) {
    const int c = idx / n, node = idx % n;
    const double* X = (c == 0 || c == 3) ? shT : slT;
    const double* Y = (c <= 1) ? shT : slT;
    const int k = node / dp1, l = node % dp1;
    for (int f = 0; f < ncomp; ++f) {
      double v = contract(X, Y, sp + f * n, dp1, k, l);
      phi.p[f][(int64_t)(c0 + c) * n + node] += v;
    }
  }
}

// L2P onto the leaf particle grid plus the direct (type 0/1) particle field:
// child_panel_copy (fmm_funcs_impl.hpp:182-187) for the reference layout,
// child_panel_interp_3_ll (:341-362) otherwise.
__global__ __launch_bounds__(128) void k_l2p(DevTree t, int ncomp, CPtr3 phi, CPtr3 part, Ptr3 sols) {
  __shared__ double sp[3 * kMaxN];
  __shared__ double sTt[kMaxN];
  const int leaf = blockIdx.x;
  const int n = t.n, dp1 = t.dp1, m = t.m, m2 = t.m2;
  for (int f = 0; f < ncomp; ++f)
    for (int i = threadIdx.x; i < n; i += blockDim.x) sp[f * n + i] = phi.p[f][(int64_t)(t.offset + leaf) * n + i];
  for (int i = threadIdx.x; i < dp1 * m; i += blockDim.x) sTt[i] = t.p2mT[i];
  __syncthreads();
  for (int o = threadIdx.x; o < m2; o += blockDim.x) {
    for (int f = 0; f < ncomp; ++f) {
      double v = contract(sTt, sTt, sp + f * n, dp1, o / m, o % m);
      int64_t u = (int64_t)leaf * m2 + o;
      if (part.p[f]) v += part.p[f][u];
      sols.p[f][u] = v;
    }
  }
}

// Identity L2P of the reference layout: sols = leaf proxy rows (+ particle field).
__global__ void k_l2p_copy(int64_t npart, int ncomp, CPtr3 leafrows, CPtr3 part, Ptr3 sols) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= npart) return;
  for (int f = 0; f < ncomp; ++f) {
    double v = leafrows.p[f][u];
    if (part.p[f]) v += part.p[f][u];
    sols.p[f][u] = v;
  }
}

__global__ void k_fill_zero(double* p, int64_t count) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u < count) p[u] = 0.0;
}

__global__ void k_copy(double* __restrict__ dst, const double* __restrict__ src, int64_t count) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u < count) dst[u] = src[u];
}

inline unsigned blocks_for(int64_t count, int bs) { return (unsigned)((count + bs - 1) / bs); }

bool is_identity_p2m(const DevTree& t) { return t.m == t.dp1; }

}  // namespace

void launch_fill_zero(double* p, int64_t count, hipStream_t s) {
  if (count > 0) hipLaunchKernelGGL(k_fill_zero, dim3(blocks_for(count, 256)), dim3(256), 0, s, p, count);
}

void launch_copy(double* dst, const double* src, int64_t count, hipStream_t s) {
  if (count > 0) hipLaunchKernelGGL(k_copy, dim3(blocks_for(count, 256)), dim3(256), 0, s, dst, src, count);
}

// fmm_funcs.cpp:341-355.  Assigns q_part and every row of q_proxy (the
// reference accumulates into caller-zeroed rows: bve_time_step.cpp:321).
void launch_upward_pass(const DevTree& t, const double* area, const double* pots, double* q_part, double* q_proxy,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_base_pots, dim3(blocks_for(t.npart, 256)), dim3(256), 0, s, t.npart, area, pots, q_part);
  if (is_identity_p2m(t))
    launch_copy(q_proxy + (int64_t)t.offset * t.n, q_part, t.npart, s);
  else
    hipLaunchKernelGGL(k_p2m, dim3(t.Pl), dim3(128), 0, s, t, (const double*)q_part, q_proxy);
  // Parents level L-2 .. 0; level l starts at 2*(4^l - 1).
  for (int l = t.levels - 2; l >= 0; --l) {
    int pbegin = 2 * ((1 << (2 * l)) - 1);
    int count = 6 << (2 * l);
    hipLaunchKernelGGL(k_m2m, dim3(count), dim3(128), 0, s, t, pbegin, q_proxy);
  }
}

// fmm_funcs.cpp:357-392.  phi_proxy rows are accumulated top-down; sols are
// assigned = L2P(leaf proxies) + phi_part (phi_part entries may be null).
void launch_downward_pass(const DevTree& t, int ncomp, double* const* phi_proxy, const double* const* phi_part,
                          double* const* sols, hipStream_t s) {
  Ptr3 phi{}; CPtr3 part{}, leafrows{}; Ptr3 out{};
  for (int f = 0; f < ncomp; ++f) {
    phi.p[f] = phi_proxy[f];
    part.p[f] = phi_part ? phi_part[f] : nullptr;
    leafrows.p[f] = phi_proxy[f] + (int64_t)t.offset * t.n;
    out.p[f] = sols[f];
  }
  for (int l = 0; l + 1 < t.levels; ++l) {
    int pbegin = 2 * ((1 << (2 * l)) - 1);
    int count = 6 << (2 * l);
    hipLaunchKernelGGL(k_l2l, dim3(count), dim3(256), 0, s, t, pbegin, ncomp, phi);
  }
  CPtr3 cphi{};
  for (int f = 0; f < ncomp; ++f) cphi.p[f] = phi_proxy[f];
  if (is_identity_p2m(t))
    hipLaunchKernelGGL(k_l2p_copy, dim3(blocks_for(t.npart, 256)), dim3(256), 0, s, t.npart, ncomp, leafrows, part,
                       out);
  else
    hipLaunchKernelGGL(k_l2p, dim3(t.Pl), dim3(128), 0, s, t, ncomp, cphi, part, out);
}

}  // namespace sbslpm
