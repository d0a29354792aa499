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

// Component f of a pointer triple as a select, so that the triple stays in
// registers (dynamic indexing of a kernel-argument array goes to scratch).
template <typename P3>
__device__ inline auto comp(const P3& a, int f) -> decltype(a.p[0]) {
  return f == 0 ? a.p[0] : (f == 1 ? a.p[1] : a.p[2]);
}

// The transfers are tensor products, out = X in Y^T with small dense X, Y;
// every kernel below evaluates them as two 1-D sweeps through LDS
// (R^2 (R + Q) instead of Q^2 R^2 multiply-adds).

// q_part = area * pots (base_pots, fmm_funcs_impl.hpp:29-33).
__global__ void k_base_pots(int64_t npart, const double* __restrict__ area, const double* __restrict__ pots,
                            double* __restrict__ q_part) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u < npart) q_part[u] = area[u] * pots[u];
}

// Leaf proxies from the particle grid: Q_leaf[a][b] = sum_kl T[a][k] T[b][l] q[k][l],
// T[a][k] = l_a(x_k) (base_pots_ll, fmm_funcs_impl.hpp:207-223, on the leaf grid).
__global__ __launch_bounds__(128) void k_p2m(DevTree t, const double* __restrict__ q_part,
                                             double* __restrict__ q_proxy) {
  __shared__ double sq[kMaxN], sv[kMaxN], sT[kMaxN];
  const int leaf = blockIdx.x;
  const int m = t.m, m2 = t.m2, dp1 = t.dp1, n = t.n;
  for (int i = threadIdx.x; i < m2; i += blockDim.x) sq[i] = q_part[(int64_t)leaf * m2 + i];
  for (int i = threadIdx.x; i < dp1 * m; i += blockDim.x) sT[i] = t.p2m[i];
  __syncthreads();
  for (int o = threadIdx.x; o < m * dp1; o += blockDim.x) {  // V[k][b] = sum_l T[b][l] q[k][l]
    const int k = o / dp1, b = o % dp1;
    double v = 0;
    for (int l = 0; l < m; ++l) v += sT[b * m + l] * sq[k * m + l];
    sv[o] = v;
  }
  __syncthreads();
  for (int o = threadIdx.x; o < n; o += blockDim.x) {  // Q[a][b] = sum_k T[a][k] V[k][b]
    const int a = o / dp1, b = o % dp1;
    double v = 0;
    for (int k = 0; k < m; ++k) v += sT[a * m + k] * sv[k * dp1 + b];
    q_proxy[(int64_t)(t.offset + leaf) * n + o] = v;
  }
}

// M2M for all parents of one level (child_to_parent, fmm_funcs_impl.hpp:47-70):
// Q_p[a][b] = sum_c sum_kl X_c[a][k] Y_c[b][l] Q_c[k][l], assigning the parent rows.
// X_c / Y_c = upper- or lower-half matrix in xi / eta (children 1,4 upper in xi; 1,2 upper in eta).
// Parents are pbegin + blockIdx.x, or list[blockIdx.x] when a list is given
// (the irregular tree's non-leaf panels of one level).
__global__ __launch_bounds__(128) void k_m2m(DevTree t, int pbegin, const int* __restrict__ list,
                                             double* __restrict__ q) {
  __shared__ double sq[4 * kMaxN], sv[4 * kMaxN];
  __shared__ double shi[kMaxN], slo[kMaxN];
  const int par = list ? list[blockIdx.x] : pbegin + blockIdx.x;
  const int n = t.n, dp1 = t.dp1;
  const int c0 = t.panels[par].child0;
  q += member_proxy(t);
  for (int i = threadIdx.x; i < 4 * n; i += blockDim.x) sq[i] = q[(int64_t)c0 * n + i];
  for (int i = threadIdx.x; i < n; i += blockDim.x) { shi[i] = t.m2m_hi[i]; slo[i] = t.m2m_lo[i]; }
  __syncthreads();
  for (int o = threadIdx.x; o < 4 * n; o += blockDim.x) {  // V_c[k][b] = sum_l Y_c[b][l] Q_c[k][l]
    const int c = o / n, k = (o % n) / dp1, b = o % dp1;
    const double* Y = (c <= 1) ? shi : slo;
    const double* Q = sq + c * n;
    double v = 0;
    for (int l = 0; l < dp1; ++l) v += Y[b * dp1 + l] * Q[k * dp1 + l];
    sv[o] = v;
  }
  __syncthreads();
  for (int o = threadIdx.x; o < n; o += blockDim.x) {  // Q_p[a][b] = sum_c sum_k X_c[a][k] V_c[k][b]
    const int a = o / dp1, b = o % dp1;
    double v = 0;
    for (int c = 0; c < 4; ++c) {
      const double* X = (c == 0 || c == 3) ? shi : slo;
      for (int k = 0; k < dp1; ++k) v += X[a * dp1 + k] * sv[c * n + k * dp1 + b];
    }
    q[(int64_t)par * n + o] = v;
  }
}

// base_pots + P2M + the first M2M level fused, one workgroup per parent of
// the leaves: q_part = area * pots is written for the pair kernels, the four
// leaves' proxy charges stay in LDS (leaf proxy rows are never interaction
// sources, so they are not written) and the parent row is assigned.
__global__ __launch_bounds__(256) void k_p2m_m2m(DevTree t, int pbegin, const double* __restrict__ area,
                                                 const double* __restrict__ pots, double* __restrict__ q_part,
                                                 double* __restrict__ q) {
  __shared__ double sq[4 * kMaxN], sv[4 * kMaxN], sQ[4 * kMaxN], sw[4 * kMaxN];
  __shared__ double shi[kMaxN], slo[kMaxN], sT[kMaxN];
  const int par = pbegin + blockIdx.x;
  const int n = t.n, dp1 = t.dp1, m = t.m, m2 = t.m2;
  const int leaf0 = t.panels[par].child0 - t.offset;
  pots += member_part(t);
  q_part += member_part(t);
  q += member_proxy(t);
  for (int i = threadIdx.x; i < 4 * m2; i += blockDim.x) {
    const int64_t u = (int64_t)leaf0 * m2 + i;
    const double v = area[u] * pots[u];
    q_part[u] = v;
    sq[i] = v;
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) { shi[i] = t.m2m_hi[i]; slo[i] = t.m2m_lo[i]; }
  for (int i = threadIdx.x; i < dp1 * m; i += blockDim.x) sT[i] = t.p2m[i];
  __syncthreads();
  const double* Q = sq;
  if (m != dp1) {  // P2M per leaf: Q_c[a][b] = sum_kl T[a][k] T[b][l] q_c[k][l]
    const int vn = m * dp1;
    for (int o = threadIdx.x; o < 4 * vn; o += blockDim.x) {
      const int c = o / vn, k = (o % vn) / dp1, b = o % dp1;
      double v = 0;
      for (int l = 0; l < m; ++l) v += sT[b * m + l] * sq[c * m2 + k * m + l];
      sv[o] = v;
    }
    __syncthreads();
    for (int o = threadIdx.x; o < 4 * n; o += blockDim.x) {
      const int c = o / n, a = (o % n) / dp1, b = o % dp1;
      double v = 0;
      for (int k = 0; k < m; ++k) v += sT[a * m + k] * sv[c * vn + k * dp1 + b];
      sQ[o] = v;
    }
    __syncthreads();
    Q = sQ;
  }
  for (int o = threadIdx.x; o < 4 * n; o += blockDim.x) {  // V_c[k][b] = sum_l Y_c[b][l] Q_c[k][l]
    const int c = o / n, k = (o % n) / dp1, b = o % dp1;
    const double* Y = (c <= 1) ? shi : slo;
    double v = 0;
    for (int l = 0; l < dp1; ++l) v += Y[b * dp1 + l] * Q[c * n + k * dp1 + l];
    sw[o] = v;
  }
  __syncthreads();
  for (int o = threadIdx.x; o < n; o += blockDim.x) {  // Q_p[a][b] = sum_c sum_k X_c[a][k] V_c[k][b]
    const int a = o / dp1, b = o % dp1;
    double v = 0;
    for (int c = 0; c < 4; ++c) {
      const double* X = (c == 0 || c == 3) ? shi : slo;
      for (int k = 0; k < dp1; ++k) v += X[a * dp1 + k] * sw[c * n + k * dp1 + b];
    }
    q[(int64_t)par * n + o] = v;
  }
}

// L2L for all parents of one level (parent_to_child_3, fmm_funcs_impl.hpp:138-171),
// accumulating into the four children:
// phi_c[k][l] += sum_a X_c[a][k] sum_b Y_c[b][l] phi_p[a][b].  The eta sweep
// depends only on the eta half, so it is done once for the upper and once
// for the lower half.
__global__ __launch_bounds__(256) void k_l2l(DevTree t, int pbegin, const int* __restrict__ list, int ncomp, Ptr3 phi,
                                             int assign) {
  __shared__ double sp[3 * kMaxN], sv[3 * 2 * kMaxN];
  __shared__ double shi[kMaxN], slo[kMaxN];
  const int par = list ? list[blockIdx.x] : pbegin + blockIdx.x;
  const int n = t.n, dp1 = t.dp1;
  const int c0 = t.panels[par].child0;
  const int64_t mx = member_proxy(t);
  for (int f = 0; f < ncomp; ++f)
    for (int i = threadIdx.x; i < n; i += blockDim.x) sp[f * n + i] = comp(phi, f)[mx + (int64_t)par * n + i];
  for (int i = threadIdx.x; i < n; i += blockDim.x) { shi[i] = t.m2m_hi[i]; slo[i] = t.m2m_lo[i]; }
  __syncthreads();
  for (int o = threadIdx.x; o < ncomp * 2 * n; o += blockDim.x) {  // V_h[a][l] = sum_b Y_h[b][l] phi[a][b]
    const int f = o / (2 * n), h = (o / n) & 1, a = (o % n) / dp1, l = o % dp1;
    const double* Y = h ? slo : shi;
    const double* P = sp + f * n;
    double v = 0;
    for (int b = 0; b < dp1; ++b) v += Y[b * dp1 + l] * P[a * dp1 + b];
    sv[o] = v;
  }
  __syncthreads();
  for (int o = threadIdx.x; o < ncomp * 4 * n; o += blockDim.x) {  // phi_c[k][l] += sum_a X_c[a][k] V_h(c)[a][l]
    const int f = o / (4 * n), c = (o / n) & 3, k = (o % n) / dp1, l = o % dp1;
    const double* X = (c == 0 || c == 3) ? shi : slo;
    const double* V = sv + (f * 2 + (c <= 1 ? 0 : 1)) * n;
    double v = 0;
    for (int a = 0; a < dp1; ++a) v += X[a * dp1 + k] * V[a * dp1 + l];
    double* dst = comp(phi, f) + mx + (int64_t)(c0 + c) * n + k * dp1 + l;
    *dst = assign ? v : *dst + v;
  }
}

// L2P onto the leaf particle grid plus the direct (type 0/1) particle field:
// child_panel_interp_3_ll (fmm_funcs_impl.hpp:341-362) on the leaf grid:
// sols[k][l] = sum_a T[a][k] sum_b T[b][l] phi[a][b] (+ part).
__global__ __launch_bounds__(128) void k_l2p(DevTree t, int ncomp, CPtr3 phi, CPtr3 part, Ptr3 sols) {
  __shared__ double sp[3 * kMaxN], sv[3 * kMaxN];
  __shared__ double sT[kMaxN];
  const int leaf = blockIdx.x;
  const int n = t.n, dp1 = t.dp1, m = t.m, m2 = t.m2;
  const int64_t mp = member_part(t), mx = member_proxy(t);
  for (int f = 0; f < ncomp; ++f)
    for (int i = threadIdx.x; i < n; i += blockDim.x)
      sp[f * n + i] = comp(phi, f)[mx + (int64_t)(t.offset + leaf) * n + i];
  for (int i = threadIdx.x; i < dp1 * m; i += blockDim.x) sT[i] = t.p2m[i];
  __syncthreads();
  const int vn = dp1 * m;
  for (int o = threadIdx.x; o < ncomp * vn; o += blockDim.x) {  // V[a][l] = sum_b T[b][l] phi[a][b]
    const int f = o / vn, a = (o % vn) / m, l = o % m;
    const double* P = sp + f * n;
    double v = 0;
    for (int b = 0; b < dp1; ++b) v += sT[b * m + l] * P[a * dp1 + b];
    sv[o] = v;
  }
  __syncthreads();
  for (int o = threadIdx.x; o < ncomp * m2; o += blockDim.x) {  // sol[k][l] = sum_a T[a][k] V[a][l]
    const int f = o / m2, k = (o % m2) / m, l = o % m;
    const double* V = sv + f * vn;
    double v = 0;
    for (int a = 0; a < dp1; ++a) v += sT[a * m + k] * V[a * m + l];
    const int64_t u = mp + (int64_t)leaf * m2 + (o % m2);
    const double* pf = comp(part, f);
    if (pf) v += pf[u];
    comp(sols, f)[u] = v;
  }
}

// Identity L2P of the reference layout: sols = leaf proxy rows (+ particle field).
__global__ void k_l2p_copy(int64_t npart, int ncomp, CPtr3 leafrows, CPtr3 part, Ptr3 sols, int64_t proxy_stride) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= npart) return;
  const int64_t ux = (int64_t)blockIdx.y * proxy_stride + u;
  u += (int64_t)blockIdx.y * npart;
  for (int f = 0; f < ncomp; ++f) {
    double v = comp(leafrows, f)[ux];
    const double* pf = comp(part, f);
    if (pf) v += pf[u];
    comp(sols, f)[u] = v;
  }
}

__global__ void k_zero_rows(const int* __restrict__ rows, int nrows, int n, int ncomp, Ptr3 a, int64_t member_stride) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)nrows * n) return;
  const int64_t o = (int64_t)blockIdx.y * member_stride + (int64_t)rows[i / n] * n + i % n;
  for (int f = 0; f < 3; ++f)
    if (f < ncomp) a.p[f][o] = 0.0;
}

__global__ void k_fill_zero(double* p, int64_t count) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u < count) p[u] = 0.0;
}

__global__ void k_fill_value(double* p, int64_t count, double v) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u < count) p[u] = v;
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

void launch_fill_value(double* p, int64_t count, double v, hipStream_t s) {
  if (count > 0) hipLaunchKernelGGL(k_fill_value, dim3(blocks_for(count, 256)), dim3(256), 0, s, p, count, v);
}

void launch_zero_rows(const int* rows, int nrows, int n, int ncomp, double* const* arrays, hipStream_t s, int nmem,
                      int64_t member_stride) {
  if (nrows <= 0) return;
  Ptr3 a{};
  for (int f = 0; f < ncomp; ++f) a.p[f] = arrays[f];
  const int64_t cnt = (int64_t)nrows * n;
  hipLaunchKernelGGL(k_zero_rows, dim3(blocks_for(cnt, 256), nmem), dim3(256), 0, s, rows, nrows, n, ncomp, a,
                     member_stride);
}

void launch_copy(double* dst, const double* src, int64_t count, hipStream_t s) {
  if (count > 0) hipLaunchKernelGGL(k_copy, dim3(blocks_for(count, 256)), dim3(256), 0, s, dst, src, count);
}

// M2M for parent levels lfirst .. 0, a launch per level (level l starts at
// panel 2 (4^l - 1) and holds 6 4^l panels).  A launch per level measured
// faster than one workgroup per cube face walking the top levels with a
// barrier between them: that kernel's waves wait on a chain of dependent
// loads with too few waves per compute unit to hide them (178 us against
// about 30 us for the four launches it replaced at 6 x 512^2).
static void launch_m2m_levels(const DevTree& t, int lfirst, double* q_proxy, hipStream_t s, int nmem) {
  for (int l = lfirst; l >= 0; --l) {
    const int pbegin = 2 * ((1 << (2 * l)) - 1), count = 6 << (2 * l);
    hipLaunchKernelGGL(k_m2m, dim3(count, nmem), dim3(128), 0, s, t, pbegin, (const int*)nullptr, q_proxy);
  }
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
  launch_m2m_levels(t, t.levels - 2, q_proxy, s, 1);
}

// Upward pass with base_pots, P2M and the first M2M level fused: q_part and
// the non-leaf rows of q_proxy are assigned, the leaf rows are not written.
void launch_upward_fused(const DevTree& t, const double* area, const double* pots, double* q_part, double* q_proxy,
                         hipStream_t s, int nmem) {
  if (t.levels < 2) {
    launch_upward_pass(t, area, pots, q_part, q_proxy, s);
    return;
  }
  {
    const int l = t.levels - 2;
    const int pbegin = 2 * ((1 << (2 * l)) - 1), count = 6 << (2 * l);
    hipLaunchKernelGGL(k_p2m_m2m, dim3(count, nmem), dim3(256), 0, s, t, pbegin, area, pots, q_part, q_proxy);
  }
  launch_m2m_levels(t, t.levels - 3, q_proxy, s, nmem);
}

// fmm_funcs.cpp:357-392.  phi_proxy rows are accumulated top-down; sols are
// assigned = L2P(leaf proxies) + phi_part (phi_part entries may be null).
void launch_l2l(const DevTree& t, int ncomp, double* const* phi_proxy, hipStream_t s, bool assign_leaf_rows,
                int nmem) {
  Ptr3 phi{};
  for (int f = 0; f < ncomp; ++f) phi.p[f] = phi_proxy[f];
  for (int l = 0; l + 1 < t.levels; ++l) {
    int pbegin = 2 * ((1 << (2 * l)) - 1);
    int count = 6 << (2 * l);
    const int assign = (assign_leaf_rows && l + 2 == t.levels) ? 1 : 0;
    hipLaunchKernelGGL(k_l2l, dim3(count, nmem), dim3(256), 0, s, t, pbegin, (const int*)nullptr, ncomp, phi, assign);
  }
}

void launch_downward_pass(const DevTree& t, int ncomp, double* const* phi_proxy, const double* const* phi_part,
                          double* const* sols, hipStream_t s, bool assign_leaf_rows, int nmem) {
  CPtr3 part{}, leafrows{}; Ptr3 out{};
  for (int f = 0; f < ncomp; ++f) {
    part.p[f] = phi_part ? phi_part[f] : nullptr;
    leafrows.p[f] = phi_proxy[f] + (int64_t)t.offset * t.n;
    out.p[f] = sols[f];
  }
  launch_l2l(t, ncomp, phi_proxy, s, assign_leaf_rows, nmem);
  CPtr3 cphi{};
  for (int f = 0; f < ncomp; ++f) cphi.p[f] = phi_proxy[f];
  if (is_identity_p2m(t))
    hipLaunchKernelGGL(k_l2p_copy, dim3(blocks_for(t.npart, 256), nmem), dim3(256), 0, s, t.npart, ncomp, leafrows,
                       part, out, (int64_t)t.P * t.n);
  else
    hipLaunchKernelGGL(k_l2p, dim3(t.Pl, nmem), dim3(128), 0, s, t, ncomp, cphi, part, out);
}

void launch_m2m_list(const DevTree& t, const int* parents, int count, double* q_proxy, hipStream_t s, int nmem) {
  if (count > 0) hipLaunchKernelGGL(k_m2m, dim3(count, nmem), dim3(128), 0, s, t, 0, parents, q_proxy);
}

void launch_l2l_list(const DevTree& t, const int* parents, int count, int ncomp, double* const* phi_proxy,
                     hipStream_t s) {
  Ptr3 phi{};
  for (int f = 0; f < ncomp; ++f) phi.p[f] = phi_proxy[f];
  if (count > 0) hipLaunchKernelGGL(k_l2l, dim3(count), dim3(256), 0, s, t, 0, parents, ncomp, phi, 0);
}

}  // namespace sbslpm
