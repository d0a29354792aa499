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

// The downward pass fused into one launch, one workgroup per parent of the
// leaves (level L-2), for trees whose leaf rows are not pair-sum targets.
// fmm_funcs.cpp:357-392 accumulates every level's rows into the next level
// (L2L) and interpolates the leaves' rows to the particles (L2P).  A panel's
// accumulated row is its own row plus each ancestor's own row evaluated at
// the panel's nodes, since each level's interpolant reproduces the degree-d
// polynomial it is given.  So the workgroup
//   1. forms the parent's accumulated row from its own row and its ancestors'
//      through the composite matrices (DevTree::anc_l2l), up to kAncChunk
//      ancestors per LDS pass;
//   2. forms the four leaves' rows by one L2L step (k_l2l's two sweeps), in LDS;
//   3. interpolates them to the leaf particle grids (k_l2p's two sweeps) and
//      adds the particle field.
// Nothing is written but sols: the proxy rows stay as the pair sums left
// them.  One launch replaces L-1 L2L levels and the L2P, and the leaf rows
// make no round trip through HBM.
constexpr int kAncChunk = 6;

__global__ __launch_bounds__(256) void k_l2l_l2p(DevTree t, int ncomp, CPtr3 phi, CPtr3 part, Ptr3 sols) {
  extern __shared__ double lds[];
  const int n = t.n, D = t.dp1, DD = D * D, m = t.m, m2 = t.m2;
  const int lp = t.levels - 2, ip = blockIdx.x;
  const int par = 2 * ((1 << (2 * lp)) - 1) + ip;
  const int64_t mp = member_part(t), mx = member_proxy(t);
  const int nc = ncomp * n;
  double* acc = lds;        // [ncomp][n] accumulated parent row
  double* R = lds + 3 * n;  // phase 1 and phase 2 scratch
  for (int i = threadIdx.x; i < nc; i += blockDim.x) acc[i] = comp(phi, i / n)[mx + (int64_t)par * n + i % n];
  for (int g0 = 0; g0 < lp; g0 += kAncChunk) {
    const int na = min(kAncChunk, lp - g0);
    double* sP = R;             // [na][ncomp][n] ancestors' own rows
    double* sV = sP + na * nc;  // [na][ncomp][n] eta sweeps
    double* sC = sV + na * nc;  // [na][xi, eta][D][D] composite matrices
    __syncthreads();  // the previous chunk's sweeps are done with sP, sV, sC
    for (int i = threadIdx.x; i < na * nc; i += blockDim.x) {
      const int g = g0 + 1 + i / nc, j = i % nc;
      const int anc = 2 * ((1 << (2 * (lp - g))) - 1) + (ip >> (2 * g));
      sP[i] = comp(phi, j / n)[mx + (int64_t)anc * n + j % n];
    }
    for (int i = threadIdx.x; i < na * 2 * DD; i += blockDim.x) {
      const int g = g0 + 1 + i / (2 * DD), eta = (i / DD) & 1;
      // position of the parent's interval in the ancestor's: bit s-1 is the
      // upper-half flag of the step s levels above the parent (quadrants
      // (hi,hi), (lo,hi), (lo,lo), (hi,lo) in (xi, eta))
      int pos = 0;
      for (int s = 1; s <= g; ++s) {
        const int c = (ip >> (2 * (s - 1))) & 3;
        pos |= (eta ? (c <= 1) : (c == 0 || c == 3)) << (s - 1);
      }
      sC[i] = t.anc_l2l[(int64_t)((1 << g) - 2 + pos) * DD + i % DD];
    }
    __syncthreads();
    for (int o = threadIdx.x; o < na * nc; o += blockDim.x) {  // V[a][l] = sum_b Ceta[l][b] P[a][b]
      const int r = o % n, a = r / D, l = r % D;
      const double* Ce = sC + ((o / nc) * 2 + 1) * DD;
      const double* P = sP + (o - r);
      double v = 0;
      for (int b = 0; b < D; ++b) v += Ce[l * D + b] * P[a * D + b];
      sV[o] = v;
    }
    __syncthreads();
    for (int o = threadIdx.x; o < nc; o += blockDim.x) {  // acc[k][l] += sum_g sum_a Cxi[k][a] V[a][l]
      const int r = o % n, k = r / D, l = r % D;
      double v = acc[o];
      for (int gi = 0; gi < na; ++gi) {
        const double* Cx = sC + gi * 2 * DD;
        const double* V = sV + gi * nc + (o - r);
        for (int a = 0; a < D; ++a) v += Cx[k * D + a] * V[a * D + l];
      }
      acc[o] = v;
    }
  }
  __syncthreads();
  double* shi = R;               // [D][D]
  double* slo = shi + DD;        // [D][D]
  double* sT = slo + DD;         // [D][m]
  double* sVh = sT + D * m;      // [ncomp][2][n]
  double* sL = sVh + 2 * nc;     // [ncomp][4][n] leaf rows
  double* sW = sL + 4 * nc;      // [ncomp][4][D][m]
  for (int i = threadIdx.x; i < DD; i += blockDim.x) { shi[i] = t.m2m_hi[i]; slo[i] = t.m2m_lo[i]; }
  for (int i = threadIdx.x; i < D * m; i += blockDim.x) sT[i] = t.p2m[i];
  __syncthreads();
  for (int o = threadIdx.x; o < 2 * nc; o += blockDim.x) {  // V_h[a][l] = sum_b Y_h[b][l] acc[a][b]
    const int f = o / (2 * n), h = (o / n) & 1, a = (o % n) / D, l = o % D;
    const double* Y = h ? slo : shi;
    const double* P = acc + f * n;
    double v = 0;
    for (int b = 0; b < D; ++b) v += Y[b * D + l] * P[a * D + b];
    sVh[o] = v;
  }
  __syncthreads();
  for (int o = threadIdx.x; o < 4 * nc; o += blockDim.x) {  // leaf_c[k][l] = sum_a X_c[a][k] V_h(c)[a][l]
    const int f = o / (4 * n), c = (o / n) & 3, k = (o % n) / D, l = o % D;
    const double* X = (c == 0 || c == 3) ? shi : slo;
    const double* V = sVh + (f * 2 + (c <= 1 ? 0 : 1)) * n;
    double v = 0;
    for (int a = 0; a < D; ++a) v += X[a * D + k] * V[a * D + l];
    sL[o] = v;
  }
  __syncthreads();
  const int64_t leaf0 = 4 * (int64_t)ip;  // leaves are the last level: panel index - offset
  if (m == D) {  // the reference layout: the particle grid is the proxy grid
    for (int o = threadIdx.x; o < 4 * nc; o += blockDim.x) {
      const int f = o / (4 * n);
      const int64_t u = mp + leaf0 * m2 + o % (4 * n);
      const double* pf = comp(part, f);
      double v = sL[o];
      if (pf) v += pf[u];
      comp(sols, f)[u] = v;
    }
    return;
  }
  const int vn = D * m;
  for (int o = threadIdx.x; o < 4 * ncomp * vn; o += blockDim.x) {  // W[a][l] = sum_b T[b][l] leaf[a][b]
    const int lc = o / vn, a = (o % vn) / m, l = o % m;
    const double* P = sL + lc * n;
    double v = 0;
    for (int b = 0; b < D; ++b) v += sT[b * m + l] * P[a * D + b];
    sW[o] = v;
  }
  __syncthreads();
  for (int o = threadIdx.x; o < 4 * ncomp * m2; o += blockDim.x) {  // sol[k][l] = sum_a T[a][k] W[a][l]
    const int lc = o / m2, f = lc / 4, c = lc % 4, k = (o % m2) / m, l = o % m;
    const double* V = sW + lc * vn;
    double v = 0;
    for (int a = 0; a < D; ++a) v += sT[a * m + k] * V[a * m + l];
    const int64_t u = mp + (leaf0 + c) * m2 + (o % m2);
    const double* pf = comp(part, f);
    if (pf) v += pf[u];
    comp(sols, f)[u] = v;
  }
}

// LDS accesses of one wave complete in order; the fences keep the compiler
// from moving a lane's LDS reads above the other lanes' writes.
__device__ inline void wave_lds_order() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// A read-only table with a wave-uniform address: a pointer into the constant
// address space, which the compiler reads with scalar loads.
typedef const __attribute__((address_space(4))) double* ConstTab;
__device__ inline ConstTab const_tab(const double* p) { return (ConstTab)p; }

// Charge sets of one upward launch (UpCharge in device.hpp), entry y taken by
// blockIdx.y; n = 0: the member slices of entry 0.
struct UpSets {
  const double* area[kMaxUpSets];
  const double* pots[kMaxUpSets];
  double* qp[kMaxUpSets];
  double* qx[kMaxUpSets];
  int n;
};

// a[y] for a wave-uniform y, as a chain of selects (no indexed copy of the
// argument array).
template <typename T>
__device__ inline T pick_set(T const (&a)[kMaxUpSets], int y) {
  T v = a[0];
#pragma unroll
  for (int i = 1; i < kMaxUpSets; ++i)
    if (y == i) v = a[i];
  return v;
}

// k_l2l_l2p for a proxy grid of D x D nodes and a leaf grid of M x M
// particles known at compile time, tiled for registers.  Every sweep is a
// batch of small products out[r][k] = sum_b in[r][b] C[k][b] whose matrix C
// is the same for all lanes of a wave: a lane owns one row r (a component and
// a node index) and keeps it in registers, and the loop over outputs k reads
// row k of C through the scalar unit, so a multiply-add costs no LDS read.
//   phase 1: wave w evaluates ancestors g = w+1, w+5, ... at the parent's
//     nodes (the eta sweep by rows, the xi sweep by columns through the
//     wave's LDS slice, its sum kept in the slice); the parent's own row and
//     the four waves' sums are added in a fixed order;
//   phase 2: waves 0 and 1 evaluate the parent's row at the particle nodes of
//     the upper and lower half in eta (l2p_hi, l2p_lo); wave c then forms
//     leaf c's particle values by the xi sweep and adds the particle field.
// The leaf rows and the L2P matrix are folded into l2p_hi / l2p_lo (the
// parent's interpolant evaluated at the leaves' particle nodes).
template <int D, int M>
__global__ __launch_bounds__(256) void k_down_fused(DevTree t, int ncomp, CPtr3 phi, CPtr3 part, Ptr3 sols) {
  constexpr int N = D * D, M2 = M * M;
  // LDS is the occupancy limit, so phase 2 reuses phase 1's buffers: the W
  // rows go where the waves' V slices were (phase 1 is over at the barrier
  // before they are written), and the summed row goes over wave 0's sum (each
  // thread reads its own entries of the four sums, then overwrites the first).
  __shared__ double sV[4][3 * N];
  __shared__ double sPart[4][3 * N];
  static_assert(2 * 3 * D * M <= 4 * 3 * N, "k_down_fused: W rows do not fit the V slices");
  double* const sAcc = sPart[0];
  double(*const sW)[3 * D * M] = reinterpret_cast<double(*)[3 * D * M]>(&sV[0][0]);
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lp = t.levels - 2, ip = blockIdx.x;
  const int par = 2 * ((1 << (2 * lp)) - 1) + ip;
  const int64_t mp = member_part(t), mx = member_proxy(t);
  static_assert(3 * N <= 512 && 3 * D <= 64 && 3 * M <= 64, "k_down_fused: grid too large for one workgroup");
  const bool row_lane = lane < ncomp * D;
  const int f = row_lane ? lane / D : 0, a = lane - f * D;
  const double* __restrict__ phif = comp(phi, f) + mx;
  // Loads whose latency would otherwise sit in the dependent chain: the
  // parent's own row (two entries per thread at most) and, for the lanes of
  // the last sweep, the particle field of leaf c = w.
  const int nall = ncomp * N;
  double own[2] = {0.0, 0.0};
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int o = threadIdx.x + 256 * i;
    if (o < nall) own[i] = comp(phi, o / N)[mx + (int64_t)par * N + o % N];
  }
  const bool col_lane = lane < ncomp * M;
  const int fc = col_lane ? lane / M : 0, lc = lane - fc * M;
  const int64_t u0 = mp + (4 * (int64_t)ip + w) * M2 + lc;
  const double* pfc = comp(part, fc);
  double pv[M];
#pragma unroll
  for (int k = 0; k < M; ++k) pv[k] = (col_lane && pfc) ? pfc[u0 + k * M] : 0.0;
  double* part_w = &sPart[w][f * N + a];  // this lane's column of the wave's sum: [k * D]
  if (row_lane)
    for (int k = 0; k < D; ++k) part_w[k * D] = 0.0;
  for (int g = w + 1; g <= lp; g += 4) {
    int px = 0, py = 0;
    for (int s = 1; s <= g; ++s) {
      const int c = (ip >> (2 * (s - 1))) & 3;
      px |= (c == 0 || c == 3) << (s - 1);
      py |= (c <= 1) << (s - 1);
    }
    const ConstTab Cx = const_tab(t.anc_l2l) + (int64_t)((1 << g) - 2 + px) * N;
    const ConstTab Cy = const_tab(t.anc_l2l) + (int64_t)((1 << g) - 2 + py) * N;
    const int64_t anc = 2 * ((1 << (2 * (lp - g))) - 1) + (ip >> (2 * g));
    if (row_lane) {  // V[a][l] = sum_b Cy[l][b] P[a][b]
      double p[D];
#pragma unroll
      for (int b = 0; b < D; ++b) p[b] = phif[anc * N + a * D + b];
#pragma unroll 3
      for (int l = 0; l < D; ++l) {
        double v = 0.0;
#pragma unroll
        for (int b = 0; b < D; ++b) v += Cy[l * D + b] * p[b];
        sV[w][f * N + a * D + l] = v;
      }
    }
    wave_lds_order();
    if (row_lane) {  // column l = a: sum[k][l] += sum_r Cx[k][r] V[r][l]
      double q[D];
#pragma unroll
      for (int r = 0; r < D; ++r) q[r] = sV[w][f * N + r * D + a];
#pragma unroll 3
      for (int k = 0; k < D; ++k) {
        double v = part_w[k * D];
#pragma unroll
        for (int r = 0; r < D; ++r) v += Cx[k * D + r] * q[r];
        part_w[k * D] = v;
      }
    }
    wave_lds_order();
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int o = threadIdx.x + 256 * i;
    if (o < nall) sAcc[o] = own[i] + sPart[0][o] + sPart[1][o] + sPart[2][o] + sPart[3][o];
  }
  __syncthreads();
  if (w < 2 && row_lane) {  // W_h[a][l] = sum_b Z_h[l][b] acc[a][b], h = w
    const ConstTab Z = const_tab(w == 0 ? t.l2p_hi : t.l2p_lo);
    double p[D];
#pragma unroll
    for (int b = 0; b < D; ++b) p[b] = sAcc[f * N + a * D + b];
#pragma unroll 3
    for (int l = 0; l < M; ++l) {
      double v = 0.0;
#pragma unroll
      for (int b = 0; b < D; ++b) v += Z[l * D + b] * p[b];
      sW[w][(f * D + a) * M + l] = v;
    }
  }
  __syncthreads();
  if (col_lane) {  // leaf c = w: sol[k][l] = sum_r Zx[k][r] W_h(c)[r][l] + part
    const int c = w;
    const ConstTab Zx = const_tab((c == 0 || c == 3) ? t.l2p_hi : t.l2p_lo);
    const double* Wc = sW[c <= 1 ? 0 : 1] + fc * D * M;
    double q[D];
#pragma unroll
    for (int r = 0; r < D; ++r) q[r] = Wc[r * M + lc];
    double* sf = comp(sols, fc);
    // unrolled for the prefetched pv[k]; the barriers keep each output's
    // coefficient row from being loaded ahead (the rows would not fit the
    // scalar registers together)
#pragma unroll
    for (int k = 0; k < M; ++k) {
      double v = 0.0;
#pragma unroll
      for (int r = 0; r < D; ++r) v += Zx[k * D + r] * q[r];
      sf[u0 + k * M] = v + pv[k];
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

template <int D, int M>
bool launch_down_fused_dm(const DevTree& t, int ncomp, CPtr3 phi, CPtr3 part, Ptr3 out, hipStream_t s, int nmem) {
  if (t.dp1 != D || t.m != M) return false;
  hipLaunchKernelGGL((k_down_fused<D, M>), dim3(6 << (2 * (t.levels - 2)), nmem), dim3(256), 0, s, t, ncomp, phi, part,
                     out);
  return true;
}

// The register-tiled form for the layouts in use (degree 8 with an 8 x 8,
// 9 x 9 or 6 x 6 leaf grid, degree 5 with 6 x 6, ...); false for the others.
bool launch_down_fused(const DevTree& t, int ncomp, CPtr3 phi, CPtr3 part, Ptr3 out, hipStream_t s, int nmem) {
  return launch_down_fused_dm<9, 8>(t, ncomp, phi, part, out, s, nmem) ||
         launch_down_fused_dm<9, 9>(t, ncomp, phi, part, out, s, nmem) ||
         launch_down_fused_dm<9, 6>(t, ncomp, phi, part, out, s, nmem) ||
         launch_down_fused_dm<6, 6>(t, ncomp, phi, part, out, s, nmem) ||
         launch_down_fused_dm<5, 5>(t, ncomp, phi, part, out, s, nmem) ||
         launch_down_fused_dm<5, 6>(t, ncomp, phi, part, out, s, nmem) ||
         launch_down_fused_dm<6, 7>(t, ncomp, phi, part, out, s, nmem) ||
         launch_down_fused_dm<7, 7>(t, ncomp, phi, part, out, s, nmem) ||
         launch_down_fused_dm<11, 11>(t, ncomp, phi, part, out, s, nmem) ||
         launch_down_fused_dm<11, 8>(t, ncomp, phi, part, out, s, nmem);
}

// Charge of a parent from its four children's M x M grids (row-major in LDS,
// sq[c * M * M + k * M + l]): Q[a][b] = sum_c sum_kl X_c[a][k] Y_c[b][l]
// q_c[k][l], with X_c = hi for children 0 and 3, Y_c = hi for children 0 and 1
// (quadrants (hi,hi), (lo,hi), (lo,lo), (hi,lo) in (xi, eta)) and the tables
// laid out [basis a][child node k], rows contiguous in k:
//   - wave c: lanes k < M form U_c[k][b] = sum_l Y_c[b][l] q_c[k][l];
//   - wave w: lanes b < D form Q[a][b] for a = w, w + 4, ..., summing the
//     children in order, and write row a of the parent.
// Tiled for registers like k_down_fused; every thread of the workgroup calls it.
template <int D, int M>
__device__ inline void parent_from_children(const double* sq, ConstTab Thi, ConstTab Tlo, double (*sU)[M * D],
                                            double* __restrict__ qrow, int lane, int w) {
  constexpr int M2 = M * M;
  if (lane < M) {
    const ConstTab Y = w <= 1 ? Thi : Tlo;
    double r[M];
#pragma unroll
    for (int l = 0; l < M; ++l) r[l] = sq[w * M2 + lane * M + l];
#pragma unroll 3
    for (int b = 0; b < D; ++b) {
      double v = 0.0;
#pragma unroll
      for (int l = 0; l < M; ++l) v += Y[b * M + l] * r[l];
      sU[w][lane * D + b] = v;
    }
  }
  __syncthreads();
  if (lane < D) {
    double u[4][M];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int k = 0; k < M; ++k) u[c][k] = sU[c][k * D + lane];
#pragma unroll 3
    for (int a = w; a < D; a += 4) {
      double v = 0.0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const ConstTab X = (c == 0 || c == 3) ? Thi : Tlo;
#pragma unroll
        for (int k = 0; k < M; ++k) v += X[a * M + k] * u[c][k];
      }
      qrow[a * D + lane] = v;
    }
  }
}

// k_p2m_m2m for a proxy grid of D x D nodes and a leaf grid of M x M
// particles known at compile time.  The P2M of each leaf and the M2M into
// the parent compose to one product per direction with l2p_hi / l2p_lo, the
// parent basis at the particles of an upper / lower half child (through their
// transposes l2pT): q_part = area * pots for the four leaves is written and
// staged in LDS, then parent_from_children.  blockIdx.y takes charge set y
// of `sets` (sets.n > 0) or member slice y of set 0.
template <int D, int M>
__global__ __launch_bounds__(256) void k_up_fused(DevTree t, int pbegin, UpSets sets) {
  constexpr int N = D * D, M2 = M * M;
  __shared__ double sq[4 * M2];
  __shared__ double sU[4][M * D];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int par = pbegin + blockIdx.x;
  const int64_t leaf0 = 4 * (int64_t)blockIdx.x;  // the parents' children are the leaves 4i .. 4i+3
  const double* __restrict__ area = sets.area[0];
  const double* __restrict__ pots = sets.pots[0] + member_part(t);
  double* __restrict__ q_part = sets.qp[0] + member_part(t);
  double* __restrict__ q = sets.qx[0] + member_proxy(t);
  if (sets.n > 0) {
    const int y = blockIdx.y;
    area = pick_set(sets.area, y);
    pots = pick_set(sets.pots, y);
    q_part = pick_set(sets.qp, y);
    q = pick_set(sets.qx, y);
  }
  for (int i = threadIdx.x; i < 4 * M2; i += blockDim.x) {
    const int64_t u = leaf0 * M2 + i;
    const double v = area[u] * pots[u];
    q_part[u] = v;
    sq[i] = v;
  }
  __syncthreads();
  parent_from_children<D, M>(sq, const_tab(t.l2pT_hi), const_tab(t.l2pT_lo), sU, q + (int64_t)par * N, lane, w);
}

// k_m2m (one level: parents pbegin .., or list[]) in the same tiled form,
// M = D, over the charge sets as k_up_fused.
template <int D>
__global__ __launch_bounds__(256) void k_m2m_tiled(DevTree t, int pbegin, const int* __restrict__ list, UpSets sets) {
  constexpr int N = D * D;
  __shared__ double sq[4 * N];
  __shared__ double sU[4][N];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int par = list ? list[blockIdx.x] : pbegin + blockIdx.x;
  const int c0 = t.panels[par].child0;
  double* __restrict__ q = sets.n > 0 ? pick_set(sets.qx, (int)blockIdx.y) : sets.qx[0] + member_proxy(t);
  for (int i = threadIdx.x; i < 4 * N; i += blockDim.x) sq[i] = q[(int64_t)c0 * N + i];
  __syncthreads();
  parent_from_children<D, D>(sq, const_tab(t.m2m_hi), const_tab(t.m2m_lo), sU, q + (int64_t)par * N, lane, w);
}

// Two M2M levels in one launch, one workgroup per grandparent g (panels
// gbegin ..): wave c forms child c's row from its four children (the 16
// grandchildren rows, consecutive in a uniform tree, staged in LDS; both
// sweeps inside the wave; the eta sweep forms the sums of both halves with
// uniform table reads and keeps the lane's own, the upper half for
// grandchildren 0 and 1; the xi sweep spreads the rows over the wave's lanes
// and reads the matrices from LDS), writes it and keeps it in LDS; parent_from_children
// then forms g's row.  The arithmetic of each row is that of k_m2m_tiled.
// Uniform trees only (every panel above the leaves has four children,
// numbered level by level).
template <int D>
__global__ __launch_bounds__(256) void k_m2m_two(DevTree t, int gbegin, UpSets sets) {
  constexpr int N = D * D;
  static_assert(4 * D <= 64, "k_m2m_two: one lane per grandchild row");
  __shared__ double sg[16 * N];
  __shared__ double sW[4][4 * N];
  __shared__ double sc[4 * N];
  __shared__ double sU[4][N];
  __shared__ double sX[2][N];  // m2m_hi, m2m_lo: the xi sweep reads a row per lane
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int g = gbegin + blockIdx.x;
  const int c0 = t.panels[g].child0;
  const int gc0 = t.panels[c0].child0;
  double* __restrict__ q = sets.n > 0 ? pick_set(sets.qx, (int)blockIdx.y) : sets.qx[0] + member_proxy(t);
  for (int i = threadIdx.x; i < 16 * N; i += blockDim.x) sg[i] = q[(int64_t)gc0 * N + i];
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    sX[0][i] = t.m2m_hi[i];
    sX[1][i] = t.m2m_lo[i];
  }
  __syncthreads();
  const ConstTab Thi = const_tab(t.m2m_hi), Tlo = const_tab(t.m2m_lo);
  if (lane < 4 * D) {  // child c = w, row k of grandchild cc: U[cc][k][b] = sum_l Y_cc[b][l] q[cc][k][l]
    const int cc = lane / D, k = lane - cc * D;
    const bool yhi = cc <= 1;
    double r[D];
#pragma unroll
    for (int l = 0; l < D; ++l) r[l] = sg[((w * 4 + cc) * D + k) * D + l];
#pragma unroll 3
    for (int b = 0; b < D; ++b) {  // both halves' sums (uniform table reads), then the lane's one
      double vh = 0.0, vl = 0.0;
#pragma unroll
      for (int l = 0; l < D; ++l) {
        vh += Thi[b * D + l] * r[l];
        vl += Tlo[b * D + l] * r[l];
      }
      sW[w][(cc * D + k) * D + b] = yhi ? vh : vl;
    }
  }
  wave_lds_order();
  constexpr int A = 64 / D;  // row groups of the xi sweep
  if (lane < A * D) {  // child c = w, column b, rows a = ai, ai + A, ...: Q[a][b] = sum_cc sum_k X_cc[a][k] U[cc][k][b]
    const int ai = lane / D, b = lane - ai * D;
    double u[4][D];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc)
#pragma unroll
      for (int k = 0; k < D; ++k) u[cc][k] = sW[w][(cc * D + k) * D + b];
#pragma unroll 1
    for (int a = ai; a < D; a += A) {
      double v = 0.0;
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) {
        const double* X = sX[(cc == 0 || cc == 3) ? 0 : 1] + a * D;
#pragma unroll
        for (int k = 0; k < D; ++k) v += X[k] * u[cc][k];
      }
      q[(int64_t)(c0 + w) * N + a * D + b] = v;
      sc[w * N + a * D + b] = v;
    }
  }
  __syncthreads();
  parent_from_children<D, D>(sc, Thi, Tlo, sU, q + (int64_t)g * N, lane, w);
}

template <int D>
bool launch_m2m_two_d(const DevTree& t, int gbegin, int count, const UpSets& sets, hipStream_t s, int nmem) {
  if (t.dp1 != D) return false;
  hipLaunchKernelGGL((k_m2m_two<D>), dim3(count, sets.n > 0 ? sets.n : nmem), dim3(256), 0, s, t, gbegin, sets);
  return true;
}

// Levels l and l - 1 of a uniform tree in one launch (grandparents at level
// l - 1); false for a degree without an instance.
bool launch_m2m_two(const DevTree& t, int gbegin, int count, const UpSets& sets, hipStream_t s, int nmem) {
#define SBSLPM_M2M2(D) launch_m2m_two_d<D>(t, gbegin, count, sets, s, nmem)
  return SBSLPM_M2M2(9) || SBSLPM_M2M2(6) || SBSLPM_M2M2(5) || SBSLPM_M2M2(7) || SBSLPM_M2M2(11);
#undef SBSLPM_M2M2
}

template <int D>
bool launch_m2m_tiled_d(const DevTree& t, int pbegin, const int* list, int count, const UpSets& sets, hipStream_t s,
                        int nmem) {
  if (t.dp1 != D) return false;
  hipLaunchKernelGGL((k_m2m_tiled<D>), dim3(count, sets.n > 0 ? sets.n : nmem), dim3(256), 0, s, t, pbegin, list,
                     sets);
  return true;
}

// The tiled M2M for the degrees in use (parents pbegin .. pbegin + count, or
// list[0 .. count)); false for the others.
bool launch_m2m_tiled(const DevTree& t, int pbegin, int count, const UpSets& sets, hipStream_t s, int nmem,
                      const int* list = nullptr) {
#define SBSLPM_M2M(D) launch_m2m_tiled_d<D>(t, pbegin, list, count, sets, s, nmem)
  return SBSLPM_M2M(9) || SBSLPM_M2M(6) || SBSLPM_M2M(5) || SBSLPM_M2M(7) || SBSLPM_M2M(11);
#undef SBSLPM_M2M
}

template <int D, int M>
bool launch_up_fused_dm(const DevTree& t, int pbegin, int count, const UpSets& sets, hipStream_t s, int nmem) {
  if (t.dp1 != D || t.m != M || !t.l2pT_hi) return false;
  hipLaunchKernelGGL((k_up_fused<D, M>), dim3(count, sets.n > 0 ? sets.n : nmem), dim3(256), 0, s, t, pbegin, sets);
  return true;
}

bool launch_up_fused(const DevTree& t, int pbegin, int count, const UpSets& sets, hipStream_t s, int nmem) {
#define SBSLPM_UP(D, M) launch_up_fused_dm<D, M>(t, pbegin, count, sets, s, nmem)
  return SBSLPM_UP(9, 8) || SBSLPM_UP(9, 9) || SBSLPM_UP(9, 6) || SBSLPM_UP(6, 6) || SBSLPM_UP(5, 5) ||
         SBSLPM_UP(5, 6) || SBSLPM_UP(6, 7) || SBSLPM_UP(7, 7) || SBSLPM_UP(11, 11) || SBSLPM_UP(11, 8);
#undef SBSLPM_UP
}

// The L2L sweep of a tree whose leaves sit at several levels (the irregular
// lat-lon tree) in one launch: a leaf's row becomes its own row plus each
// ancestor's own row evaluated at the leaf's nodes through a composite matrix
// (tree.cpp composite_l2l), one workgroup per leaf.  The ancestors of leaf i
// are anc[3e .. 3e+2] = (panel, xi matrix, eta matrix) for e in
// [anc_first[i], anc_first[i+1]).  Wave w takes entries w, w + 4, ... with the
// sweeps of k_down_fused's first phase; the leaf's row and the four waves'
// sums are added in a fixed order.  Only the leaf rows are written.
template <int D>
__global__ __launch_bounds__(256) void k_leaf_rows(const int* __restrict__ leaves, const int* __restrict__ anc_first,
                                                   const int* __restrict__ anc, const double* __restrict__ anc_tab,
                                                   int ncomp, Ptr3 phi) {
  constexpr int N = D * D;
  static_assert(3 * N <= 512 && 3 * D <= 64, "k_leaf_rows: grid too large for one workgroup");
  __shared__ double sV[4][3 * N];
  __shared__ double sPart[4][3 * N];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int leaf = leaves[blockIdx.x];
  const int e0 = anc_first[blockIdx.x], e1 = anc_first[blockIdx.x + 1];
  const bool row_lane = lane < ncomp * D;
  const int f = row_lane ? lane / D : 0, a = lane - f * D;
  const double* __restrict__ phif = comp(phi, f);
  double* part_w = &sPart[w][f * N + a];
  if (row_lane)
    for (int k = 0; k < D; ++k) part_w[k * D] = 0.0;
  for (int e = e0 + w; e < e1; e += 4) {
    const int64_t ap = anc[3 * e];
    const ConstTab Cx = const_tab(anc_tab) + (int64_t)anc[3 * e + 1] * N;
    const ConstTab Cy = const_tab(anc_tab) + (int64_t)anc[3 * e + 2] * N;
    if (row_lane) {  // V[a][l] = sum_b Cy[l][b] P[a][b]
      double p[D];
#pragma unroll
      for (int b = 0; b < D; ++b) p[b] = phif[ap * N + a * D + b];
#pragma unroll 3
      for (int l = 0; l < D; ++l) {
        double v = 0.0;
#pragma unroll
        for (int b = 0; b < D; ++b) v += Cy[l * D + b] * p[b];
        sV[w][f * N + a * D + l] = v;
      }
    }
    wave_lds_order();
    if (row_lane) {  // column l = a: sum[k][l] += sum_r Cx[k][r] V[r][l]
      double q[D];
#pragma unroll
      for (int r = 0; r < D; ++r) q[r] = sV[w][f * N + r * D + a];
#pragma unroll 3
      for (int k = 0; k < D; ++k) {
        double v = part_w[k * D];
#pragma unroll
        for (int r = 0; r < D; ++r) v += Cx[k * D + r] * q[r];
        part_w[k * D] = v;
      }
    }
    wave_lds_order();
  }
  __syncthreads();
  for (int o = threadIdx.x; o < ncomp * N; o += blockDim.x) {
    const int ff = o / N;
    double* row = comp(phi, ff) + (int64_t)leaf * N;
    const int j = o - ff * N;
    row[j] = row[j] + sPart[0][o] + sPart[1][o] + sPart[2][o] + sPart[3][o];
  }
}

template <int D>
bool launch_leaf_rows_d(const DevTree& t, const int* leaves, int nleaves, const int* anc_first, const int* anc,
                        const double* anc_tab, int ncomp, Ptr3 phi, hipStream_t s) {
  if (t.dp1 != D) return false;
  if (nleaves > 0)
    hipLaunchKernelGGL((k_leaf_rows<D>), dim3(nleaves), dim3(256), 0, s, leaves, anc_first, anc, anc_tab, ncomp, phi);
  return true;
}

// Dynamic LDS of k_l2l_l2p (bytes).
size_t l2l_l2p_lds(const DevTree& t, int ncomp) {
  const size_t n = t.n, D = t.dp1, m = t.m, nc = ncomp * n;
  const size_t na = std::min(kAncChunk, std::max(t.levels - 2, 0));
  const size_t ph1 = na * (2 * nc + 2 * D * D);
  const size_t ph2 = 2 * D * D + D * m + 2 * nc + 4 * nc + 4 * ncomp * D * m;
  return (3 * n + std::max(ph1, ph2)) * sizeof(double);
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
static void launch_m2m_levels(const DevTree& t, int lfirst, double* q_proxy, hipStream_t s, int nmem,
                              const UpSets* sets = nullptr) {
  UpSets one{};
  one.qx[0] = q_proxy;
  const UpSets& us = sets ? *sets : one;
  for (int l = lfirst; l >= 0; --l) {
    if (l >= 1) {  // levels l and l - 1 in one launch
      const int gbegin = 2 * ((1 << (2 * (l - 1))) - 1), gcount = 6 << (2 * (l - 1));
      if (launch_m2m_two(t, gbegin, gcount, us, s, nmem)) {
        --l;
        continue;
      }
    }
    const int pbegin = 2 * ((1 << (2 * l)) - 1), count = 6 << (2 * l);
    if (launch_m2m_tiled(t, pbegin, count, us, s, nmem)) continue;
    if (us.n == 0)
      hipLaunchKernelGGL(k_m2m, dim3(count, nmem), dim3(128), 0, s, t, pbegin, (const int*)nullptr, us.qx[0]);
    for (int i = 0; i < us.n; ++i)
      hipLaunchKernelGGL(k_m2m, dim3(count, 1), dim3(128), 0, s, t, pbegin, (const int*)nullptr, us.qx[i]);
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
// The fused upward pass for the charge sets `us` (us.n > 0) or for nmem
// member slices of us entry 0.
static void upward_fused(const DevTree& t, const UpSets& us, hipStream_t s, int nmem) {
  const int ns = us.n > 0 ? us.n : 1;
  if (t.levels < 2) {
    for (int i = 0; i < ns; ++i) launch_upward_pass(t, us.area[i], us.pots[i], us.qp[i], us.qx[i], s);
    return;
  }
  const int l = t.levels - 2;
  const int pbegin = 2 * ((1 << (2 * l)) - 1), count = 6 << (2 * l);
  if (!launch_up_fused(t, pbegin, count, us, s, nmem)) {
    for (int i = 0; i < ns; ++i)
      hipLaunchKernelGGL(k_p2m_m2m, dim3(count, us.n > 0 ? 1 : nmem), dim3(256), 0, s, t, pbegin, us.area[i],
                         us.pots[i], us.qp[i], us.qx[i]);
  }
  launch_m2m_levels(t, t.levels - 3, us.qx[0], s, nmem, &us);
}

void launch_upward_fused(const DevTree& t, const double* area, const double* pots, double* q_part, double* q_proxy,
                         hipStream_t s, int nmem) {
  UpSets us{};
  us.area[0] = area;
  us.pots[0] = pots;
  us.qp[0] = q_part;
  us.qx[0] = q_proxy;
  upward_fused(t, us, s, nmem);
}

void launch_upward_sets(const DevTree& t, const UpCharge* sets, int nsets, hipStream_t s) {
  for (int b = 0; b < nsets; b += kMaxUpSets) {
    UpSets us{};
    us.n = std::min(kMaxUpSets, nsets - b);
    for (int i = 0; i < us.n; ++i) {
      us.area[i] = sets[b + i].area;
      us.pots[i] = sets[b + i].pots;
      us.qp[i] = sets[b + i].q_part;
      us.qx[i] = sets[b + i].q_proxy;
    }
    upward_fused(t, us, s, 1);
  }
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
  CPtr3 cphi{};
  for (int f = 0; f < ncomp; ++f) cphi.p[f] = phi_proxy[f];
  if (assign_leaf_rows && t.levels >= 2 && t.anc_l2l) {
    if (launch_down_fused(t, ncomp, cphi, part, out, s, nmem)) return;
    const size_t lds = l2l_l2p_lds(t, ncomp);
    if (lds <= 65536) {
      hipLaunchKernelGGL(k_l2l_l2p, dim3(6 << (2 * (t.levels - 2)), nmem), dim3(256), lds, s, t, ncomp, cphi, part,
                         out);
      return;
    }
  }
  launch_l2l(t, ncomp, phi_proxy, s, assign_leaf_rows, nmem);
  if (is_identity_p2m(t))
    hipLaunchKernelGGL(k_l2p_copy, dim3(blocks_for(t.npart, 256), nmem), dim3(256), 0, s, t.npart, ncomp, leafrows,
                       part, out, (int64_t)t.P * t.n);
  else
    hipLaunchKernelGGL(k_l2p, dim3(t.Pl, nmem), dim3(128), 0, s, t, ncomp, cphi, part, out);
}

void launch_m2m_list(const DevTree& t, const int* parents, int count, double* q_proxy, hipStream_t s, int nmem) {
  if (count <= 0) return;
  UpSets one{};
  one.qx[0] = q_proxy;
  if (!launch_m2m_tiled(t, 0, count, one, s, nmem, parents))
    hipLaunchKernelGGL(k_m2m, dim3(count, nmem), dim3(128), 0, s, t, 0, parents, q_proxy);
}

void launch_l2l_list(const DevTree& t, const int* parents, int count, int ncomp, double* const* phi_proxy,
                     hipStream_t s) {
  Ptr3 phi{};
  for (int f = 0; f < ncomp; ++f) phi.p[f] = phi_proxy[f];
  if (count > 0) hipLaunchKernelGGL(k_l2l, dim3(count), dim3(256), 0, s, t, 0, parents, ncomp, phi, 0);
}

bool launch_leaf_rows(const DevTree& t, const int* leaves, int nleaves, const int* anc_first, const int* anc,
                      const double* anc_tab, int ncomp, double* const* phi_proxy, hipStream_t s) {
  if (!anc_tab) return false;
  Ptr3 phi{};
  for (int f = 0; f < ncomp; ++f) phi.p[f] = phi_proxy[f];
#define SBSLPM_LR(D) launch_leaf_rows_d<D>(t, leaves, nleaves, anc_first, anc, anc_tab, ncomp, phi, s)
  return SBSLPM_LR(11) || SBSLPM_LR(9) || SBSLPM_LR(7) || SBSLPM_LR(6) || SBSLPM_LR(5);
#undef SBSLPM_LR
}

}  // namespace sbslpm
