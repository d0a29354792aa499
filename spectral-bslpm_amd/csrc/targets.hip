// targets.hip -- FMM evaluation at arbitrary points of the sphere, FP64.
//
// A target x inside leaf L receives
//   * the far field of L: the L2L-swept proxy potentials of L interpolated at
//     x with the degree-d tensor BLI basis (child_panel_interp_3_ll,
//     fmm_funcs_impl.hpp:318-364), and
//   * the near field: the kernel summed over L's own source sets, i.e. the
//     particle-target interaction types 0 and 1 (swe_vel_impl.hpp:188-255),
//     with the same source records and kernels as the pair kernels.
// One wave per target: lanes split the basis elements and the sources, and a
// butterfly reduction gives the value.  Targets are independent; a target
// set of any size and order is handled in one launch.
#include "device.hpp"

namespace sbslpm {
namespace {

struct CPtr3 {
  const double* p[3];
};
struct Ptr3 {
  double* p[3];
};

// Weight i of bli_weights_1d(t, lo, hi, deg, cheb, .), computed alone so that
// each lane forms only the weights of its own basis elements.
__device__ inline double bli_weight(double t, double lo, double hi, int deg, const double* __restrict__ cheb, int i) {
  if (deg == 0) return 1.0;
  const double range = 0.5 * (hi - lo), off = 0.5 * (hi + lo);
  int hit = -1;
  double den = 0, wi = 0;
  for (int j = 0; j <= deg; ++j) {
    const double node = cheb[j] * range + off;
    if (hit < 0 && fabs(t - node) < kSnapTol) hit = j;
    double bw = (j == 0 || j == deg) ? 0.5 : 1.0;
    if (j & 1) bw = -bw;
    const double v = bw / (t - node);
    den += v;
    if (j == i) wi = v;
  }
  if (hit >= 0) return hit == i ? 1.0 : 0.0;
  return wi * (1.0 / den);
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

template <int KIND>
__global__ __launch_bounds__(256) void k_eval_targets(DevTree t, const int* __restrict__ set_start,
                                                      const SourceSet* __restrict__ sets,
                                                      const SrcRec64* __restrict__ rec, CPtr3 far, double eps,
                                                      int64_t nt, const double* __restrict__ tx,
                                                      const double* __restrict__ ty,
                                                      const double* __restrict__ tz, Ptr3 out) {
  constexpr int NC = KIND == 0 ? 1 : 3;
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= nt) return;
  const double x = tx[i], y = ty[i], z = tz[i];
  double xe[2];
  const int leaf = locate_leaf(t, x, y, z, xe);
  if (leaf < 0) {
    if (lane == 0) {
      atomicOr(t.err, kErrLocate);
      for (int f = 0; f < NC; ++f) out.p[f][i] = 0.0;
    }
    return;
  }
  const DevPanel& P = t.panels[leaf];
  const bool inside = xe[0] >= P.min_xi - kSnapTol && xe[0] <= P.max_xi + kSnapTol &&
                      xe[1] >= P.min_eta - kSnapTol && xe[1] <= P.max_eta + kSnapTol;
  if (!inside && lane == 0) atomicOr(t.err, kErrBliBounds);

  // far field: sum_k wx[k / (d+1)] we[k % (d+1)] phi[L][k]
  double acc[NC];
#pragma unroll
  for (int f = 0; f < NC; ++f) acc[f] = 0.0;
  const int deg = t.dp1 - 1;
  const int64_t row = (int64_t)leaf * t.n;
  for (int k = lane; k < t.n; k += 64) {
    const int a = k / t.dp1, b = k - a * t.dp1;
    const double w = bli_weight(xe[0], P.min_xi, P.max_xi, deg, t.cheb_d, a) *
                     bli_weight(xe[1], P.min_eta, P.max_eta, deg, t.cheb_d, b);
#pragma unroll
    for (int f = 0; f < NC; ++f) acc[f] = fma(w, far.p[f][row + k], acc[f]);
  }

  // near field over the leaf's source sets
  double g0 = 0, g1 = 0, g2 = 0;
  for (int e = set_start[leaf]; e < set_start[leaf + 1]; ++e) {
    const SourceSet S = sets[e];
    const SrcRec64* __restrict__ r = rec + S.first;
    for (int k = lane; k < S.count; k += 64) {
      const SrcRec64 s = r[k];
      const double dp = x * s.x + y * s.y + z * s.z;
      if (KIND == 0) {
        if (dp < 1.0 - 1e-15) g0 += log(1.0 - dp) * s.q;
      } else {
        const double g = s.q / (1.0 - dp + eps);
        g0 = fma(g, s.x, g0);
        g1 = fma(g, s.y, g1);
        g2 = fma(g, s.z, g2);
      }
    }
  }
  if (KIND == 0) {
    const double v = wave_sum(acc[0] + g0);
    if (lane == 0) out.p[0][i] = v;
  } else {
    g0 = wave_sum(g0);
    g1 = wave_sum(g1);
    g2 = wave_sum(g2);
    const double f0 = wave_sum(acc[0]), f1 = wave_sum(acc[NC > 1 ? 1 : 0]), f2 = wave_sum(acc[NC > 2 ? 2 : 0]);
    if (lane == 0) {
      out.p[0][i] = f0 + (y * g2 - z * g1);
      out.p[1][i] = f1 + (z * g0 - x * g2);
      out.p[2][i] = f2 + (x * g1 - y * g0);
    }
  }
}

}  // namespace

void launch_eval_targets(const DevTree& t, int kind, const int* set_start, const SourceSet* sets,
                         const SrcRec64* rec, const double* const* far, double eps, int64_t nt, const double* tx,
                         const double* ty, const double* tz, double* const* out, hipStream_t s) {
  if (nt <= 0) return;
  CPtr3 f{};
  Ptr3 o{};
  const int nc = kind == 0 ? 1 : 3;
  for (int c = 0; c < nc; ++c) {
    f.p[c] = far[c];
    o.p[c] = out[c];
  }
  const unsigned blocks = (unsigned)((nt + 3) / 4);
  if (kind == 0)
    hipLaunchKernelGGL(k_eval_targets<0>, dim3(blocks), dim3(256), 0, s, t, set_start, sets, rec, f, eps, nt, tx,
                       ty, tz, o);
  else
    hipLaunchKernelGGL(k_eval_targets<1>, dim3(blocks), dim3(256), 0, s, t, set_start, sets, rec, f, eps, nt, tx,
                       ty, tz, o);
}

}  // namespace sbslpm
