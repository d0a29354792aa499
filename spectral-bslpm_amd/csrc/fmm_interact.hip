// fmm_interact.hip -- pair kernels over the target-major interaction lists.
//
// Reference: bve_vel_panel_interaction (fmm_interactions/bve_vel_impl.hpp:34-70)
// and poisson_panel_interaction (inverse_poisson_impl.hpp:23-52) run one
// thread per interaction, recompute every source node's map per pair and
// scatter three atomics per pair.  Here:
//   * one wave64 owns up to 64 target points of one target panel and walks
//     that panel's source sets, so each target is accumulated in registers by
//     one lane (no atomics, fixed summation order);
//   * displaced, re-projected source positions and scaled charges are built
//     once per evaluation into a record array (SrcRec64 / SrcRec32);
//   * every lane of a wave reads the same source record, so the records are
//     wave-uniform loads (scalar cache) with no LDS and no barriers;
//   * u = sum_j g_j q_j (x cross y_j) is accumulated as x cross (sum_j g_j q_j y_j),
//     saving the per-pair cross product (six multiplies) of the reference form.
// Workgroups of four waves; the workgroup id is remapped so that consecutive
// tiles (spatially adjacent target panels sharing sources) land on the same
// XCD and its L2.
#include <cmath>

#include "device.hpp"
#include "lapkernel.hpp"

namespace sbslpm {
namespace {

struct Ptr3 {
  double* p[3];
};
struct CPtr3 {
  const double* p[3];
};

// A uniform constant moved into a vector register: the first step of a Horner
// scheme whose multiplier is such a register is one three-operand fma (with
// two constants in scalar registers the instruction set needs a copy).
__device__ inline double in_vgpr(double v) {
  __asm__ volatile("" : "+v"(v));
  return v;
}
__device__ inline float in_vgpr_f(float v) {
  __asm__ volatile("" : "+v"(v));
  return v;
}


// Member slices of a batched launch (gridDim.y members): charges and
// displacements of the sources, outputs of the targets.
__device__ inline void member_sources(const DevTree& t, const double* __restrict__& q_part,
                                      const double* __restrict__& q_proxy, CPtr3& dpart,
                                      CPtr3& dprox) {
  const int64_t mp = member_part(t), mx = member_proxy(t);
  q_part += mp;
  q_proxy += mx;
  for (int f = 0; f < 3; ++f) {
    if (dpart.p[f]) dpart.p[f] += mp;
    if (dprox.p[f]) dprox.p[f] += mx;
  }
}
__device__ inline void member_targets(const DevTree& t, Ptr3& part, Ptr3& proxy) {
  const int64_t mp = member_part(t), mx = member_proxy(t);
  for (int f = 0; f < 3; ++f) {
    if (part.p[f]) part.p[f] += mp;
    if (proxy.p[f]) proxy.p[f] += mx;
  }
}

// Target-major ownership: every target point is written by exactly one lane,
// so `assign` may replace the accumulate of the reference functors.
__device__ inline void store_u(const DevTree& t, int64_t u, Ptr3 part, Ptr3 proxy, double a, double b, double c,
                               int assign) {
  double* o[3];
  int64_t v = u;
  if (u < t.npart) {
    o[0] = part.p[0]; o[1] = part.p[1]; o[2] = part.p[2];
  } else {
    v = u - t.npart;
    o[0] = proxy.p[0]; o[1] = proxy.p[1]; o[2] = proxy.p[2];
  }
  if (assign) {
    o[0][v] = a; o[1][v] = b; o[2][v] = c;
  } else {
    o[0][v] += a; o[1][v] += b; o[2][v] += c;
  }
}

__device__ inline void store_s(const DevTree& t, int64_t u, double* part, double* proxy, double a, int assign) {
  double* o = u < t.npart ? part + u : proxy + (u - t.npart);
  *o = assign ? a : *o + a;
}

// Source records, FP64: y = P(node + disp) (bve_vel_impl.hpp:54-60), q' = -q/(4 pi).
// Leaf proxy rows are never sources and are left untouched.
__global__ void k_src64(DevTree t, const double* __restrict__ q_part, const double* __restrict__ q_proxy,
                        CPtr3 dpart, CPtr3 dprox, int project, SrcRec64* __restrict__ rec) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= t.U) return;
  member_sources(t, q_part, q_proxy, dpart, dprox);
  rec += (int64_t)blockIdx.y * t.U;
  double x = t.px[u], y = t.py[u], z = t.pz[u], q;
  if (u < t.npart) {
    if (dpart.p[0]) { x += dpart.p[0][u]; y += dpart.p[1][u]; z += dpart.p[2][u]; }
    q = q_part[u];
  } else {
    int64_t v = u - t.npart;
    if (v / t.n >= t.offset) return;
    if (dprox.p[0]) { x += dprox.p[0][v]; y += dprox.p[1][v]; z += dprox.p[2][v]; }
    q = q_proxy[v];
  }
  if (project) project_to_sphere(x, y, z);
  rec[u] = SrcRec64{x, y, z, kNegInv4Pi * q};
}

// Source records, FP32: position relative to the owning panel's centre.
__global__ void k_src32(DevTree t, const double* __restrict__ q_part, const double* __restrict__ q_proxy,
                        CPtr3 dpart, CPtr3 dprox, int project, SrcRec32* __restrict__ rec) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= t.U) return;
  member_sources(t, q_part, q_proxy, dpart, dprox);
  rec += (int64_t)blockIdx.y * t.U;
  double x = t.px[u], y = t.py[u], z = t.pz[u], q;
  int panel;
  if (u < t.npart) {
    if (dpart.p[0]) { x += dpart.p[0][u]; y += dpart.p[1][u]; z += dpart.p[2][u]; }
    q = q_part[u];
    panel = t.offset + (int)(u / t.m2);
  } else {
    int64_t v = u - t.npart;
    panel = (int)(v / t.n);
    if (panel >= t.offset) return;
    if (dprox.p[0]) { x += dprox.p[0][v]; y += dprox.p[1][v]; z += dprox.p[2][v]; }
    q = q_proxy[v];
  }
  if (project) project_to_sphere(x, y, z);
  const DevPanel& P = t.panels[panel];
  rec[u] = SrcRec32{(float)(x - P.cx), (float)(y - P.cy), (float)(z - P.cz), (float)(kNegInv4Pi * q)};
}

// FP64 Biot-Savart pair kernel, the reference's arithmetic:
// g = -q/(4 pi (1 - x.y + eps)).
__global__ __launch_bounds__(256) void k_bve64(DevTree t, const TargetTile* __restrict__ tiles, int ntiles,
                                               const SourceSet* __restrict__ sets,
                                               const SrcRec64* __restrict__ rec, double eps, Ptr3 part,
                                               Ptr3 proxy, int assign, int xcd) {
  member_targets(t, part, proxy);
  rec += (int64_t)blockIdx.y * t.U;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int tile = shaped_tile(t.xcd_start ? t.xcd_start + 9 : nullptr, wave, xcd);
  if (tile >= ntiles) return;
  const TargetTile T = tiles[tile];
  const int64_t u = (int64_t)T.first + (lane < T.count ? lane : T.count - 1);
  const double x = t.px[u], y = t.py[u], z = t.pz[u];
  double w0 = 0, w1 = 0, w2 = 0;
  for (int e = T.src_begin; e < T.src_end; ++e) {
    const SourceSet S = sets[e];
    const SrcRec64* __restrict__ r = rec + S.first;
#pragma unroll 3
    for (int k = 0; k < S.count; ++k) {
      const SrcRec64 s = r[k];
      const double dp = x * s.x + y * s.y + z * s.z;
      const double g = rcp64_pair((1.0 - dp) + eps) * s.q;
      w0 = fma(g, s.x, w0);
      w1 = fma(g, s.y, w1);
      w2 = fma(g, s.z, w2);
    }
  }
  if (lane < T.count) store_u(t, u, part, proxy, y * w2 - z * w1, z * w0 - x * w2, x * w1 - y * w0, assign);
}

// FP64 log pair kernel: psi += -1/(4 pi) log(1 - x.y) q, pairs with
// x.y >= 1 - 1e-15 skipped (inverse_poisson_impl.hpp:46-49).  The logarithm
// is log_tab (fasttab.hpp, within 1.1 ulp of the correctly rounded one) on
// the context's table, staged in LDS.  A skipped pair (a source that coincides
// with a target: the target's own node, or a node on a shared leaf edge) is
// rare, so the wave tests the whole source at once (a ballot, one compare per
// pair) and only a wave holding such a pair takes the guarded form.
__global__ __launch_bounds__(256) void k_log64(DevTree t, const TargetTile* __restrict__ tiles, int ntiles,
                                               const SourceSet* __restrict__ sets,
                                               const SrcRec64* __restrict__ rec, double* __restrict__ part,
                                               double* __restrict__ proxy, int assign) {
  __shared__ LogTabEntry ltab[kLogTabSize];
  for (int i = threadIdx.x; i < kLogTabSize; i += blockDim.x) ltab[i] = t.logtab[i];
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int tile = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + wave;
  if (tile >= ntiles) return;
  const TargetTile T = tiles[tile];
  const int64_t u = (int64_t)T.first + (lane < T.count ? lane : T.count - 1);
  const double x = t.px[u], y = t.py[u], z = t.pz[u];
  constexpr double kSkip = 1.0 - 1e-15;
  const double c6 = in_vgpr(-1.0 / 6.0);
  double acc = 0;
  for (int e = T.src_begin; e < T.src_end; ++e) {
    const SourceSet S = sets[e];
    const SrcRec64* __restrict__ r = rec + S.first;
    int k = 0;
    for (; k + 1 < S.count; k += 2) {
      const SrcRec64 s0 = r[k], s1 = r[k + 1];
      const double dp0 = x * s0.x + y * s0.y + z * s0.z;
      const double dp1 = x * s1.x + y * s1.y + z * s1.z;
      double l0 = log_tab(1.0 - dp0, ltab, c6), l1 = log_tab(1.0 - dp1, ltab, c6);
      const bool sk0 = dp0 >= kSkip, sk1 = dp1 >= kSkip;
      if ((__builtin_amdgcn_ballot_w64(sk0) | __builtin_amdgcn_ballot_w64(sk1)) != 0) {
        // a real branch (the empty asm keeps the compiler from turning it into
        // selects on every pair); log_tab's index is masked, so the discarded
        // values were read in bounds
        __asm__ volatile("");
        l0 = sk0 ? 0.0 : l0;
        l1 = sk1 ? 0.0 : l1;
      }
      acc = fma(l0, s0.q, acc);
      acc = fma(l1, s1.q, acc);
    }
    if (k < S.count) {
      const SrcRec64 s0 = r[k];
      const double dp0 = x * s0.x + y * s0.y + z * s0.z;
      const double l0 = log_tab(1.0 - dp0, ltab, c6);
      acc = fma(dp0 >= kSkip ? 0.0 : l0, s0.q, acc);
    }
  }
  if (lane < T.count) store_s(t, u, part, proxy, acc, assign);
}

// Horner multipliers of the FP64 regularised-Laplacian kernel in vector registers.
struct LapV64 {
  double a1, b2, c24;
};

// K(t) of k_lap64 where t >= t_sat (erf factor exactly 1): A + E B.
__device__ inline double lap_pair(double dp, const LapCoef64& kp, const LapV64& kv, const double* etab) {
  const double t2 = dp * dp;
  const double A = dp * fma(kv.a1, t2, kp.a0);
  const double B = fma(fma(kv.b2, t2, kp.b1), t2, kp.b0);
  return fma(exp_neg(fma(t2, kp.inv_eps, -kp.inv_eps), etab, kv.c24), B, A);
}
// K(t) for any t: A + f E B, zero at or below the cut.
__device__ inline double lap_pair_slow(double dp, const LapCoef64& kp, const LapV64& kv, const double* etab) {
  const double t2 = dp * dp;
  const double A = dp * fma(kv.a1, t2, kp.a0);
  const double B = fma(fma(kv.b2, t2, kp.b1), t2, kp.b0);
  const double sv = dp * kp.inv_eps;
  const double f = sv >= 6.0 ? 1.0 : (sv <= -6.0 ? 0.0 : 0.5 * (1.0 + erf(sv)));
  const double K = fma(exp_neg(fma(t2, kp.inv_eps, -kp.inv_eps), etab, kv.c24) * f, B, A);
  return dp > kp.cut ? K : 0.0;
}

// Regularised-Laplacian pair kernel (lap_panel_interaction,
// laplacian_impl.hpp:46-66, eps = 0.1): phi += K(x.y) q with K in the folded
// form of lapkernel.hpp (A + E B where t >= 6 eps, erf only below).  Sources
// are displaced and projected; the records hold q' = -q/(4 pi), so the
// coefficients arrive multiplied by -4 pi / eps^{5/2}.  Over 99% of the pair
// work has t > 0.6, so a wave tests the whole source at once (a ballot) and
// evaluates erf only when one of its pairs needs it.  The exponential is
// exp_neg (fasttab.hpp: its argument lies in [-1/eps, 0]) on the context's
// table, staged in LDS.
__global__ __launch_bounds__(256) void k_lap64(DevTree t, const TargetTile* __restrict__ tiles, int ntiles,
                                               const SourceSet* __restrict__ sets,
                                               const SrcRec64* __restrict__ rec, LapCoef64 kp,
                                               double* __restrict__ part, double* __restrict__ proxy, int assign) {
  __shared__ double etab[kExpNegTabSize];
  for (int i = threadIdx.x; i < kExpNegTabSize; i += blockDim.x) etab[i] = t.expnegtab[i];
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int tile = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + wave;
  if (tile >= ntiles) return;
  const TargetTile T = tiles[tile];
  const int64_t u = (int64_t)T.first + (lane < T.count ? lane : T.count - 1);
  const double x = t.px[u], y = t.py[u], z = t.pz[u];
  // every pair passes the cut when 1 - 30 eps < -1 (eps = 0.1)
  const bool no_cut = kp.cut < -1.0;
  const LapV64 kv{in_vgpr(kp.a1), in_vgpr(kp.b2), in_vgpr(1.0 / 24.0)};
  double acc = 0;
  for (int e = T.src_begin; e < T.src_end; ++e) {
    const SourceSet S = sets[e];
    const SrcRec64* __restrict__ r = rec + S.first;
    int k = 0;
    for (; k + 1 < S.count; k += 2) {
      const SrcRec64 s0 = r[k], s1 = r[k + 1];
      const double dp0 = x * s0.x + y * s0.y + z * s0.z;
      const double dp1 = x * s1.x + y * s1.y + z * s1.z;
      double K0 = lap_pair(dp0, kp, kv, etab), K1 = lap_pair(dp1, kp, kv, etab);
      if (!no_cut || (__builtin_amdgcn_ballot_w64(dp0 < kp.t_sat) | __builtin_amdgcn_ballot_w64(dp1 < kp.t_sat)) != 0) {
        // a real branch (the empty asm keeps the compiler from evaluating erf
        // on every pair)
        __asm__ volatile("");
        K0 = lap_pair_slow(dp0, kp, kv, etab);
        K1 = lap_pair_slow(dp1, kp, kv, etab);
      }
      acc = fma(K0, s0.q, acc);
      acc = fma(K1, s1.q, acc);
    }
    if (k < S.count) {
      const SrcRec64 s0 = r[k];
      const double dp0 = x * s0.x + y * s0.y + z * s0.z;
      acc = fma(lap_pair_slow(dp0, kp, kv, etab), s0.q, acc);
    }
  }
  if (lane < T.count) store_s(t, u, part, proxy, acc, assign);
}

// Mixed-precision regularised-Laplacian pairs in the folded form of
// lapkernel.hpp: the squared chord h2 = |x - y|^2 = 2(1 - t) in FP32 from
// panel-centred coordinates, t^2 - 1 = h2 (h2/4 - 1) (no cancellation), the
// exponential on the hardware base-2 exponential of u = (t^2 - 1) log2(e)/eps
// (u in [-2/(eps ln 2), 0]), the polynomials in u; erf only when a wave-wide
// ballot finds a pair with t < 0.6 (elsewhere it is +-1 in FP32 too).  FP32
// sums per source set; c3 (times -4 pi, undoing the record's -1/(4 pi))
// applied per set in FP64.
__device__ inline float lap32_K(float h2, const LapCoef32& kp, float a1v, float b2v) {
  const float tt = fmaf(h2, -0.5f, 1.0f);
  const float u = h2 * fmaf(h2, 0.25f * kp.k, -kp.k);
  const float A = tt * fmaf(a1v, u, kp.a0);
  const float B = fmaf(fmaf(b2v, u, kp.b1), u, kp.b0);
  return fmaf(__builtin_amdgcn_exp2f(u), B, A);
}
__device__ inline float lap32_K_slow(float h2, const LapCoef32& kp, float a1v, float b2v) {
  const float tt = fmaf(h2, -0.5f, 1.0f);
  const float u = h2 * fmaf(h2, 0.25f * kp.k, -kp.k);
  const float A = tt * fmaf(a1v, u, kp.a0);
  const float B = fmaf(fmaf(b2v, u, kp.b1), u, kp.b0);
  const float sv = tt * kp.inv_eps;
  const float f = sv >= 6.0f ? 1.0f : (sv <= -6.0f ? 0.0f : 0.5f * (1.0f + erff(sv)));
  const float K = fmaf(__builtin_amdgcn_exp2f(u) * f, B, A);
  return tt > kp.cut ? K : 0.0f;
}

__global__ __launch_bounds__(256) void k_lap_mixed(DevTree t, const TargetTile* __restrict__ tiles, int ntiles,
                                                   const SourceSet* __restrict__ sets,
                                                   const SrcRec32* __restrict__ rec, LapCoef32 kp, double c3,
                                                   double* __restrict__ part, double* __restrict__ proxy,
                                                   int assign) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int tile = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + wave;
  if (tile >= ntiles) return;
  const TargetTile T = tiles[tile];
  const DevPanel& PT = t.panels[T.panel];
  const int64_t u = (int64_t)T.first + (lane < T.count ? lane : T.count - 1);
  const float ax = (float)(t.px[u] - PT.cx), ay = (float)(t.py[u] - PT.cy), az = (float)(t.pz[u] - PT.cz);
  const bool no_cut = kp.cut < -1.0f;
  // t < t_sat  <=>  h2 > h_sat
  const float h_sat = 2.0f * (1.0f - kp.t_sat);
  const float a1v = in_vgpr_f(kp.a1), b2v = in_vgpr_f(kp.b2);
  double acc = 0;
  for (int e = T.src_begin; e < T.src_end; ++e) {
    const SourceSet S = sets[e];
    const float px = ax + S.ox, py = ay + S.oy, pz = az + S.oz;
    const SrcRec32* __restrict__ r = rec + S.first;
    float v = 0.f;
    int k = 0;
    for (; k + 1 < S.count; k += 2) {
      const SrcRec32 s0 = r[k], s1 = r[k + 1];
      const float d0 = s0.x - px, d1 = s0.y - py, d2 = s0.z - pz;
      const float g0 = s1.x - px, g1 = s1.y - py, g2 = s1.z - pz;
      const float h0 = fmaf(d0, d0, fmaf(d1, d1, d2 * d2));
      const float h1 = fmaf(g0, g0, fmaf(g1, g1, g2 * g2));
      float K0 = lap32_K(h0, kp, a1v, b2v), K1 = lap32_K(h1, kp, a1v, b2v);
      if (!no_cut || (__builtin_amdgcn_ballot_w64(h0 > h_sat) | __builtin_amdgcn_ballot_w64(h1 > h_sat)) != 0) {
        // a real branch (the empty asm keeps the compiler from evaluating erf on every pair)
        __asm__ volatile("");
        K0 = lap32_K_slow(h0, kp, a1v, b2v);
        K1 = lap32_K_slow(h1, kp, a1v, b2v);
      }
      v = fmaf(K0, s0.q, v);
      v = fmaf(K1, s1.q, v);
    }
    if (k < S.count) {
      const SrcRec32 s0 = r[k];
      const float d0 = s0.x - px, d1 = s0.y - py, d2 = s0.z - pz;
      v = fmaf(lap32_K_slow(fmaf(d0, d0, fmaf(d1, d1, d2 * d2)), kp, a1v, b2v), s0.q, v);
    }
    acc += c3 * (double)v;
  }
  if (lane < T.count) store_s(t, u, part, proxy, acc, assign);
}

// k_lap64 and k_lap_mixed in lane groups (RowGroups): each 36-row tile at
// m = 6 runs as two 18-row sub-tiles in three copies per wave, every copy
// taking every third source of a set from records staged in LDS; one pair per
// iteration behind the same wave-wide ballot for the erf form.
__global__ __launch_bounds__(256) void k_lap64_gr(DevTree t, const TargetTile* __restrict__ tiles, int ntiles,
                                                  RowGroups gs, const SourceSet* __restrict__ sets,
                                                  const SrcRec64* __restrict__ rec, LapCoef64 kp,
                                                  double* __restrict__ part, double* __restrict__ proxy, int assign) {
  __shared__ double etab[kExpNegTabSize];
  __shared__ SrcRec64 stage[kWavesPerBlock][64];
  for (int i = threadIdx.x; i < kExpNegTabSize; i += blockDim.x) etab[i] = t.expnegtab[i];
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int tile, row0, cg, ci;
  if (!group_tile<true>(gs, ntiles, wave, 1, tile, row0, cg, ci)) return;
  const TargetTile T = tiles[tile];
  const int row = row0 + ci;
  const int64_t u = (int64_t)T.first + (row < T.count ? row : T.count - 1);
  const bool copy_ok = cg < gs.G;
  const double x = t.px[u], y = t.py[u], z = t.pz[u];
  const bool no_cut = kp.cut < -1.0;
  const LapV64 kv{in_vgpr(kp.a1), in_vgpr(kp.b2), in_vgpr(1.0 / 24.0)};
  double acc = 0;
  for (int e = T.src_begin; e < T.src_end; ++e) {
    const SourceSet S = sets[e];
    const SrcRec64* __restrict__ r = rec + S.first;
    for (int c0 = 0; c0 < S.count; c0 += 64) {
      const int cn = min(64, S.count - c0);
      stage_records(stage[wave], r + c0, cn);
      const int trips = (cn + gs.G - 1) / gs.G;
      for (int kk = 0; kk < trips; ++kk) {
        const int k = kk * gs.G + cg;
        if (!(copy_ok && k < cn)) continue;
        const SrcRec64 s0 = stage[wave][k];
        const double dp0 = x * s0.x + y * s0.y + z * s0.z;
        double K0 = lap_pair(dp0, kp, kv, etab);
        if (!no_cut || __builtin_amdgcn_ballot_w64(dp0 < kp.t_sat) != 0) {
          // a real branch (the empty asm keeps the compiler from evaluating erf on every pair)
          __asm__ volatile("");
          K0 = lap_pair_slow(dp0, kp, kv, etab);
        }
        acc = fma(K0, s0.q, acc);
      }
    }
  }
  acc = group_sum(acc, gs, ci);
  if (cg == 0 && row < T.count) store_s(t, u, part, proxy, acc, assign);
}

__global__ __launch_bounds__(256) void k_lap_mixed_gr(DevTree t, const TargetTile* __restrict__ tiles, int ntiles,
                                                      RowGroups gs, const SourceSet* __restrict__ sets,
                                                      const SrcRec32* __restrict__ rec, LapCoef32 kp, double c3,
                                                      double* __restrict__ part, double* __restrict__ proxy,
                                                      int assign) {
  __shared__ SrcRec32 stage[kWavesPerBlock][64];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int tile, row0, cg, ci;
  if (!group_tile<true>(gs, ntiles, wave, 1, tile, row0, cg, ci)) return;
  const TargetTile T = tiles[tile];
  const DevPanel& PT = t.panels[T.panel];
  const int row = row0 + ci;
  const int64_t u = (int64_t)T.first + (row < T.count ? row : T.count - 1);
  const bool copy_ok = cg < gs.G;
  const float ax = (float)(t.px[u] - PT.cx), ay = (float)(t.py[u] - PT.cy), az = (float)(t.pz[u] - PT.cz);
  const bool no_cut = kp.cut < -1.0f;
  const float h_sat = 2.0f * (1.0f - kp.t_sat);
  const float a1v = in_vgpr_f(kp.a1), b2v = in_vgpr_f(kp.b2);
  double acc = 0;
  for (int e = T.src_begin; e < T.src_end; ++e) {
    const SourceSet S = sets[e];
    const float px = ax + S.ox, py = ay + S.oy, pz = az + S.oz;
    const SrcRec32* __restrict__ r = rec + S.first;
    float v = 0.f;
    for (int c0 = 0; c0 < S.count; c0 += 64) {
      const int cn = min(64, S.count - c0);
      stage_records(stage[wave], r + c0, cn);
      const int trips = (cn + gs.G - 1) / gs.G;
      for (int kk = 0; kk < trips; ++kk) {
        const int k = kk * gs.G + cg;
        if (!(copy_ok && k < cn)) continue;
        const SrcRec32 s0 = stage[wave][k];
        const float d0 = s0.x - px, d1 = s0.y - py, d2 = s0.z - pz;
        const float h0 = fmaf(d0, d0, fmaf(d1, d1, d2 * d2));
        float K0 = lap32_K(h0, kp, a1v, b2v);
        if (!no_cut || __builtin_amdgcn_ballot_w64(h0 > h_sat) != 0) {
          // a real branch (the empty asm keeps the compiler from evaluating erf on every pair)
          __asm__ volatile("");
          K0 = lap32_K_slow(h0, kp, a1v, b2v);
        }
        v = fmaf(K0, s0.q, v);
      }
    }
    acc += c3 * (double)v;
  }
  acc = group_sum(acc, gs, ci);
  if (cg == 0 && row < T.count) store_s(t, u, part, proxy, acc, assign);
}

// Mixed precision Biot-Savart: FP32 pair arithmetic in panel-centred
// coordinates with 1 - x.y computed as |x - y|^2 / 2, FP32 partial sums per
// source set flushed into FP64 accumulators, cross product in FP64.
__global__ __launch_bounds__(256) void k_bve_mixed(DevTree t, const TargetTile* __restrict__ tiles, int ntiles,
                                                   const SourceSet* __restrict__ sets,
                                                   const SrcRec32* __restrict__ rec, float eps, Ptr3 part,
                                                   Ptr3 proxy, int assign, int xcd) {
  member_targets(t, part, proxy);
  rec += (int64_t)blockIdx.y * t.U;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int tile = shaped_tile(t.xcd_start ? t.xcd_start + 9 : nullptr, wave, xcd);
  if (tile >= ntiles) return;
  const TargetTile T = tiles[tile];
  const DevPanel& PT = t.panels[T.panel];
  const int64_t u = (int64_t)T.first + (lane < T.count ? lane : T.count - 1);
  const double x = t.px[u], y = t.py[u], z = t.pz[u];
  const float ax = (float)(x - PT.cx), ay = (float)(y - PT.cy), az = (float)(z - PT.cz);
  double w0 = 0, w1 = 0, w2 = 0;
  for (int e = T.src_begin; e < T.src_end; ++e) {
    const SourceSet S = sets[e];
    // target relative to the source panel centre: a + (c_t - c_s)
    const float px = ax + S.ox, py = ay + S.oy, pz = az + S.oz;
    const SrcRec32* __restrict__ r = rec + S.first;
    float v0 = 0.f, v1 = 0.f, v2 = 0.f;
#pragma unroll 4
    for (int k = 0; k < S.count; ++k) {
      const SrcRec32 s = r[k];
      const float d0 = s.x - px, d1 = s.y - py, d2 = s.z - pz;
      const float h = fmaf(0.5f, fmaf(d0, d0, fmaf(d1, d1, d2 * d2)), eps);
      const float g = __builtin_amdgcn_rcpf(h) * s.q;
      v0 = fmaf(g, d0, v0);
      v1 = fmaf(g, d1, v1);
      v2 = fmaf(g, d2, v2);
    }
    w0 += v0; w1 += v1; w2 += v2;
  }
  if (lane < T.count) store_u(t, u, part, proxy, y * w2 - z * w1, z * w0 - x * w2, x * w1 - y * w0, assign);
}

// Mixed precision log kernel: log(|x - y|^2 / 2) in FP32, same guard.  The
// pair takes the hardware base-2 logarithm (v_log_f32; h >= 1e-15 is far from
// the denormal range, so the library's range scaling is not needed) and the
// factor ln 2 is applied once per source set in FP64.
__global__ __launch_bounds__(256) void k_log_mixed(DevTree t, const TargetTile* __restrict__ tiles, int ntiles,
                                                   const SourceSet* __restrict__ sets,
                                                   const SrcRec32* __restrict__ rec, double* __restrict__ part,
                                                   double* __restrict__ proxy, int assign) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int tile = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + wave;
  if (tile >= ntiles) return;
  const TargetTile T = tiles[tile];
  const DevPanel& PT = t.panels[T.panel];
  const int64_t u = (int64_t)T.first + (lane < T.count ? lane : T.count - 1);
  const float ax = (float)(t.px[u] - PT.cx), ay = (float)(t.py[u] - PT.cy), az = (float)(t.pz[u] - PT.cz);
  double acc = 0;
  for (int e = T.src_begin; e < T.src_end; ++e) {
    const SourceSet S = sets[e];
    const float px = ax + S.ox, py = ay + S.oy, pz = az + S.oz;
    const SrcRec32* __restrict__ r = rec + S.first;
    float v = 0.f;
#pragma unroll 4
    for (int k = 0; k < S.count; ++k) {
      const SrcRec32 s = r[k];
      const float d0 = s.x - px, d1 = s.y - py, d2 = s.z - pz;
      const float h = 0.5f * fmaf(d0, d0, fmaf(d1, d1, d2 * d2));
      const float l2 = __builtin_amdgcn_logf(h);
      v += (h > 1e-15f) ? l2 * s.q : 0.f;
    }
    acc += 0.69314718055994530942 * (double)v;
  }
  if (lane < T.count) store_s(t, u, part, proxy, acc, assign);
}

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v4bf __attribute__((ext_vector_type(4)));

// The kernel products of NTT 16 x 16 tiles, D_ij = beta_i - p_j.b_i: FP32
// (v_mfma_f32_16x16x4_f32, lane i + 16 k holding component k), or, for an M2L
// set on a context with bf16 M2L, bf16 operands with FP32 accumulation
// (v_mfma_f32_16x16x16_bf16: lane i + 16 k holds K slots 4k..4k+3, the
// component in the first and zeros in the others, the same pairing of A and B).
// One branch for all NTT products.
template <int NTT>
__device__ inline void kernel_products(float a1, const float (&b1)[NTT], bool lowp, v4f (&d)[NTT]) {
  if (lowp) {
    const v4bf a = v4bf{(__bf16)a1, (__bf16)0.f, (__bf16)0.f, (__bf16)0.f};
#pragma unroll
    for (int tt = 0; tt < NTT; ++tt)
      d[tt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(
          a, v4bf{(__bf16)b1[tt], (__bf16)0.f, (__bf16)0.f, (__bf16)0.f}, v4f{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
  } else {
#pragma unroll
    for (int tt = 0; tt < NTT; ++tt)
      d[tt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1[tt], v4f{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
  }
}

// Records of the MFMA kernel (16-padded per source set; padding slots zero).
__global__ void k_src_mfma(DevTree t, const double* __restrict__ q_part, const double* __restrict__ q_proxy,
                           CPtr3 dpart, CPtr3 dprox, SrcRecMfma* __restrict__ rec) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= t.U16) return;
  member_sources(t, q_part, q_proxy, dpart, dprox);
  rec += (int64_t)blockIdx.y * t.U16;
  const int64_t lp = (int64_t)t.Pl * t.m2p;
  int panel, k;
  int64_t u;
  bool valid;
  if (s < lp) {
    const int leaf = (int)(s / t.m2p);
    k = (int)(s % t.m2p);
    panel = t.offset + leaf;
    valid = k < t.m2;
    u = (int64_t)leaf * t.m2 + k;
  } else {
    const int64_t v = s - lp;
    panel = (int)(v / t.np16);
    k = (int)(v % t.np16);
    valid = k < t.n;
    u = t.npart + (int64_t)panel * t.n + k;
  }
  float v[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (valid) {
    double x = t.px[u], y = t.py[u], z = t.pz[u], q;
    if (u < t.npart) {
      if (dpart.p[0]) { x += dpart.p[0][u]; y += dpart.p[1][u]; z += dpart.p[2][u]; }
      q = q_part[u];
    } else {
      const int64_t w = u - t.npart;
      if (dprox.p[0]) { x += dprox.p[0][w]; y += dprox.p[1][w]; z += dprox.p[2][w]; }
      q = q_proxy[w];
    }
    project_to_sphere(x, y, z);
    const DevPanel& P = t.panels[panel];
    const float bx = (float)(x - P.cx), by = (float)(y - P.cy), bz = (float)(z - P.cz);
    const double qq = kNegInv4Pi * q;
    v[0] = bx; v[1] = by; v[2] = bz;
    v[3] = (float)(0.5 * ((double)bx * bx + (double)by * by + (double)bz * bz));
    v[4] = (float)(qq * bx); v[5] = (float)(qq * by); v[6] = (float)(qq * bz); v[7] = (float)qq;
  }
  float* tile = reinterpret_cast<float*>(rec) + (s >> 4) * 128;
  const int i = (int)(s & 15);
#pragma unroll
  for (int k = 0; k < 4; ++k) tile[k * 16 + i] = v[k];
#pragma unroll
  for (int c = 0; c < 4; ++c) tile[64 + ((i >> 2) * 4 + c) * 4 + (i & 3)] = v[4 + c];
}

// Mixed-precision Biot-Savart pair kernel on the matrix cores.
//
// Per wave: NTT (1..4) target tiles of 16 (up to 64 targets of one panel).
// For every source set, coordinates are taken relative to the source-panel
// centre c_s: target p = (x - c_t) + (c_t - c_s), source b (record).  For a
// 16 x 16 tile
//   D = alpha_j + beta_i - p_j.b_i = eps + |p_j - b_i|^2 / 2  (= 1 - x.y + eps)
// is one v_mfma_f32_16x16x4_f32 (A = (b, beta) per source, B = (-p, 1) per
// target) plus alpha per target column; R = 1/D on the vector unit; then four
// v_mfma_f32_4x4x1_16b_f32 accumulate W_j = sum_i R_ij (q b_i, q), 4
// components x 16 targets per quarter of the source tile.  After each source
// set every lane adds its quarter's w_j += W_xyz - p_j W_q into FP64; the four
// quarters are summed across lanes once at the end, and u = x cross w in FP64
// (x cross y = x cross (y - x)).
// Lane layouts: A[i][k] from lane i + 16 k, B[k][j] from lane j + 16 k,
// D[i][j] in lane j + 16 (i / 4), register i % 4 (16x16x4); block b of the
// 4x4x1 form uses lanes 4 b .. 4 b + 3 (A row / B column = lane % 4, D row =
// register).  The next source tile's operands are loaded before the current
// tile's arithmetic (every index stays inside the set's padded records).
template <int NTT, bool M2L>
__device__ inline void bve_mfma_tiles(const DevTree& t, const TargetTile& T, const SourceSet* __restrict__ sets,
                                      const float* __restrict__ rec, float eps, int lane, double (&w)[4][3]) {
  const int col = lane & 15, grp = lane >> 4;
  const DevPanel& PT = t.panels[T.panel];
  float ax[NTT], ay[NTT], az[NTT];
#pragma unroll
  for (int tt = 0; tt < NTT; ++tt) {
    const int j = min(tt * 16 + col, T.count - 1);
    const int64_t u = (int64_t)T.first + j;
    ax[tt] = (float)(t.px[u] - PT.cx);
    ay[tt] = (float)(t.py[u] - PT.cy);
    az[tt] = (float)(t.pz[u] - PT.cz);
  }
  const int off_b = 64 + 4 * (grp * 4 + (lane & 3));  // float4: (q b, q)[lane % 4] of sources 4 grp + 0..3
  if (T.src_begin >= T.src_end) return;
  // b1 = (-p_x, -p_y, -p_z, 1)[lane / 16] without branches
  const float sx = grp == 0 ? -1.f : 0.f, sy = grp == 1 ? -1.f : 0.f, sz = grp == 2 ? -1.f : 0.f;
  const float s1 = grp == 3 ? 1.f : 0.f;
  float px[NTT], py[NTT], pz[NTT], b1[NTT], alpha[NTT];
  v4f acc[NTT];
  auto begin_set = [&](const SourceSet& S) {
#pragma unroll
    for (int tt = 0; tt < NTT; ++tt) {
      px[tt] = ax[tt] + S.ox;
      py[tt] = ay[tt] + S.oy;
      pz[tt] = az[tt] + S.oz;
      alpha[tt] = fmaf(0.5f, fmaf(px[tt], px[tt], fmaf(py[tt], py[tt], pz[tt] * pz[tt])), eps);
      b1[tt] = fmaf(sx, px[tt], fmaf(sy, py[tt], fmaf(sz, pz[tt], s1)));
      acc[tt] = v4f{0.f, 0.f, 0.f, 0.f};
    }
  };
  // One flat loop over the 16-source tiles of all source sets, so that the
  // operands of the next tile are loaded one tile ahead across set boundaries
  // and the set records two sets ahead (the loads stay in the loop body).
  const int last = T.src_end - 1;
  int e = T.src_begin;
  SourceSet S = sets[e], Sn = sets[min(e + 1, last)], Snn = sets[min(e + 2, last)];
  begin_set(S);
  int nst = (S.count + 15) >> 4, st = 0;
  const float* __restrict__ r = rec + (int64_t)S.first16 * 8;
  float a1 = r[lane];
  v4f bq = *reinterpret_cast<const v4f*>(r + off_b);
  while (true) {
    const float* __restrict__ rn = st + 1 < nst ? r + 128 * (st + 1) : rec + (int64_t)Sn.first16 * 8;
    const float na1 = rn[lane];
    const v4f nbq = *reinterpret_cast<const v4f*>(rn + off_b);
    v4f d[NTT];
    kernel_products<NTT>(a1, b1, M2L && S.m2l, d);
#pragma unroll
    for (int tt = 0; tt < NTT; ++tt)
#pragma unroll
      for (int q = 0; q < 4; ++q) d[tt][q] = __builtin_amdgcn_rcpf(d[tt][q] + alpha[tt]);
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int tt = 0; tt < NTT; ++tt)
        acc[tt] = __builtin_amdgcn_mfma_f32_4x4x1f32(bq[q], d[tt][q], acc[tt], 0, 0, 0);
    a1 = na1;
    bq = nbq;
    if (++st == nst) {
#pragma unroll
      for (int tt = 0; tt < NTT; ++tt) {
        w[tt][0] += (double)fmaf(-px[tt], acc[tt][3], acc[tt][0]);
        w[tt][1] += (double)fmaf(-py[tt], acc[tt][3], acc[tt][1]);
        w[tt][2] += (double)fmaf(-pz[tt], acc[tt][3], acc[tt][2]);
      }
      if (++e > last) break;
      S = Sn;
      Sn = Snn;
      Snn = sets[min(e + 2, last)];
      begin_set(S);
      r = rec + (int64_t)S.first16 * 8;
      nst = (S.count + 15) >> 4;
      st = 0;
    }
  }
}

template <bool M2L>
__global__ __launch_bounds__(256) void k_bve_mfma(DevTree t, const TargetTile* __restrict__ tiles, int ntiles,
                                                  const SourceSet* __restrict__ sets,
                                                  const float* __restrict__ rec, float eps, Ptr3 part,
                                                  Ptr3 proxy, int assign, int xcd) {
  member_targets(t, part, proxy);
  rec += (int64_t)blockIdx.y * t.U16 * 8;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int col = lane & 15, grp = lane >> 4;
  const int tile = shaped_tile(t.xcd_start, wave, xcd);
  if (tile >= ntiles) return;
  const TargetTile T = tiles[tile];
  const int ntt = (T.count + 15) >> 4;
  double w[4][3];
#pragma unroll
  for (int tt = 0; tt < 4; ++tt) w[tt][0] = w[tt][1] = w[tt][2] = 0.0;
  switch (ntt) {
    case 4: bve_mfma_tiles<4, M2L>(t, T, sets, rec, eps, lane, w); break;
    case 3: bve_mfma_tiles<3, M2L>(t, T, sets, rec, eps, lane, w); break;
    case 2: bve_mfma_tiles<2, M2L>(t, T, sets, rec, eps, lane, w); break;
    default: bve_mfma_tiles<1, M2L>(t, T, sets, rec, eps, lane, w); break;
  }
  // sum the four source quarters (lanes col, col+16, col+32, col+48)
#pragma unroll
  for (int tt = 0; tt < 4; ++tt)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      w[tt][k] += __shfl_xor(w[tt][k], 16);
      w[tt][k] += __shfl_xor(w[tt][k], 32);
    }
  // lane group g stores target tile g
  const int j = grp * 16 + col;
  if (grp < ntt && j < T.count) {
    const double w0 = grp == 0 ? w[0][0] : grp == 1 ? w[1][0] : grp == 2 ? w[2][0] : w[3][0];
    const double w1 = grp == 0 ? w[0][1] : grp == 1 ? w[1][1] : grp == 2 ? w[2][1] : w[3][1];
    const double w2 = grp == 0 ? w[0][2] : grp == 1 ? w[1][2] : grp == 2 ? w[2][2] : w[3][2];
    const int64_t u = (int64_t)T.first + j;
    const double x = t.px[u], y = t.py[u], z = t.pz[u];
    store_u(t, u, part, proxy, y * w2 - z * w1, z * w0 - x * w2, x * w1 - y * w0, assign);
  }
}

// ---- multi-right-hand-side form (ensembles with shared positions)
//
// When every member of an ensemble has its sources and targets at the same
// positions (the undisplaced first RK4 stage), R_ij = 1/(eps + |p_j - b_i|^2/2)
// is the same matrix for all members and the pair sums are the product
// W[j][n] = sum_i R_ij Q[i][n] with one column n per member and component
// (q b_x, q b_y, q b_z, q).  R comes from the 16x16x4 product of the
// single-member kernel once per wave and is used again as the A operand of
// 16x16x4 products against 16 charge columns (four members): register r of
// lane l holds R[4 (l/16) + r][l % 16], i.e. A[j][a] = R[4 a + r][j], so
// product r takes sources {4 a + r}, and its B operand is Q[4 a + r][n] from
// lane n + 16 a.  The result W[j][n] lands in lane n + 16 (j/4), register
// j % 4.
//
// Charge columns: for 16-source tile g and chunk k of four members, 64 lanes x
// float4,  qcol[((g * nchunk + k) * 64 + n + 16 a) * 4 + r] = Q[4 a + r][n],
// n = 4 * (member % 4) + c.  Members beyond the ensemble (padding of the last
// chunk group) have zero columns.
__global__ void k_src_mrhs(DevTree t, const double* __restrict__ q_part, const double* __restrict__ q_proxy,
                           const float* __restrict__ geo, int nmem, int nchunk, float* __restrict__ qcol) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= t.U16) return;
  const int mem = (int)blockIdx.y;
  const int64_t lp = (int64_t)t.Pl * t.m2p;
  int64_t u;
  bool valid;
  if (s < lp) {
    const int k = (int)(s % t.m2p);
    valid = k < t.m2;
    u = (s / t.m2p) * t.m2 + k;
  } else {
    const int64_t v = s - lp;
    const int k = (int)(v % t.np16);
    valid = k < t.n;
    u = t.npart + (v / t.np16) * t.n + k;
  }
  double q = 0.0;
  if (valid && mem < nmem)
    q = u < t.npart ? q_part[(int64_t)mem * t.npart + u]
                    : q_proxy[(int64_t)mem * t.P * t.n + (u - t.npart)];
  const int64_t g = s >> 4;
  const int i = (int)(s & 15);
  const float* tile = geo + g * 128;  // (b_x, b_y, b_z) of source i at [16 k + i]
  const double qq = kNegInv4Pi * q;
  const float v[4] = {(float)(qq * tile[i]), (float)(qq * tile[16 + i]), (float)(qq * tile[32 + i]), (float)qq};
  float* o = qcol + ((g * nchunk + (mem >> 2)) * 64 + 16 * (i >> 2) + 4 * (mem & 3)) * 4 + (i & 3);
#pragma unroll
  for (int c = 0; c < 4; ++c) o[4 * c] = v[c];
}

// Lane SEL of each quad of lanes (DPP quad_perm, all four lanes read SEL).
template <int SEL>
__device__ inline float quad_bcast(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), SEL * 0x55, 0xF, 0xF, false));
}
template <int SEL>
__device__ inline double quad_bcast(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), SEL * 0x55, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), SEL * 0x55, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Pair sums of NTT target tiles of 16 for CB chunks of four members.  Lane l
// accumulates, per source set, w[tt][cb][r] += W_c - p_c W_q for target
// tt*16 + 4 (l/16) + r, member 4 (chunk0 + cb) + (l%16)/4, component c = l % 4
// (lanes with c = 3 carry W_q and are discarded).
template <int NTT, int CB, bool M2L>
__device__ inline void bve_mrhs_tiles(const DevTree& t, const TargetTile& T, const SourceSet* __restrict__ sets,
                                      const float* __restrict__ geo, const float* __restrict__ qcol, int nchunk,
                                      int chunk0, float eps, int lane, double (&w)[4][CB][4]) {
  const int col = lane & 15, grp = lane >> 4, comp = lane & 3;
  const DevPanel& PT = t.panels[T.panel];
  const double* __restrict__ pc_arr = comp == 0 ? t.px : comp == 1 ? t.py : t.pz;
  const double pc_ctr = comp == 0 ? PT.cx : comp == 1 ? PT.cy : PT.cz;
  float ax[NTT], ay[NTT], az[NTT], ac[NTT][4];
#pragma unroll
  for (int tt = 0; tt < NTT; ++tt) {
    const int j = min(tt * 16 + col, T.count - 1);
    const int64_t u = (int64_t)T.first + j;
    ax[tt] = (float)(t.px[u] - PT.cx);
    ay[tt] = (float)(t.py[u] - PT.cy);
    az[tt] = (float)(t.pz[u] - PT.cz);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int jj = min(tt * 16 + 4 * grp + r, T.count - 1);
      ac[tt][r] = comp == 3 ? 0.f : (float)(pc_arr[(int64_t)T.first + jj] - pc_ctr);
    }
  }
  if (T.src_begin >= T.src_end) return;
  const float sx = grp == 0 ? -1.f : 0.f, sy = grp == 1 ? -1.f : 0.f, sz = grp == 2 ? -1.f : 0.f;
  const float s1 = grp == 3 ? 1.f : 0.f;
  float b1[NTT], alpha[NTT], pc[NTT][4];
  v4f acc[NTT][CB];
  auto begin_set = [&](const SourceSet& S) {
    const float so = comp == 0 ? S.ox : comp == 1 ? S.oy : comp == 2 ? S.oz : 0.f;
#pragma unroll
    for (int tt = 0; tt < NTT; ++tt) {
      const float px = ax[tt] + S.ox, py = ay[tt] + S.oy, pz = az[tt] + S.oz;
      alpha[tt] = fmaf(0.5f, fmaf(px, px, fmaf(py, py, pz * pz)), eps);
      b1[tt] = fmaf(sx, px, fmaf(sy, py, fmaf(sz, pz, s1)));
#pragma unroll
      for (int r = 0; r < 4; ++r) pc[tt][r] = comp == 3 ? 0.f : ac[tt][r] + so;
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) acc[tt][cb] = v4f{0.f, 0.f, 0.f, 0.f};
    }
  };
  auto qptr = [&](int64_t g) { return qcol + ((g * nchunk + chunk0) * 64 + lane) * 4; };
  const int last = T.src_end - 1;
  int e = T.src_begin;
  SourceSet S = sets[e], Sn = sets[min(e + 1, last)], Snn = sets[min(e + 2, last)];
  begin_set(S);
  int nst = (S.count + 15) >> 4, st = 0;
  int64_t g = S.first16 >> 4;
  float a1 = geo[g * 128 + lane];
  v4f qv[CB];
#pragma unroll
  for (int cb = 0; cb < CB; ++cb) qv[cb] = *reinterpret_cast<const v4f*>(qptr(g) + cb * 256);
  while (true) {
    const int64_t gn = st + 1 < nst ? g + 1 : (int64_t)(Sn.first16 >> 4);
    const float na1 = geo[gn * 128 + lane];
    v4f nqv[CB];
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) nqv[cb] = *reinterpret_cast<const v4f*>(qptr(gn) + cb * 256);
    v4f d[NTT];
    kernel_products<NTT>(a1, b1, M2L && S.m2l, d);
#pragma unroll
    for (int tt = 0; tt < NTT; ++tt)
#pragma unroll
      for (int r = 0; r < 4; ++r) d[tt][r] = __builtin_amdgcn_rcpf(d[tt][r] + alpha[tt]);
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int tt = 0; tt < NTT; ++tt)
#pragma unroll
        for (int cb = 0; cb < CB; ++cb)
          acc[tt][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(d[tt][r], qv[cb][r], acc[tt][cb], 0, 0, 0);
    a1 = na1;
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) qv[cb] = nqv[cb];
    g = gn;
    if (++st == nst) {
#pragma unroll
      for (int tt = 0; tt < NTT; ++tt)
#pragma unroll
        for (int cb = 0; cb < CB; ++cb)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float wq = quad_bcast<3>(acc[tt][cb][r]);
            w[tt][cb][r] += (double)fmaf(-pc[tt][r], wq, acc[tt][cb][r]);
          }
      if (++e > last) break;
      S = Sn;
      Sn = Snn;
      Snn = sets[min(e + 2, last)];
      begin_set(S);
      nst = (S.count + 15) >> 4;
      st = 0;
    }
  }
}

// One wave per target tile and group of CB member chunks (blockIdx.y).  Every
// (target, member, component) is written by exactly one lane.
template <int CB, bool M2L>
__global__ __launch_bounds__(256) void k_bve_mrhs(DevTree t, const TargetTile* __restrict__ tiles, int ntiles,
                                                  const SourceSet* __restrict__ sets, const float* __restrict__ geo,
                                                  const float* __restrict__ qcol, int nchunk, int nmem, float eps,
                                                  Ptr3 part, Ptr3 proxy, int assign, int xcd) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int col = lane & 15, grp = lane >> 4, comp = lane & 3;
  const int tile = shaped_tile(t.xcd_start, wave, xcd);
  if (tile >= ntiles) return;
  const TargetTile T = tiles[tile];
  const int ntt = (T.count + 15) >> 4;
  const int chunk0 = (int)blockIdx.y * CB;
  double w[4][CB][4];
#pragma unroll
  for (int tt = 0; tt < 4; ++tt)
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
      for (int r = 0; r < 4; ++r) w[tt][cb][r] = 0.0;
  switch (ntt) {
    case 4: bve_mrhs_tiles<4, CB, M2L>(t, T, sets, geo, qcol, nchunk, chunk0, eps, lane, w); break;
    case 3: bve_mrhs_tiles<3, CB, M2L>(t, T, sets, geo, qcol, nchunk, chunk0, eps, lane, w); break;
    case 2: bve_mrhs_tiles<2, CB, M2L>(t, T, sets, geo, qcol, nchunk, chunk0, eps, lane, w); break;
    default: bve_mrhs_tiles<1, CB, M2L>(t, T, sets, geo, qcol, nchunk, chunk0, eps, lane, w); break;
  }
  // u = x cross w: the three components of one (target, member) sit in one quad
#pragma unroll
  for (int tt = 0; tt < 4; ++tt) {
    if (tt >= ntt) break;
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double w0 = quad_bcast<0>(w[tt][cb][r]), w1 = quad_bcast<1>(w[tt][cb][r]),
                     w2 = quad_bcast<2>(w[tt][cb][r]);
        const int j = tt * 16 + 4 * grp + r;
        const int mem = 4 * (chunk0 + cb) + (col >> 2);
        if (comp < 3 && j < T.count && mem < nmem) {
          const int64_t u = (int64_t)T.first + j;
          const double x = t.px[u], y = t.py[u], z = t.pz[u];
          const double val = comp == 0 ? y * w2 - z * w1 : comp == 1 ? z * w0 - x * w2 : x * w1 - y * w0;
          double* o = u < t.npart ? part.p[comp] + (int64_t)mem * t.npart + u
                                  : proxy.p[comp] + (int64_t)mem * t.P * t.n + (u - t.npart);
          if (assign) *o = val;
          else *o += val;
        }
      }
  }
}

inline unsigned blocks_for(int64_t count, int bs) { return (unsigned)((count + bs - 1) / bs); }

inline CPtr3 cptr3(const double* const* a) {
  CPtr3 r{};
  if (a)
    for (int f = 0; f < 3; ++f) r.p[f] = a[f];
  return r;
}
inline Ptr3 ptr3(double* const* a) {
  Ptr3 r{};
  for (int f = 0; f < 3; ++f) r.p[f] = a[f];
  return r;
}

}  // namespace

// The matrix-core kernels in their FP32 form or with bf16 M2L products.
static auto mfma_kernel(const DevTree& t) { return t.bf16_m2l ? k_bve_mfma<true> : k_bve_mfma<false>; }
template <int CB>
static auto mrhs_kernel(const DevTree& t) { return t.bf16_m2l ? k_bve_mrhs<CB, true> : k_bve_mrhs<CB, false>; }

// Mapping 2 needs the tree's XCD ranges and the grid the context computed for
// its full list; otherwise the launch uses mapping 0.
static PairShape checked_shape(const DevTree& t, PairShape shape) {
  if (shape.xcd == 2 && (!t.xcd_start || shape.grid <= 0)) shape.xcd = 0;
  return shape;
}

void launch_source_records64(const DevTree& t, const double* q_part, const double* q_proxy,
                             const double* const* disp_part, const double* const* disp_proxy, bool project,
                             SrcRec64* rec, hipStream_t s, int nmem) {
  hipLaunchKernelGGL(k_src64, dim3(blocks_for(t.U, 256), nmem), dim3(256), 0, s, t, q_part, q_proxy, cptr3(disp_part),
                     cptr3(disp_proxy), project ? 1 : 0, rec);
}

void launch_source_records32(const DevTree& t, const double* q_part, const double* q_proxy,
                             const double* const* disp_part, const double* const* disp_proxy, SrcRec32* rec,
                             hipStream_t s, int nmem) {
  hipLaunchKernelGGL(k_src32, dim3(blocks_for(t.U, 256), nmem), dim3(256), 0, s, t, q_part, q_proxy, cptr3(disp_part),
                     cptr3(disp_proxy), 1, rec);
}

void launch_bve_interact64(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* sets,
                           const SrcRec64* rec, double eps, double* const* out_part, double* const* out_proxy,
                           hipStream_t s, bool assign, int nmem, PairShape shape) {
  if (ntiles == 0) return;
  shape = checked_shape(t, shape);
  hipLaunchKernelGGL(k_bve64, dim3(pair_grid(ntiles, shape), nmem), dim3(64 * shape.waves), 0, s, t, tiles,
                     ntiles, sets, rec, eps, ptr3(out_part), ptr3(out_proxy), assign ? 1 : 0, shape.xcd);
}

void launch_poisson_interact64(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* sets,
                               const SrcRec64* rec, double* out_part, double* out_proxy, hipStream_t s,
                               bool assign) {
  if (ntiles == 0) return;
  hipLaunchKernelGGL(k_log64, dim3(blocks_for(ntiles, kWavesPerBlock)), dim3(64 * kWavesPerBlock), 0, s, t, tiles,
                     ntiles, sets, rec, out_part, out_proxy, assign ? 1 : 0);
}

void launch_laplacian_interact64(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* sets,
                                 const SrcRec64* rec, double* out_part, double* out_proxy, hipStream_t s,
                                 bool assign) {
  if (ntiles == 0) return;
  const double eps = 0.1, pi = 3.14159265358979323846;
  // the records carry -q/(4 pi): undo it with the kernel's eps^{-5/2}
  const LapCoef64 kp = lap_coef64(eps, -4.0 * pi / std::pow(eps, 2.5));
  const RowGroups gs = row_groups(t);
  if (gs.G > 1)
    hipLaunchKernelGGL(k_lap64_gr, dim3(blocks_for((int64_t)ntiles * gs.H, kWavesPerBlock)), dim3(64 * kWavesPerBlock),
                       0, s, t, tiles, ntiles, gs, sets, rec, kp, out_part, out_proxy, assign ? 1 : 0);
  else
    hipLaunchKernelGGL(k_lap64, dim3(blocks_for(ntiles, kWavesPerBlock)), dim3(64 * kWavesPerBlock), 0, s, t, tiles,
                       ntiles, sets, rec, kp, out_part, out_proxy, assign ? 1 : 0);
}

void launch_laplacian_interact_mixed(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* sets,
                                     const SrcRec32* rec, double* out_part, double* out_proxy, hipStream_t s,
                                     bool assign) {
  if (ntiles == 0) return;
  const double eps = 0.1, pi = 3.14159265358979323846;
  const double c3 = -4.0 * pi / std::pow(eps, 2.5);
  const RowGroups gs = row_groups(t);
  if (gs.G > 1)
    hipLaunchKernelGGL(k_lap_mixed_gr, dim3(blocks_for((int64_t)ntiles * gs.H, kWavesPerBlock)),
                       dim3(64 * kWavesPerBlock), 0, s, t, tiles, ntiles, gs, sets, rec, lap_coef32(eps, 1.0), c3,
                       out_part, out_proxy, assign ? 1 : 0);
  else
    hipLaunchKernelGGL(k_lap_mixed, dim3(blocks_for(ntiles, kWavesPerBlock)), dim3(64 * kWavesPerBlock), 0, s, t,
                       tiles, ntiles, sets, rec, lap_coef32(eps, 1.0), c3, out_part, out_proxy, assign ? 1 : 0);
}

void launch_bve_interact_mixed(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* sets,
                               const SrcRec32* rec, double eps, double* const* out_part,
                               double* const* out_proxy, hipStream_t s, bool assign, int nmem, PairShape shape) {
  if (ntiles == 0) return;
  shape = checked_shape(t, shape);
  hipLaunchKernelGGL(k_bve_mixed, dim3(pair_grid(ntiles, shape), nmem), dim3(64 * shape.waves), 0, s, t,
                     tiles, ntiles, sets, rec, (float)eps, ptr3(out_part), ptr3(out_proxy), assign ? 1 : 0, shape.xcd);
}

void launch_source_records_mfma(const DevTree& t, const double* q_part, const double* q_proxy,
                                const double* const* disp_part, const double* const* disp_proxy, SrcRecMfma* rec,
                                hipStream_t s, int nmem) {
  hipLaunchKernelGGL(k_src_mfma, dim3(blocks_for(t.U16, 256), nmem), dim3(256), 0, s, t, q_part, q_proxy,
                     cptr3(disp_part), cptr3(disp_proxy), rec);
}

void launch_bve_interact_mfma(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* sets,
                              const SrcRecMfma* rec, double eps, double* const* out_part, double* const* out_proxy,
                              hipStream_t s, bool assign, int nmem, PairShape shape) {
  if (ntiles == 0) return;
  shape = checked_shape(t, shape);
  hipLaunchKernelGGL(mfma_kernel(t), dim3(pair_grid(ntiles, shape), nmem), dim3(64 * shape.waves), 0, s, t,
                     tiles, ntiles, sets, (const float*)rec, (float)eps, ptr3(out_part), ptr3(out_proxy),
                     assign ? 1 : 0, shape.xcd);
}

void launch_source_columns_mrhs(const DevTree& t, const double* q_part, const double* q_proxy,
                                const SrcRecMfma* geo, int nmem, int cb, float* qcol, hipStream_t s) {
  const int nchunk = mrhs_chunks(nmem, cb);
  hipLaunchKernelGGL(k_src_mrhs, dim3(blocks_for(t.U16, 256), 4 * nchunk), dim3(256), 0, s, t, q_part, q_proxy,
                     (const float*)geo, nmem, nchunk, qcol);
}

void launch_bve_interact_mrhs(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* sets,
                              const SrcRecMfma* geo, const float* qcol, int nmem, int cb, double eps,
                              double* const* out_part, double* const* out_proxy, hipStream_t s, bool assign,
                              PairShape shape) {
  if (ntiles == 0) return;
  const int nchunk = mrhs_chunks(nmem, cb);
  shape = checked_shape(t, shape);
  const dim3 grid(pair_grid(ntiles, shape), nchunk / cb), block(64 * shape.waves);
  if (cb == 2)
    hipLaunchKernelGGL(mrhs_kernel<2>(t), grid, block, 0, s, t, tiles, ntiles, sets, (const float*)geo, qcol, nchunk,
                       nmem, (float)eps, ptr3(out_part), ptr3(out_proxy), assign ? 1 : 0, shape.xcd);
  else
    hipLaunchKernelGGL(mrhs_kernel<1>(t), grid, block, 0, s, t, tiles, ntiles, sets, (const float*)geo, qcol, nchunk,
                       nmem, (float)eps, ptr3(out_part), ptr3(out_proxy), assign ? 1 : 0, shape.xcd);
}

void launch_poisson_interact_mixed(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* sets,
                                   const SrcRec32* rec, double* out_part, double* out_proxy, hipStream_t s,
                                   bool assign) {
  if (ntiles == 0) return;
  hipLaunchKernelGGL(k_log_mixed, dim3(blocks_for(ntiles, kWavesPerBlock)), dim3(64 * kWavesPerBlock), 0, s, t,
                     tiles, ntiles, sets, rec, out_part, out_proxy, assign ? 1 : 0);
}

}  // namespace sbslpm
