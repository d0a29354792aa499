// swe_adjoint.hip -- exact transposes of the shallow-water step's pair sums
// with respect to their source charges and displaced source positions, over
// the interaction list seen from its sources (one wave per tile of at most 64
// source rows, each lane owning one row, target records read wave-uniformly;
// tiles that leave lanes idle run in lane groups, see RowGroups):
//   * the shallow-water velocity sums, variant 1 (swe_vel_panel_interaction,
//     swe_vel_impl.hpp:12-76): per pair
//       u_t += r (Qz (x cross y) + Qd (y - x (x.y))),  r = 1 / (1 - x.y + eps),
//     Qz, Qd the charges times -1/(4 pi).  For a velocity cotangent psi at the
//     target, psi.(x cross y) = y.c1 and psi.(y - x (x.y)) = y.c2 with
//     c1 = psi cross x, c2 = psi - x (x.psi), so
//       Qz_bar += r (y.c1),  Qd_bar += r (y.c2),
//       y_bar  += r (r (Qz y.c1 + Qd y.c2) x + Qz c1 + Qd c2)     (dr/dy = r^2 x);
//   * the regularised-Laplacian sums (lap_panel_interaction,
//     laplacian_impl.hpp:46-66): phi_t += K(x.y) q, so
//       q_bar += psi K(t),  y_bar += q psi K'(t) x,
//     K' from the derivative of each factor of K.
// Source-side finish as in the Biot-Savart adjoint: the position cotangent
// goes back through the projection y = P(node + d).  FP64 (k_*_adj64) and,
// for mixed-precision contexts, FP32 pair arithmetic in panel-centred
// coordinates with FP32 partial sums per target set flushed into FP64
// (k_*_adj32).
#include <algorithm>
#include <cmath>

#include "device.hpp"
#include "lapkernel.hpp"

namespace sbslpm {
namespace {

struct CPtr3 {
  const double* p[3];
};
struct Ptr3 {
  double* p[3];
};

// ---- shallow-water velocity sums ----------------------------------------------

// Target record: node, c1 = psi cross x, c2 = psi - x (x.psi) (72 bytes).
struct SweTgt {
  double x, y, z, a0, a1, a2, b0, b1, b2;
};

__global__ void k_swe_tgt(DevTree t, CPtr3 psi_part, CPtr3 psi_prox, SweTgt* __restrict__ rec) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= t.U) return;
  double a, b, d;
  if (u < t.npart) {
    a = psi_part.p[0][u]; b = psi_part.p[1][u]; d = psi_part.p[2][u];
  } else {
    const int64_t v = u - t.npart;
    if (v / t.n >= t.offset) return;  // leaf proxy rows are never targets
    a = psi_prox.p[0][v]; b = psi_prox.p[1][v]; d = psi_prox.p[2][v];
  }
  const double x = t.px[u], y = t.py[u], z = t.pz[u];
  const double xp = x * a + y * b + z * d;
  rec[u] = SweTgt{x, y, z, b * z - d * y, d * x - a * z, a * y - b * x, a - x * xp, b - y * xp, d - z * xp};
}

// Stores of one source row: charge cotangents (times -1/(4 pi): the records
// carry the scaled charges), and the displacement cotangent through the
// projection.  acc: add the displacement cotangent to what is there.
template <bool POS>
__device__ inline void swe_store(const DevTree& t, int64_t u, const double* S3, double az, double ad, const double* pb,
                                 CPtr3 dpart, CPtr3 dprox, double* qz_part, double* qz_prox, double* qd_part,
                                 double* qd_prox, Ptr3 dbar_part, Ptr3 dbar_prox, int acc) {
  const bool part = u < t.npart;
  const int64_t v = part ? u : u - t.npart;
  (part ? qz_part : qz_prox)[v] = kNegInv4Pi * az;
  (part ? qd_part : qd_prox)[v] = kNegInv4Pi * ad;
  if (POS) {
    const double w0 = t.px[u] + (part ? dpart.p[0] : dprox.p[0])[v];
    const double w1 = t.py[u] + (part ? dpart.p[1] : dprox.p[1])[v];
    const double w2 = t.pz[u] + (part ? dpart.p[2] : dprox.p[2])[v];
    const double ir = 1.0 / sqrt(w0 * w0 + w1 * w1 + w2 * w2);
    const double yy = S3[0] * pb[0] + S3[1] * pb[1] + S3[2] * pb[2];
    // one select per component (selecting the pointer array itself puts it in scratch)
    double* o[3] = {part ? dbar_part.p[0] : dbar_prox.p[0], part ? dbar_part.p[1] : dbar_prox.p[1],
                    part ? dbar_part.p[2] : dbar_prox.p[2]};
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const double d = (pb[f] - S3[f] * yy) * ir;
      o[f][v] = acc ? o[f][v] + d : d;
    }
  }
}

template <bool POS, bool GR>
__global__ __launch_bounds__(256) void k_swe_adj64(DevTree t, const TargetTile* __restrict__ tiles, int ntiles,
                                                   RowGroups gs, const SourceSet* __restrict__ targets,
                                                   const SweTgt* __restrict__ trec,
                                                   const SrcRecSwe* __restrict__ srec, double eps, CPtr3 dpart,
                                                   CPtr3 dprox, double* __restrict__ qz_part,
                                                   double* __restrict__ qz_prox, double* __restrict__ qd_part,
                                                   double* __restrict__ qd_prox, Ptr3 dbar_part, Ptr3 dbar_prox,
                                                   int acc, int xcd) {
  __shared__ SweTgt stage[GR ? kWavesPerBlock : 1][64];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int tile, row0, cg, ci;
  if (!group_tile<GR>(gs, ntiles, wave, xcd, tile, row0, cg, ci)) return;
  const TargetTile T = tiles[tile];
  const int row = row0 + ci;
  const int64_t u = (int64_t)T.first + (row < T.count ? row : T.count - 1);
  const int ng = GR ? gs.G : 1;
  const bool copy_ok = !GR || cg < ng;
  const SrcRecSwe S = srec[u];
  double az = 0, ad = 0, p0 = 0, p1 = 0, p2 = 0;
  for (int e = T.src_begin; e < T.src_end; ++e) {
    const SourceSet G = targets[e];
    const SweTgt* __restrict__ r = trec + G.first;
    for (int c0 = 0; c0 < G.count; c0 += GR ? 64 : G.count) {
      const int cn = GR ? min(64, G.count - c0) : G.count;
      if (GR) stage_records(stage[wave], r + c0, cn);
      const int trips = GR ? (cn + ng - 1) / ng : cn;
#pragma unroll 2
      for (int kk = 0; kk < trips; ++kk) {
        const int k = GR ? kk * ng + cg : kk;
        if (GR && !(copy_ok && k < cn)) continue;
        const SweTgt R = GR ? stage[wave][k] : r[k];
        const double g = rcp64_pair((1.0 - (R.x * S.x + R.y * S.y + R.z * S.z)) + eps);
        const double a1 = S.x * R.a0 + S.y * R.a1 + S.z * R.a2;
        const double a2 = S.x * R.b0 + S.y * R.b1 + S.z * R.b2;
        az = fma(g, a1, az);
        ad = fma(g, a2, ad);
        if (POS) {
          const double w = g * fma(S.qz, a1, S.qd * a2);
          p0 = fma(g, fma(w, R.x, fma(S.qz, R.a0, S.qd * R.b0)), p0);
          p1 = fma(g, fma(w, R.y, fma(S.qz, R.a1, S.qd * R.b1)), p1);
          p2 = fma(g, fma(w, R.z, fma(S.qz, R.a2, S.qd * R.b2)), p2);
        }
      }
    }
  }
  if (GR) {
    az = group_sum(az, gs, ci);
    ad = group_sum(ad, gs, ci);
    if (POS) {
      p0 = group_sum(p0, gs, ci);
      p1 = group_sum(p1, gs, ci);
      p2 = group_sum(p2, gs, ci);
    }
  }
  if (cg != 0 || row >= T.count) return;
  const double S3[3] = {S.x, S.y, S.z}, pb[3] = {p0, p1, p2};
  swe_store<POS>(t, u, S3, az, ad, pb, dpart, dprox, qz_part, qz_prox, qd_part, qd_prox, dbar_part, dbar_prox, acc);
}

// FP32 target record: node relative to its panel centre, absolute node, c1, c2.
struct SweTgt32 {
  float ax, ay, az, x, y, z, a0, a1, a2, b0, b1, b2;
};

__global__ void k_swe_tgt32(DevTree t, CPtr3 psi_part, CPtr3 psi_prox, SweTgt32* __restrict__ rec) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= t.U) return;
  double a, b, d;
  int panel;
  if (u < t.npart) {
    a = psi_part.p[0][u]; b = psi_part.p[1][u]; d = psi_part.p[2][u];
    panel = t.offset + (int)(u / t.m2);
  } else {
    const int64_t v = u - t.npart;
    panel = (int)(v / t.n);
    if (panel >= t.offset) return;
    a = psi_prox.p[0][v]; b = psi_prox.p[1][v]; d = psi_prox.p[2][v];
  }
  const double x = t.px[u], y = t.py[u], z = t.pz[u];
  const DevPanel& P = t.panels[panel];
  const double xp = x * a + y * b + z * d;
  rec[u] = SweTgt32{(float)(x - P.cx), (float)(y - P.cy), (float)(z - P.cz), (float)x, (float)y, (float)z,
                    (float)(b * z - d * y), (float)(d * x - a * z), (float)(a * y - b * x),
                    (float)(a - x * xp), (float)(b - y * xp), (float)(d - z * xp)};
}

// FP32 pairs: d = y - x from panel-centred coordinates, h = |d|^2/2 = 1 - x.y,
// and since c1, c2 are orthogonal to x, y.c1 = d.c1 and y.c2 = d.c2.
template <bool POS, bool GR>
__global__ __launch_bounds__(256) void k_swe_adj32(DevTree t, const TargetTile* __restrict__ tiles, int ntiles,
                                                   RowGroups gs, const SourceSet* __restrict__ targets,
                                                   const SweTgt32* __restrict__ trec,
                                                   const SrcRecSwe* __restrict__ srec, float eps, CPtr3 dpart,
                                                   CPtr3 dprox, double* __restrict__ qz_part,
                                                   double* __restrict__ qz_prox, double* __restrict__ qd_part,
                                                   double* __restrict__ qd_prox, Ptr3 dbar_part, Ptr3 dbar_prox,
                                                   int acc, int xcd) {
  __shared__ SweTgt32 stage[GR ? kWavesPerBlock : 1][64];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int tile, row0, cg, ci;
  if (!group_tile<GR>(gs, ntiles, wave, xcd, tile, row0, cg, ci)) return;
  const TargetTile T = tiles[tile];
  const int row = row0 + ci;
  const int64_t u = (int64_t)T.first + (row < T.count ? row : T.count - 1);
  const int ng = GR ? gs.G : 1;
  const bool copy_ok = !GR || cg < ng;
  const SrcRecSwe S = srec[u];
  const DevPanel& PS = t.panels[T.panel];
  const float bx = (float)(S.x - PS.cx), by = (float)(S.y - PS.cy), bz = (float)(S.z - PS.cz);
  const float qz = (float)S.qz, qd = (float)S.qd;
  double az = 0, ad = 0, p0 = 0, p1 = 0, p2 = 0;
  for (int e = T.src_begin; e < T.src_end; ++e) {
    const SourceSet G = targets[e];
    const SweTgt32* __restrict__ r = trec + G.first;
    // source relative to the target panel's centre
    const float sx = bx - G.ox, sy = by - G.oy, sz = bz - G.oz;
    float fz = 0.f, fd = 0.f, f0 = 0.f, f1 = 0.f, f2 = 0.f;
    for (int c0 = 0; c0 < G.count; c0 += GR ? 64 : G.count) {
      const int cn = GR ? min(64, G.count - c0) : G.count;
      if (GR) stage_records(stage[wave], r + c0, cn);
      const int trips = GR ? (cn + ng - 1) / ng : cn;
#pragma unroll 2
      for (int kk = 0; kk < trips; ++kk) {
        const int k = GR ? kk * ng + cg : kk;
        if (GR && !(copy_ok && k < cn)) continue;
        const SweTgt32 R = GR ? stage[wave][k] : r[k];
        const float d0 = sx - R.ax, d1 = sy - R.ay, d2 = sz - R.az;  // y_s - x_t
        const float g = __builtin_amdgcn_rcpf(fmaf(0.5f, fmaf(d0, d0, fmaf(d1, d1, d2 * d2)), eps));
        const float a1 = fmaf(d0, R.a0, fmaf(d1, R.a1, d2 * R.a2));
        const float a2 = fmaf(d0, R.b0, fmaf(d1, R.b1, d2 * R.b2));
        fz = fmaf(g, a1, fz);
        fd = fmaf(g, a2, fd);
        if (POS) {
          const float w = g * fmaf(qz, a1, qd * a2);
          f0 = fmaf(g, fmaf(w, R.x, fmaf(qz, R.a0, qd * R.b0)), f0);
          f1 = fmaf(g, fmaf(w, R.y, fmaf(qz, R.a1, qd * R.b1)), f1);
          f2 = fmaf(g, fmaf(w, R.z, fmaf(qz, R.a2, qd * R.b2)), f2);
        }
      }
    }
    az += fz;
    ad += fd;
    if (POS) {
      p0 += f0; p1 += f1; p2 += f2;
    }
  }
  if (GR) {
    az = group_sum(az, gs, ci);
    ad = group_sum(ad, gs, ci);
    if (POS) {
      p0 = group_sum(p0, gs, ci);
      p1 = group_sum(p1, gs, ci);
      p2 = group_sum(p2, gs, ci);
    }
  }
  if (cg != 0 || row >= T.count) return;
  const double S3[3] = {S.x, S.y, S.z}, pb[3] = {p0, p1, p2};
  swe_store<POS>(t, u, S3, az, ad, pb, dpart, dprox, qz_part, qz_prox, qd_part, qd_prox, dbar_part, dbar_prox, acc);
}

// ---- regularised-Laplacian sums ------------------------------------------------

constexpr double kLapEps = 0.1;
constexpr double kPiL = 3.14159265358979323846;

struct LapTgt {
  double x, y, z, psi;
};

__global__ void k_lap_tgt(DevTree t, const double* __restrict__ psi_part, const double* __restrict__ psi_prox,
                          LapTgt* __restrict__ rec) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= t.U) return;
  double p;
  if (u < t.npart) {
    p = psi_part[u];
  } else {
    const int64_t v = u - t.npart;
    if (v / t.n >= t.offset) return;
    p = psi_prox[v];
  }
  rec[u] = LapTgt{t.px[u], t.py[u], t.pz[u], p};
}

// Horner multipliers of the FP64 transposed Laplacian in vector registers (the
// first step of each Horner scheme is then one three-operand fma).
__device__ inline double in_vgpr(double v) {
  __asm__ volatile("" : "+v"(v));
  return v;
}
__device__ inline float in_vgpr_f(float v) {
  __asm__ volatile("" : "+v"(v));
  return v;
}

// K(t) and K'(t) eps^{5/2} in the folded form of lapkernel.hpp.  fast: t >= 6 eps
// (K = A + E B, K' = A' + E D); otherwise the erf factor f and the e^{-s^2}
// term g, and zero at or below the cut.
template <bool POS>
__device__ inline void lap_kd(double tt, const LapCoef64& kp, double a1v, double b2v, double e1v, double d2v,
                              double c24v, const double* etab, bool fast, double& k, double& kd) {
  const double t2 = tt * tt;
  const double A = tt * fma(a1v, t2, kp.a0);
  const double B = fma(fma(b2v, t2, kp.b1), t2, kp.b0);
  const double E = exp_neg(fma(t2, kp.inv_eps, -kp.inv_eps), etab, c24v);
  double Ad = 0, D = 0;
  if (POS) {
    Ad = fma(e1v, t2, kp.e0);
    D = tt * fma(fma(d2v, t2, kp.d1), t2, kp.d0);
  }
  if (fast) {
    k = fma(E, B, A);
    if (POS) kd = fma(E, D, Ad);
    return;
  }
  const double sv = tt * kp.inv_eps;
  const bool mid = fabs(sv) < 6.0;
  const double f = sv >= 6.0 ? 1.0 : (sv <= -6.0 ? 0.0 : 0.5 * (1.0 + erf(sv)));
  const bool in = tt > kp.cut;
  k = in ? fma(E * f, B, A) : 0.0;
  if (POS) {
    const double g = mid ? 2.0 / (sqrt(kPiL) * kLapEps) * exp(-sv * sv) : 0.0;
    kd = in ? fma(E, fma(f, D, 0.5 * g * B), Ad) : 0.0;
  }
}

// Records are the forward's (SrcRec64, q times -1/(4 pi)); c3 = -4 pi / eps^{5/2}
// undoes the scaling.  q_bar = eps^{-5/2} sum psi K;  y_bar = q' c3 sum psi K' x.
// A wave tests each source at once (a ballot) and takes the erf form only when
// one of its pairs has t < 6 eps.
template <bool POS, bool GR>
__global__ __launch_bounds__(256) void k_lap_adj64(DevTree t, const TargetTile* __restrict__ tiles, int ntiles,
                                                   RowGroups gs, const SourceSet* __restrict__ targets,
                                                   const LapTgt* __restrict__ trec,
                                                   const SrcRec64* __restrict__ srec, LapCoef64 kp, double c3,
                                                   CPtr3 dpart, CPtr3 dprox, double* __restrict__ q_part,
                                                   double* __restrict__ q_prox, Ptr3 dbar_part, Ptr3 dbar_prox,
                                                   int acc, int xcd) {
  __shared__ double etab[kExpNegTabSize];
  for (int i = threadIdx.x; i < kExpNegTabSize; i += blockDim.x) etab[i] = t.expnegtab[i];
  __syncthreads();
  __shared__ LapTgt stage[GR ? kWavesPerBlock : 1][64];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int tile, row0, cg, ci;
  if (!group_tile<GR>(gs, ntiles, wave, xcd, tile, row0, cg, ci)) return;
  const TargetTile T = tiles[tile];
  const int row = row0 + ci;
  const int64_t u = (int64_t)T.first + (row < T.count ? row : T.count - 1);
  const int ng = GR ? gs.G : 1;
  const bool copy_ok = !GR || cg < ng;
  const SrcRec64 S = srec[u];
  const bool no_cut = kp.cut < -1.0;
  const double a1v = in_vgpr(kp.a1), b2v = in_vgpr(kp.b2), e1v = in_vgpr(kp.e1), d2v = in_vgpr(kp.d2);
  const double c24v = in_vgpr(1.0 / 24.0);
  double aq = 0, p0 = 0, p1 = 0, p2 = 0;
  for (int e = T.src_begin; e < T.src_end; ++e) {
    const SourceSet G = targets[e];
    const LapTgt* __restrict__ r = trec + G.first;
    for (int c0 = 0; c0 < G.count; c0 += GR ? 64 : G.count) {
      const int cn = GR ? min(64, G.count - c0) : G.count;
      if (GR) stage_records(stage[wave], r + c0, cn);
      const int trips = GR ? (cn + ng - 1) / ng : cn;
      for (int kk = 0; kk < trips; ++kk) {
        const int k = GR ? kk * ng + cg : kk;
        if (GR && !(copy_ok && k < cn)) continue;
        const LapTgt R = GR ? stage[wave][k] : r[k];
        const double tt = R.x * S.x + R.y * S.y + R.z * S.z;
        double kv, kd = 0;
        if (no_cut && __builtin_amdgcn_ballot_w64(tt < kp.t_sat) == 0) {
          lap_kd<POS>(tt, kp, a1v, b2v, e1v, d2v, c24v, etab, true, kv, kd);
        } else {
          // a real branch (the empty asm keeps the compiler from evaluating erf on every pair)
          __asm__ volatile("");
          lap_kd<POS>(tt, kp, a1v, b2v, e1v, d2v, c24v, etab, false, kv, kd);
        }
        aq = fma(R.psi, kv, aq);
        if (POS) {
          const double w = R.psi * kd;
          p0 = fma(w, R.x, p0);
          p1 = fma(w, R.y, p1);
          p2 = fma(w, R.z, p2);
        }
      }
    }
  }
  if (GR) {
    aq = group_sum(aq, gs, ci);
    if (POS) {
      p0 = group_sum(p0, gs, ci);
      p1 = group_sum(p1, gs, ci);
      p2 = group_sum(p2, gs, ci);
    }
  }
  if (cg != 0 || row >= T.count) return;
  const bool part = u < t.npart;
  const int64_t v = part ? u : u - t.npart;
  (part ? q_part : q_prox)[v] = c3 * kNegInv4Pi * aq;
  if (POS) {
    const double w0 = t.px[u] + (part ? dpart.p[0] : dprox.p[0])[v];
    const double w1 = t.py[u] + (part ? dpart.p[1] : dprox.p[1])[v];
    const double w2 = t.pz[u] + (part ? dpart.p[2] : dprox.p[2])[v];
    const double ir = 1.0 / sqrt(w0 * w0 + w1 * w1 + w2 * w2);
    const double sc = S.q * c3;
    const double yb[3] = {sc * p0, sc * p1, sc * p2};
    const double yy = S.x * yb[0] + S.y * yb[1] + S.z * yb[2];
    double* const* o = part ? dbar_part.p : dbar_prox.p;
    const double s3[3] = {S.x, S.y, S.z};
    for (int f = 0; f < 3; ++f) {
      const double d = (yb[f] - s3[f] * yy) * ir;
      o[f][v] = acc ? o[f][v] + d : d;
    }
  }
}

// FP32 target record for the Laplacian: node relative to its panel centre,
// absolute node, psi.
struct LapTgt32 {
  float ax, ay, az, x, y, z, psi, pad;
};

__global__ void k_lap_tgt32(DevTree t, const double* __restrict__ psi_part, const double* __restrict__ psi_prox,
                            LapTgt32* __restrict__ rec) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= t.U) return;
  double p;
  int panel;
  if (u < t.npart) {
    p = psi_part[u];
    panel = t.offset + (int)(u / t.m2);
  } else {
    const int64_t v = u - t.npart;
    panel = (int)(v / t.n);
    if (panel >= t.offset) return;
    p = psi_prox[v];
  }
  const double x = t.px[u], y = t.py[u], z = t.pz[u];
  const DevPanel& P = t.panels[panel];
  rec[u] = LapTgt32{(float)(x - P.cx), (float)(y - P.cy), (float)(z - P.cz), (float)x, (float)y, (float)z, (float)p,
                    0.f};
}

// FP32 pairs as k_lap_mixed (lapkernel.hpp's polynomials in u, the base-2
// exponent of E, from the squared chord h2), erf and the e^{-s^2} term only
// when a ballot finds a pair with t < 6 eps.
template <bool POS>
__device__ inline void lap_kd32(float h2, const LapCoef32& kp, float a1v, float b2v, float e1v, float d2v, bool fast,
                                float& k, float& kd) {
  const float tt = fmaf(h2, -0.5f, 1.0f);
  const float uu = h2 * fmaf(h2, 0.25f * kp.k, -kp.k);
  const float A = tt * fmaf(a1v, uu, kp.a0);
  const float B = fmaf(fmaf(b2v, uu, kp.b1), uu, kp.b0);
  const float E = __builtin_amdgcn_exp2f(uu);
  float Ad = 0.f, D = 0.f;
  if (POS) {
    Ad = fmaf(e1v, uu, kp.e0);
    D = tt * fmaf(fmaf(d2v, uu, kp.d1), uu, kp.d0);
  }
  if (fast) {
    k = fmaf(E, B, A);
    if (POS) kd = fmaf(E, D, Ad);
    return;
  }
  const float sv = tt * kp.inv_eps;
  const bool mid = fabsf(sv) < 6.0f;
  const float f = sv >= 6.0f ? 1.0f : (sv <= -6.0f ? 0.0f : 0.5f * (1.0f + erff(sv)));
  const bool in = tt > kp.cut;
  k = in ? fmaf(E * f, B, A) : 0.0f;
  if (POS) {
    const float g = mid ? (float)(2.0 / (1.7724538509055160273 * kLapEps)) *
                              __builtin_amdgcn_exp2f(-sv * sv * 1.44269504088896341f)
                        : 0.0f;
    kd = in ? fmaf(E, fmaf(f, D, 0.5f * g * B), Ad) : 0.0f;
  }
}

template <bool POS, bool GR>
__global__ __launch_bounds__(256) void k_lap_adj32(DevTree t, const TargetTile* __restrict__ tiles, int ntiles,
                                                   RowGroups gs, const SourceSet* __restrict__ targets,
                                                   const LapTgt32* __restrict__ trec,
                                                   const SrcRec64* __restrict__ srec, LapCoef32 kp, double c3,
                                                   CPtr3 dpart, CPtr3 dprox, double* __restrict__ q_part,
                                                   double* __restrict__ q_prox, Ptr3 dbar_part, Ptr3 dbar_prox,
                                                   int acc, int xcd) {
  __shared__ LapTgt32 stage[GR ? kWavesPerBlock : 1][64];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int tile, row0, cg, ci;
  if (!group_tile<GR>(gs, ntiles, wave, xcd, tile, row0, cg, ci)) return;
  const TargetTile T = tiles[tile];
  const int row = row0 + ci;
  const int64_t u = (int64_t)T.first + (row < T.count ? row : T.count - 1);
  const int ng = GR ? gs.G : 1;
  const bool copy_ok = !GR || cg < ng;
  const SrcRec64 S = srec[u];
  const DevPanel& PS = t.panels[T.panel];
  const float bx = (float)(S.x - PS.cx), by = (float)(S.y - PS.cy), bz = (float)(S.z - PS.cz);
  const bool no_cut = kp.cut < -1.0f;
  // t < t_sat  <=>  h2 > h_sat
  const float h_sat = 2.0f * (1.0f - kp.t_sat);
  const float a1v = in_vgpr_f(kp.a1), b2v = in_vgpr_f(kp.b2), e1v = in_vgpr_f(kp.e1), d2v = in_vgpr_f(kp.d2);
  double aq = 0, p0 = 0, p1 = 0, p2 = 0;
  for (int e = T.src_begin; e < T.src_end; ++e) {
    const SourceSet G = targets[e];
    const LapTgt32* __restrict__ r = trec + G.first;
    const float sx = bx - G.ox, sy = by - G.oy, sz = bz - G.oz;
    float fq = 0.f, f0 = 0.f, f1 = 0.f, f2 = 0.f;
    for (int c0 = 0; c0 < G.count; c0 += GR ? 64 : G.count) {
      const int cn = GR ? min(64, G.count - c0) : G.count;
      if (GR) stage_records(stage[wave], r + c0, cn);
      const int trips = GR ? (cn + ng - 1) / ng : cn;
      for (int kk = 0; kk < trips; ++kk) {
        const int k = GR ? kk * ng + cg : kk;
        if (GR && !(copy_ok && k < cn)) continue;
        const LapTgt32 R = GR ? stage[wave][k] : r[k];
        const float d0 = sx - R.ax, d1 = sy - R.ay, d2 = sz - R.az;
        const float h2 = fmaf(d0, d0, fmaf(d1, d1, d2 * d2));
        float kv, kd = 0.f;
        if (no_cut && __builtin_amdgcn_ballot_w64(h2 > h_sat) == 0) {
          lap_kd32<POS>(h2, kp, a1v, b2v, e1v, d2v, true, kv, kd);
        } else {
          // a real branch (the empty asm keeps the compiler from evaluating erf on every pair)
          __asm__ volatile("");
          lap_kd32<POS>(h2, kp, a1v, b2v, e1v, d2v, false, kv, kd);
        }
        fq = fmaf(R.psi, kv, fq);
        if (POS) {
          const float w = R.psi * kd;
          f0 = fmaf(w, R.x, f0);
          f1 = fmaf(w, R.y, f1);
          f2 = fmaf(w, R.z, f2);
        }
      }
    }
    aq += fq;
    if (POS) {
      p0 += f0; p1 += f1; p2 += f2;
    }
  }
  if (GR) {
    aq = group_sum(aq, gs, ci);
    if (POS) {
      p0 = group_sum(p0, gs, ci);
      p1 = group_sum(p1, gs, ci);
      p2 = group_sum(p2, gs, ci);
    }
  }
  if (cg != 0 || row >= T.count) return;
  const bool part = u < t.npart;
  const int64_t v = part ? u : u - t.npart;
  (part ? q_part : q_prox)[v] = c3 * kNegInv4Pi * aq;
  if (POS) {
    const double w0 = t.px[u] + (part ? dpart.p[0] : dprox.p[0])[v];
    const double w1 = t.py[u] + (part ? dpart.p[1] : dprox.p[1])[v];
    const double w2 = t.pz[u] + (part ? dpart.p[2] : dprox.p[2])[v];
    const double ir = 1.0 / sqrt(w0 * w0 + w1 * w1 + w2 * w2);
    const double sc = S.q * c3;
    const double yb[3] = {sc * p0, sc * p1, sc * p2};
    const double yy = S.x * yb[0] + S.y * yb[1] + S.z * yb[2];
    double* const* o = part ? dbar_part.p : dbar_prox.p;
    const double s3[3] = {S.x, S.y, S.z};
    for (int f = 0; f < 3; ++f) {
      const double d = (yb[f] - s3[f] * yy) * ir;
      o[f][v] = acc ? o[f][v] + d : d;
    }
  }
}

CPtr3 cptr(const double* const* p) {
  CPtr3 r{};
  for (int i = 0; i < 3; ++i) r.p[i] = p ? p[i] : nullptr;
  return r;
}
Ptr3 ptr(double* const* p) {
  Ptr3 r{};
  for (int i = 0; i < 3; ++i) r.p[i] = p ? p[i] : nullptr;
  return r;
}

}  // namespace

size_t swe_adj_target_record_bytes() {
  size_t b = sizeof(SweTgt);
  b = std::max(b, sizeof(SweTgt32));
  b = std::max(b, sizeof(LapTgt));
  return std::max(b, sizeof(LapTgt32));
}

void launch_swe_adj_target_records(const DevTree& t, bool fp32, const double* const* psi_part,
                                   const double* const* psi_prox, void* rec, hipStream_t s) {
  const dim3 grid((unsigned)((t.U + 255) / 256)), block(256);
  if (fp32)
    hipLaunchKernelGGL(k_swe_tgt32, grid, block, 0, s, t, cptr(psi_part), cptr(psi_prox), static_cast<SweTgt32*>(rec));
  else
    hipLaunchKernelGGL(k_swe_tgt, grid, block, 0, s, t, cptr(psi_part), cptr(psi_prox), static_cast<SweTgt*>(rec));
}

void launch_swe_adjoint(const DevTree& t, bool fp32, const TargetTile* tiles, int ntiles, const SourceSet* targets,
                        const void* trec, const SrcRecSwe* srec, double eps, const double* const* dpart,
                        const double* const* dprox, double* qz_part, double* qz_prox, double* qd_part,
                        double* qd_prox, double* const* dbar_part, double* const* dbar_prox, bool accumulate,
                        hipStream_t s, PairShape shape) {
  if (ntiles == 0) return;
  const RowGroups gs = row_groups(t);
  const bool gr = gs.G > 1;
  const int tasks = ntiles * gs.H;
  const dim3 grid((unsigned)((tasks + shape.waves - 1) / shape.waves)), block(64 * shape.waves);
  const int acc = accumulate ? 1 : 0;
  const bool pos = dpart != nullptr;
  const CPtr3 dp = pos ? cptr(dpart) : CPtr3{}, dx = pos ? cptr(dprox) : CPtr3{};
  const Ptr3 bp = pos ? ptr(dbar_part) : Ptr3{}, bx = pos ? ptr(dbar_prox) : Ptr3{};
#define SBSLPM_SWE_ADJ(KERNEL, REC, EPSV, P, G)                                                                   \
  hipLaunchKernelGGL((KERNEL<P, G>), grid, block, 0, s, t, tiles, ntiles, gs, targets, static_cast<const REC*>(trec), \
                     srec, EPSV, dp, dx, qz_part, qz_prox, qd_part, qd_prox, bp, bx, acc, shape.xcd)
  if (fp32) {
    if (pos && gr) SBSLPM_SWE_ADJ(k_swe_adj32, SweTgt32, (float)eps, true, true);
    else if (pos) SBSLPM_SWE_ADJ(k_swe_adj32, SweTgt32, (float)eps, true, false);
    else if (gr) SBSLPM_SWE_ADJ(k_swe_adj32, SweTgt32, (float)eps, false, true);
    else SBSLPM_SWE_ADJ(k_swe_adj32, SweTgt32, (float)eps, false, false);
  } else {
    if (pos && gr) SBSLPM_SWE_ADJ(k_swe_adj64, SweTgt, eps, true, true);
    else if (pos) SBSLPM_SWE_ADJ(k_swe_adj64, SweTgt, eps, true, false);
    else if (gr) SBSLPM_SWE_ADJ(k_swe_adj64, SweTgt, eps, false, true);
    else SBSLPM_SWE_ADJ(k_swe_adj64, SweTgt, eps, false, false);
  }
#undef SBSLPM_SWE_ADJ
}

void launch_lap_adj_target_records(const DevTree& t, bool fp32, const double* psi_part, const double* psi_prox,
                                   void* rec, hipStream_t s) {
  const dim3 grid((unsigned)((t.U + 255) / 256)), block(256);
  if (fp32)
    hipLaunchKernelGGL(k_lap_tgt32, grid, block, 0, s, t, psi_part, psi_prox, static_cast<LapTgt32*>(rec));
  else
    hipLaunchKernelGGL(k_lap_tgt, grid, block, 0, s, t, psi_part, psi_prox, static_cast<LapTgt*>(rec));
}

void launch_laplacian_adjoint(const DevTree& t, bool fp32, const TargetTile* tiles, int ntiles,
                              const SourceSet* targets, const void* trec, const SrcRec64* srec,
                              const double* const* dpart, const double* const* dprox, double* q_part, double* q_prox,
                              double* const* dbar_part, double* const* dbar_prox, bool accumulate, hipStream_t s,
                              PairShape shape) {
  if (ntiles == 0) return;
  const double eps = kLapEps;
  const double c3 = -4.0 * kPiL / std::pow(eps, 2.5);
  const RowGroups gs = row_groups(t);
  const bool gr = gs.G > 1;
  const int tasks = ntiles * gs.H;
  const dim3 grid((unsigned)((tasks + shape.waves - 1) / shape.waves)), block(64 * shape.waves);
  const int acc = accumulate ? 1 : 0;
  const bool pos = dpart != nullptr;
  const CPtr3 dp = pos ? cptr(dpart) : CPtr3{}, dx = pos ? cptr(dprox) : CPtr3{};
  const Ptr3 bp = pos ? ptr(dbar_part) : Ptr3{}, bx = pos ? ptr(dbar_prox) : Ptr3{};
#define SBSLPM_LAP_ADJ(KERNEL, REC, COEF, P, G)                                                                   \
  hipLaunchKernelGGL((KERNEL<P, G>), grid, block, 0, s, t, tiles, ntiles, gs, targets, static_cast<const REC*>(trec), \
                     srec, COEF, c3, dp, dx, q_part, q_prox, bp, bx, acc, shape.xcd)
  if (fp32) {
    const LapCoef32 kp = lap_coef32(eps, 1.0);
    if (pos && gr) SBSLPM_LAP_ADJ(k_lap_adj32, LapTgt32, kp, true, true);
    else if (pos) SBSLPM_LAP_ADJ(k_lap_adj32, LapTgt32, kp, true, false);
    else if (gr) SBSLPM_LAP_ADJ(k_lap_adj32, LapTgt32, kp, false, true);
    else SBSLPM_LAP_ADJ(k_lap_adj32, LapTgt32, kp, false, false);
  } else {
    const LapCoef64 kp = lap_coef64(eps, 1.0);
    if (pos && gr) SBSLPM_LAP_ADJ(k_lap_adj64, LapTgt, kp, true, true);
    else if (pos) SBSLPM_LAP_ADJ(k_lap_adj64, LapTgt, kp, true, false);
    else if (gr) SBSLPM_LAP_ADJ(k_lap_adj64, LapTgt, kp, false, true);
    else SBSLPM_LAP_ADJ(k_lap_adj64, LapTgt, kp, false, false);
  }
#undef SBSLPM_LAP_ADJ
}

}  // namespace sbslpm
