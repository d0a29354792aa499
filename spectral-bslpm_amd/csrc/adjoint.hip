// adjoint.hip -- exact transpose of the FP64 Biot-Savart FMM evaluation
// (velocity: upward pass, pair sums, downward pass; bve_vel_impl.hpp:34-70,
// fmm_funcs.cpp:341-392) with respect to the source charges and the source
// displacements.
//
// The forward pass computes, for every target row t of the interaction list
// (leaf particles, proxies of non-leaf target panels),
//   phi_t = x_t cross sum_s g(x_t, y_s) q'_s y_s,  g(x, y) = 1 / (1 - x.y + eps),
// with y_s = P(node_s + d_s) and q'_s = -q_s / (4 pi), and the particle
// velocity u = phi_part + L2P(L2L(phi_proxy)).  For a velocity cotangent ubar:
//   * the target-side cotangent is psi = ubar at particles and its
//     anterpolation (P2M, M2M: the transposes of L2P, L2L) at proxy rows;
//   * since psi_t . (x_t cross y) = y . c_t with c_t = psi_t cross x_t, one
//     pair contributes  g q' (y . c_t)  to the loss, whence
//       qbar'_s = sum_t g (y_s . c_t),
//       ybar_s  = q'_s sum_t g (c_t + g (y_s . c_t) x_t)      (dg/dy = g^2 x);
//   * the source-side cotangents go back through the projection
//     (dbar = (I - y y^T) ybar / |node + d|) and, for proxy sources, through
//     the anterpolation of the charges (downward pass) and the interpolation
//     of their displacements (scattered back to the leaf particles).
// The pair sums run over the interaction list seen from its sources: one
// wave per tile of at most 64 source points, each lane owning one source
// row (assigned, no atomics), the targets' records read wave-uniformly.
#include "device.hpp"

namespace sbslpm {
namespace {

struct CPtr3 {
  const double* p[3];
};
struct Ptr3 {
  double* p[3];
};

// Target record: undisplaced target node and c = psi cross x (48 bytes, read
// with wave-uniform scalar loads).
struct TgtRec {
  double x, y, z, c0, c1, c2;
};

__global__ void k_tgt_rec(DevTree t, CPtr3 psi_part, CPtr3 psi_prox, TgtRec* __restrict__ rec) {
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
  rec[u] = TgtRec{x, y, z, b * z - d * y, d * x - a * z, a * y - b * x};
}

// Transposed pair sums.  POS: also the position cotangent (displaced stages).
template <bool POS, bool GR>
__global__ __launch_bounds__(256) void k_bve_adj64(DevTree t, const TargetTile* __restrict__ tiles, int ntiles,
                                                   RowGroups gs, const SourceSet* __restrict__ targets,
                                                   const TgtRec* __restrict__ trec, const SrcRec64* __restrict__ srec,
                                                   double eps, CPtr3 dpart, CPtr3 dprox, double* __restrict__ qbar_part,
                                                   double* __restrict__ qbar_prox, Ptr3 dbar_part, Ptr3 dbar_prox, int xcd) {
  __shared__ TgtRec stage[GR ? kWavesPerBlock : 1][64];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int tile, row0, cg, ci;
  if (!group_tile<GR>(gs, ntiles, wave, xcd, tile, row0, cg, ci)) return;
  const TargetTile T = tiles[tile];
  const int row = row0 + ci;
  const int64_t u = (int64_t)T.first + (row < T.count ? row : T.count - 1);
  const int ng = GR ? gs.G : 1;
  const bool copy_ok = !GR || cg < ng;
  const SrcRec64 S = srec[u];
  double aq = 0, a0 = 0, a1 = 0, a2 = 0;
  for (int e = T.src_begin; e < T.src_end; ++e) {
    const SourceSet G = targets[e];
    const TgtRec* __restrict__ r = trec + G.first;
    for (int c0 = 0; c0 < G.count; c0 += GR ? 64 : G.count) {
      const int cn = GR ? min(64, G.count - c0) : G.count;
      if (GR) stage_records(stage[wave], r + c0, cn);
      const int trips = GR ? (cn + ng - 1) / ng : cn;
#pragma unroll 4
      for (int kk = 0; kk < trips; ++kk) {
        const int k = GR ? kk * ng + cg : kk;
        if (GR && !(copy_ok && k < cn)) continue;
        const TgtRec R = GR ? stage[wave][k] : r[k];
        // the forward kernel's denominator, x = target, y = source
        const double g = rcp64_pair((1.0 - (R.x * S.x + R.y * S.y + R.z * S.z)) + eps);
        const double dc = S.x * R.c0 + S.y * R.c1 + S.z * R.c2;
        aq = fma(g, dc, aq);
        if (POS) {
          const double h = g * dc;
          a0 = fma(g, fma(h, R.x, R.c0), a0);
          a1 = fma(g, fma(h, R.y, R.c1), a1);
          a2 = fma(g, fma(h, R.z, R.c2), a2);
        }
      }
    }
  }
  if (GR) {
    aq = group_sum(aq, gs, ci);
    if (POS) {
      a0 = group_sum(a0, gs, ci);
      a1 = group_sum(a1, gs, ci);
      a2 = group_sum(a2, gs, ci);
    }
  }
  if (cg != 0 || row >= T.count) return;
  const bool part = u < t.npart;
  const int64_t v = part ? u : u - t.npart;
  (part ? qbar_part : qbar_prox)[v] = kNegInv4Pi * aq;
  if (POS) {
    const double w0 = t.px[u] + (part ? dpart.p[0] : dprox.p[0])[v];
    const double w1 = t.py[u] + (part ? dpart.p[1] : dprox.p[1])[v];
    const double w2 = t.pz[u] + (part ? dpart.p[2] : dprox.p[2])[v];
    const double ir = 1.0 / sqrt(w0 * w0 + w1 * w1 + w2 * w2);
    const double yb0 = S.q * a0, yb1 = S.q * a1, yb2 = S.q * a2;
    const double yy = S.x * yb0 + S.y * yb1 + S.z * yb2;
    (part ? dbar_part.p[0] : dbar_prox.p[0])[v] = (yb0 - S.x * yy) * ir;
    (part ? dbar_part.p[1] : dbar_prox.p[1])[v] = (yb1 - S.y * yy) * ir;
    (part ? dbar_part.p[2] : dbar_prox.p[2])[v] = (yb2 - S.z * yy) * ir;
  }
}

// FP32 target record: node relative to its panel centre and c = psi cross x.
struct TgtRec32 {
  float ax, ay, az, c0, c1, c2;
};

__global__ void k_tgt_rec32(DevTree t, CPtr3 psi_part, CPtr3 psi_prox, TgtRec32* __restrict__ rec) {
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
  rec[u] = TgtRec32{(float)(x - P.cx), (float)(y - P.cy), (float)(z - P.cz), (float)(b * z - d * y),
                    (float)(d * x - a * z), (float)(a * y - b * x)};
}

// Transposed pair sums in FP32 (mixed-precision contexts): 1 - x.y as
// |x - y|^2 / 2 in panel-centred coordinates like the forward mixed kernel,
// FP32 partial sums per target set flushed into FP64 accumulators, the
// source-side finish in FP64.
template <bool POS, bool GR>
__global__ __launch_bounds__(256) void k_bve_adj32(DevTree t, const TargetTile* __restrict__ tiles, int ntiles,
                                                   RowGroups gs, const SourceSet* __restrict__ targets,
                                                   const TgtRec32* __restrict__ trec, const SrcRec64* __restrict__ srec,
                                                   float eps, CPtr3 dpart, CPtr3 dprox, double* __restrict__ qbar_part,
                                                   double* __restrict__ qbar_prox, Ptr3 dbar_part, Ptr3 dbar_prox, int xcd) {
  __shared__ TgtRec32 stage[GR ? kWavesPerBlock : 1][64];
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
  const float yx = (float)S.x, yy = (float)S.y, yz = (float)S.z;
  double aq = 0, a0 = 0, a1 = 0, a2 = 0;
  for (int e = T.src_begin; e < T.src_end; ++e) {
    const SourceSet G = targets[e];
    const TgtRec32* __restrict__ r = trec + G.first;
    // source relative to the target panel's centre
    const float qx = bx - G.ox, qy = by - G.oy, qz = bz - G.oz;
    float sq = 0.f, s0 = 0.f, s1 = 0.f, s2 = 0.f;
    for (int c0 = 0; c0 < G.count; c0 += GR ? 64 : G.count) {
      const int cn = GR ? min(64, G.count - c0) : G.count;
      if (GR) stage_records(stage[wave], r + c0, cn);
      const int trips = GR ? (cn + ng - 1) / ng : cn;
#pragma unroll 4
      for (int kk = 0; kk < trips; ++kk) {
        const int k = GR ? kk * ng + cg : kk;
        if (GR && !(copy_ok && k < cn)) continue;
        const TgtRec32 R = GR ? stage[wave][k] : r[k];
        const float d0 = R.ax - qx, d1 = R.ay - qy, d2 = R.az - qz;  // x_t - y_s
        const float g = __builtin_amdgcn_rcpf(fmaf(0.5f, fmaf(d0, d0, fmaf(d1, d1, d2 * d2)), eps));
        const float dc = fmaf(yx, R.c0, fmaf(yy, R.c1, yz * R.c2));
        sq = fmaf(g, dc, sq);
        if (POS) {
          const float h = g * dc;
          s0 = fmaf(g, fmaf(h, d0 + yx, R.c0), s0);
          s1 = fmaf(g, fmaf(h, d1 + yy, R.c1), s1);
          s2 = fmaf(g, fmaf(h, d2 + yz, R.c2), s2);
        }
      }
    }
    aq += sq;
    if (POS) {
      a0 += s0; a1 += s1; a2 += s2;
    }
  }
  if (GR) {
    aq = group_sum(aq, gs, ci);
    if (POS) {
      a0 = group_sum(a0, gs, ci);
      a1 = group_sum(a1, gs, ci);
      a2 = group_sum(a2, gs, ci);
    }
  }
  if (cg != 0 || row >= T.count) return;
  const bool part = u < t.npart;
  const int64_t v = part ? u : u - t.npart;
  (part ? qbar_part : qbar_prox)[v] = kNegInv4Pi * aq;
  if (POS) {
    const double w0 = t.px[u] + (part ? dpart.p[0] : dprox.p[0])[v];
    const double w1 = t.py[u] + (part ? dpart.p[1] : dprox.p[1])[v];
    const double w2 = t.pz[u] + (part ? dpart.p[2] : dprox.p[2])[v];
    const double ir = 1.0 / sqrt(w0 * w0 + w1 * w1 + w2 * w2);
    const double yb0 = S.q * a0, yb1 = S.q * a1, yb2 = S.q * a2;
    const double yy2 = S.x * yb0 + S.y * yb1 + S.z * yb2;
    (part ? dbar_part.p[0] : dbar_prox.p[0])[v] = (yb0 - S.x * yy2) * ir;
    (part ? dbar_part.p[1] : dbar_prox.p[1])[v] = (yb1 - S.y * yy2) * ir;
    (part ? dbar_part.p[2] : dbar_prox.p[2])[v] = (yb2 - S.z * yy2) * ir;
  }
}

}  // namespace

size_t adj_target_record_bytes() { return sizeof(TgtRec) > sizeof(TgtRec32) ? sizeof(TgtRec) : sizeof(TgtRec32); }

void launch_adj_target_records32(const DevTree& t, const double* const* psi_part, const double* const* psi_prox,
                                 void* rec, hipStream_t s) {
  CPtr3 pp{{psi_part[0], psi_part[1], psi_part[2]}}, px{{psi_prox[0], psi_prox[1], psi_prox[2]}};
  hipLaunchKernelGGL(k_tgt_rec32, dim3((unsigned)((t.U + 255) / 256)), dim3(256), 0, s, t, pp, px,
                     static_cast<TgtRec32*>(rec));
}

void launch_bve_adjoint32(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* targets,
                          const void* trec, const SrcRec64* srec, double eps, const double* const* dpart,
                          const double* const* dprox, double* qbar_part, double* qbar_prox, double* const* dbar_part,
                          double* const* dbar_prox, hipStream_t s, PairShape shape) {
  if (ntiles == 0) return;
  const RowGroups gs = row_groups(t);
  const bool gr = gs.G > 1;
  const dim3 grid((unsigned)(((int64_t)ntiles * gs.H + shape.waves - 1) / shape.waves)), block(64 * shape.waves);
  const TgtRec32* r = static_cast<const TgtRec32*>(trec);
  CPtr3 dp{}, dx{};
  Ptr3 bp{}, bx{};
  if (dpart) {
    dp = CPtr3{{dpart[0], dpart[1], dpart[2]}};
    dx = CPtr3{{dprox[0], dprox[1], dprox[2]}};
    bp = Ptr3{{dbar_part[0], dbar_part[1], dbar_part[2]}};
    bx = Ptr3{{dbar_prox[0], dbar_prox[1], dbar_prox[2]}};
  }
#define SBSLPM_BVE_ADJ(P, G)                                                                                        \
  hipLaunchKernelGGL((k_bve_adj32<P, G>), grid, block, 0, s, t, tiles, ntiles, gs, targets, r, srec, (float)eps, dp, dx,     \
                     qbar_part, qbar_prox, bp, bx, shape.xcd)
  if (dpart && gr) SBSLPM_BVE_ADJ(true, true);
  else if (dpart) SBSLPM_BVE_ADJ(true, false);
  else if (gr) SBSLPM_BVE_ADJ(false, true);
  else SBSLPM_BVE_ADJ(false, false);
#undef SBSLPM_BVE_ADJ
}

void launch_adj_target_records(const DevTree& t, const double* const* psi_part, const double* const* psi_prox,
                               void* rec, hipStream_t s) {
  CPtr3 pp{{psi_part[0], psi_part[1], psi_part[2]}}, px{{psi_prox[0], psi_prox[1], psi_prox[2]}};
  hipLaunchKernelGGL(k_tgt_rec, dim3((unsigned)((t.U + 255) / 256)), dim3(256), 0, s, t, pp, px,
                     static_cast<TgtRec*>(rec));
}

void launch_bve_adjoint64(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* targets,
                          const void* trec, const SrcRec64* srec, double eps, const double* const* dpart,
                          const double* const* dprox, double* qbar_part, double* qbar_prox, double* const* dbar_part,
                          double* const* dbar_prox, hipStream_t s, PairShape shape) {
  if (ntiles == 0) return;
  const RowGroups gs = row_groups(t);
  const bool gr = gs.G > 1;
  const dim3 grid((unsigned)(((int64_t)ntiles * gs.H + shape.waves - 1) / shape.waves)), block(64 * shape.waves);
  const TgtRec* r = static_cast<const TgtRec*>(trec);
  CPtr3 dp{}, dx{};
  Ptr3 bp{}, bx{};
  if (dpart) {
    dp = CPtr3{{dpart[0], dpart[1], dpart[2]}};
    dx = CPtr3{{dprox[0], dprox[1], dprox[2]}};
    bp = Ptr3{{dbar_part[0], dbar_part[1], dbar_part[2]}};
    bx = Ptr3{{dbar_prox[0], dbar_prox[1], dbar_prox[2]}};
  }
#define SBSLPM_BVE_ADJ(P, G)                                                                                        \
  hipLaunchKernelGGL((k_bve_adj64<P, G>), grid, block, 0, s, t, tiles, ntiles, gs, targets, r, srec, eps, dp, dx,     \
                     qbar_part, qbar_prox, bp, bx, shape.xcd)
  if (dpart && gr) SBSLPM_BVE_ADJ(true, true);
  else if (dpart) SBSLPM_BVE_ADJ(true, false);
  else if (gr) SBSLPM_BVE_ADJ(false, true);
  else SBSLPM_BVE_ADJ(false, false);
#undef SBSLPM_BVE_ADJ
}

}  // namespace sbslpm
