// swe.hip -- shallow-water velocity pair sums (fmm_interactions/swe_vel.cpp:9-67,
// swe_vel_impl.hpp:12-147), FP64, on the same target-major tiles as the
// Biot-Savart kernels.
//
// Per pair, with x the target, y = P(node + disp) the displaced source and
// y0 the undisplaced node:
//   variant 1 (swe_vel_panel_interaction):
//     u += g qz (x cross y) + g qd (y - x (x.y)),      g = -1/(4 pi (1 - x.y + eps))
//   variant 2 (swe_vel_panel_interaction_2):
//     u += g qz (x cross y) + g0 qd (y0 - x (x.y0)),   g0 = +1/(4 pi (1 - x.y0 + eps))
//   variant 3 (swe_vel_panel_interaction_ll, swe_vel_impl.hpp:265-276, no displacement):
//     u += g qz (x cross y) - g qd (y - x (x.y)),      g = -1/(4 pi (1 - x.y + eps))
// The kernel accumulates W = sum g qz y, D = sum gd y', S = sum gd (x.y')
// (y' = y or y0) and forms u = x cross W + D - x S once per target, the same
// sums regrouped.
#include "device.hpp"

namespace sbslpm {
namespace {

struct CPtr3 {
  const double* p[3];
};
struct Ptr3 {
  double* p[3];
};

// Source records: displaced re-projected position, undisplaced node, and the
// charges pre-multiplied by the variant's constants.
__global__ void k_src_swe(DevTree t, const double* __restrict__ qz_part, const double* __restrict__ qz_proxy,
                          const double* __restrict__ qd_part, const double* __restrict__ qd_proxy, CPtr3 dpart,
                          CPtr3 dprox, double qd_scale, SrcRecSwe* __restrict__ rec) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= t.U) return;
  const double x0 = t.px[u], y0 = t.py[u], z0 = t.pz[u];
  double x = x0, y = y0, z = z0, qz, qd;
  if (u < t.npart) {
    if (dpart.p[0]) { x += dpart.p[0][u]; y += dpart.p[1][u]; z += dpart.p[2][u]; }
    qz = qz_part[u];
    qd = qd_part[u];
  } else {
    const int64_t v = u - t.npart;
    if (v / t.n >= t.offset) return;  // leaf proxy rows are never sources
    if (dprox.p[0]) { x += dprox.p[0][v]; y += dprox.p[1][v]; z += dprox.p[2][v]; }
    qz = qz_proxy[v];
    qd = qd_proxy[v];
  }
  project_to_sphere(x, y, z);
  rec[u] = SrcRecSwe{x, y, z, kNegInv4Pi * qz, x0, y0, z0, qd_scale * qd};
}

template <int VARIANT, bool GR>
__global__ __launch_bounds__(256) void k_swe64(DevTree t, const TargetTile* __restrict__ tiles, int ntiles,
                                               RowGroups gs, const SourceSet* __restrict__ sets,
                                               const SrcRecSwe* __restrict__ rec, double eps, Ptr3 part, Ptr3 proxy,
                                               int assign) {
  __shared__ SrcRecSwe stage[GR ? kWavesPerBlock : 1][64];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int tile, row0, cg, ci;
  if (!group_tile<GR>(gs, ntiles, wave, 1, tile, row0, cg, ci)) return;
  const TargetTile T = tiles[tile];
  const int row = row0 + ci;
  const int64_t u = (int64_t)T.first + (row < T.count ? row : T.count - 1);
  const int ng = GR ? gs.G : 1;
  const bool copy_ok = !GR || cg < ng;
  const double x = t.px[u], y = t.py[u], z = t.pz[u];
  double w0 = 0, w1 = 0, w2 = 0, d0 = 0, d1 = 0, d2 = 0, sd = 0;
  for (int e = T.src_begin; e < T.src_end; ++e) {
    const SourceSet S = sets[e];
    const SrcRecSwe* __restrict__ r = rec + S.first;
    for (int c0 = 0; c0 < S.count; c0 += GR ? 64 : S.count) {
      const int cn = GR ? min(64, S.count - c0) : S.count;
      if (GR) stage_records(stage[wave], r + c0, cn);
      const int trips = GR ? (cn + ng - 1) / ng : cn;
#pragma unroll 2
      for (int kk = 0; kk < trips; ++kk) {
        const int k = GR ? kk * ng + cg : kk;
        if (GR && !(copy_ok && k < cn)) continue;
        const SrcRecSwe s = GR ? stage[wave][k] : r[k];
        const double dp = x * s.x + y * s.y + z * s.z;
        const double g = rcp64_pair((1.0 - dp) + eps);
        const double gz = g * s.qz;
        w0 = fma(gz, s.x, w0);
        w1 = fma(gz, s.y, w1);
        w2 = fma(gz, s.z, w2);
        if (VARIANT == 1) {
          const double gd = g * s.qd;
          d0 = fma(gd, s.x, d0);
          d1 = fma(gd, s.y, d1);
          d2 = fma(gd, s.z, d2);
          sd = fma(gd, dp, sd);
        } else {
          const double dp0 = x * s.x0 + y * s.y0 + z * s.z0;
          const double gd = rcp64_pair((1.0 - dp0) + eps) * s.qd;
          d0 = fma(gd, s.x0, d0);
          d1 = fma(gd, s.y0, d1);
          d2 = fma(gd, s.z0, d2);
          sd = fma(gd, dp0, sd);
        }
      }
    }
  }
  if (GR) {
    w0 = group_sum(w0, gs, ci); w1 = group_sum(w1, gs, ci); w2 = group_sum(w2, gs, ci);
    d0 = group_sum(d0, gs, ci); d1 = group_sum(d1, gs, ci); d2 = group_sum(d2, gs, ci);
    sd = group_sum(sd, gs, ci);
  }
  if (cg != 0 || row >= T.count) return;
  const double a = (y * w2 - z * w1) + (d0 - x * sd);
  const double b = (z * w0 - x * w2) + (d1 - y * sd);
  const double c = (x * w1 - y * w0) + (d2 - z * sd);
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

// k_swe64 with the source sets of one tile shared by the workgroup's four waves
// (wave w takes sets w, w+4, ...), the four partial sums added in wave order
// through LDS.  For tile lists too short to fill the chip with one wave per
// tile (the lat-lon path's irregular tree: ~2,000 tiles of long source lists),
// this gives each SIMD four times the waves to hide the record loads behind.
// Workgroup b takes tile b: a list sorted heaviest first is dispatched in that
// order (longest-processing-time first), which evens the CUs' loads.
template <int VARIANT>
__global__ __launch_bounds__(256) void k_swe64_shared(DevTree t, const TargetTile* __restrict__ tiles, int ntiles,
                                                      const SourceSet* __restrict__ sets,
                                                      const SrcRecSwe* __restrict__ rec, double eps, Ptr3 part,
                                                      Ptr3 proxy, int assign) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int tile = blockIdx.x;
  if (tile >= ntiles) return;  // uniform over the workgroup
  const TargetTile T = tiles[tile];
  const int64_t u = (int64_t)T.first + (lane < T.count ? lane : T.count - 1);
  const double x = t.px[u], y = t.py[u], z = t.pz[u];
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};  // W (3), D (3), S
  for (int e = T.src_begin + wave; e < T.src_end; e += kWavesPerBlock) {
    const SourceSet S = sets[e];
    const SrcRecSwe* __restrict__ r = rec + S.first;
#pragma unroll 2
    for (int k = 0; k < S.count; ++k) {
      const SrcRecSwe s = r[k];
      const double dp = x * s.x + y * s.y + z * s.z;
      const double g = rcp64_pair((1.0 - dp) + eps);
      const double gz = g * s.qz;
      acc[0] = fma(gz, s.x, acc[0]);
      acc[1] = fma(gz, s.y, acc[1]);
      acc[2] = fma(gz, s.z, acc[2]);
      if (VARIANT == 1) {
        const double gd = g * s.qd;
        acc[3] = fma(gd, s.x, acc[3]);
        acc[4] = fma(gd, s.y, acc[4]);
        acc[5] = fma(gd, s.z, acc[5]);
        acc[6] = fma(gd, dp, acc[6]);
      } else {
        const double dp0 = x * s.x0 + y * s.y0 + z * s.z0;
        const double gd = rcp64_pair((1.0 - dp0) + eps) * s.qd;
        acc[3] = fma(gd, s.x0, acc[3]);
        acc[4] = fma(gd, s.y0, acc[4]);
        acc[5] = fma(gd, s.z0, acc[5]);
        acc[6] = fma(gd, dp0, acc[6]);
      }
    }
  }
  __shared__ double red[7][kWavesPerBlock - 1][64];
  if (wave > 0)
#pragma unroll
    for (int q = 0; q < 7; ++q) red[q][wave - 1][lane] = acc[q];
  __syncthreads();
  if (wave > 0 || lane >= T.count) return;
#pragma unroll
  for (int q = 0; q < 7; ++q)
    for (int w = 0; w < kWavesPerBlock - 1; ++w) acc[q] += red[q][w][lane];
  const double a = (y * acc[2] - z * acc[1]) + (acc[3] - x * acc[6]);
  const double b = (z * acc[0] - x * acc[2]) + (acc[4] - y * acc[6]);
  const double c = (x * acc[1] - y * acc[0]) + (acc[5] - z * acc[6]);
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

// Mixed-precision records: position relative to the source panel centre (FP32)
// and the two charges with the variant's constants, FP32.
__global__ void k_src_swe32(DevTree t, const double* __restrict__ qz_part, const double* __restrict__ qz_proxy,
                            const double* __restrict__ qd_part, const double* __restrict__ qd_proxy, CPtr3 dpart,
                            CPtr3 dprox, double qd_scale, SrcRecSwe32* __restrict__ rec) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= t.U) return;
  double x = t.px[u], y = t.py[u], z = t.pz[u], qz, qd;
  int panel;
  if (u < t.npart) {
    if (dpart.p[0]) { x += dpart.p[0][u]; y += dpart.p[1][u]; z += dpart.p[2][u]; }
    qz = qz_part[u];
    qd = qd_part[u];
    panel = t.offset + (int)(u / t.m2);
  } else {
    const int64_t v = u - t.npart;
    panel = (int)(v / t.n);
    if (panel >= t.offset) return;  // leaf proxy rows are never sources
    if (dprox.p[0]) { x += dprox.p[0][v]; y += dprox.p[1][v]; z += dprox.p[2][v]; }
    qz = qz_proxy[v];
    qd = qd_proxy[v];
  }
  project_to_sphere(x, y, z);
  const DevPanel& P = t.panels[panel];
  rec[u] = SrcRecSwe32{(float)(x - P.cx), (float)(y - P.cy), (float)(z - P.cz), (float)(kNegInv4Pi * qz),
                       (float)(qd_scale * qd), 0.f, 0.f, 0.f};
}

// Mixed-precision shallow-water pairs (variants 1 and 3): with d = y - x and
// h = |d|^2 / 2 = 1 - x.y computed in FP32 from panel-centred coordinates,
// x cross y = x cross d and y - x (x.y) = d + x h, so
//   u = x cross sum(g qz d) + sum(g qd d) + x sum(g qd h),   g = 1/(h + eps).
// FP32 partial sums per source set are added into FP64; the cross product and
// the output are FP64.
template <bool GR>
__global__ __launch_bounds__(256) void k_swe_mixed(DevTree t, const TargetTile* __restrict__ tiles, int ntiles,
                                                   RowGroups gs, const SourceSet* __restrict__ sets,
                                                   const SrcRecSwe32* __restrict__ rec, float eps, Ptr3 part,
                                                   Ptr3 proxy, int assign) {
  __shared__ SrcRecSwe32 stage[GR ? kWavesPerBlock : 1][64];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int tile, row0, cg, ci;
  if (!group_tile<GR>(gs, ntiles, wave, 1, tile, row0, cg, ci)) return;
  const TargetTile T = tiles[tile];
  const DevPanel& PT = t.panels[T.panel];
  const int row = row0 + ci;
  const int64_t u = (int64_t)T.first + (row < T.count ? row : T.count - 1);
  const int ng = GR ? gs.G : 1;
  const bool copy_ok = !GR || cg < ng;
  const double x = t.px[u], y = t.py[u], z = t.pz[u];
  const float ax = (float)(x - PT.cx), ay = (float)(y - PT.cy), az = (float)(z - PT.cz);
  double w0 = 0, w1 = 0, w2 = 0, d0 = 0, d1 = 0, d2 = 0, sd = 0;
  for (int e = T.src_begin; e < T.src_end; ++e) {
    const SourceSet S = sets[e];
    const float px = ax + S.ox, py = ay + S.oy, pz = az + S.oz;
    const SrcRecSwe32* __restrict__ r = rec + S.first;
    float v0 = 0.f, v1 = 0.f, v2 = 0.f, e0 = 0.f, e1 = 0.f, e2 = 0.f, eh = 0.f;
    for (int c0 = 0; c0 < S.count; c0 += GR ? 64 : S.count) {
      const int cn = GR ? min(64, S.count - c0) : S.count;
      if (GR) stage_records(stage[wave], r + c0, cn);
      const int trips = GR ? (cn + ng - 1) / ng : cn;
#pragma unroll 4
      for (int kk = 0; kk < trips; ++kk) {
        const int k = GR ? kk * ng + cg : kk;
        if (GR && !(copy_ok && k < cn)) continue;
        const SrcRecSwe32 s = GR ? stage[wave][k] : r[k];
        const float c0f = s.x - px, c1f = s.y - py, c2f = s.z - pz;
        const float h = 0.5f * fmaf(c0f, c0f, fmaf(c1f, c1f, c2f * c2f));
        const float g = __builtin_amdgcn_rcpf(h + eps);
        const float gz = g * s.qz, gd = g * s.qd;
        v0 = fmaf(gz, c0f, v0);
        v1 = fmaf(gz, c1f, v1);
        v2 = fmaf(gz, c2f, v2);
        e0 = fmaf(gd, c0f, e0);
        e1 = fmaf(gd, c1f, e1);
        e2 = fmaf(gd, c2f, e2);
        eh = fmaf(gd, h, eh);
      }
    }
    w0 += v0; w1 += v1; w2 += v2;
    d0 += e0; d1 += e1; d2 += e2;
    sd += eh;
  }
  if (GR) {
    w0 = group_sum(w0, gs, ci); w1 = group_sum(w1, gs, ci); w2 = group_sum(w2, gs, ci);
    d0 = group_sum(d0, gs, ci); d1 = group_sum(d1, gs, ci); d2 = group_sum(d2, gs, ci);
    sd = group_sum(sd, gs, ci);
  }
  if (cg != 0 || row >= T.count) return;
  const double a = (y * w2 - z * w1) + (d0 + x * sd);
  const double b = (z * w0 - x * w2) + (d1 + y * sd);
  const double c = (x * w1 - y * w0) + (d2 + z * sd);
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

}  // namespace

void launch_source_records_swe32(const DevTree& t, int variant, const double* qz_part, const double* qz_proxy,
                                 const double* qd_part, const double* qd_proxy, const double* const* disp_part,
                                 const double* const* disp_proxy, SrcRecSwe32* rec, hipStream_t s) {
  CPtr3 dp{}, dx{};
  for (int f = 0; f < 3; ++f) {
    dp.p[f] = disp_part ? disp_part[f] : nullptr;
    dx.p[f] = disp_proxy ? disp_proxy[f] : nullptr;
  }
  const double qd_scale = variant == 1 ? kNegInv4Pi : -kNegInv4Pi;
  hipLaunchKernelGGL(k_src_swe32, dim3((unsigned)((t.U + 255) / 256)), dim3(256), 0, s, t, qz_part, qz_proxy,
                     qd_part, qd_proxy, dp, dx, qd_scale, rec);
}

void launch_swe_interact_mixed(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* sets,
                               const SrcRecSwe32* rec, double eps, double* const* out_part,
                               double* const* out_proxy, hipStream_t s, bool assign) {
  if (ntiles == 0) return;
  Ptr3 op{}, ox{};
  for (int f = 0; f < 3; ++f) {
    op.p[f] = out_part[f];
    ox.p[f] = out_proxy[f];
  }
  const unsigned blocks = (unsigned)((ntiles + kWavesPerBlock - 1) / kWavesPerBlock);
  const RowGroups gs = row_groups(t);
  if (gs.G > 1)
    hipLaunchKernelGGL(k_swe_mixed<true>, dim3((((int64_t)ntiles * gs.H + kWavesPerBlock - 1) / kWavesPerBlock)),
                       dim3(64 * kWavesPerBlock), 0, s, t, tiles, ntiles, gs, sets, rec, (float)eps, op, ox,
                       assign ? 1 : 0);
  else
    hipLaunchKernelGGL(k_swe_mixed<false>, dim3(blocks), dim3(64 * kWavesPerBlock), 0, s, t, tiles, ntiles, gs, sets,
                       rec, (float)eps, op, ox, assign ? 1 : 0);
}

void launch_source_records_swe(const DevTree& t, int variant, const double* qz_part, const double* qz_proxy,
                               const double* qd_part, const double* qd_proxy, const double* const* disp_part,
                               const double* const* disp_proxy, SrcRecSwe* rec, hipStream_t s) {
  CPtr3 dp{}, dx{};
  for (int f = 0; f < 3; ++f) {
    dp.p[f] = disp_part ? disp_part[f] : nullptr;
    dx.p[f] = disp_proxy ? disp_proxy[f] : nullptr;
  }
  // variant 1: g qd with g = -1/(4 pi ...); variants 2 and 3: +1/(4 pi ...)
  const double qd_scale = variant == 1 ? kNegInv4Pi : -kNegInv4Pi;
  hipLaunchKernelGGL(k_src_swe, dim3((unsigned)((t.U + 255) / 256)), dim3(256), 0, s, t, qz_part, qz_proxy, qd_part,
                     qd_proxy, dp, dx, qd_scale, rec);
}

void launch_swe_interact64(const DevTree& t, int variant, const TargetTile* tiles, int ntiles, const SourceSet* sets,
                           const SrcRecSwe* rec, double eps, double* const* out_part, double* const* out_proxy,
                           hipStream_t s, bool assign, bool shared_tiles) {
  if (ntiles == 0) return;
  Ptr3 op{}, ox{};
  for (int f = 0; f < 3; ++f) {
    op.p[f] = out_part[f];
    ox.p[f] = out_proxy[f];
  }
  if (shared_tiles) {
    if (variant != 2)
      hipLaunchKernelGGL(k_swe64_shared<1>, dim3(ntiles), dim3(64 * kWavesPerBlock), 0, s, t, tiles, ntiles, sets, rec,
                         eps, op, ox, assign ? 1 : 0);
    else
      hipLaunchKernelGGL(k_swe64_shared<2>, dim3(ntiles), dim3(64 * kWavesPerBlock), 0, s, t, tiles, ntiles, sets, rec,
                         eps, op, ox, assign ? 1 : 0);
    return;
  }
  const RowGroups gs = row_groups(t);
  const unsigned blocks = (unsigned)(((int64_t)ntiles * gs.H + kWavesPerBlock - 1) / kWavesPerBlock);
  const dim3 block(64 * kWavesPerBlock);
  const int as = assign ? 1 : 0;
  if (variant != 2) {
    if (gs.G > 1)
      hipLaunchKernelGGL((k_swe64<1, true>), dim3(blocks), block, 0, s, t, tiles, ntiles, gs, sets, rec, eps, op, ox, as);
    else
      hipLaunchKernelGGL((k_swe64<1, false>), dim3(blocks), block, 0, s, t, tiles, ntiles, gs, sets, rec, eps, op, ox, as);
  } else {
    if (gs.G > 1)
      hipLaunchKernelGGL((k_swe64<2, true>), dim3(blocks), block, 0, s, t, tiles, ntiles, gs, sets, rec, eps, op, ox, as);
    else
      hipLaunchKernelGGL((k_swe64<2, false>), dim3(blocks), block, 0, s, t, tiles, ntiles, gs, sets, rec, eps, op, ox, as);
  }
}

}  // namespace sbslpm
