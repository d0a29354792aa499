// device.hpp -- device-side view of the tree and the kernel launchers.
#pragma once
#include "fasttab.hpp"
#include <hip/hip_runtime.h>
#include <cstdint>

#include "geometry.hpp"
#include "tree.hpp"

namespace sbslpm {

// Device error word bits (Kokkos::abort sites of the reference become flags).
enum : int {
  kErrBliBounds = 1,   // interp_vals_bli point outside its panel (interp_funcs_impl.hpp:22-33)
  kErrLocate = 2,      // point location failed, e.g. NaN (interp_funcs_impl.hpp:138)
};

// Everything a kernel needs to know about the tree; all pointers are device
// memory owned by the context.
struct DevTree {
  const DevPanel* panels;
  const double* px;  // unified point coordinates [U]: leaf particles then proxies
  const double* py;
  const double* pz;
  const double* cheb_d;  // [d+1]
  const double* cheb_m;  // [m]
  const double* m2m_hi;  // [d+1][d+1] parent basis a at upper-half child node k
  const double* m2m_lo;
  const double* m2m_hiT;  // transposes
  const double* m2m_loT;
  const double* p2m;   // [d+1][m]
  const double* p2mT;  // [m][d+1]
  const double* anc_l2l;  // composite L2L matrices (tree.hpp), or null: no fused downward pass
  const double* l2p_hi;   // [m][d+1] parent basis at the particle nodes of an upper-half child
  const double* l2p_lo;
  const double* l2pT_hi;  // [d+1][m] transposes
  const double* l2pT_lo;
  const int* proxy_leaf;      // [offset*n]
  const double* proxy_xieta;  // [offset*n][2]
  const int* proxy_order;     // [offset*n] proxy nodes sorted by leaf
  const LogTabEntry* logtab;  // [kLogTabSize] table of log_tab (fasttab.hpp)
  const double* exptab;       // [kExpTabSize] table of exp_tab
  const double* expnegtab;    // [kExpNegTabSize] table of exp_neg
  const int* xcd_start;       // [2][9] cost-balanced XCD ranges of the target-major tile list (or null)
  int* err;
  int bf16_m2l;               // matrix-core kernels take bf16 kernel-product operands for M2L sets
  int P, Pl, offset, n, dp1, m, m2, levels;
  int64_t npart, U;
  int m2p, np16;  // m2, n rounded up to 16
  int64_t U16;    // padded record count of the MFMA kernels
};

// Ensembles: kernels launched with gridDim.y = members take member
// blockIdx.y's slice of every per-member array (particle arrays [Pl][m2],
// proxy arrays [P][n], source records), stacked member after member.  With
// gridDim.y = 1 the offsets are zero.
__device__ inline int64_t member_part(const DevTree& t) { return (int64_t)blockIdx.y * t.npart; }
__device__ inline int64_t member_proxy(const DevTree& t) { return (int64_t)blockIdx.y * t.P * t.n; }

// Pair kernels run one wave per target tile, four waves per workgroup.
constexpr int kWavesPerBlock = 4;

// Launch shape of the Biot-Savart pair kernels, chosen by the context's
// autotuner: waves (tiles) per workgroup and how workgroups map to tiles:
//   xcd 0: workgroup b takes tiles b*waves.. (the hardware deals workgroups
//          round-robin over the XCDs, so neighbouring tiles land on different XCDs);
//   xcd 1: each XCD takes a contiguous range of equally many tiles;
//   xcd 2: each XCD takes a contiguous range of equal estimated work
//          (DevTree::xcd_start), for the context's full target-major list
//          only; `grid` is then the launch's workgroup count, 8 x the largest
//          range in workgroups.
// The default (four waves, mapping 0) needs no ranges; at 6 x 512^2 the
// autotuner picks mapping 2.
struct PairShape {
  int waves = kWavesPerBlock;
  int xcd = 0;
  int grid = 0;
};

// Hardware dispatches workgroup b to XCD b % 8.  Give each XCD a contiguous
// range of logical workgroups (neighbouring target tiles share sources, which
// then stay in one XCD's L2).
__device__ inline int xcd_remap(int b, int nb) {
  const int per = nb / 8, rem = nb % 8, x = b % 8, y = b / 8;
  return x < rem ? x * (per + 1) + y : rem * (per + 1) + (x - rem) * per + y;
}

// Tile of this wave under a PairShape launch (blockDim.x = 64 * waves) with
// mapping 0 or 1 (a mapping-2 request on a list without ranges launches as 0).
__device__ inline int shaped_tile(int wave, int xcd) {
  const int b = xcd == 1 ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  return b * (int)(blockDim.x >> 6) + wave;
}
// The same for the context's target-major list, where mapping 2 takes the
// y-th group of `waves` tiles of XCD x's range, x = b % 8, y = b / 8 (a
// workgroup past its range gets no tile: INT_MAX).
__device__ inline int shaped_tile(const int* __restrict__ xcd_start, int wave, int xcd) {
  if (xcd == 2) {
    const int x = (int)blockIdx.x & 7, y = (int)blockIdx.x >> 3;
    const int lo = xcd_start[x], hi = xcd_start[x + 1];
    const int tile = lo + y * (int)(blockDim.x >> 6) + wave;
    return tile < hi ? tile : 0x7fffffff;
  }
  return shaped_tile(wave, xcd);
}
// Workgroups of a pair-kernel launch over ntiles tiles.
inline int pair_grid(int ntiles, const PairShape& shape) {
  return shape.xcd == 2 && shape.grid > 0 ? shape.grid : (ntiles + shape.waves - 1) / shape.waves;
}

// Lane groups.  A tile of R rows (targets of a target-major tile, sources of a
// source-major one) leaves 64 - R lanes of its wave idle (28 of 64 at m = 6).
// With groups, the tile is cut into H sub-tiles of at most T rows, one per
// wave, and the wave's lanes hold G = 64 / T copies of the sub-tile (lane =
// g T + i: row h T + i, copy g); copy g takes records g, g + G, ... of every
// set, and the copies' sums are added in copy order at the end.  Records are
// then read per lane, so grouped waves stage them in LDS first.  {1, 64, 1} is
// the plain one-tile-per-wave form.
struct RowGroups {
  int H, T, G;
};

// Groups for a context's tiles: rows per tile follow the tile builders (a
// panel's m2 particles or n proxy rows, split into equal tiles of at most 64).
// Grouping is used only where it fills more lanes.
inline RowGroups row_groups(const DevTree& t) {
  auto tile_rows = [](int c) {
    if (c <= 0) return 0;
    const int chunks = (c + 63) / 64;
    return (c + chunks - 1) / chunks;
  };
  const int rows = tile_rows(t.m2) > tile_rows(t.n) ? tile_rows(t.m2) : tile_rows(t.n);
  if (rows <= 0) return RowGroups{1, 64, 1};
  const int H = rows <= 32 ? 1 : 2, T = (rows + H - 1) / H, G = 64 / T;
  if (G * T <= rows) return RowGroups{1, 64, 1};
  return RowGroups{H, T, G};
}

// The wave's tile, sub-tile row offset, copy and row-in-sub-tile for a launch
// over ntiles tiles (mapping xcd 0 or 1 over the ntiles * H sub-tiles when
// grouped); false if the wave has no work.
template <bool GR>
__device__ inline bool group_tile(RowGroups gs, int ntiles, int wave, int xcd, int& tile, int& row0, int& g, int& i) {
  const int lane = threadIdx.x & 63;
  if (!GR) {
    tile = shaped_tile(wave, xcd);
    row0 = 0;
    g = 0;
    i = lane;
    return tile < ntiles;
  }
  const int task = shaped_tile(wave, xcd);
  if (task >= ntiles * gs.H) return false;
  tile = task / gs.H;
  row0 = (task - tile * gs.H) * gs.T;
  g = lane / gs.T;
  i = lane - g * gs.T;
  return true;
}

// A grouped wave stages a chunk of up to 64 records in its LDS slice (one
// coalesced record per lane) before reading them per lane; LDS operations of
// one wave complete in order, so the fences only keep the compiler from moving
// the accesses across each other.
template <typename R>
__device__ inline void stage_records(R* dst, const R* __restrict__ src, int cn) {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  const int lane = threadIdx.x & 63;
  if (lane < cn) dst[lane] = src[lane];
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// Sum of a lane's value over the G copies of its row, in copy order (valid in
// copy 0's lanes; every lane must call it).
__device__ inline double group_sum(double v, RowGroups gs, int i) {
  double s = v;
  for (int q = 1; q < gs.G; ++q) s += __shfl(v, i + q * gs.T, 64);
  return s;
}

// 1/d in FP64: hardware reciprocal estimate plus two Newton-Raphson steps
// (each step squares the relative error; the estimate is good to ~2^-26), so
// the result is within an ulp of the correctly rounded quotient.
__device__ inline double rcp64(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}

// 1/d in FP64 with one Newton-Raphson step on the hardware estimate (relative
// error about the square of the estimate's, i.e. an ulp or two), for the pair
// kernels where d = (1 - x.y) + eps >= eps.
__device__ inline double rcp64_pair(double d) {
  const double r = __builtin_amdgcn_rcp(d);
  return fma(r, fma(-d, r, 1.0), r);
}

// interp_funcs_impl.hpp:114-142: leaf containing (x,y,z) or -1; xe gets the
// label coordinates on the located face (locate_uniform: no panel reads).
__device__ inline int locate_leaf(const DevTree& t, double x, double y, double z, double* xe) {
  return locate_uniform(t.levels, x, y, z, xe);
}

// Shallow-water physics of a context: topography kind (0 none, 1 the reference's
// cone, 2 Gaussian), the Gaussian mountain (height, width, unit centre), the
// coefficient of the free-surface Laplacian in the divergence tendency, and
// whether the height correction subtracts the free surface.
struct SwePhys {
  int topo;
  int surface_correction;
  double gravity;
  double mtn_h, mtn_w, mc[3];
};
// Topography height at a unit vector (0 for topo 0) and, when grad is given, its
// gradient in the point (unit-sphere components; the caller projects).
__host__ __device__ inline double topo_height(const SwePhys& ph, double a, double b, double c, double* grad = nullptr) {
  if (grad) grad[0] = grad[1] = grad[2] = 0.0;
  if (ph.topo == 2) {
    // r = atan2(|p x c|, p.c), h = H exp(-(r/w)^2)
    const double cx = b * ph.mc[2] - c * ph.mc[1], cy = c * ph.mc[0] - a * ph.mc[2], cz = a * ph.mc[1] - b * ph.mc[0];
    const double sn = sqrt(cx * cx + cy * cy + cz * cz), cs = a * ph.mc[0] + b * ph.mc[1] + c * ph.mc[2];
    const double r = atan2(sn, cs), q = r / ph.mtn_w, e = ph.mtn_h * exp(-q * q);
    if (grad && sn > 0.0) {
      // dr/dp = (cs (p - c cs) / sn - sn c) / (sn^2 + cs^2)
      const double k = -2.0 * e * r / (ph.mtn_w * ph.mtn_w) / (sn * sn + cs * cs);
      const double p[3] = {a, b, c};
      for (int i = 0; i < 3; ++i) grad[i] = k * (cs * (p[i] - ph.mc[i] * cs) / sn - sn * ph.mc[i]);
    }
    return e;
  }
  if (ph.topo == 1) {
    // cone_mountain (topography.cpp:6-45): h0 exp(-16 r^2), r^2 = (lon - pi/2)^2 + (lat - pi/6)^2
    const double pi = 3.14159265358979323846;
    const double rho2 = a * a + b * b, rho = sqrt(rho2);
    const double lat = pi / 2.0 - atan2(rho, c);
    double lon = atan2(b, a);
    lon += 2 * pi;
    lon = fmod(lon, 2 * pi);
    const double dl = lon - 0.5 * pi, dtt = lat - pi / 6.0;
    const double e = (2000.0 / 6371000.0) * exp(-16.0 * (dl * dl + dtt * dtt));
    if (grad && rho2 > 0.0) {
      // d lon = (-b, a, 0) / rho^2;  d lat = -(c d rho - rho d c) / (rho^2 + c^2), d rho = (a, b, 0) / rho
      const double n2 = rho2 + c * c, k = -32.0 * e;
      const double glon[3] = {-b / rho2, a / rho2, 0.0};
      const double glat[3] = {-c * a / (rho * n2), -c * b / (rho * n2), rho / n2};
      for (int i = 0; i < 3; ++i) grad[i] = k * (dl * glon[i] + dtt * glat[i]);
    }
    return e;
  }
  return 0.0;
}
// Source record of the FP64 pair kernels: displaced, re-projected source
// position and its charge pre-multiplied by -1/(4 pi).
struct SrcRec64 {
  double x, y, z, q;
};

// Source record of the FP64 shallow-water pair kernels: displaced, re-projected
// position with its vorticity charge x -1/(4 pi), the undisplaced node and the
// divergence charge x -+1/(4 pi) (variant 1 / 2).
struct SrcRecSwe {
  double x, y, z, qz, x0, y0, z0, qd;
};

// Source record of the mixed-precision pair kernels: position relative to the
// source panel centre and the scaled charge, FP32.
struct SrcRec32 {
  float x, y, z, q;
};

// Source record of the mixed-precision shallow-water pair kernel: position
// relative to the source panel centre and the two scaled charges, FP32
// (padded to 32 bytes for whole scalar loads).
struct SrcRecSwe32 {
  float x, y, z, qz, qd, pad0, pad1, pad2;
};

// Source records of the MFMA pair kernel, in tiles of 16 sources (a source
// set is padded to whole tiles with zero records).  A tile is 128 floats:
//   [k*16 + i]                    k = 0..3: (b_x, b_y, b_z, beta) of source i,
//   [64 + ((i/4)*4 + c)*4 + i%4]  c = 0..3: (q b_x, q b_y, q b_z, q) of source i,
// with b the position relative to the source-panel centre, beta = |b|^2/2 and
// q = -charge/(4 pi).  Lane l of a wave then reads its 16x16x4 A operand at
// [l] and its four 4x4x1 A operands as one float4 at 64 + 4*((l/16)*4 + l%4).
struct SrcRecMfma {
  float v[8];
};

struct LaunchCfg {
  hipStream_t stream;
};

// ---- passes (fmm_passes.hip)
void launch_upward_pass(const DevTree& t, const double* area, const double* pots, double* q_part,
                        double* q_proxy, hipStream_t s);
// assign_leaf_rows: the last L2L level overwrites the leaf proxy rows instead of
// adding to them (valid when nothing else writes those rows).
// L2L over every level (no L2P); the last level assigns the leaf rows when
// assign_leaf_rows.
void launch_l2l(const DevTree& t, int ncomp, double* const* phi_proxy, hipStream_t s, bool assign_leaf_rows,
                int nmem = 1);
// Upward pass with base_pots + P2M + first M2M fused; the leaf rows of q_proxy
// are not written (they are never interaction sources).
void launch_upward_fused(const DevTree& t, const double* area, const double* pots, double* q_part, double* q_proxy,
                         hipStream_t s, int nmem = 1);
// launch_upward_fused for several charges over the same points (e.g. the
// shallow-water vorticity and divergence, and the adjoint's cotangent
// components), one launch per level for up to kMaxUpSets of them.
constexpr int kMaxUpSets = 5;
struct UpCharge {
  const double* area;
  const double* pots;
  double* q_part;
  double* q_proxy;
};
void launch_upward_sets(const DevTree& t, const UpCharge* sets, int nsets, hipStream_t s);
void launch_downward_pass(const DevTree& t, int ncomp, double* const* phi_proxy, const double* const* phi_part,
                          double* const* sols, hipStream_t s, bool assign_leaf_rows = false, int nmem = 1);
void launch_fill_zero(double* p, int64_t count, hipStream_t s);
void launch_fill_value(double* p, int64_t count, double v, hipStream_t s);
// Zeroes rows[0..nrows) (panel ids) of ncomp [P][n] arrays.
void launch_zero_rows(const int* rows, int nrows, int n, int ncomp, double* const* arrays, hipStream_t s,
                      int nmem = 1, int64_t member_stride = 0);

// ---- interactions (fmm_interact.hip)
void launch_source_records64(const DevTree& t, const double* q_part, const double* q_proxy,
                             const double* const* disp_part, const double* const* disp_proxy, bool project,
                             SrcRec64* rec, hipStream_t s, int nmem = 1);
// The pair kernels accumulate into their outputs like the reference functors;
// with assign = true they overwrite them (each target written by one lane).
void launch_bve_interact64(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* sets,
                           const SrcRec64* rec, double eps, double* const* out_part, double* const* out_proxy,
                           hipStream_t s, bool assign = false, int nmem = 1, PairShape shape = {});
void launch_poisson_interact64(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* sets,
                               const SrcRec64* rec, double* out_part, double* out_proxy, hipStream_t s,
                               bool assign = false);
void launch_source_records32(const DevTree& t, const double* q_part, const double* q_proxy,
                             const double* const* disp_part, const double* const* disp_proxy, SrcRec32* rec,
                             hipStream_t s, int nmem = 1);
void launch_bve_interact_mixed(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* sets,
                               const SrcRec32* rec, double eps, double* const* out_part,
                               double* const* out_proxy, hipStream_t s, bool assign = false, int nmem = 1,
                               PairShape shape = {});
void launch_poisson_interact_mixed(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* sets,
                                   const SrcRec32* rec, double* out_part, double* out_proxy, hipStream_t s,
                                   bool assign = false);
void launch_source_records_mfma(const DevTree& t, const double* q_part, const double* q_proxy,
                                const double* const* disp_part, const double* const* disp_proxy, SrcRecMfma* rec,
                                hipStream_t s, int nmem = 1);
void launch_bve_interact_mfma(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* sets,
                              const SrcRecMfma* rec, double eps, double* const* out_part, double* const* out_proxy,
                              hipStream_t s, bool assign = false, int nmem = 1, PairShape shape = {});

// Multi-right-hand-side MFMA pair kernel for nmem members whose sources and
// targets share positions: one kernel matrix per wave, applied to the charge
// columns of 4 * cb members (cb = 1 or 2 chunks of four per wave).  The columns
// hold mrhs_chunks(nmem, cb) chunks of four members (the last chunk group
// zero-padded), 16 floats per record slot and chunk:
// U16 * 16 * mrhs_chunks(nmem, cb) floats, at most U16 * 16 * mrhs_chunks(nmem, 2).
// geo is member 0's SrcRecMfma array (positions only are read).
inline int mrhs_chunks_per_wave(int nmem) { return nmem > 4 ? 2 : 1; }
inline int mrhs_chunks(int nmem, int cb) { return (nmem + 4 * cb - 1) / (4 * cb) * cb; }
void launch_source_columns_mrhs(const DevTree& t, const double* q_part, const double* q_proxy,
                                const SrcRecMfma* geo, int nmem, int cb, float* qcol, hipStream_t s);
void launch_bve_interact_mrhs(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* sets,
                              const SrcRecMfma* geo, const float* qcol, int nmem, int cb, double eps,
                              double* const* out_part, double* const* out_proxy, hipStream_t s, bool assign,
                              PairShape shape);

// ---- shallow-water pair sums (swe.hip); variant 1 or 2, 3 = the lat-lon form
// (swe_vel_impl.hpp:265-276: variant 1's kernel with the divergence term's sign reversed)
void launch_source_records_swe(const DevTree& t, int variant, const double* qz_part, const double* qz_proxy,
                               const double* qd_part, const double* qd_proxy, const double* const* disp_part,
                               const double* const* disp_proxy, SrcRecSwe* rec, hipStream_t s);
void launch_swe_interact64(const DevTree& t, int variant, const TargetTile* tiles, int ntiles, const SourceSet* sets,
                           const SrcRecSwe* rec, double eps, double* const* out_part, double* const* out_proxy,
                           hipStream_t s, bool assign = false, bool shared_tiles = false);

// ---- adjoints (adjoint.hip: FMM transpose; remap_adjoint.hip: remap side)
size_t adj_target_record_bytes();
// Target records (x_t, psi_t cross x_t) over the unified index space.
void launch_adj_target_records(const DevTree& t, const double* const* psi_part, const double* const* psi_prox,
                               void* rec, hipStream_t s);
// Transposed pair sums over the source-major list: assigns qbar (dL/dq, every
// source row) and, when dpart is given, dbar (dL/d displacement) of every
// source row.
void launch_bve_adjoint64(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* targets,
                          const void* trec, const SrcRec64* srec, double eps, const double* const* dpart,
                          const double* const* dprox, double* qbar_part, double* qbar_prox, double* const* dbar_part,
                          double* const* dbar_prox, hipStream_t s, PairShape shape = {});
// FP32 variants for mixed-precision contexts (panel-centred |x - y|^2 / 2 form).
void launch_adj_target_records32(const DevTree& t, const double* const* psi_part, const double* const* psi_prox,
                                 void* rec, hipStream_t s);
void launch_bve_adjoint32(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* targets,
                          const void* trec, const SrcRec64* srec, double eps, const double* const* dpart,
                          const double* const* dprox, double* qbar_part, double* qbar_prox, double* const* dbar_part,
                          double* const* dbar_prox, hipStream_t s, PairShape shape = {});
void launch_remap_adjoint(const DevTree& t, const double* dx, const double* dy, const double* dz, const double* newbar,
                          double* wbar, hipStream_t s);
// field_bar [Pl][m2] += B^T out_bar, B the BLI interpolation at n points.
void launch_interp_points_adjoint(const DevTree& t, int64_t n, const double* x, const double* y, const double* z,
                                  const double* out_bar, double* field_bar, hipStream_t s);
void launch_departure_adjoint(const DevTree& t, const double* x, const double* y, const double* z,
                              const double* const* vels /*12*/, const double* vor, const double* new_vor_bar,
                              double dt, int combine, double omega, double* const* vel_bar /*12*/, double* vor_bar,
                              hipStream_t s);
void launch_disp_interp_adjoint(const DevTree& t, const double* const* dbar_proxy, double* const* dbar_part,
                                hipStream_t s);
// Reverse of the departure search and the three-field remap of the shallow-water step
// (vel_bar as launch_departure_adjoint; field_bar[3] accumulated).
void launch_departure_adjoint3(const DevTree& t, const double* x, const double* y, const double* z,
                               const double* const* vels /*12*/, const double* const* fields /*3*/,
                               const double* const* new_bar /*3*/, double dt, int combine,
                               double* const* vel_bar /*12*/, double* const* field_bar /*3*/, hipStream_t s);

// ---- shallow-water step adjoint (swe_adjoint.hip, swe_step.hip)
size_t swe_adj_target_record_bytes();
void launch_swe_adj_target_records(const DevTree& t, bool fp32, const double* const* psi_part,
                                   const double* const* psi_prox, void* rec, hipStream_t s);
// Transposed shallow-water pair sums (variant 1): charge cotangents of the vorticity and
// divergence charges assigned at every source row; with dpart the displacement cotangents
// assigned (or added with accumulate) as well.
void launch_swe_adjoint(const DevTree& t, bool fp32, const TargetTile* tiles, int ntiles, const SourceSet* targets,
                        const void* trec, const SrcRecSwe* srec, double eps, const double* const* dpart,
                        const double* const* dprox, double* qz_part, double* qz_prox, double* qd_part,
                        double* qd_prox, double* const* dbar_part, double* const* dbar_prox, bool accumulate,
                        hipStream_t s, PairShape shape = {});
void launch_lap_adj_target_records(const DevTree& t, bool fp32, const double* psi_part, const double* psi_prox,
                                   void* rec, hipStream_t s);
// Transposed regularised-Laplacian pair sums, likewise.
void launch_laplacian_adjoint(const DevTree& t, bool fp32, const TargetTile* tiles, int ntiles,
                              const SourceSet* targets, const void* trec, const SrcRec64* srec,
                              const double* const* dpart, const double* const* dprox, double* q_part, double* q_prox,
                              double* const* dbar_part, double* const* dbar_prox, bool accumulate, hipStream_t s,
                              PairShape shape = {});
void launch_swe_tendency_adjoint(const DevTree& t, const double* x, const double* y, const double* z,
                                 const double* const* vel, const double* vor, const double* div, const double* h,
                                 const double* const* tbar, double omega, int rule, double gravity,
                                 double* const* vbar, double* lapbar, double* const* sbar, hipStream_t s);
void launch_panel_laplacian_adjoint(const DevTree& t, const double* lapbar, double* fbar, hipStream_t s);
// lap -= 8/0.1 (h + h_s(x)) at the nodes (h_s with the surface correction only).
void launch_node_correction(int64_t n, const double* x, const double* y, const double* z, const double* h,
                            const SwePhys& ph, double* lap, hipStream_t s);
// out = x + alpha y (out may alias x or y).
void launch_lincomb(int64_t n, const double* x, double alpha, const double* y, double* out, hipStream_t s);
// Small fused element-wise steps of the reverse passes: zero fill of n arrays
// in one launch; out = x + alpha y with zero[] cleared; a += src, b += beta
// src; out1 = f1 + alpha1 r, out2 = f2 + alpha2 r for three components.
void launch_fill_zero_sets(double* const* arrays, const int64_t* counts, int n, hipStream_t s);
void launch_lincomb_zero(int64_t n, const double* x, double alpha, const double* y, double* out, double* zero,
                         hipStream_t s);
void launch_add2(int64_t n, const double* src, double* a, double beta, double* b, hipStream_t s);
void launch_apply_two(int64_t n, const double* const* r, const double* const* f1, double alpha1, double* const* out1,
                      const double* const* f2, double alpha2, double* const* out2, hipStream_t s);
void launch_height_correction_adjoint(const DevTree& t, const double* x, const double* y, const double* z,
                                      const double* const* vel, const double* hh, double tau, const SwePhys& ph,
                                      const double* lapbar, double* hscat, double* const* vbar, hipStream_t s);
void launch_swe_stage_adjoint(const DevTree& t, const double* x, const double* y, const double* z,
                              const double* const* s0, const double* const* tprev, const double* area, double taua,
                              const double* qz, const double* qd, const double* qe, const double* const* dpart,
                              const double* const* dprox, int rule, const SwePhys& ph, double* const* sbar,
                              double* const* tbar, double* const* dbar_part, double* const* dbar_prox,
                              hipStream_t s);
void launch_stage_adjoint(const DevTree& t, const double* x, const double* y, const double* z,
                          const double* const* vel, double tau, double omega, const double* area, const double* qbar,
                          const double* const* dbar, double* vor_bar, double* const* vbar, hipStream_t s);

// ---- irregular particle tree (irreg.hip; M2M / L2L over panel lists in fmm_passes.hip)
// nmem proxy arrays, P*n apart
// The L2L sweep of a tree with leaves at several levels in one launch: the
// leaf rows (leaves[0 .. nleaves)) += each ancestor's row evaluated at the
// leaf's nodes (ancestor triples anc[anc_first[i] ..), composite matrices
// anc_tab); the non-leaf rows are left as they are.  False (nothing launched)
// for a degree without a compiled instance.
bool launch_leaf_rows(const DevTree& t, const int* leaves, int nleaves, const int* anc_first, const int* anc,
                      const double* anc_tab, int ncomp, double* const* phi_proxy, hipStream_t s);
void launch_m2m_list(const DevTree& t, const int* parents, int count, double* q_proxy, hipStream_t s, int nmem = 1);
void launch_l2l_list(const DevTree& t, const int* parents, int count, int ncomp, double* const* phi_proxy,
                     hipStream_t s);
void launch_gather_charges(int64_t n, const int* perm, const double* area, const double* pots, double* q,
                           hipStream_t s);
void launch_scatter(int64_t n, int ncomp, const int* perm, const double* const* in, double* const* out,
                    bool accumulate, hipStream_t s);
void launch_add(int64_t n, int ncomp, double* const* a, const double* const* b, hipStream_t s);
// w [npart][2 dp1]: each sorted point's 1-D BLI basis at xi, then at eta, on its leaf
void launch_point_weights(const DevTree& t, const int* point_leaf, double* w, hipStream_t s);
// nq (1 or 2) charge arrays q + j*qstride -> proxy rows proxy + j*pstride
void launch_p2m_ll(const DevTree& t, const int* leaves, int nleaves, const int* leaf_first, const int* leaf_count,
                   const double* w, int nq, const double* q, int64_t qstride, double* proxy, int64_t pstride,
                   hipStream_t s);
void launch_l2p_ll(const DevTree& t, const int* point_leaf, const double* w, int ncomp, const double* const* phi,
                   const double* const* part, double* const* out, hipStream_t s);

// ---- lat-lon shallow-water grid operators (swe_ll.hip); fields [nlat][nlon]
// row-major on the cell-centred grid, spacing dx degrees (dxr = dx*pi/180)
struct LlGrid {
  int nlat, nlon;
  double dx, dxr;
};
LlGrid ll_grid(int nlat, int nlon, double dx_deg);
// vor_t, div_t, h_t (h_t may be null) from the state and vel[3]
void launch_ll_tendencies(const LlGrid& g, double omega, const double* x, const double* y, const double* z,
                          const double* vor, const double* div, const double* h, const double* topo,
                          const double* const* vel, double* vor_t, double* div_t, double* h_t, hipStream_t s);
// means[4] = {south vor, south h, north vor, north h} of the first and last rows
void launch_ll_pole_means(const LlGrid& g, double dt, const double* vor, const double* div, const double* h,
                          const double* vor_t, double* means, hipStream_t s);
void launch_ll_arrival(const LlGrid& g, double dt, const double* x, const double* y, const double* z,
                       const double* vor, const double* div, const double* h, const double* const* vel,
                       const double* vor_t, const double* div_t, const double* means, double* a_vor, double* a_div,
                       double* a_h, hipStream_t s);
void launch_ll_average(const LlGrid& g, const double* x, const double* y, const double* z, const double* f,
                       double* out, double eps, hipStream_t s);
void launch_ll_diffuse(const LlGrid& g, double dt, const double* x, const double* y, const double* z,
                       const double* f, double* out, hipStream_t s);
// div = diffuse(average(a_div)), vor = a_vor, h = a_h
void launch_ll_smooth(const LlGrid& g, double dt, const double* x, const double* y, const double* z,
                      const double* a_vor, const double* a_div, const double* a_h, double eps, double* vor,
                      double* div, double* h, hipStream_t s);
void launch_ll_polar_filter(const LlGrid& g, const double* x, const double* y, const double* z, double* f,
                            hipStream_t s);

// ---- output operators (output.hip)
int max_interp_fields();
// out[q][i] = BLI of leaf field q [Pl][m2] at point i (n points, nf <= max_interp_fields()).
void launch_interp_points(const DevTree& t, int64_t n, const double* px, const double* py, const double* pz, int nf,
                          const double* const* fields, double* const* out, hipStream_t s);
void launch_xyz_vel_to_uv(int64_t n, const double* x, const double* y, const double* z, const double* vx,
                          const double* vy, const double* vz, double* u, double* v, hipStream_t s);

// ---- evaluation at arbitrary targets (targets.hip)
// kind 0: log kernel (psi), kind 1: Biot-Savart (3 components).  far[f] are the
// [P][n] proxy potentials after the L2L sweep (leaf rows = far field of each
// leaf); set_start[P+1] is the per-target-panel CSR into sets.
void launch_eval_targets(const DevTree& t, int kind, const int* set_start, const SourceSet* sets,
                         const SrcRec64* rec, const double* const* far, double eps, int64_t nt, const double* tx,
                         const double* ty, const double* tz, double* const* out, hipStream_t s);

// ---- displacement, departure points, remap (remap.hip)
void launch_bve_displacement(const DevTree& t, const double* x, const double* y, const double* z,
                             const double* const* vel, const double* vor, double tau, double omega,
                             double* new_vor, double* const* disp_part, double* const* disp_proxy, hipStream_t s,
                             int nmem = 1);
void launch_bve_departure(const DevTree& t, const double* x, const double* y, const double* z,
                          const double* const* vels /*12*/, double dt, int combine, double* dx, double* dy,
                          double* dz, hipStream_t s, int nmem = 1);
void launch_bve_departure_to_target(const DevTree& t, const double* z, const double* dx, const double* dy,
                                    const double* dz, const double* vor, const double* tracers, int ntr,
                                    double omega, double* new_vor, double* new_tracers, hipStream_t s, int nmem = 1);
// swe_departure_to_target: fields = {vor, div, h} [Pl][m2], no Coriolis terms.
void launch_swe_departure_to_target(const DevTree& t, const double* dx, const double* dy, const double* dz,
                                    const double* const* fields, const double* tracers, int ntr,
                                    double* const* new_fields, double* new_tracers, hipStream_t s);
void launch_copy(double* dst, const double* src, int64_t count, hipStream_t s);

// ---- shallow-water step (swe_step.hip, fmm_interact.hip)
void launch_laplacian_interact64(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* sets,
                                 const SrcRec64* rec, double* out_part, double* out_proxy, hipStream_t s,
                                 bool assign = false);
// Mixed precision (FP32 pairs, FP64 per-set accumulation) forms of the
// shallow-water (variants 1 and 3) and regularised-Laplacian pair sums.
void launch_source_records_swe32(const DevTree& t, int variant, const double* qz_part, const double* qz_proxy,
                                 const double* qd_part, const double* qd_proxy, const double* const* disp_part,
                                 const double* const* disp_proxy, SrcRecSwe32* rec, hipStream_t s);
void launch_swe_interact_mixed(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* sets,
                               const SrcRecSwe32* rec, double eps, double* const* out_part,
                               double* const* out_proxy, hipStream_t s, bool assign);
void launch_laplacian_interact_mixed(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* sets,
                                     const SrcRec32* rec, double* out_part, double* out_proxy, hipStream_t s,
                                     bool assign = false);
void launch_panel_laplacian(const DevTree& t, const double* f, double* lap, hipStream_t s);
void launch_swe_tendency(const DevTree& t, const double* x, const double* y, const double* z,
                         const double* const* vel, const double* vor, const double* div, const double* h,
                         const double* lap, double omega, int rule, double gravity, double* const* out,
                         hipStream_t s);
void launch_topography(const DevTree& t, const double* x, const double* y, const double* z,
                       const double* const* disp_part, const double* const* disp_proxy, const double* h, int rule,
                       const SwePhys& ph, double* eff, hipStream_t s);
void launch_apply_tendencies(int64_t n, const double* const* f, const double* const* ft, const double* area,
                             double tau, double* const* out, double* new_area, hipStream_t s);
void launch_height_correction(const DevTree& t, const double* x, const double* y, const double* z,
                              const double* const* vel, const double* h, const double* ht, double tau,
                              const SwePhys& ph, double* lap, hipStream_t s);
// Mean over the copies of every unique node; sum1d [n1d] is scratch.
void launch_unify_boundary_vals(int64_t n, const int* map, const int* counts, int64_t n1d, double* sum1d,
                                double* vals, hipStream_t s);

// ---- direct sums (direct.hip)
int64_t direct_partials_size(int64_t nt, int64_t ns, int ncomp);
// Doubles in the fixed device scratch of the allocation-free direct sum (8 MiB).
constexpr int64_t kNoAllocScratchDoubles = int64_t(1) << 20;
// Looks up the fixed scratch's address on `device` (which must be current);
// 0 on success, else the HIP error code.  Idempotent, thread-safe.
int direct_noalloc_init(int device);
// The address direct_noalloc_init found for `device`, or null before it ran.
double* direct_noalloc_scratch(int device);
void launch_direct_inv_lap(int64_t nt, const double* tx, const double* ty, const double* tz, int64_t ns,
                           const double* sx, const double* sy, const double* sz, const double* area,
                           const double* pots, int64_t self_offset, bool guard_coincident, double* partials,
                           double* soln, hipStream_t s);
void launch_direct_bve_vel(int64_t nt, const double* tx, const double* ty, const double* tz, int64_t ns,
                           const double* sx, const double* sy, const double* sz, const double* area,
                           const double* vor, const double* const* disp, double eps, double* partials,
                           double* const* vel, hipStream_t s);

}  // namespace sbslpm
