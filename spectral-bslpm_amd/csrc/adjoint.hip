// adjoint.hip -- adjoints (transposes) of the BVE velocity evaluation and of
// the semi-Lagrangian remap, FP64.
//
// Velocity.  With undisplaced sources the Biot-Savart sum is
//   u_t = x_t cross sum_s g(x_t, x_s) q_s x_s,  q_s = -area_s w_s / (4 pi),
//   g(x, y) = 1 / (1 - x.y + eps)                     (bve_vel_impl.hpp:90-103)
// so for a given cotangent ubar the transpose is
//   wbar_s = -(area_s / (4 pi)) x_s . V_s,  V_s = sum_t g(x_s, x_t) c_t,
//   c_t = ubar_t cross x_t,
// because g is symmetric.  V is a sum of the same kernel over three scalar
// charge fields, evaluated with the same tree, interaction list and passes:
// anterpolation (P2M/M2M) is the transpose of the interpolation (L2L/L2P) and
// the proxy-proxy kernel is symmetric, so the FMM evaluation of V is the
// transpose of the FMM velocity operator up to the FMM's own truncation
// (the interaction list is not symmetric under exchanging targets and sources).
//
// Remap.  new_w(x_i) = sum_k B_ik w_k (+ Coriolis terms that do not depend on
// w), B_ik the BLI weights of node k of the leaf that contains departure
// point i (bve_time_step.cpp:38-66).  The transpose scatters
// wbar_k += B_ik newbar_i with FP64 atomic adds.
#include "device.hpp"

namespace sbslpm {
namespace {

struct CPtr3 {
  const double* p[3];
};
struct Ptr3 {
  double* p[3];
};

struct SrcRecVec {
  double x, y, z, pad0, c0, c1, c2, pad1;
};

// c = ubar cross x on the particles.
__global__ void k_cross_charges(DevTree t, CPtr3 ub, Ptr3 c) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= t.npart) return;
  const double x = t.px[u], y = t.py[u], z = t.pz[u];
  const double a = ub.p[0][u], b = ub.p[1][u], d = ub.p[2][u];
  c.p[0][u] = b * z - d * y;
  c.p[1][u] = d * x - a * z;
  c.p[2][u] = a * y - b * x;
}

// Records: re-projected node and its three charges (particles, then the
// proxies of non-leaf panels).
__global__ void k_src_vec(DevTree t, CPtr3 cpart, CPtr3 cprox, SrcRecVec* __restrict__ rec) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= t.U) return;
  double x = t.px[u], y = t.py[u], z = t.pz[u];
  double c[3];
  if (u < t.npart) {
    for (int f = 0; f < 3; ++f) c[f] = cpart.p[f][u];
  } else {
    const int64_t v = u - t.npart;
    if (v / t.n >= t.offset) return;
    for (int f = 0; f < 3; ++f) c[f] = cprox.p[f][v];
  }
  project_to_sphere(x, y, z);
  rec[u] = SrcRecVec{x, y, z, 0.0, c[0], c[1], c[2], 0.0};
}

// V_t = sum_s g(x_t, y_s) c_s, assigned (every target row is owned by one lane).
__global__ __launch_bounds__(256) void k_gvec64(DevTree t, const TargetTile* __restrict__ tiles, int ntiles,
                                                const SourceSet* __restrict__ sets,
                                                const SrcRecVec* __restrict__ rec, double eps, Ptr3 part,
                                                Ptr3 proxy) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int tile = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + wave;
  if (tile >= ntiles) return;
  const TargetTile T = tiles[tile];
  const int64_t u = (int64_t)T.first + (lane < T.count ? lane : T.count - 1);
  const double x = t.px[u], y = t.py[u], z = t.pz[u];
  double w0 = 0, w1 = 0, w2 = 0;
  for (int e = T.src_begin; e < T.src_end; ++e) {
    const SourceSet S = sets[e];
    const SrcRecVec* __restrict__ r = rec + S.first;
#pragma unroll 2
    for (int k = 0; k < S.count; ++k) {
      const SrcRecVec s = r[k];
      const double g = rcp64_pair((1.0 - (x * s.x + y * s.y + z * s.z)) + eps);
      w0 = fma(g, s.c0, w0);
      w1 = fma(g, s.c1, w1);
      w2 = fma(g, s.c2, w2);
    }
  }
  if (lane >= T.count) return;
  if (u < t.npart) {
    part.p[0][u] = w0; part.p[1][u] = w1; part.p[2][u] = w2;
  } else {
    const int64_t v = u - t.npart;
    proxy.p[0][v] = w0; proxy.p[1][v] = w1; proxy.p[2][v] = w2;
  }
}

// wbar = -(area / (4 pi)) x . V
__global__ void k_adj_finish(DevTree t, const double* __restrict__ area, CPtr3 V, double* __restrict__ wbar) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= t.npart) return;
  wbar[u] = kNegInv4Pi * area[u] * (t.px[u] * V.p[0][u] + t.py[u] * V.p[1][u] + t.pz[u] * V.p[2][u]);
}

// Transpose of the remap interpolation: wbar[leaf node] += B newbar.
template <int M>
__global__ __launch_bounds__(128) void k_remap_adj(DevTree t, const double* __restrict__ dx,
                                                   const double* __restrict__ dy, const double* __restrict__ dz,
                                                   const double* __restrict__ newbar, double* __restrict__ wbar) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= t.npart) return;
  double xe[2];
  const int leaf = locate_leaf(t, dx[u], dy[u], dz[u], xe);
  if (leaf < 0) {
    atomicOr(t.err, kErrLocate);
    return;
  }
  const DevPanel& P = t.panels[leaf];
  double wx[M], we[M];
  bool ok = bli_weights_1d(xe[0], P.min_xi, P.max_xi, M - 1, t.cheb_m, wx);
  ok = bli_weights_1d(xe[1], P.min_eta, P.max_eta, M - 1, t.cheb_m, we) && ok;
  if (!ok) atomicOr(t.err, kErrBliBounds);
  const double g = newbar[u];
  double* dst = wbar + (int64_t)(leaf - t.offset) * (M * M);
#pragma unroll
  for (int a = 0; a < M; ++a)
#pragma unroll
    for (int c = 0; c < M; ++c) atomicAdd(dst + a * M + c, wx[a] * we[c] * g);
}

template <int M>
void run_remap_adj(const DevTree& t, const double* dx, const double* dy, const double* dz, const double* newbar,
                   double* wbar, hipStream_t s) {
  hipLaunchKernelGGL(k_remap_adj<M>, dim3((unsigned)((t.npart + 127) / 128)), dim3(128), 0, s, t, dx, dy, dz, newbar,
                     wbar);
}

}  // namespace

void launch_cross_charges(const DevTree& t, const double* const* ubar, double* const* c, hipStream_t s) {
  CPtr3 u{{ubar[0], ubar[1], ubar[2]}};
  Ptr3 o{{c[0], c[1], c[2]}};
  hipLaunchKernelGGL(k_cross_charges, dim3((unsigned)((t.npart + 255) / 256)), dim3(256), 0, s, t, u, o);
}

void launch_gvec_interact64(const DevTree& t, const TargetTile* tiles, int ntiles, const SourceSet* sets,
                            const double* const* c_part, const double* const* c_proxy, void* rec, double eps,
                            double* const* out_part, double* const* out_proxy, hipStream_t s) {
  CPtr3 cp{{c_part[0], c_part[1], c_part[2]}}, cx{{c_proxy[0], c_proxy[1], c_proxy[2]}};
  SrcRecVec* r = static_cast<SrcRecVec*>(rec);
  hipLaunchKernelGGL(k_src_vec, dim3((unsigned)((t.U + 255) / 256)), dim3(256), 0, s, t, cp, cx, r);
  if (ntiles == 0) return;
  Ptr3 op{{out_part[0], out_part[1], out_part[2]}}, ox{{out_proxy[0], out_proxy[1], out_proxy[2]}};
  hipLaunchKernelGGL(k_gvec64, dim3((unsigned)((ntiles + kWavesPerBlock - 1) / kWavesPerBlock)),
                     dim3(64 * kWavesPerBlock), 0, s, t, tiles, ntiles, sets, (const SrcRecVec*)r, eps, op, ox);
}

size_t gvec_record_bytes() { return sizeof(SrcRecVec); }

void launch_adjoint_finish(const DevTree& t, const double* area, const double* const* V, double* wbar,
                           hipStream_t s) {
  CPtr3 v{{V[0], V[1], V[2]}};
  hipLaunchKernelGGL(k_adj_finish, dim3((unsigned)((t.npart + 255) / 256)), dim3(256), 0, s, t, area, v, wbar);
}

void launch_remap_adjoint(const DevTree& t, const double* dx, const double* dy, const double* dz, const double* newbar,
                          double* wbar, hipStream_t s) {
  switch (t.m) {
    case 2: run_remap_adj<2>(t, dx, dy, dz, newbar, wbar, s); break;
    case 3: run_remap_adj<3>(t, dx, dy, dz, newbar, wbar, s); break;
    case 4: run_remap_adj<4>(t, dx, dy, dz, newbar, wbar, s); break;
    case 5: run_remap_adj<5>(t, dx, dy, dz, newbar, wbar, s); break;
    case 6: run_remap_adj<6>(t, dx, dy, dz, newbar, wbar, s); break;
    case 7: run_remap_adj<7>(t, dx, dy, dz, newbar, wbar, s); break;
    case 8: run_remap_adj<8>(t, dx, dy, dz, newbar, wbar, s); break;
    case 9: run_remap_adj<9>(t, dx, dy, dz, newbar, wbar, s); break;
    case 10: run_remap_adj<10>(t, dx, dy, dz, newbar, wbar, s); break;
    default: run_remap_adj<11>(t, dx, dy, dz, newbar, wbar, s); break;
  }
}

}  // namespace sbslpm
