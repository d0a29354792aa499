// direct.hip -- O(N^2) direct sums (direct_sum.cpp:8-15, direct_sum_impl.hpp:18-27)
// and the Biot-Savart direct sum the reference lacks (same pair kernel as
// bve_vel_impl.hpp:54-66).
//
// One wave64 per 64 targets and one source chunk per grid row; sources are
// read as wave-uniform loads.  Each (target tile, chunk) writes a partial sum
// and a second kernel adds the chunks in a fixed order, so the result is
// deterministic and needs no atomics (the reference does one atomic per pair).
#include <algorithm>
#include <mutex>

#include "device.hpp"

namespace sbslpm {
namespace {

constexpr int kWaves = 4;

struct Plan {
  int64_t tiles, chunks, chunk_len;
};

Plan plan(int64_t nt, int64_t ns) {
  Plan p;
  p.tiles = (nt + 63) / 64;
  int64_t want = (4096 + p.tiles - 1) / p.tiles;  // aim for >= 4096 waves
  int64_t maxc = (ns + 255) / 256;
  p.chunks = std::max<int64_t>(1, std::min(want, maxc));
  p.chunk_len = (ns + p.chunks - 1) / p.chunks;
  p.chunks = (ns + p.chunk_len - 1) / p.chunk_len;
  if (p.chunks < 1) p.chunks = 1;
  return p;
}

// guard 0: skip j == self_offset + i (the reference's i != j);
// guard 1: skip x.y >= 1 - 1e-15 (the FMM kernel's coincidence guard).
__global__ __launch_bounds__(256) void k_direct_log(int64_t nt, const double* __restrict__ tx,
                                                    const double* __restrict__ ty, const double* __restrict__ tz,
                                                    int64_t ns, const double* __restrict__ sx,
                                                    const double* __restrict__ sy, const double* __restrict__ sz,
                                                    const double* __restrict__ area, const double* __restrict__ pots,
                                                    int64_t self_offset, int guard, int64_t chunk_len,
                                                    double* __restrict__ partials) {
  // x.y rounded exactly as written in direct_sum_impl.hpp:23 (no fused
  // multiply-add): for nearly coincident points log(1 - x.y) amplifies a
  // one-ulp change of x.y by 1/(1 - x.y).
#pragma clang fp contract(off)
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int64_t tile = (int64_t)blockIdx.x * kWaves + wave;
  if (tile * 64 >= nt) return;
  const int64_t i = std::min<int64_t>(tile * 64 + lane, nt - 1);
  const int64_t self = self_offset + i;
  const double x = tx[i], y = ty[i], z = tz[i];
  const int64_t j0 = (int64_t)blockIdx.y * chunk_len, j1 = std::min<int64_t>(ns, j0 + chunk_len);
  double acc = 0;
  for (int64_t j = j0; j < j1; ++j) {
    const double dp = x * sx[j] + y * sy[j] + z * sz[j];
    const double l = log(fmax(1.0 - dp, 1e-300)) * (area[j] * pots[j]);
    const bool skip = guard ? !(dp < 1.0 - 1e-15) : (j == self);
    acc += skip ? 0.0 : l;
  }
  if (tile * 64 + lane < nt) partials[(int64_t)blockIdx.y * nt + i] = kNegInv4Pi * acc;
}

__global__ __launch_bounds__(256) void k_direct_bs(int64_t nt, const double* __restrict__ tx,
                                                   const double* __restrict__ ty, const double* __restrict__ tz,
                                                   int64_t ns, const double* __restrict__ sx,
                                                   const double* __restrict__ sy, const double* __restrict__ sz,
                                                   const double* __restrict__ area, const double* __restrict__ vor,
                                                   double eps, int64_t chunk_len, double* __restrict__ partials) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int64_t tile = (int64_t)blockIdx.x * kWaves + wave;
  if (tile * 64 >= nt) return;
  const int64_t i = std::min<int64_t>(tile * 64 + lane, nt - 1);
  const double x = tx[i], y = ty[i], z = tz[i];
  const int64_t j0 = (int64_t)blockIdx.y * chunk_len, j1 = std::min<int64_t>(ns, j0 + chunk_len);
  double w0 = 0, w1 = 0, w2 = 0;
  for (int64_t j = j0; j < j1; ++j) {
    const double a = sx[j], b = sy[j], c = sz[j];
    const double dp = x * a + y * b + z * c;
    const double g = rcp64((1.0 - dp) + eps) * (area[j] * vor[j]);
    w0 = fma(g, a, w0);
    w1 = fma(g, b, w1);
    w2 = fma(g, c, w2);
  }
  if (tile * 64 + lane < nt) {
    double* p = partials + (int64_t)blockIdx.y * 3 * nt;
    p[i] = kNegInv4Pi * (y * w2 - z * w1);
    p[nt + i] = kNegInv4Pi * (z * w0 - x * w2);
    p[2 * nt + i] = kNegInv4Pi * (x * w1 - y * w0);
  }
}

struct Ptr3 {
  double* p[3];
};

__global__ void k_reduce(int64_t nt, int ncomp, int64_t chunks, const double* __restrict__ partials, Ptr3 out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  for (int f = 0; f < ncomp; ++f) {
    double s = 0;
    for (int64_t c = 0; c < chunks; ++c) s += partials[(c * ncomp + f) * nt + i];
    out.p[f][i] += s;
  }
}

// Displaced, re-projected source positions (bve_vel_impl.hpp:57-60).
__global__ void k_displace(int64_t ns, const double* __restrict__ sx, const double* __restrict__ sy,
                           const double* __restrict__ sz, const double* __restrict__ dx,
                           const double* __restrict__ dy, const double* __restrict__ dz, double* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ns) return;
  double a = sx[j] + dx[j], b = sy[j] + dy[j], c = sz[j] + dz[j];
  project_to_sphere(a, b, c);
  out[j] = a; out[ns + j] = b; out[2 * ns + j] = c;
}

}  // namespace

// Fixed device scratch of the allocation-free direct sum
// (sbslpm_direct_sum_inv_lap_noalloc): the per-chunk partial sums of
// k_direct_log.  Each device holds its own copy until the process ends.  In
// host code `scratch` only names the symbol; its device address comes from
// hipGetSymbolAddress, once per device, in direct_noalloc_init.
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int direct_noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return -1;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return 0;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return (int)(e != hipSuccess ? e : hipErrorInvalidSymbol);
  g_scratch_ptr[device] = static_cast<double*>(p);
  return 0;
}

double* direct_noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

// Scratch doubles needed by a direct sum with nt targets, ns sources.
int64_t direct_partials_size(int64_t nt, int64_t ns, int ncomp) {
  Plan p = plan(nt, ns);
  return p.chunks * ncomp * nt + (ncomp == 3 ? 3 * ns : 0);
}

void launch_direct_inv_lap(int64_t nt, const double* tx, const double* ty, const double* tz, int64_t ns,
                           const double* sx, const double* sy, const double* sz, const double* area,
                           const double* pots, int64_t self_offset, bool guard_coincident, double* partials,
                           double* soln, hipStream_t s) {
  if (nt <= 0) return;
  Plan p = plan(nt, ns);
  dim3 grid((unsigned)((p.tiles + kWaves - 1) / kWaves), (unsigned)p.chunks);
  hipLaunchKernelGGL(k_direct_log, grid, dim3(64 * kWaves), 0, s, nt, tx, ty, tz, ns, sx, sy, sz, area, pots,
                     self_offset, guard_coincident ? 1 : 0, p.chunk_len, partials);
  Ptr3 out{{soln, nullptr, nullptr}};
  hipLaunchKernelGGL(k_reduce, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, nt, 1, p.chunks,
                     (const double*)partials, out);
}

void launch_direct_bve_vel(int64_t nt, const double* tx, const double* ty, const double* tz, int64_t ns,
                           const double* sx, const double* sy, const double* sz, const double* area,
                           const double* vor, const double* const* disp, double eps, double* partials,
                           double* const* vel, hipStream_t s) {
  if (nt <= 0) return;
  Plan p = plan(nt, ns);
  const double *ax = sx, *ay = sy, *az = sz;
  if (disp && disp[0]) {
    double* moved = partials + p.chunks * 3 * nt;
    hipLaunchKernelGGL(k_displace, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, s, ns, sx, sy, sz, disp[0],
                       disp[1], disp[2], moved);
    ax = moved; ay = moved + ns; az = moved + 2 * ns;
  }
  dim3 grid((unsigned)((p.tiles + kWaves - 1) / kWaves), (unsigned)p.chunks);
  hipLaunchKernelGGL(k_direct_bs, grid, dim3(64 * kWaves), 0, s, nt, tx, ty, tz, ns, ax, ay, az, area, vor, eps,
                     p.chunk_len, partials);
  Ptr3 out{{vel[0], vel[1], vel[2]}};
  hipLaunchKernelGGL(k_reduce, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, nt, 3, p.chunks,
                     (const double*)partials, out);
}

}  // namespace sbslpm
