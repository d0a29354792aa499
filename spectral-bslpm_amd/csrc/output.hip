// output.hip -- output-side operators on the device:
//   * BLI interpolation of leaf fields [Pl][m2] at arbitrary points, the
//     functor of latlon_interp_host (interp_funcs.hpp:22-60) for any point
//     set (lat-lon grids, observation points);
//   * Cartesian to zonal/meridional velocity, xyz_vel_to_uv_vel
//     (general_utils_impl.hpp:195-229).
#include "bli.hpp"

namespace sbslpm {
namespace {

using namespace bli;

constexpr int kMaxOutFields = 8;

struct OutFields {
  const double* in[kMaxOutFields];
  double* out[kMaxOutFields];
};

template <int M>
__global__ __launch_bounds__(128) void k_interp_points(DevTree t, int64_t n, const double* __restrict__ px,
                                                       const double* __restrict__ py, const double* __restrict__ pz,
                                                       int nf, OutFields f) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  double xe[2];
  const int leaf = locate_leaf(t, px[u], py[u], pz[u], xe);
  if (leaf < 0) {
    flag(t, kErrLocate);
    for (int q = 0; q < nf; ++q) f.out[q][u] = 0;
    return;
  }
  const DevPanel& P = t.panels[leaf];
  double wx[M], we[M];
  bool ok = bli_w<M>(xe[0], P.min_xi, P.max_xi, t.cheb_m, wx);
  ok = bli_w<M>(xe[1], P.min_eta, P.max_eta, t.cheb_m, we) && ok;
  if (!ok) flag(t, kErrBliBounds);
  const int64_t base = (int64_t)(leaf - t.offset) * (M * M);
  for (int q = 0; q < nf; ++q) f.out[q][u] = tensor_gather<M>(f.in[q] + base, wx, we);
}

// general_utils_impl.hpp:211-228: u = (-y vx + x vy)/rho, v = -((x vx + y vy) z - rho^2 vz)/rho,
// rho = sqrt(x^2 + y^2); zero within 1e-16 of the poles.
__global__ void k_uv(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                     const double* __restrict__ z, const double* __restrict__ vx, const double* __restrict__ vy,
                     const double* __restrict__ vz, double* __restrict__ u, double* __restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double X = x[i], Y = y[i], Z = z[i], a = vx[i], b = vy[i], c = vz[i];
  if (fabs(Z) < 1 - 1e-16) {
    const double r2 = X * X + Y * Y, r = sqrt(r2);
    u[i] = (-Y * a + X * b) / r;
    v[i] = -((X * a + Y * b) * Z - r2 * c) / r;
  } else {
    u[i] = 0;
    v[i] = 0;
  }
}

template <int M>
struct InterpPoints {
  static void run(const DevTree& t, int64_t n, const double* px, const double* py, const double* pz, int nf,
                  OutFields f, hipStream_t s) {
    hipLaunchKernelGGL(k_interp_points<M>, dim3(blocks_for(n, 128)), dim3(128), 0, s, t, n, px, py, pz, nf, f);
  }
};

}  // namespace

int max_interp_fields() { return kMaxOutFields; }

void launch_interp_points(const DevTree& t, int64_t n, const double* px, const double* py, const double* pz, int nf,
                          const double* const* fields, double* const* out, hipStream_t s) {
  if (n <= 0) return;
  OutFields f{};
  for (int q = 0; q < nf; ++q) {
    f.in[q] = fields[q];
    f.out[q] = out[q];
  }
  dispatch_m<InterpPoints>(t.m, t, n, px, py, pz, nf, f, s);
}

void launch_xyz_vel_to_uv(int64_t n, const double* x, const double* y, const double* z, const double* vx,
                          const double* vy, const double* vz, double* u, double* v, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_uv, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, x, y, z, vx, vy, vz, u, v);
}

}  // namespace sbslpm
