// swe_ll.hip -- grid operators of the lat-lon shallow-water step
// (swe_time_step_ll.cpp): finite-difference tendencies, quadratic
// semi-Lagrangian arrival values, the divergence's polar averaging and
// Laplacian diffusion, and the polar Fourier filter.
//
// Fields are [lat][lon] row-major on the cell-centred grid of latlon_point
// (general_utils_impl.hpp:252-276).  Every operator is a stencil over at most
// the 3 x 3 neighbourhood of a point (the ring means of the first and last rows
// aside), so each runs one thread per point and is bound by HBM and launch
// latency; the FMM velocity in front of them dominates a step.  Multiply-adds
// are not contracted, so each value is rounded as the reference writes it.
#include "bli.hpp"

#pragma clang fp contract(off)

namespace sbslpm {
namespace {

constexpr int kLlBlock = 256;
constexpr int kPoleChunk = 512;  // ring values staged per wave in the pole-mean kernel
constexpr int kSpecMax = 29;     // polar filter: at most 29 Fourier terms (:615)

__device__ inline int64_t ll_at(const LlGrid& g, int i, int j) { return (int64_t)i * g.nlon + j; }
__device__ inline int ll_jp(const LlGrid& g, int j) { return j + 1 == g.nlon ? 0 : j + 1; }
__device__ inline int ll_jm(const LlGrid& g, int j) { return j == 0 ? g.nlon - 1 : j - 1; }

// xyz_to_latlon (general_utils_impl.hpp:232-236).
__device__ inline double ll_lat(double x, double y, double z) { return M_PI / 2.0 - atan2(sqrt(x * x + y * y), z); }

// xyz_vel_to_uv_vel's per-point body (general_utils_impl.hpp:214-226).
__device__ inline void ll_uv(double x, double y, double z, double vx, double vy, double vz, double& u, double& v) {
  if (fabs(z) < 1 - 1e-16) {
    const double r = sqrt(x * x + y * y);
    u = (-y * vx + x * vy) / r;
    v = -((x * vx + y * vy) * z - (x * x + y * y) * vz) / r;
  } else {
    u = 0;
    v = 0;
  }
}

// A grid index from a floored coordinate, kept inside [0, hi) (NaN -> 0); equal
// to the reference's wherever the reference's stays inside the grid.
__device__ inline int ll_index(double fl, int hi) {
  if (!(fl >= 0.0)) return 0;
  if (fl > double(hi - 1)) return hi - 1;
  return (int)fl;
}

struct LlIn {
  const double *x, *y, *z, *vor, *div, *h, *topo, *vx, *vy, *vz;
};

// swe_tendency_computation_ll (swe_time_step_ll.cpp:365-387) at one point:
// vor_tend_ll (:14-32), height_tend_ll (:34-95), div_tend_ll (:97-231), then
// filter_vals(div, 2e-12) (general_utils_impl.hpp:278-293).  u, v and |vel|^2
// of the neighbours are recomputed in place of the reference's arrays (same
// values).  h_t may be null (the step does not use the height tendency).
__global__ __launch_bounds__(kLlBlock) void k_ll_tend(LlGrid g, double omega, LlIn a, double* __restrict__ vor_t,
                                                      double* __restrict__ div_t, double* __restrict__ h_t) {
  const int64_t k = (int64_t)blockIdx.x * kLlBlock + threadIdx.x;
  if (k >= (int64_t)g.nlat * g.nlon) return;
  const int i = (int)(k / g.nlon), j = (int)(k - (int64_t)i * g.nlon);
  const int64_t kp = ll_at(g, i, ll_jp(g, j)), km = ll_at(g, i, ll_jm(g, j));
  auto uv = [&](int64_t q, double& u, double& v) { ll_uv(a.x[q], a.y[q], a.z[q], a.vx[q], a.vy[q], a.vz[q], u, v); };
  const double grav = 9.81 / 6371000.0;
  auto bern = [&](int64_t q) {
    const double dp = a.vx[q] * a.vx[q] + a.vy[q] * a.vy[q] + a.vz[q] * a.vz[q];
    return grav * (a.h[q] + a.topo[q]) + 0.5 * dp;
  };
  const double lat = ll_lat(a.x[k], a.y[k], a.z[k]);
  const double cl = cos(lat);

  const double avor = a.vor[k] + 2.0 * omega * a.z[k];
  vor_t[k] = -avor * a.div[k] - 2.0 * omega * a.vz[k];

  double up, vp, um, vm;
  uv(kp, up, vp);
  uv(km, um, vm);
  if (h_t) {
    const double d1 = (a.h[kp] * up - a.h[km] * um) / (2.0 * g.dx);
    double d2;
    if (i == 0) {
      const int64_t kn = ll_at(g, 1, j);
      double u0, v0, un, vn;
      uv(k, u0, v0);
      uv(kn, un, vn);
      d2 = (3.0 * a.h[k] * v0 * cl + a.h[kn] * vn * cos(lat + g.dxr)) / (3.0 * g.dxr);
    } else if (i == g.nlat - 1) {
      const int64_t ks = ll_at(g, i - 1, j);
      double u0, v0, us, vs;
      uv(k, u0, v0);
      uv(ks, us, vs);
      d2 = (-a.h[ks] * vs * cos(lat - g.dxr) - 3.0 * a.h[k] * v0 * cl) / (3.0 * g.dxr);
    } else {
      const int64_t kn = ll_at(g, i + 1, j), ks = ll_at(g, i - 1, j);
      double un, vn, us, vs;
      uv(kn, un, vn);
      uv(ks, us, vs);
      d2 = (a.h[kn] * vn * cos(lat + g.dxr) - a.h[ks] * vs * cos(lat - g.dxr)) / (2.0 * g.dxr);
    }
    h_t[k] = -(d1 + d2) / cl;
  }

  const double e1 = (a.vor[kp] + 2.0 * omega * a.z[kp]) * vp;
  const double e2 = (a.vor[km] + 2.0 * omega * a.z[km]) * vm;
  const double b0 = bern(k);
  double curl = (e1 - e2) / (2.0 * g.dxr);
  double lap = (bern(kp) - 2.0 * b0 + bern(km)) / (g.dxr * g.dxr * cl * cl);
  if (i == 0 || i == g.nlat - 1) {
    // one-sided rows (the pole node carries weight 0); the product is
    // 2 Omega z u here, not (zeta + 2 Omega z) u as in the interior
    const int64_t kq = ll_at(g, i == 0 ? 1 : i - 1, j);
    const double lq = i == 0 ? lat + g.dxr : lat - g.dxr;
    double u0, v0, uq, vq;
    uv(k, u0, v0);
    uv(kq, uq, vq);
    const double w1 = (a.vor[k] + 2.0 * omega * a.z[k] * u0) * cl;
    const double w2 = (a.vor[kq] + 2.0 * omega * a.z[kq] * uq) * cos(lq);
    if (i == 0)
      curl -= (3.0 * w1 + w2) / (3.0 * g.dxr);
    else
      curl -= (-3.0 * w1 - w2) / (3.0 * g.dxr);
    lap += curl / cl;
  } else {
    const int64_t kn = ll_at(g, i + 1, j), ks = ll_at(g, i - 1, j);
    double un, vn, us, vs;
    uv(kn, un, vn);
    uv(ks, us, vs);
    const double w1 = (a.vor[kn] + 2.0 * omega * a.z[kn]) * un * cos(lat + g.dxr);
    const double w2 = (a.vor[ks] + 2.0 * omega * a.z[ks]) * us * cos(lat - g.dxr);
    curl -= (w1 - w2) / (2.0 * g.dxr);
    const double bn = bern(kn), bs = bern(ks);
    lap += -tan(lat) * (bn - bs) / (2.0 * g.dxr) + (bn - 2.0 * b0 + bs) / (g.dxr * g.dxr);
  }
  double dtk = curl / cl - lap;
  if (fabs(lat) * 180.0 / M_PI > 60) dtk *= cl * cl;
  if (fabs(dtk) < 2e-12) dtk = 0.0;
  div_t[k] = dtk;
}

// Ring means of the first and last rows, the values standing at the poles in the
// arrival interpolation (swe_time_step_ll.cpp:532-547, :565-580).  One wave per
// mean: out = {south vor, south h, north vor, north h}; each is summed in column
// order by one lane from values staged in LDS, as the reference's loop does.
__global__ __launch_bounds__(kLlBlock) void k_ll_pole_means(LlGrid g, double dt, const double* __restrict__ vor,
                                                            const double* __restrict__ div,
                                                            const double* __restrict__ h,
                                                            const double* __restrict__ vor_t,
                                                            double* __restrict__ out) {
  __shared__ double buf[4][kPoleChunk];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = w < 2 ? 0 : g.nlat - 1;
  const bool is_vor = (w & 1) == 0;
  double acc = 0.0;
  for (int c0 = 0; c0 < g.nlon; c0 += kPoleChunk) {
    const int cn = min(kPoleChunk, g.nlon - c0);
    for (int m = lane; m < cn; m += 64) {
      const int64_t k = ll_at(g, row, c0 + m);
      buf[w][m] = is_vor ? vor[k] + 1.0 * dt * vor_t[k] : h[k] - 1.0 * h[k] * div[k] * dt;
    }
    __syncthreads();
    if (lane == 0)
      for (int m = 0; m < cn; ++m) acc += buf[w][m];
    __syncthreads();
  }
  if (lane == 0) out[w] = acc / g.nlon;
}

__device__ inline void ll_lon_coeffs(double lon_d, int li, double dx, double c[3]) {
  c[0] = (lon_d - dx * (li + 0.5)) * (lon_d - dx * (li + 1.5)) / (2.0 * dx * dx);
  c[1] = -(lon_d - dx * (li - 0.5)) * (lon_d - dx * (li + 1.5)) / (dx * dx);
  c[2] = (lon_d - dx * (li - 0.5)) * (lon_d - dx * (li + 0.5)) / (2.0 * dx * dx);
}

__device__ inline int ll_lon_index(double& lon_d, const LlGrid& g) {
  if (lon_d < 0.0) lon_d += 360.0;
  return ll_index(floor(lon_d / g.dx), g.nlon);
}

// swe_compute_arrival_values_ll (swe_time_step_ll.cpp:389-598) at one point.
__global__ __launch_bounds__(kLlBlock) void k_ll_arrival(LlGrid g, double dt, LlIn a, const double* __restrict__ vor_t,
                                                         const double* __restrict__ div_t,
                                                         const double* __restrict__ pole,
                                                         double* __restrict__ a_vor, double* __restrict__ a_div,
                                                         double* __restrict__ a_h) {
  const int64_t k = (int64_t)blockIdx.x * kLlBlock + threadIdx.x;
  if (k >= (int64_t)g.nlat * g.nlon) return;
  const int i = (int)(k / g.nlon);
  const double dx = g.dx;
  auto vval = [&](int r, int c) {
    const int64_t q = ll_at(g, r, c);
    return a.vor[q] + 1.0 * dt * vor_t[q];
  };
  auto hval = [&](int r, int c) {
    const int64_t q = ll_at(g, r, c);
    return a.h[q] - dt * 1.0 * a.h[q] * a.div[q];
  };
  double u, v;
  ll_uv(a.x[k], a.y[k], a.z[k], a.vx[k], a.vy[k], a.vz[k], u, v);
  if (fabs(v) < 1e-16) {
    // meridional velocity negligible: 1-D interpolation along the row at lon - u dt
    double lon = atan2(a.y[k], a.x[k]);
    lon -= u * dt;
    double lon_d = 180.0 / M_PI * lon;
    lon_d = fmod(lon_d + 360.0, 360.0);
    const int li = ll_lon_index(lon_d, g);
    const int lm = ((li - 1) + g.nlon) % g.nlon, lp = (li + 1) % g.nlon;
    double c[3];
    ll_lon_coeffs(lon_d, li, dx, c);
    const int cols[3] = {lm, li, lp};
    double av = 0, ah = 0;
    for (int l = 0; l < 3; ++l) {
      av += vval(i, cols[l]) * c[l];
      ah += hval(i, cols[l]) * c[l];
    }
    a_vor[k] = av;
    a_h[k] = ah;
  } else {
    double px = a.x[k] - dt * a.vx[k], py = a.y[k] - dt * a.vy[k], pz = a.z[k] - dt * a.vz[k];
    const double nrm = sqrt(px * px + py * py + pz * pz);
    px /= nrm;
    py /= nrm;
    pz /= nrm;
    const double lat_d = 180.0 / M_PI * ll_lat(px, py, pz);
    double lon_d = 180.0 / M_PI * atan2(py, px);
    lon_d = fmod(lon_d + 360.0, 360.0);
    int ti = ll_index(floor((lat_d + 90.0) / dx), g.nlat);
    const int li = ll_lon_index(lon_d, g);
    double r[3];
    bool done = false;
    if (ti == 0) {
      if (lat_d > -90.0 + 0.5 * dx) {
        ti += 1;
      } else {  // nodes: south pole, row 0, row 1
        r[0] = (lat_d - (-90.0 + dx * (ti + 0.5))) * (lat_d - (-90.0 + dx * (ti + 1.5))) / (dx * dx * 0.5 * 1.5);
        r[1] = -(lat_d - (-90.0)) * (lat_d - (-90.0 + dx * (ti + 1.5))) / (dx * dx * 0.5);
        r[2] = (lat_d - (-90.0)) * (lat_d - (-90.0 + dx * (ti + 0.5))) / (dx * dx * 1.5);
        done = true;
      }
    }
    if (ti == g.nlat - 1) {
      if (lat_d < 90.0 - 0.5 * dx) {
        ti -= 1;
      } else {  // nodes: row nlat-2, row nlat-1, north pole
        r[0] = (lat_d - (-90.0 + dx * (ti + 0.5))) * (lat_d - 90.0) / (dx * dx * 1.5);
        r[1] = -(lat_d - (-90.0 + dx * (ti - 0.5))) * (lat_d - 90.0) / (dx * dx * 0.5);
        r[2] = (lat_d - (-90.0 + dx * (ti - 0.5))) * (lat_d - (-90.0 + dx * (ti + 0.5))) / (dx * dx * 0.5 * 1.5);
        done = true;
      }
    }
    if (!done) {
      r[0] = (lat_d - (-90.0 + dx * (ti + 0.5))) * (lat_d - (-90.0 + dx * (ti + 1.5))) / (2.0 * dx * dx);
      r[1] = -(lat_d - (-90.0 + dx * (ti - 0.5))) * (lat_d - (-90.0 + dx * (ti + 1.5))) / (dx * dx);
      r[2] = (lat_d - (-90.0 + dx * (ti - 0.5))) * (lat_d - (-90.0 + dx * (ti + 0.5))) / (2.0 * dx * dx);
    }
    double c[3];
    ll_lon_coeffs(lon_d, li, dx, c);
    const int cols[3] = {((li - 1) + g.nlon) % g.nlon, li, (li + 1) % g.nlon};
    double av = 0, ah = 0;
    for (int row = 0; row < 3; ++row) {
      const int rr = ti - 1 + row;
      const bool south = row == 0 && ti == 0, north = row == 2 && ti == g.nlat - 1;
      for (int l = 0; l < 3; ++l) {
        const double vv = south ? pole[0] : north ? pole[2] : vval(rr, cols[l]);
        const double hh = south ? pole[1] : north ? pole[3] : hval(rr, cols[l]);
        av += vv * r[row] * c[l];
        ah += hh * r[row] * c[l];
      }
    }
    a_vor[k] = av;
    a_h[k] = ah;
  }
  a_div[k] = a.div[k] + 1.0 * dt * div_t[k];
}

// field_averaging (swe_time_step_ll.cpp:270-363) at (i, j): the 5-point average
// with weight eps on rows with |lat| > 60 degrees (lat of column 0), times
// cos^2(lat); weight 0 elsewhere.  The longitude neighbours are the periodic
// j+1 and j-1 (the reference reads them before setting them, :322-323).
__device__ inline double ll_average(const LlGrid& g, const double* __restrict__ x, const double* __restrict__ y,
                                    const double* __restrict__ z, const double* __restrict__ f, int i, int j,
                                    double eps) {
  const int64_t k0 = ll_at(g, i, 0);
  const double lat = ll_lat(x[k0], y[k0], z[k0]);
  const bool polar = fabs(lat) > M_PI / 3.0;
  const double e = polar ? eps : 0.00;
  const int64_t k = ll_at(g, i, j);
  const double f1 = e * f[ll_at(g, i, ll_jp(g, j))], f2 = e * f[ll_at(g, i, ll_jm(g, j))];
  double f3, f4, coeff;
  if (i > 0 && i < g.nlat - 1) {
    f3 = e * f[ll_at(g, i + 1, j)];
    f4 = e * f[ll_at(g, i - 1, j)];
    coeff = 4;
  } else {
    f3 = e * f[ll_at(g, i == 0 ? 1 : i - 1, j)];
    f4 = 0;
    coeff = 3;
  }
  double val = (1.0 - coeff * e) * f[k] + f1 + f2 + f3 + f4;
  if (polar) val *= cos(lat) * cos(lat);
  return val;
}

// laplacian_diffusion (swe_time_step_ll.cpp:233-268) at an interior point, from
// the five stencil values: f + dt * 0.89/6371000 * (spherical 5-point Laplacian).
__device__ inline double ll_diffuse(const LlGrid& g, double dt, double lat, double c, double e, double w, double nn,
                                    double ss) {
  const double coeff = 0.89 / 6371000.0;
  const double l1 = 1.0 / (g.dxr * g.dxr * cos(lat) * cos(lat)) * (e - 2.0 * c + w);
  const double l2 = tan(lat) * (nn - ss) / (2.0 * g.dxr);
  const double l3 = (nn - 2.0 * c + ss) / (g.dxr * g.dxr);
  return c + (l1 - l2 + l3) * dt * coeff;
}

__global__ __launch_bounds__(kLlBlock) void k_ll_average(LlGrid g, const double* __restrict__ x,
                                                         const double* __restrict__ y, const double* __restrict__ z,
                                                         const double* __restrict__ f, double* __restrict__ out,
                                                         double eps) {
  const int64_t k = (int64_t)blockIdx.x * kLlBlock + threadIdx.x;
  if (k >= (int64_t)g.nlat * g.nlon) return;
  const int i = (int)(k / g.nlon), j = (int)(k - (int64_t)i * g.nlon);
  out[k] = ll_average(g, x, y, z, f, i, j, eps);
}

__global__ __launch_bounds__(kLlBlock) void k_ll_diffuse(LlGrid g, double dt, const double* __restrict__ x,
                                                         const double* __restrict__ y, const double* __restrict__ z,
                                                         const double* __restrict__ f, double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * kLlBlock + threadIdx.x;
  if (k >= (int64_t)g.nlat * g.nlon) return;
  const int i = (int)(k / g.nlon), j = (int)(k - (int64_t)i * g.nlon);
  if (i == 0 || i == g.nlat - 1) {
    out[k] = f[k];
    return;
  }
  out[k] = ll_diffuse(g, dt, ll_lat(x[k], y[k], z[k]), f[k], f[ll_at(g, i, ll_jp(g, j))], f[ll_at(g, i, ll_jm(g, j))],
                      f[ll_at(g, i + 1, j)], f[ll_at(g, i - 1, j)]);
}

// The step's tail in one pass: div = diffusion(averaging(a_div)) with the
// averaged values of the five stencil points recomputed (the same operations as
// the two kernels above, so the same values), and vor, h = a_vor, a_h.
__global__ __launch_bounds__(kLlBlock) void k_ll_smooth(LlGrid g, double dt, const double* __restrict__ x,
                                                        const double* __restrict__ y, const double* __restrict__ z,
                                                        const double* __restrict__ a_vor,
                                                        const double* __restrict__ a_div,
                                                        const double* __restrict__ a_h, double eps,
                                                        double* __restrict__ vor, double* __restrict__ div,
                                                        double* __restrict__ h) {
  const int64_t k = (int64_t)blockIdx.x * kLlBlock + threadIdx.x;
  if (k >= (int64_t)g.nlat * g.nlon) return;
  const int i = (int)(k / g.nlon), j = (int)(k - (int64_t)i * g.nlon);
  vor[k] = a_vor[k];
  h[k] = a_h[k];
  const double c = ll_average(g, x, y, z, a_div, i, j, eps);
  if (i == 0 || i == g.nlat - 1) {
    div[k] = c;
    return;
  }
  div[k] = ll_diffuse(g, dt, ll_lat(x[k], y[k], z[k]), c, ll_average(g, x, y, z, a_div, i, ll_jp(g, j), eps),
                      ll_average(g, x, y, z, a_div, i, ll_jm(g, j), eps),
                      ll_average(g, x, y, z, a_div, i + 1, j, eps), ll_average(g, x, y, z, a_div, i - 1, j, eps));
}

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// polar_filtering (swe_time_step_ll.cpp:600-664), one workgroup per row, in
// place; rows with |lat| <= 60 degrees (lat of column 0) return at once.  The
// Fourier coefficients start at zero (the reference leaves all but the first
// uninitialised, :628-631).  Values are clamped as the reference does (above
// 1e-11 and below -1e-11 both to +1e-11), the sums over the row are reduced
// across the workgroup, then each point is rebuilt from the coefficients.
// cos(m lat) and sin(m lat) come from one sincos and the angle-addition
// recurrence (error growing by about an ulp per term, 28 terms at most) in place
// of a library call per term.
__global__ __launch_bounds__(kLlBlock) void k_ll_polar(LlGrid g, const double* __restrict__ x,
                                                       const double* __restrict__ y, const double* __restrict__ z,
                                                       double* __restrict__ f) {
  const int i = blockIdx.x;
  const int64_t k0 = ll_at(g, i, 0);
  const double lat0 = ll_lat(x[k0], y[k0], z[k0]);
  if (!(fabs(lat0) > M_PI / 3.0)) return;  // uniform over the workgroup
  int spec = (int)floor(90.0 - fabs(lat0 * 180.0 / M_PI));
  spec = min(spec + 1, kSpecMax);
  double cs[kSpecMax], sn[kSpecMax];
#pragma unroll
  for (int m = 0; m < kSpecMax; ++m) cs[m] = sn[m] = 0.0;
  for (int j = threadIdx.x; j < g.nlon; j += kLlBlock) {
    const int64_t k = ll_at(g, i, j);
    const double lat = ll_lat(x[k], y[k], z[k]);
    double fv = f[k];
    if (fv > 1e-11)
      fv = 1e-11;
    else if (fv < -1e-11)
      fv = 1e-11;
    cs[0] += fv;
    double s1, c1;
    sincos(lat, &s1, &c1);
    double sm = s1, cm = c1;  // sin(m lat), cos(m lat) by the angle-addition recurrence
#pragma unroll
    for (int m = 1; m < kSpecMax; ++m) {
      if (m < spec) {
        cs[m] += fv * cm * g.dxr;
        sn[m] += fv * sm * g.dxr;
      }
      const double cn = cm * c1 - sm * s1;
      sm = sm * c1 + cm * s1;
      cm = cn;
    }
  }
  __shared__ double part[kLlBlock / 64][2 * kSpecMax];
  __shared__ double coef[2 * kSpecMax];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int m = 0; m < kSpecMax; ++m) {
    const double a = wave_sum(cs[m]), b = wave_sum(sn[m]);
    if (lane == 0) {
      part[w][m] = a;
      part[w][kSpecMax + m] = b;
    }
  }
  __syncthreads();
  if (threadIdx.x < 2 * kSpecMax) {
    double t = 0.0;
    for (int q = 0; q < kLlBlock / 64; ++q) t += part[q][threadIdx.x];
    t /= g.nlon;
    if (t != t) t = 0;
    coef[threadIdx.x] = t;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < g.nlon; j += kLlBlock) {
    const int64_t k = ll_at(g, i, j);
    const double lat = ll_lat(x[k], y[k], z[k]);
    double val = coef[0];
    double s1, c1;
    sincos(lat, &s1, &c1);
    double sm = s1, cm = c1;
    for (int m = 1; m < spec; ++m) {
      val += cm * coef[m] + sm * coef[kSpecMax + m];
      const double cn = cm * c1 - sm * s1;
      sm = sm * c1 + cm * s1;
      cm = cn;
    }
    f[k] = val;
  }
}

LlIn ll_in(const double* x, const double* y, const double* z, const double* vor, const double* div, const double* h,
           const double* topo, const double* const* vel) {
  return LlIn{x, y, z, vor, div, h, topo, vel ? vel[0] : nullptr, vel ? vel[1] : nullptr, vel ? vel[2] : nullptr};
}

unsigned ll_blocks(const LlGrid& g) { return bli::blocks_for((int64_t)g.nlat * g.nlon, kLlBlock); }

}  // namespace

LlGrid ll_grid(int nlat, int nlon, double dx_deg) { return LlGrid{nlat, nlon, dx_deg, dx_deg * M_PI / 180.0}; }

void launch_ll_tendencies(const LlGrid& g, double omega, const double* x, const double* y, const double* z,
                          const double* vor, const double* div, const double* h, const double* topo,
                          const double* const* vel, double* vor_t, double* div_t, double* h_t, hipStream_t s) {
  hipLaunchKernelGGL(k_ll_tend, dim3(ll_blocks(g)), dim3(kLlBlock), 0, s, g, omega,
                     ll_in(x, y, z, vor, div, h, topo, vel), vor_t, div_t, h_t);
}

void launch_ll_pole_means(const LlGrid& g, double dt, const double* vor, const double* div, const double* h,
                          const double* vor_t, double* means, hipStream_t s) {
  hipLaunchKernelGGL(k_ll_pole_means, dim3(1), dim3(kLlBlock), 0, s, g, dt, vor, div, h, vor_t, means);
}

void launch_ll_arrival(const LlGrid& g, double dt, const double* x, const double* y, const double* z,
                       const double* vor, const double* div, const double* h, const double* const* vel,
                       const double* vor_t, const double* div_t, const double* means, double* a_vor, double* a_div,
                       double* a_h, hipStream_t s) {
  hipLaunchKernelGGL(k_ll_arrival, dim3(ll_blocks(g)), dim3(kLlBlock), 0, s, g, dt,
                     ll_in(x, y, z, vor, div, h, nullptr, vel), vor_t, div_t, means, a_vor, a_div, a_h);
}

void launch_ll_average(const LlGrid& g, const double* x, const double* y, const double* z, const double* f,
                       double* out, double eps, hipStream_t s) {
  hipLaunchKernelGGL(k_ll_average, dim3(ll_blocks(g)), dim3(kLlBlock), 0, s, g, x, y, z, f, out, eps);
}

void launch_ll_diffuse(const LlGrid& g, double dt, const double* x, const double* y, const double* z,
                       const double* f, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_ll_diffuse, dim3(ll_blocks(g)), dim3(kLlBlock), 0, s, g, dt, x, y, z, f, out);
}

void launch_ll_smooth(const LlGrid& g, double dt, const double* x, const double* y, const double* z,
                      const double* a_vor, const double* a_div, const double* a_h, double eps, double* vor,
                      double* div, double* h, hipStream_t s) {
  hipLaunchKernelGGL(k_ll_smooth, dim3(ll_blocks(g)), dim3(kLlBlock), 0, s, g, dt, x, y, z, a_vor, a_div, a_h, eps,
                     vor, div, h);
}

void launch_ll_polar_filter(const LlGrid& g, const double* x, const double* y, const double* z, double* f,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_ll_polar, dim3(g.nlat), dim3(kLlBlock), 0, s, g, x, y, z, f);
}

}  // namespace sbslpm
