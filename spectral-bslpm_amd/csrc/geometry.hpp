// geometry.hpp -- cubed-sphere maps and 1-D barycentric Lagrange weights,
// shared by the host set-up code and the gfx950 kernels.
//
// Equiangular gnomonic cubed sphere: face f in 0..5, label coordinates
// (xi, eta) in radians, |xi|,|eta| <= pi/4.  Face orientation and the
// face-selection tie rule follow cubed_sphere_transforms_impl.hpp:8-123.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

namespace sbslpm {

constexpr double kPi = 3.14159265358979323846;
constexpr double kNegInv4Pi = -0.07957747154594767;  // -1/(4 pi), bve_vel_impl.hpp:63
constexpr int kMaxDeg = 10;                          // interp_degree clamp, io_utils.cpp:36-38
constexpr double kSnapTol = 1e-15;                   // node snap / bounds tolerance, interp_funcs_impl.hpp:22-33, 60

// cubed_sphere_transforms_impl.hpp:8-32
__host__ __device__ inline int face_from_xyz(double x, double y, double z) {
  double ax = fabs(x), ay = fabs(y), az = fabs(z);
  if ((ax >= ay) && (ax >= az)) return x >= 0 ? 0 : 2;
  if ((ay >= ax) && (ay >= az)) return y >= 0 ? 1 : 3;
  return z >= 0 ? 4 : 5;
}

// cubed_sphere_transforms_impl.hpp:34-65.  Returns false for a face outside 0..5.
__host__ __device__ inline bool xyz_from_xieta(double xi, double eta, int face, double* xyz) {
  double X = tan(xi), Y = tan(eta);
  double r = 1.0 / sqrt(1 + X * X + Y * Y);
  switch (face) {
    case 0: xyz[0] = r;  xyz[1] = X * r;  xyz[2] = Y * r;  return true;
    case 1: xyz[1] = r;  xyz[0] = -X * r; xyz[2] = Y * r;  return true;
    case 2: xyz[0] = -r; xyz[1] = -X * r; xyz[2] = Y * r;  return true;
    case 3: xyz[1] = -r; xyz[0] = X * r;  xyz[2] = Y * r;  return true;
    case 4: xyz[2] = r;  xyz[0] = -Y * r; xyz[1] = X * r;  return true;
    case 5: xyz[2] = -r; xyz[0] = Y * r;  xyz[1] = X * r;  return true;
    default: return false;
  }
}

// cubed_sphere_transforms_impl.hpp:100-123 (face given)
__host__ __device__ inline void xieta_from_xyz_face(double x, double y, double z, int face, double* xe) {
  switch (face) {
    case 0: xe[0] = atan(y / x);  xe[1] = atan(z / x);  break;
    case 1: xe[0] = atan(-x / y); xe[1] = atan(z / y);  break;
    case 2: xe[0] = atan(y / x);  xe[1] = atan(-z / x); break;
    case 3: xe[0] = atan(-x / y); xe[1] = atan(-z / y); break;
    case 4: xe[0] = atan(y / z);  xe[1] = atan(-x / z); break;
    default: xe[0] = atan(-y / z); xe[1] = atan(-x / z); break;
  }
}

// Point location in the uniform quadtree, the descent of point_locate_panel
// (interp_funcs_impl.hpp:114-142) done arithmetically instead of by reading
// panels.  The level-l panel with integer coordinates (i, j) spans label
// units [-1 + i h, -1 + (i+1) h], h = 2^(1-l), so max + min = (2i+1) h - 2
// exactly and its mid-line (pi/8)(max + min) rounds exactly as the stored
// tree's.  Children are numbered 4k + c below their parent's level-relative
// index k in the order (hi,hi), (lo,hi), (lo,lo), (hi,lo), with the
// reference's comparison sequence deciding ties.  Returns the leaf panel id
// (level_start[levels-1] + k) or -1 when no child takes the point (NaN);
// xe receives the label coordinates on the located face.
__host__ __device__ inline int locate_uniform(int levels, double x, double y, double z, double* xe) {
  const int face = face_from_xyz(x, y, z);
  xieta_from_xyz_face(x, y, z, face, xe);
  int k = face, i = 0, j = 0;
  for (int l = 0; l + 1 < levels; ++l) {
    const double h = ldexp(1.0, 1 - l);
    const double mx = 0.5 * kPi / 4.0 * ((2 * i + 1) * h - 2.0);
    const double me = 0.5 * kPi / 4.0 * ((2 * j + 1) * h - 2.0);
    int c;
    if (xe[0] >= mx && xe[1] >= me) c = 0;
    else if (xe[0] <= mx && xe[1] >= me) c = 1;
    else if (xe[0] <= mx && xe[1] <= me) c = 2;
    else if (xe[0] >= mx && xe[1] <= me) c = 3;
    else return -1;
    i = 2 * i + ((c == 0 || c == 3) ? 1 : 0);
    j = 2 * j + (c <= 1 ? 1 : 0);
    k = 4 * k + c;
  }
  return 2 * ((1 << (2 * (levels - 1))) - 1) + k;
}

// general_utils_impl.hpp:179-185
__host__ __device__ inline void project_to_sphere(double& x, double& y, double& z) {
  double n = sqrt(x * x + y * y + z * z);
  x /= n; y /= n; z /= n;
}

// One-dimensional factor of interp_vals_bli (interp_funcs_impl.hpp:15-103):
// Lagrange weights w[0..deg] of the Chebyshev (2nd kind) nodes mapped onto
// [lo, hi], second barycentric form, snapping to a node within 1e-15.
// `cheb` holds cos(pi*i/deg) as the reference evaluates it.  Returns false
// if t lies outside [lo, hi] by more than 1e-15 (the reference aborts there).
__host__ __device__ inline bool bli_weights_1d(double t, double lo, double hi, int deg, const double* cheb,
                                               double* w) {
  double range = 0.5 * (hi - lo), off = 0.5 * (hi + lo);
  bool ok = (t >= lo - kSnapTol) && (t <= hi + kSnapTol);  // false for NaN as well
  if (deg == 0) { w[0] = 1.0; return ok; }
  int hit = -1;
  for (int i = 0; i <= deg; ++i) {
    double node = cheb[i] * range + off;
    if (hit < 0 && fabs(t - node) < kSnapTol) hit = i;
  }
  if (hit >= 0) {
    for (int i = 0; i <= deg; ++i) w[i] = (i == hit) ? 1.0 : 0.0;
    return ok;
  }
  double den = 0;
  for (int i = 0; i <= deg; ++i) {
    double bw = (i == 0 || i == deg) ? 0.5 : 1.0;
    if (i & 1) bw = -bw;
    double v = bw / (t - (cheb[i] * range + off));
    w[i] = v;
    den += v;
  }
  double inv = 1.0 / den;
  for (int i = 0; i <= deg; ++i) w[i] *= inv;
  return ok;
}

}  // namespace sbslpm
