// sbslpm_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// FP64 CPU restatement of the spectral-BSLPM hot path (cubed-sphere BLI FMM,
// direct sum, backward semi-Lagrangian BVE step).  It is the checker for the
// HIP library: only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg may load it.  The product path never links or calls it.
//
// Every function cites the reference file:line it restates.  The loop
// structure of the reference is kept on purpose (per-pair recomputation of
// source node coordinates, per-node barycentric basis evaluation), so that
// timing this file is timing the reference algorithm on the host cores.
// Two deliberate differences, both summation-order only:
//   * accumulation is target-major and sequential per target row instead of
//     Kokkos::atomic_add (the reference order is non-deterministic anyway);
//   * the caller-side zeroing of accumulated arrays (bve_time_step.cpp:321,
//     341-343, 351-354) is done inside the pass that accumulates.
//
// Generalisation ("layout A-prime"): each leaf carries an m x m Chebyshev
// (2nd kind) particle grid of degree m-1, proxies have degree d.  For m == d+1
// this is exactly the reference layout (particles == leaf proxy nodes,
// P2M/L2P identities, fmm_funcs_impl.hpp:17-34, 174-188).  For m != d+1 the
// leaf <-> proxy transfers and the type-dependent interactions follow the
// reference's particle-tree functions (base_pots_ll / child_panel_interp_3_ll,
// fmm_funcs_impl.hpp:190-225, 318-364; swe_vel_panel_interaction_ll,
// swe_vel_impl.hpp:179-287), with particle label coordinates taken from the
// leaf grid.
//
// How each part is pinned (checks that do not restate it) is in the tests:
// tests/test_oracle_*.py (closed forms, brute force, invariants).  The
// shallow-water step's reference indexing rule (orc_velocity_gradient rule 0:
// the leaf-local face index, the metric applied twice, the pole shortcut of
// xyzvec_from_xietavec) is pinned through its parts only -- the barycentric
// differentiation and the vector conversion are checked against closed forms
// and the two rules agree where the reference's is exact -- so its composition
// is "partially pinned": it follows the reference's lines.
//
// Build: oracle/build.sh (g++ -O2 -fopenmp -ffp-contract=off).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <queue>
#include <deque>
#include <vector>
#include <algorithm>
#include <omp.h>

namespace {

const double kPi = M_PI;
const double kNegInv4Pi = -0.07957747154594767;  // bve_vel_impl.hpp:63, inverse_poisson_impl.hpp:48

enum { ORC_OK = 0, ORC_E_GEOMETRY = 1, ORC_E_ARG = 2 };

// ---------------------------------------------------------------- geometry

// cubed_sphere_transforms_impl.hpp:8-32
int face_from_xyz(double x, double y, double z) {
  double ax = std::fabs(x), ay = std::fabs(y), az = std::fabs(z);
  if ((ax >= ay) && (ax >= az)) return x >= 0 ? 0 : 2;
  if ((ay >= ax) && (ay >= az)) return y >= 0 ? 1 : 3;
  return z >= 0 ? 4 : 5;
}

// cubed_sphere_transforms_impl.hpp:34-65 (abort -> error code)
int xyz_from_xieta(double xi, double eta, int face, double* xyz) {
  double X = std::tan(xi), Y = std::tan(eta);
  if (face == 0) {
    xyz[0] = 1.0 / std::sqrt(1 + X * X + Y * Y); xyz[1] = X * xyz[0]; xyz[2] = Y * xyz[0];
  } else if (face == 1) {
    xyz[1] = 1.0 / std::sqrt(1 + X * X + Y * Y); xyz[0] = -X * xyz[1]; xyz[2] = Y * xyz[1];
  } else if (face == 2) {
    xyz[0] = -1.0 / std::sqrt(1 + X * X + Y * Y); xyz[1] = X * xyz[0]; xyz[2] = -Y * xyz[0];
  } else if (face == 3) {
    xyz[1] = -1.0 / std::sqrt(1 + X * X + Y * Y); xyz[0] = -X * xyz[1]; xyz[2] = -Y * xyz[1];
  } else if (face == 4) {
    xyz[2] = 1.0 / std::sqrt(1 + X * X + Y * Y); xyz[0] = -Y * xyz[2]; xyz[1] = X * xyz[2];
  } else if (face == 5) {
    xyz[2] = -1.0 / std::sqrt(1 + X * X + Y * Y); xyz[0] = -Y * xyz[2]; xyz[1] = -X * xyz[2];
  } else {
    return ORC_E_GEOMETRY;
  }
  return ORC_OK;
}

// cubed_sphere_transforms_impl.hpp:67-98
void xieta_from_xyz(double x, double y, double z, double* xieta) {
  double ax = std::fabs(x), ay = std::fabs(y), az = std::fabs(z);
  if ((ax >= ay) && (ax >= az)) {
    if (x >= 0) { xieta[0] = std::atan(y / x); xieta[1] = std::atan(z / x); }
    else        { xieta[0] = std::atan(y / x); xieta[1] = std::atan(-z / x); }
  } else if ((ay >= ax) && (ay >= az)) {
    if (y >= 0) { xieta[0] = std::atan(-x / y); xieta[1] = std::atan(z / y); }
    else        { xieta[0] = std::atan(-x / y); xieta[1] = std::atan(-z / y); }
  } else {
    if (z >= 0) { xieta[0] = std::atan(y / z); xieta[1] = std::atan(-x / z); }
    else        { xieta[0] = std::atan(-y / z); xieta[1] = std::atan(-x / z); }
  }
}

// cubed_sphere_transforms_impl.hpp:100-123
int xieta_from_xyz_face(double x, double y, double z, int face, double* xieta) {
  if (face == 0)      { xieta[0] = std::atan(y / x);  xieta[1] = std::atan(z / x); }
  else if (face == 1) { xieta[0] = std::atan(-x / y); xieta[1] = std::atan(z / y); }
  else if (face == 2) { xieta[0] = std::atan(y / x);  xieta[1] = std::atan(-z / x); }
  else if (face == 3) { xieta[0] = std::atan(-x / y); xieta[1] = std::atan(-z / y); }
  else if (face == 4) { xieta[0] = std::atan(y / z);  xieta[1] = std::atan(-x / z); }
  else if (face == 5) { xieta[0] = std::atan(-y / z); xieta[1] = std::atan(-x / z); }
  else return ORC_E_GEOMETRY;
  return ORC_OK;
}

// general_utils_impl.hpp:179-185
inline void project_to_sphere(double& x, double& y, double& z) {
  double n = std::sqrt(x * x + y * y + z * z);
  x /= n; y /= n; z /= n;
}

// general_utils.cpp:23-26
double gcdist(const double* p1, const double* p2) {
  double dp = std::max(std::min(p1[0] * p2[0] + p1[1] * p2[1] + p1[2] * p2[2], 1.0), -1.0);
  return std::acos(dp);
}

// general_utils.cpp:5-21 (l'Huilier)
double sphere_tri_area(const double* p1, const double* p2, const double* p3) {
  double p1n[3], p2n[3], p3n[3];
  std::memcpy(p1n, p1, 3 * sizeof(double));
  std::memcpy(p2n, p2, 3 * sizeof(double));
  std::memcpy(p3n, p3, 3 * sizeof(double));
  p2n[0] -= p3[0], p2n[1] -= p3[1], p2n[2] -= p3[2];
  p3n[0] -= p1[0], p3n[1] -= p1[1], p3n[2] -= p1[2];
  p1n[0] -= p2[0], p1n[1] -= p2[1], p1n[2] -= p2[2];
  double a = std::acos(1 - 0.5 * (p2n[0] * p2n[0] + p2n[1] * p2n[1] + p2n[2] * p2n[2]));
  double b = std::acos(1 - 0.5 * (p3n[0] * p3n[0] + p3n[1] * p3n[1] + p3n[2] * p3n[2]));
  double c = std::acos(1 - 0.5 * (p1n[0] * p1n[0] + p1n[1] * p1n[1] + p1n[2] * p1n[2]));
  double s = (a + b + c) / 2;
  double z = std::tan(s / 2) * std::tan((s - a) / 2) * std::tan((s - b) / 2) * std::tan((s - c) / 2);
  return 4 * std::atan(std::sqrt(z));
}

// ---------------------------------------------------------------- interpolation

// interp_funcs.cpp:11-22
double bli_coeff(int j, int degree) {
  if (degree == 0) return 1;
  if (j == 0) return 0.5;
  if (j == degree) return 0.5 * std::pow(-1, j);
  return std::pow(-1, j);
}

// Clenshaw-Curtis weights on [-1,1] at cos(k pi/N) (Trefethen, "Spectral
// Methods in MATLAB", clencurt.m).  Not in the reference: used for the particle
// grid when its degree is odd, where the reference formula does not integrate
// constants (its weights then do not sum to 2).
void cc_weights_textbook(int N, double* w) {
  if (N == 0) { w[0] = 2.0; return; }
  std::vector<double> v(N + 1, 1.0);
  for (int i = 0; i <= N; ++i) w[i] = 0;
  double theta;
  if (N % 2 == 0) {
    w[0] = 1.0 / (N * N - 1.0); w[N] = w[0];
    for (int k = 1; k <= N / 2 - 1; ++k)
      for (int i = 1; i < N; ++i) { theta = kPi * i / N; v[i] -= 2.0 * std::cos(2.0 * k * theta) / (4.0 * k * k - 1.0); }
    for (int i = 1; i < N; ++i) { theta = kPi * i / N; v[i] -= std::cos(N * theta) / (N * N - 1.0); }
  } else {
    w[0] = 1.0 / ((double)N * N); w[N] = w[0];
    for (int k = 1; k <= (N - 1) / 2; ++k)
      for (int i = 1; i < N; ++i) { theta = kPi * i / N; v[i] -= 2.0 * std::cos(2.0 * k * theta) / (4.0 * k * k - 1.0); }
  }
  for (int i = 1; i < N; ++i) w[i] = 2.0 * v[i] / N;
}

// interp_funcs.cpp:24-73.  tab[(d+1)^2][4] = {xi node, eta node, bary weight
// product, quadrature weight product}.  weight_rule 0 keeps the reference's
// quadrature formula (:42-60), 1 uses textbook Clenshaw-Curtis, 2 leaves the
// reference formula in column 3 (equal-area weights are applied by spec_points).
void interp_init(int degree, int weight_rule, double* tab) {
  int pps = degree + 1;
  std::vector<double> cheb(pps, 0), bw(pps, 0);
  if (degree == 0) {
    cheb[0] = 0; bw[0] = 1;
  } else {
    for (int i = 0; i < pps; i++) {
      cheb[i] = std::cos(kPi / degree * i);
      if (std::fabs(cheb[i]) < 1e-16) cheb[i] = 0;
      bw[i] = bli_coeff(i, degree);
    }
  }
  std::vector<double> cc(pps, 0);
  if (weight_rule == 1) {
    cc_weights_textbook(degree, cc.data());
  } else {
    std::vector<double> vec(degree / 2 + 1, 2);
    vec[0] = 1;
    for (int i = 1; i < degree / 2 + 1; i++) vec[i] = 2.0 / (1 - 4 * i * i);
    for (int i = 0; i < degree / 2 + 1; i++) {
      double sum = 0;
      for (int j = 0; j < degree / 2 + 1; j++) sum += 2 * std::cos(2 * i * j * kPi / degree) / degree * vec[j];
      cc[i] = sum;
    }
    cc[0] *= 0.5;
    for (int i = 0; i < degree / 2; i++) cc[degree - i] = cc[i];
  }
  for (int i = 0; i < pps; i++)
    for (int j = 0; j < pps; j++) {
      int index = i * pps + j;
      tab[index * 4 + 0] = cheb[i];
      tab[index * 4 + 1] = cheb[j];
      tab[index * 4 + 2] = bw[i] * bw[j];
      tab[index * 4 + 3] = cc[i] * cc[j];
    }
}

// interp_funcs_impl.hpp:11-112 (Kokkos::abort -> error code, basis zeroed).
int interp_vals_bli(double* basis, double xi, double eta, double min_xi, double max_xi,
                    double min_eta, double max_eta, int deg) {
  double xi_range = 0.5 * (max_xi - min_xi), xi_offset = 0.5 * (max_xi + min_xi);
  double eta_range = 0.5 * (max_eta - min_eta), eta_offset = 0.5 * (max_eta + min_eta);
  double cheb_xi[11], cheb_eta[11], bw[11], xf[11], ef[11], val;
  for (int i = 0; i < (deg + 1) * (deg + 1); i++) basis[i] = 0.0;
  if (xi < min_xi - 1e-15 || xi > max_xi + 1e-15 || eta < min_eta - 1e-15 || eta > max_eta + 1e-15 ||
      !(xi == xi) || !(eta == eta))
    return ORC_E_GEOMETRY;
  if (deg == 0) {
    cheb_xi[0] = xi_offset; cheb_eta[0] = eta_offset; bw[0] = 1;
  } else {
    for (int i = 0; i < deg + 1; i++) {
      cheb_xi[i] = std::cos(kPi * i / deg) * xi_range + xi_offset;
      cheb_eta[i] = std::cos(kPi * i / deg) * eta_range + eta_offset;
      if (i == 0) bw[i] = 0.5;
      else if (i == deg) bw[i] = 0.5 * std::pow(-1, i);
      else bw[i] = std::pow(-1, i);
    }
  }
  bool found = false;
  for (int i = 0; i < deg + 1; i++) {
    if (std::fabs(xi - cheb_xi[i]) < 1e-15) {
      found = true;
      for (int j = 0; j < deg + 1; j++) xf[j] = 0;
      xf[i] = 1;
      break;
    }
  }
  if (!found) {
    double den = 0;
    for (int i = 0; i < deg + 1; i++) { val = bw[i] / (xi - cheb_xi[i]); xf[i] = val; den += val; }
    for (int i = 0; i < deg + 1; i++) xf[i] /= den;
  }
  found = false;
  for (int i = 0; i < deg + 1; i++) {
    if (std::fabs(eta - cheb_eta[i]) < 1e-15) {
      found = true;
      for (int j = 0; j < deg + 1; j++) ef[j] = 0;
      ef[i] = 1;
      break;
    }
  }
  if (!found) {
    double den = 0;
    for (int i = 0; i < deg + 1; i++) { val = bw[i] / (eta - cheb_eta[i]); ef[i] = val; den += val; }
    for (int i = 0; i < deg + 1; i++) ef[i] /= den;
  }
  for (int i = 0; i < deg + 1; i++)
    for (int j = 0; j < deg + 1; j++) basis[i * (deg + 1) + j] = xf[i] * ef[j];
  return ORC_OK;
}

// ---------------------------------------------------------------- tree

struct Panel {  // initialize_cubed_sphere.hpp:8-29 (edge neighbours omitted: off the path)
  int face, level, parent, child[4], is_leaf;
  double min_xi, max_xi, min_eta, max_eta, radius, area;
};

struct Ctx {
  int levels, d, md, weight_rule;  // md = particle grid degree m-1
  double theta;
  int P, Pl, offset, n, m2;
  std::vector<int> level_start;
  std::vector<Panel> pn;
  std::vector<double> tab_d, tab_m;  // interp tables, proxies and particle grid
  std::vector<int> ilist;            // [I][3] = target, source, type (reference BFS order)
  std::vector<int> by_target;        // interaction indices grouped by target, list order kept
  std::vector<int> tgt_start;        // [P+1]
  // shallow-water physics beyond the reference: the coefficient of the
  // free-surface Laplacian in the divergence tendency (the reference's 1, or
  // g/a^2 for heights in metres) and the Gaussian mountain of topography 2
  // (height, width in radians, centre latitude and longitude)
  double gravity = 1.0;
  double mtn[4] = {0, 1, 0, 0};
  // shallow-water step form: bit 0, the stage-1 surface Laplacian by the
  // regularised kernel with the correction at the nodes (the stages-2-3
  // operator); bit 1, the height correction subtracts the free surface
  int swe_form = 0;
  int topo_kind = 0;  // topography of the running step (for the free-surface correction)
};

// initialize_cubed_sphere.cpp:73-140 (children and bounds)
void refine(Ctx& c, int level, int i) {
  int level_start = c.level_start[level + 1];
  int start = level_start + 4 * i;
  int prev_start = c.level_start[level];
  Panel& par = c.pn[i + prev_start];
  for (int j = 0; j < 4; j++) {
    c.pn[start + j].face = par.face;
    c.pn[start + j].parent = i + prev_start;
    c.pn[start + j].level = par.level + 1;
  }
  double max_xi = par.max_xi, min_xi = par.min_xi, max_eta = par.max_eta, min_eta = par.min_eta;
  double mid_xi = 0.5 * (max_xi + min_xi), mid_eta = 0.5 * (max_eta + min_eta);
  c.pn[start].max_xi = max_xi;     c.pn[start].min_xi = mid_xi;
  c.pn[start].max_eta = max_eta;   c.pn[start].min_eta = mid_eta;
  c.pn[start + 1].max_xi = mid_xi; c.pn[start + 1].min_xi = min_xi;
  c.pn[start + 1].max_eta = max_eta; c.pn[start + 1].min_eta = mid_eta;
  c.pn[start + 2].max_xi = mid_xi; c.pn[start + 2].min_xi = min_xi;
  c.pn[start + 2].max_eta = mid_eta; c.pn[start + 2].min_eta = min_eta;
  c.pn[start + 3].max_xi = max_xi; c.pn[start + 3].min_xi = mid_xi;
  c.pn[start + 3].max_eta = mid_eta; c.pn[start + 3].min_eta = min_eta;
  for (int j = 0; j < 4; j++) par.child[j] = start + j;
}

// initialize_cubed_sphere.cpp:151-182
void init_2(Ctx& c, int i) {
  Panel& p = c.pn[i];
  p.is_leaf = (p.level == c.levels - 1) ? 1 : 0;
  double max_xi = p.max_xi * kPi / 4.0, min_xi = p.min_xi * kPi / 4.0;
  double max_eta = p.max_eta * kPi / 4.0, min_eta = p.min_eta * kPi / 4.0;
  double mid_xi = 0.5 * (max_xi + min_xi), mid_eta = 0.5 * (max_eta + min_eta);
  double p1[3], p2[3], p3[3], p4[3], pc[3];
  xyz_from_xieta(max_xi, max_eta, p.face, p1);
  xyz_from_xieta(max_xi, min_eta, p.face, p2);
  xyz_from_xieta(min_xi, min_eta, p.face, p3);
  xyz_from_xieta(min_xi, max_eta, p.face, p4);
  xyz_from_xieta(mid_xi, mid_eta, p.face, pc);
  double d1 = gcdist(p1, pc), d2 = gcdist(p2, pc), d3 = gcdist(p3, pc), d4 = gcdist(p4, pc);
  p.radius = std::fmax(d1, std::fmax(d2, std::fmax(d3, d4)));
  p.area = sphere_tri_area(p1, p2, p3) + sphere_tri_area(p1, p3, p4);
}

// initialize_cubed_sphere.cpp:185-234
void build_tree(Ctx& c) {
  int L = c.levels;
  c.P = 2 * ((int)std::pow(4, L) - 1);
  c.Pl = 6 * (int)std::pow(4, L - 1);
  c.offset = c.P - c.Pl;
  c.pn.assign(c.P, Panel());
  for (auto& p : c.pn) { p.parent = -1; p.child[0] = p.child[1] = p.child[2] = p.child[3] = -1; }
  for (int i = 0; i < 6; i++) {
    c.pn[i].face = i; c.pn[i].min_xi = -1; c.pn[i].max_xi = 1;
    c.pn[i].min_eta = -1; c.pn[i].max_eta = 1; c.pn[i].level = 0; c.pn[i].parent = -1;
  }
  c.level_start.assign(L + 1, 0);
  for (int i = 0; i < L; i++) c.level_start[i + 1] = c.level_start[i] + 6 * (int)std::pow(4, i);
  for (int i = 0; i < L - 1; i++)
    for (int j = 0; j < 6 * (int)std::pow(4, i); j++) refine(c, i, j);
  for (int i = 0; i < c.P; i++) init_2(c, i);
}

// interp_funcs_impl.hpp:114-142
int point_locate_panel(const Ctx& c, double x, double y, double z) {
  double xieta[2];
  int cur = face_from_xyz(x, y, z);
  xieta_from_xyz_face(x, y, z, cur, xieta);
  double xi = xieta[0], eta = xieta[1];
  while (!c.pn[cur].is_leaf) {
    double mid_xi = 0.5 * kPi / 4.0 * (c.pn[cur].max_xi + c.pn[cur].min_xi);
    double mid_eta = 0.5 * kPi / 4.0 * (c.pn[cur].max_eta + c.pn[cur].min_eta);
    if ((xi >= mid_xi) && (eta >= mid_eta)) cur = c.pn[cur].child[0];
    else if ((xi <= mid_xi) && (eta >= mid_eta)) cur = c.pn[cur].child[1];
    else if ((xi <= mid_xi) && (eta <= mid_eta)) cur = c.pn[cur].child[2];
    else if ((xi >= mid_xi) && (eta <= mid_eta)) cur = c.pn[cur].child[3];
    else return -1;  // NaN input
  }
  return cur;
}

// fmm_funcs.cpp:17-123 (literal BFS; the pair order is the reference's)
void dual_tree_traversal(Ctx& c) {
  std::queue<int> tq, sq;
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) { tq.push(i); sq.push(j); }
  c.ilist.clear();
  double c1[3], c2[3];
  while (!tq.empty()) {
    int t = tq.front(), s = sq.front();
    tq.pop(); sq.pop();
    const Panel &pt = c.pn[t], &ps = c.pn[s];
    double xi1 = kPi / 4.0 * 0.5 * (pt.min_xi + pt.max_xi);
    double xi2 = kPi / 4.0 * 0.5 * (ps.min_xi + ps.max_xi);
    double eta1 = kPi / 4.0 * 0.5 * (pt.min_eta + pt.max_eta);
    double eta2 = kPi / 4.0 * 0.5 * (ps.min_eta + ps.max_eta);
    xyz_from_xieta(xi1, eta1, pt.face, c1);
    xyz_from_xieta(xi2, eta2, ps.face, c2);
    double dist = gcdist(c1, c2);
    double sep = 100;
    if (dist > 1e-16) sep = (pt.radius + ps.radius) / dist;
    if (sep < c.theta) {
      int type = 0;
      if (pt.level < c.levels - 1) type += 2;
      if (ps.level < c.levels - 1) type += 1;
      c.ilist.push_back(t); c.ilist.push_back(s); c.ilist.push_back(type);
    } else if (pt.is_leaf && ps.is_leaf) {
      c.ilist.push_back(t); c.ilist.push_back(s); c.ilist.push_back(0);
    } else if (pt.is_leaf) {
      for (int k = 0; k < 4; k++) { tq.push(t); sq.push(ps.child[k]); }
    } else if (ps.is_leaf) {
      for (int k = 0; k < 4; k++) { sq.push(s); tq.push(pt.child[k]); }
    } else if (pt.level <= ps.level) {
      for (int k = 0; k < 4; k++) { sq.push(s); tq.push(pt.child[k]); }
    } else {
      for (int k = 0; k < 4; k++) { tq.push(t); sq.push(ps.child[k]); }
    }
  }
  int I = (int)c.ilist.size() / 3;
  c.tgt_start.assign(c.P + 1, 0);
  for (int i = 0; i < I; i++) c.tgt_start[c.ilist[3 * i] + 1]++;
  for (int p = 0; p < c.P; p++) c.tgt_start[p + 1] += c.tgt_start[p];
  c.by_target.assign(I, 0);
  std::vector<int> fill(c.tgt_start.begin(), c.tgt_start.end() - 1);
  for (int i = 0; i < I; i++) c.by_target[fill[c.ilist[3 * i]]++] = i;
}

// Panel label-space scale/offset in radians, as the kernels compute them
// (bve_vel_impl.hpp:38-46).
inline void panel_frame(const Panel& p, double& xo, double& xs, double& eo, double& es) {
  xo = kPi / 4 * 0.5 * (p.max_xi + p.min_xi);
  xs = kPi / 4 * 0.5 * (p.max_xi - p.min_xi);
  eo = kPi / 4 * 0.5 * (p.max_eta + p.min_eta);
  es = kPi / 4 * 0.5 * (p.max_eta - p.min_eta);
}

inline int bli_in_panel(const Panel& p, double xi, double eta, int deg, double* basis) {
  return interp_vals_bli(basis, xi, eta, kPi / 4.0 * p.min_xi, kPi / 4.0 * p.max_xi,
                         kPi / 4.0 * p.min_eta, kPi / 4.0 * p.max_eta, deg);
}

}  // namespace

// ================================================================ C ABI
extern "C" {

void* orc_create(int levels, int degree, int leaf_degree, double theta, int weight_rule) {
  if (levels < 1 || levels > 9 || degree < 1 || degree > 10 || leaf_degree < 1 || leaf_degree > 10) return nullptr;
  Ctx* c = new Ctx();
  c->levels = levels; c->d = degree; c->md = leaf_degree; c->theta = theta; c->weight_rule = weight_rule;
  c->n = (degree + 1) * (degree + 1);
  c->m2 = (leaf_degree + 1) * (leaf_degree + 1);
  c->tab_d.resize(c->n * 4);
  c->tab_m.resize(c->m2 * 4);
  interp_init(degree, 0, c->tab_d.data());
  interp_init(leaf_degree, weight_rule == 1 ? 1 : 0, c->tab_m.data());
  build_tree(*c);
  dual_tree_traversal(*c);
  return c;
}

void orc_destroy(void* h) { delete static_cast<Ctx*>(h); }

int orc_panel_count(void* h) { return static_cast<Ctx*>(h)->P; }
int orc_leaf_count(void* h) { return static_cast<Ctx*>(h)->Pl; }
int orc_interaction_count(void* h) { return (int)static_cast<Ctx*>(h)->ilist.size() / 3; }

// ints[P][8] = face, level, parent, child1..4, is_leaf; dbls[P][6] = min/max xi, min/max eta, radius, area
void orc_get_panels(void* h, int* ints, double* dbls) {
  Ctx& c = *static_cast<Ctx*>(h);
  for (int i = 0; i < c.P; i++) {
    const Panel& p = c.pn[i];
    int* r = ints + 8 * i;
    r[0] = p.face; r[1] = p.level; r[2] = p.parent;
    r[3] = p.child[0]; r[4] = p.child[1]; r[5] = p.child[2]; r[6] = p.child[3]; r[7] = p.is_leaf;
    double* q = dbls + 6 * i;
    q[0] = p.min_xi; q[1] = p.max_xi; q[2] = p.min_eta; q[3] = p.max_eta; q[4] = p.radius; q[5] = p.area;
  }
}

void orc_get_interactions(void* h, int* out) {
  Ctx& c = *static_cast<Ctx*>(h);
  std::memcpy(out, c.ilist.data(), c.ilist.size() * sizeof(int));
}

void orc_interp_init(int degree, int weight_rule, double* tab) { interp_init(degree, weight_rule, tab); }
void orc_cc_weights_textbook(int degree, double* w) { cc_weights_textbook(degree, w); }

int orc_interp_vals_bli(double* basis, double xi, double eta, double min_xi, double max_xi,
                        double min_eta, double max_eta, int deg) {
  return interp_vals_bli(basis, xi, eta, min_xi, max_xi, min_eta, max_eta, deg);
}
int orc_face_from_xyz(double x, double y, double z) { return face_from_xyz(x, y, z); }
int orc_xyz_from_xieta(double xi, double eta, int face, double* xyz) { return xyz_from_xieta(xi, eta, face, xyz); }
void orc_xieta_from_xyz(double x, double y, double z, double* xieta) { xieta_from_xyz(x, y, z, xieta); }
int orc_xieta_from_xyz_face(double x, double y, double z, int face, double* xieta) {
  return xieta_from_xyz_face(x, y, z, face, xieta);
}
double orc_sphere_tri_area(const double* a, const double* b, const double* c) { return sphere_tri_area(a, b, c); }
int orc_point_locate_panel(void* h, double x, double y, double z) {
  return point_locate_panel(*static_cast<Ctx*>(h), x, y, z);
}

// initialize_cubed_sphere.cpp:254-277 on the particle grid (degree m-1).
// Weights: cc_i*cc_j*panel_area/4 (:275), or 4 pi/N for weight_rule 2.
// Outputs [Pl][m2].
void orc_spec_points(void* h, double* x, double* y, double* z, double* area) {
  Ctx& c = *static_cast<Ctx*>(h);
  double N = (double)c.Pl * c.m2;
  for (int i = c.offset; i < c.P; i++) {
    const Panel& p = c.pn[i];
    double max_xi = kPi / 4.0 * p.max_xi, min_xi = kPi / 4.0 * p.min_xi;
    double max_eta = kPi / 4.0 * p.max_eta, min_eta = kPi / 4.0 * p.min_eta;
    double xr = 0.5 * (max_xi - min_xi), xo = 0.5 * (max_xi + min_xi);
    double er = 0.5 * (max_eta - min_eta), eo = 0.5 * (max_eta + min_eta);
    for (int j = 0; j < c.m2; j++) {
      double xyz[3];
      xyz_from_xieta(xo + xr * c.tab_m[4 * j], eo + er * c.tab_m[4 * j + 1], p.face, xyz);
      size_t o = (size_t)(i - c.offset) * c.m2 + j;
      x[o] = xyz[0]; y[o] = xyz[1]; z[o] = xyz[2];
      area[o] = c.weight_rule == 2 ? 4.0 * kPi / N : c.tab_m[4 * j + 3] * p.area / 4.0;
    }
  }
}

// Proxy node positions of every panel, [P][n] (the coordinates the pair
// kernels recompute per pair, bve_vel_impl.hpp:50-56).
void orc_proxy_points(void* h, double* x, double* y, double* z) {
  Ctx& c = *static_cast<Ctx*>(h);
  for (int i = 0; i < c.P; i++) {
    double xo, xs, eo, es, xyz[3];
    panel_frame(c.pn[i], xo, xs, eo, es);
    for (int j = 0; j < c.n; j++) {
      xyz_from_xieta(c.tab_d[4 * j] * xs + xo, c.tab_d[4 * j + 1] * es + eo, c.pn[i].face, xyz);
      x[(size_t)i * c.n + j] = xyz[0]; y[(size_t)i * c.n + j] = xyz[1]; z[(size_t)i * c.n + j] = xyz[2];
    }
  }
}

// fmm_funcs.cpp:341-355 with base_pots (fmm_funcs_impl.hpp:29-33) and
// child_to_parent (:47-70).  For m != d+1 the leaf proxies come from BLI
// anterpolation of the particle charges (base_pots_ll, :207-223).
// q_part[Pl][m2] (= area*pots), q_proxy[P][n]; every row is assigned.
int orc_upward_pass(void* h, const double* area, const double* pots, double* q_part, double* q_proxy) {
  Ctx& c = *static_cast<Ctx*>(h);
  int err = 0;
  for (size_t k = 0; k < (size_t)c.Pl * c.m2; k++) q_part[k] = area[k] * pots[k];
  std::memset(q_proxy, 0, sizeof(double) * (size_t)c.P * c.n);
  if (c.md == c.d) {
    std::memcpy(q_proxy + (size_t)c.offset * c.n, q_part, sizeof(double) * (size_t)c.Pl * c.n);
  } else {
#pragma omp parallel for schedule(dynamic, 16) reduction(| : err)
    for (int i = 0; i < c.Pl; i++) {
      const Panel& p = c.pn[c.offset + i];
      double xo, xs, eo, es, basis[121];
      panel_frame(p, xo, xs, eo, es);
      double* out = q_proxy + (size_t)(c.offset + i) * c.n;
      for (int j = 0; j < c.m2; j++) {
        double xi = c.tab_m[4 * j] * xs + xo, eta = c.tab_m[4 * j + 1] * es + eo;
        err |= bli_in_panel(p, xi, eta, c.d, basis);
        double qj = q_part[(size_t)i * c.m2 + j];
        for (int k = 0; k < c.n; k++) out[k] += basis[k] * qj;
      }
    }
  }
  // M2M, finest children first: children at level L-1 .. 1.
  for (int lev = c.levels - 1; lev >= 1; lev--) {
    int pb = c.level_start[lev - 1], pe = c.level_start[lev];
#pragma omp parallel for schedule(dynamic, 16) reduction(| : err)
    for (int par = pb; par < pe; par++) {
      const Panel& pp = c.pn[par];
      double basis[121];
      double* out = q_proxy + (size_t)par * c.n;
      for (int cc = 0; cc < 4; cc++) {
        int k = pp.child[cc];
        double xo, xs, eo, es;
        panel_frame(c.pn[k], xo, xs, eo, es);
        for (int i = 0; i < c.n; i++) {
          double xi = c.tab_d[4 * i] * xs + xo, eta = c.tab_d[4 * i + 1] * es + eo;
          err |= bli_in_panel(pp, xi, eta, c.d, basis);
          double qi = q_proxy[(size_t)k * c.n + i];
          for (int j = 0; j < c.n; j++) out[j] += basis[j] * qi;
        }
      }
    }
  }
  return err ? ORC_E_GEOMETRY : ORC_OK;
}

// parent_to_child_3 (fmm_funcs_impl.hpp:138-171) over every level
// (fmm_funcs.cpp:378-389): ncomp fields phi_proxy[ncomp][P][n] accumulated top-down.
static int l2l_sweep(Ctx& c, int ncomp, double* const* phi_proxy) {
  int err = 0;
  for (int lev = 0; lev < c.levels - 1; lev++) {
    int pb = c.level_start[lev], pe = c.level_start[lev + 1];
#pragma omp parallel for schedule(dynamic, 16) reduction(| : err)
    for (int i = pb; i < pe; i++) {
      const Panel& pp = c.pn[i];
      double basis[121];
      for (int cc = 0; cc < 4; cc++) {
        int child = pp.child[cc];
        double xo, xs, eo, es;
        panel_frame(c.pn[child], xo, xs, eo, es);
        for (int k = 0; k < c.n; k++) {
          double xi = c.tab_d[4 * k] * xs + xo, eta = c.tab_d[4 * k + 1] * es + eo;
          err |= bli_in_panel(pp, xi, eta, c.d, basis);
          for (int f = 0; f < ncomp; f++) {
            double acc = 0;
            const double* par = phi_proxy[f] + (size_t)i * c.n;
            for (int l = 0; l < c.n; l++) acc += basis[l] * par[l];
            phi_proxy[f][(size_t)child * c.n + k] += acc;
          }
        }
      }
    }
  }
  return err;
}

// fmm_funcs.cpp:374-392 (downward_pass_3): the L2L sweep, then child_panel_copy
// (fmm_funcs_impl.hpp:182-187).  sols[ncomp][Pl][m2] are assigned
// = L2P(leaf proxies) + phi_part (phi_part may be null).
int orc_downward_pass(void* h, int ncomp, double* const* phi_proxy, const double* const* phi_part,
                      double* const* sols) {
  Ctx& c = *static_cast<Ctx*>(h);
  int err = l2l_sweep(c, ncomp, phi_proxy);
#pragma omp parallel for schedule(dynamic, 16) reduction(| : err)
  for (int i = 0; i < c.Pl; i++) {
    const Panel& p = c.pn[c.offset + i];
    double xo, xs, eo, es, basis[121];
    panel_frame(p, xo, xs, eo, es);
    for (int j = 0; j < c.m2; j++) {
      size_t o = (size_t)i * c.m2 + j;
      if (c.md == c.d) {
        for (int f = 0; f < ncomp; f++) sols[f][o] = phi_proxy[f][(size_t)(c.offset + i) * c.n + j];
      } else {
        double xi = c.tab_m[4 * j] * xs + xo, eta = c.tab_m[4 * j + 1] * es + eo;
        err |= bli_in_panel(p, xi, eta, c.d, basis);
        for (int f = 0; f < ncomp; f++) {
          double acc = 0;
          const double* lp = phi_proxy[f] + (size_t)(c.offset + i) * c.n;
          for (int k = 0; k < c.n; k++) acc += basis[k] * lp[k];
          sols[f][o] = acc;
        }
      }
      if (phi_part)
        for (int f = 0; f < ncomp; f++) sols[f][o] += phi_part[f][o];
    }
  }
  return err ? ORC_E_GEOMETRY : ORC_OK;
}

// Point set of one side of an interaction: particles of a leaf (types 0/1 for
// targets, 0/2 for sources) or the proxies of a panel.
struct Side {
  int count, deg;
  const double* tab;
  bool particles;
};

static inline Side side_of(const Ctx& c, int panel, bool particles) {
  Side s;
  s.particles = particles;
  s.count = particles ? c.m2 : c.n;
  s.deg = particles ? c.md : c.d;
  s.tab = particles ? c.tab_m.data() : c.tab_d.data();
  (void)panel;
  return s;
}

// fmm_interactions/bve_vel.cpp:9-37 + bve_vel_impl.hpp:34-70, with the
// type-dependent point sets of swe_vel_impl.hpp:188-255.  Sources sit at
// P(node + disp) (bve_vel_impl.hpp:54-60); targets are undisplaced.
// phi_part[3][Pl][m2] and phi_proxy[3][P][n] are accumulated.
// Target panels t with t % stride == 0 only (stride 1: all); used to time a
// bounded sample of the interaction phase.
void orc_bve_vel_interactions_sample(void* h, const double* const* disp_part, const double* const* disp_proxy,
                                     const double* q_part, const double* q_proxy, double eps,
                                     double* const* phi_part, double* const* phi_proxy, int stride) {
  Ctx& c = *static_cast<Ctx*>(h);
#pragma omp parallel for schedule(dynamic, 1)
  for (int t = 0; t < c.P; t++) {
    if (t % stride) continue;
    for (int e = c.tgt_start[t]; e < c.tgt_start[t + 1]; e++) {
      int ii = c.by_target[e];
      int source = c.ilist[3 * ii + 1], type = c.ilist[3 * ii + 2];
      bool tpart = (type == 0 || type == 1), spart = (type == 0 || type == 2);
      Side ts = side_of(c, t, tpart), ss = side_of(c, source, spart);
      const Panel &pt = c.pn[t], &ps = c.pn[source];
      double xot, xst, eot, est, xos, xss, eos, ess;
      panel_frame(pt, xot, xst, eot, est);
      panel_frame(ps, xos, xss, eos, ess);
      double* o1; double* o2; double* o3;
      if (tpart) {
        size_t b = (size_t)(t - c.offset) * c.m2;
        o1 = phi_part[0] + b; o2 = phi_part[1] + b; o3 = phi_part[2] + b;
      } else {
        size_t b = (size_t)t * c.n;
        o1 = phi_proxy[0] + b; o2 = phi_proxy[1] + b; o3 = phi_proxy[2] + b;
      }
      const double *sq, *dx, *dy, *dz;
      if (spart) {
        size_t b = (size_t)(source - c.offset) * c.m2;
        sq = q_part + b; dx = disp_part[0] + b; dy = disp_part[1] + b; dz = disp_part[2] + b;
      } else {
        size_t b = (size_t)source * c.n;
        sq = q_proxy + b; dx = disp_proxy[0] + b; dy = disp_proxy[1] + b; dz = disp_proxy[2] + b;
      }
      double xt[3], xs_[3];
      for (int j = 0; j < ts.count; j++) {
        double xi_t = ts.tab[4 * j] * xst + xot, eta_t = ts.tab[4 * j + 1] * est + eot;
        xyz_from_xieta(xi_t, eta_t, pt.face, xt);
        double a1 = 0, a2 = 0, a3 = 0;
        for (int k = 0; k < ss.count; k++) {
          double xi_s = ss.tab[4 * k] * xss + xos, eta_s = ss.tab[4 * k + 1] * ess + eos;
          xyz_from_xieta(xi_s, eta_s, ps.face, xs_);
          double x_s = xs_[0] + dx[k], y_s = xs_[1] + dy[k], z_s = xs_[2] + dz[k];
          project_to_sphere(x_s, y_s, z_s);
          double dp = xt[0] * x_s + xt[1] * y_s + xt[2] * z_s;
          double g = kNegInv4Pi / (1.0 - dp + eps) * sq[k];
          a1 += g * (xt[1] * z_s - xt[2] * y_s);
          a2 += g * (xt[2] * x_s - xt[0] * z_s);
          a3 += g * (xt[0] * y_s - xt[1] * x_s);
        }
        o1[j] += a1; o2[j] += a2; o3[j] += a3;
      }
    }
  }
}

void orc_bve_vel_interactions(void* h, const double* const* disp_part, const double* const* disp_proxy,
                              const double* q_part, const double* q_proxy, double eps,
                              double* const* phi_part, double* const* phi_proxy) {
  orc_bve_vel_interactions_sample(h, disp_part, disp_proxy, q_part, q_proxy, eps, phi_part, phi_proxy, 1);
}

// fmm_interactions/swe_vel.cpp:9-67 with the pair kernels of
// swe_vel_impl.hpp:12-76 (variant 1, swe_vel_panel_interaction) and :78-147
// (variant 2, swe_vel_panel_interaction_2), on the type-dependent point sets.
// Source y = P(node + disp), y0 = the undisplaced node.
//   variant 1: g = -1/(4 pi (1 - x.y + eps));
//              u += g qz (x cross y) + g qd (y - x (x.y))
//   variant 2: u += -g' qz (x cross y) + g'' qd (y0 - x (x.y0)),
//              g' = 1/(4 pi (1 - x.y + eps)), g'' = 1/(4 pi (1 - x.y0 + eps))
// phi_part[3][Pl][m2] and phi_proxy[3][P][n] are accumulated.
void orc_swe_vel_interactions(void* h, int variant, const double* const* disp_part, const double* const* disp_proxy,
                              const double* qz_part, const double* qz_proxy, const double* qd_part,
                              const double* qd_proxy, double eps, double* const* phi_part, double* const* phi_proxy) {
  Ctx& c = *static_cast<Ctx*>(h);
  const double kInv4Pi = -kNegInv4Pi;
#pragma omp parallel for schedule(dynamic, 1)
  for (int t = 0; t < c.P; t++) {
    for (int e = c.tgt_start[t]; e < c.tgt_start[t + 1]; e++) {
      int ii = c.by_target[e];
      int source = c.ilist[3 * ii + 1], type = c.ilist[3 * ii + 2];
      bool tpart = (type == 0 || type == 1), spart = (type == 0 || type == 2);
      Side ts = side_of(c, t, tpart), ss = side_of(c, source, spart);
      const Panel &pt = c.pn[t], &ps = c.pn[source];
      double xot, xst, eot, est, xos, xss, eos, ess;
      panel_frame(pt, xot, xst, eot, est);
      panel_frame(ps, xos, xss, eos, ess);
      double *o1, *o2, *o3;
      if (tpart) {
        size_t b = (size_t)(t - c.offset) * c.m2;
        o1 = phi_part[0] + b; o2 = phi_part[1] + b; o3 = phi_part[2] + b;
      } else {
        size_t b = (size_t)t * c.n;
        o1 = phi_proxy[0] + b; o2 = phi_proxy[1] + b; o3 = phi_proxy[2] + b;
      }
      const double *qz, *qd, *dx, *dy, *dz;
      if (spart) {
        size_t b = (size_t)(source - c.offset) * c.m2;
        qz = qz_part + b; qd = qd_part + b;
        dx = disp_part[0] + b; dy = disp_part[1] + b; dz = disp_part[2] + b;
      } else {
        size_t b = (size_t)source * c.n;
        qz = qz_proxy + b; qd = qd_proxy + b;
        dx = disp_proxy[0] + b; dy = disp_proxy[1] + b; dz = disp_proxy[2] + b;
      }
      double xt[3], y0[3];
      for (int j = 0; j < ts.count; j++) {
        xyz_from_xieta(ts.tab[4 * j] * xst + xot, ts.tab[4 * j + 1] * est + eot, pt.face, xt);
        double a1 = 0, a2 = 0, a3 = 0;
        for (int k = 0; k < ss.count; k++) {
          xyz_from_xieta(ss.tab[4 * k] * xss + xos, ss.tab[4 * k + 1] * ess + eos, ps.face, y0);
          double x_s = y0[0] + dx[k], y_s = y0[1] + dy[k], z_s = y0[2] + dz[k];
          project_to_sphere(x_s, y_s, z_s);
          const double dp = xt[0] * x_s + xt[1] * y_s + xt[2] * z_s;
          double g1, g2, v0, v1, v2;
          if (variant == 1) {
            const double g = kNegInv4Pi / (1.0 - dp + eps);
            g1 = g * qz[k];
            g2 = g * qd[k];
            v0 = x_s - xt[0] * dp; v1 = y_s - xt[1] * dp; v2 = z_s - xt[2] * dp;
          } else {
            const double dp2 = xt[0] * y0[0] + xt[1] * y0[1] + xt[2] * y0[2];
            g1 = -(kInv4Pi / (1.0 - dp + eps)) * qz[k];
            g2 = kInv4Pi / (1.0 - dp2 + eps) * qd[k];
            v0 = y0[0] - xt[0] * dp2; v1 = y0[1] - xt[1] * dp2; v2 = y0[2] - xt[2] * dp2;
          }
          a1 += g1 * (xt[1] * z_s - xt[2] * y_s) + g2 * v0;
          a2 += g1 * (xt[2] * x_s - xt[0] * z_s) + g2 * v1;
          a3 += g1 * (xt[0] * y_s - xt[1] * x_s) + g2 * v2;
        }
        o1[j] += a1; o2[j] += a2; o3[j] += a3;
      }
    }
  }
}

// fmm_interactions/inverse_poisson.cpp:9-33 + inverse_poisson_impl.hpp:23-52,
// with the type-dependent point sets.  psi_part[Pl][m2], psi_proxy[P][n] accumulated.
void orc_poisson_fmm_interactions(void* h, const double* q_part, const double* q_proxy,
                                  double* psi_part, double* psi_proxy) {
  Ctx& c = *static_cast<Ctx*>(h);
#pragma omp parallel for schedule(dynamic, 4)
  for (int t = 0; t < c.P; t++) {
    for (int e = c.tgt_start[t]; e < c.tgt_start[t + 1]; e++) {
      int ii = c.by_target[e];
      int source = c.ilist[3 * ii + 1], type = c.ilist[3 * ii + 2];
      bool tpart = (type == 0 || type == 1), spart = (type == 0 || type == 2);
      Side ts = side_of(c, t, tpart), ss = side_of(c, source, spart);
      const Panel &pt = c.pn[t], &ps = c.pn[source];
      double xot, xst, eot, est, xos, xss, eos, ess;
      panel_frame(pt, xot, xst, eot, est);
      panel_frame(ps, xos, xss, eos, ess);
      double* o = tpart ? psi_part + (size_t)(t - c.offset) * c.m2 : psi_proxy + (size_t)t * c.n;
      const double* sq = spart ? q_part + (size_t)(source - c.offset) * c.m2 : q_proxy + (size_t)source * c.n;
      double xt[3], xs_[3];
      for (int j = 0; j < ts.count; j++) {
        xyz_from_xieta(ts.tab[4 * j] * xst + xot, ts.tab[4 * j + 1] * est + eot, pt.face, xt);
        double acc = 0;
        for (int k = 0; k < ss.count; k++) {
          xyz_from_xieta(ss.tab[4 * k] * xss + xos, ss.tab[4 * k + 1] * ess + eos, ps.face, xs_);
          double dp = xt[0] * xs_[0] + xt[1] * xs_[1] + xt[2] * xs_[2];
          if (dp < 1.0 - 1e-15) acc += kNegInv4Pi * std::log(1 - dp) * sq[k];
        }
        o[j] += acc;
      }
    }
  }
}

// direct_sum.cpp:8-15 + direct_sum_impl.hpp:18-27: soln[i] += sum_{j != i}
// -1/(4 pi) log(1 - x_i.x_j) area_j pots_j for i in [lb, ub).
void orc_direct_sum_inv_lap(int64_t N, const double* x, const double* y, const double* z,
                            const double* area, const double* pots, double* soln, int64_t lb, int64_t ub) {
#pragma omp parallel for schedule(static)
  for (int64_t i = lb; i < ub; i++) {
    double acc = 0;
    for (int64_t j = 0; j < N; j++) {
      if (i != j) {
        double dp = x[i] * x[j] + y[i] * y[j] + z[i] * z[j];
        acc += kNegInv4Pi * std::log(1 - dp) * area[j] * pots[j];
      }
    }
    soln[i] += acc;
  }
}

// Biot-Savart direct sum with the pair kernel of bve_vel_impl.hpp:54-66 (no
// reference counterpart: the reference has only the log direct sum).  Sources
// optionally displaced (disp may be null); vel[3][N] accumulated.
void orc_direct_sum_bve_vel(int64_t N, const double* x, const double* y, const double* z,
                            const double* area, const double* vor, const double* const* disp, double eps,
                            double* const* vel, int64_t lb, int64_t ub) {
  std::vector<double> sx(N), sy(N), sz(N);
  for (int64_t j = 0; j < N; j++) {
    double a = x[j], b = y[j], cc = z[j];
    if (disp) { a += disp[0][j]; b += disp[1][j]; cc += disp[2][j]; project_to_sphere(a, b, cc); }
    sx[j] = a; sy[j] = b; sz[j] = cc;
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = lb; i < ub; i++) {
    double a1 = 0, a2 = 0, a3 = 0;
    for (int64_t j = 0; j < N; j++) {
      double dp = x[i] * sx[j] + y[i] * sy[j] + z[i] * sz[j];
      double g = kNegInv4Pi / (1.0 - dp + eps) * area[j] * vor[j];
      a1 += g * (y[i] * sz[j] - z[i] * sy[j]);
      a2 += g * (z[i] * sx[j] - x[i] * sz[j]);
      a3 += g * (x[i] * sy[j] - y[i] * sx[j]);
    }
    vel[0][i] += a1; vel[1][i] += a2; vel[2][i] += a3;
  }
}

// Log direct sum at separate target points; pairs with x.y >= 1 - 1e-15 are
// skipped (the FMM kernel's guard, inverse_poisson_impl.hpp:47).
void orc_direct_sum_inv_lap_targets(int64_t NT, const double* tx, const double* ty, const double* tz,
                                    int64_t NS, const double* sx, const double* sy, const double* sz,
                                    const double* area, const double* pots, double* soln) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < NT; i++) {
    double acc = 0;
    for (int64_t j = 0; j < NS; j++) {
      double dp = tx[i] * sx[j] + ty[i] * sy[j] + tz[i] * sz[j];
      if (dp < 1.0 - 1e-15) acc += kNegInv4Pi * std::log(1 - dp) * area[j] * pots[j];
    }
    soln[i] += acc;
  }
}

// Biot-Savart direct sum at separate target points over undisplaced sources,
// the pair kernel of bve_vel_impl.hpp:63-66; vel[3][NT] accumulated.
void orc_direct_sum_bve_vel_targets(int64_t NT, const double* tx, const double* ty, const double* tz, int64_t NS,
                                    const double* sx, const double* sy, const double* sz, const double* area,
                                    const double* vor, double eps, double* const* vel) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < NT; i++) {
    double a1 = 0, a2 = 0, a3 = 0;
    for (int64_t j = 0; j < NS; j++) {
      double dp = tx[i] * sx[j] + ty[i] * sy[j] + tz[i] * sz[j];
      double g = kNegInv4Pi / (1.0 - dp + eps) * area[j] * vor[j];
      a1 += g * (ty[i] * sz[j] - tz[i] * sy[j]);
      a2 += g * (tz[i] * sx[j] - tx[i] * sz[j]);
      a3 += g * (tx[i] * sy[j] - ty[i] * sx[j]);
    }
    vel[0][i] += a1; vel[1][i] += a2; vel[2][i] += a3;
  }
}

// FMM evaluation at arbitrary points: the `_ll` pass structure
// (fmm_funcs.cpp:394-419) with targets that are not sources.  A target in
// leaf L gets
//   far:  the degree-d BLI of L's proxy potentials after the L2L sweep
//         (child_panel_interp_3_ll, fmm_funcs_impl.hpp:318-364), and
//   near: the kernel over the sources of L's own interactions, types 0 and 1
//         (the particle-target point sets of swe_vel_impl.hpp:188-255).
// kind 0: Poisson, -1/(4 pi) log(1 - x.y) with the dp < 1 - 1e-15 guard on
// unprojected nodes (inverse_poisson_impl.hpp:42-48); out[0][NT].
// kind 1: Biot-Savart with undisplaced, re-projected sources
// (bve_vel_impl.hpp:54-66); out[0..2][NT].
int orc_eval_targets(void* h, int kind, const double* area, const double* f, double eps, int64_t NT,
                     const double* tx, const double* ty, const double* tz, double* const* out) {
  Ctx& c = *static_cast<Ctx*>(h);
  const size_t NP = (size_t)c.Pl * c.m2, Pn = (size_t)c.P * c.n;
  const int nc = kind == 0 ? 1 : 3;
  std::vector<double> qp(NP), qx(Pn);
  int err = orc_upward_pass(h, area, f, qp.data(), qx.data()) != ORC_OK;
  std::vector<double> pp(3 * NP, 0.0), px(3 * Pn, 0.0);
  double* ppp[3] = {pp.data(), pp.data() + NP, pp.data() + 2 * NP};
  double* pxp[3] = {px.data(), px.data() + Pn, px.data() + 2 * Pn};
  if (kind == 0) {
    orc_poisson_fmm_interactions(h, qp.data(), qx.data(), ppp[0], pxp[0]);
  } else {
    std::vector<double> zp(NP, 0.0), zx(Pn, 0.0);
    const double* dp[3] = {zp.data(), zp.data(), zp.data()};
    const double* dx[3] = {zx.data(), zx.data(), zx.data()};
    orc_bve_vel_interactions(h, dp, dx, qp.data(), qx.data(), eps, ppp, pxp);
  }
  err |= l2l_sweep(c, nc, pxp);
#pragma omp parallel for schedule(dynamic, 8) reduction(| : err)
  for (int64_t i = 0; i < NT; i++) {
    const double xt[3] = {tx[i], ty[i], tz[i]};
    double xieta[2], basis[121], xs_[3];
    xieta_from_xyz(xt[0], xt[1], xt[2], xieta);
    const int leaf = point_locate_panel(c, xt[0], xt[1], xt[2]);
    for (int k = 0; k < nc; k++) out[k][i] = 0;
    if (leaf < 0) {
      err |= 1;
      continue;
    }
    err |= bli_in_panel(c.pn[leaf], xieta[0], xieta[1], c.d, basis);
    double acc[3] = {0, 0, 0};
    for (int k = 0; k < c.n; k++)
      for (int q = 0; q < nc; q++) acc[q] += basis[k] * pxp[q][(size_t)leaf * c.n + k];
    for (int e = c.tgt_start[leaf]; e < c.tgt_start[leaf + 1]; e++) {
      const int ii = c.by_target[e];
      const int source = c.ilist[3 * ii + 1], type = c.ilist[3 * ii + 2];
      const bool spart = type == 0 || type == 2;
      const Side ss = side_of(c, source, spart);
      double xos, xss, eos, ess;
      panel_frame(c.pn[source], xos, xss, eos, ess);
      const double* sq = spart ? qp.data() + (size_t)(source - c.offset) * c.m2 : qx.data() + (size_t)source * c.n;
      for (int k = 0; k < ss.count; k++) {
        xyz_from_xieta(ss.tab[4 * k] * xss + xos, ss.tab[4 * k + 1] * ess + eos, c.pn[source].face, xs_);
        if (kind == 0) {
          const double dp = xt[0] * xs_[0] + xt[1] * xs_[1] + xt[2] * xs_[2];
          if (dp < 1.0 - 1e-15) acc[0] += kNegInv4Pi * std::log(1 - dp) * sq[k];
        } else {
          double x_s = xs_[0], y_s = xs_[1], z_s = xs_[2];
          project_to_sphere(x_s, y_s, z_s);
          const double dp = xt[0] * x_s + xt[1] * y_s + xt[2] * z_s;
          const double g = kNegInv4Pi / (1.0 - dp + eps) * sq[k];
          acc[0] += g * (xt[1] * z_s - xt[2] * y_s);
          acc[1] += g * (xt[2] * x_s - xt[0] * z_s);
          acc[2] += g * (xt[0] * y_s - xt[1] * x_s);
        }
      }
    }
    for (int q = 0; q < nc; q++) out[q][i] = acc[q];
  }
  return err ? ORC_E_GEOMETRY : ORC_OK;
}

// bve_time_step.cpp:91-104 (bve_disp_update) and :131-159 (bve_disp_interp).
// Leaf particles: new_vor = vor + 2 Omega z - 2 Omega z', disp_part = tau v.
// Non-leaf proxies: disp = BLI (particle-grid degree) of the leaf particle
// displacement at the undisplaced proxy node, in the leaf found by point location.
int orc_bve_displacement(void* h, const double* x, const double* y, const double* z,
                         const double* const* vel, const double* vor, double tau, double omega,
                         double* new_vor, double* const* disp_part, double* const* disp_proxy) {
  Ctx& c = *static_cast<Ctx*>(h);
  int err = 0;
  for (size_t k = 0; k < (size_t)c.Pl * c.m2; k++) {
    double a = x[k] + vel[0][k] * tau, b = y[k] + vel[1][k] * tau, cc = z[k] + vel[2][k] * tau;
    project_to_sphere(a, b, cc);
    new_vor[k] = vor[k] + 2 * omega * z[k] - 2 * omega * cc;
    disp_part[0][k] = tau * vel[0][k];
    disp_part[1][k] = tau * vel[1][k];
    disp_part[2][k] = tau * vel[2][k];
  }
#pragma omp parallel for schedule(dynamic, 16) reduction(| : err)
  for (int i = 0; i < c.offset; i++) {
    double xo, xs, eo, es, xyz[3], basis[121];
    panel_frame(c.pn[i], xo, xs, eo, es);
    for (int j = 0; j < c.n; j++) {
      double xi = c.tab_d[4 * j] * xs + xo, eta = c.tab_d[4 * j + 1] * es + eo;
      xyz_from_xieta(xi, eta, c.pn[i].face, xyz);
      int leaf = point_locate_panel(c, xyz[0], xyz[1], xyz[2]);
      double a[3] = {0, 0, 0};
      if (leaf < 0) { err |= ORC_E_GEOMETRY; }
      else {
        err |= bli_in_panel(c.pn[leaf], xi, eta, c.md, basis);
        size_t b = (size_t)(leaf - c.offset) * c.m2;
        for (int k = 0; k < c.m2; k++)
          for (int f = 0; f < 3; f++) a[f] += disp_part[f][b + k] * basis[k];
      }
      for (int f = 0; f < 3; f++) disp_proxy[f][(size_t)i * c.n + j] = a[f];
    }
  }
  return err ? ORC_E_GEOMETRY : ORC_OK;
}

// Locate + BLI-interpolate nf fields [Pl][m2] at a point (bve_time_step.cpp:220-237).
static int interp_at(const Ctx& c, double x, double y, double z, int nf, const double* const* fields,
                     double* out) {
  double xieta[2], basis[121];
  xieta_from_xyz(x, y, z, xieta);
  int leaf = point_locate_panel(c, x, y, z);
  for (int f = 0; f < nf; f++) out[f] = 0;
  if (leaf < 0) return ORC_E_GEOMETRY;
  int err = bli_in_panel(c.pn[leaf], xieta[0], xieta[1], c.md, basis);
  size_t b = (size_t)(leaf - c.offset) * c.m2;
  for (int k = 0; k < c.m2; k++)
    for (int f = 0; f < nf; f++) out[f] += fields[f][b + k] * basis[k];
  return err;
}

// bve_time_step.cpp:206-308.  vels = {v0x,v0y,v0z, v1x.., v2x.., v3x..}.
// combine 0: the reference's sequential combine (:288-303: x - dt/6 k1, then
// + dt/3 k2, + dt/3 k3, + dt/6 k4, projecting after each); 1: textbook
// x - dt/6 (k1 + 2 k2 + 2 k3 + k4), projected once.
int orc_bve_find_departure_points(void* h, const double* x, const double* y, const double* z,
                                  const double* const* vels, double dt, int combine,
                                  double* dx, double* dy, double* dz) {
  Ctx& c = *static_cast<Ctx*>(h);
  int err = 0;
  size_t N = (size_t)c.Pl * c.m2;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : err)
  for (size_t k = 0; k < N; k++) {
    double k1[3] = {vels[9][k], vels[10][k], vels[11][k]}, k2[3], k3[3], k4[3];
    double a = x[k] - 0.5 * dt * k1[0], b = y[k] - 0.5 * dt * k1[1], cc = z[k] - 0.5 * dt * k1[2];
    project_to_sphere(a, b, cc);
    err |= interp_at(c, a, b, cc, 3, vels + 6, k2);
    a = x[k] - 0.5 * dt * k2[0]; b = y[k] - 0.5 * dt * k2[1]; cc = z[k] - 0.5 * dt * k2[2];
    project_to_sphere(a, b, cc);
    err |= interp_at(c, a, b, cc, 3, vels + 3, k3);
    a = x[k] - dt * k3[0]; b = y[k] - dt * k3[1]; cc = z[k] - dt * k3[2];
    project_to_sphere(a, b, cc);
    err |= interp_at(c, a, b, cc, 3, vels + 0, k4);
    if (combine == 0) {
      a = x[k] - 1.0 / 6.0 * dt * k1[0]; b = y[k] - 1.0 / 6.0 * dt * k1[1]; cc = z[k] - 1.0 / 6.0 * dt * k1[2];
      project_to_sphere(a, b, cc);
      a -= -1.0 / 3.0 * dt * k2[0]; b -= -1.0 / 3.0 * dt * k2[1]; cc -= -1.0 / 3.0 * dt * k2[2];
      project_to_sphere(a, b, cc);
      a -= -1.0 / 3.0 * dt * k3[0]; b -= -1.0 / 3.0 * dt * k3[1]; cc -= -1.0 / 3.0 * dt * k3[2];
      project_to_sphere(a, b, cc);
      a -= -1.0 / 6.0 * dt * k4[0]; b -= -1.0 / 6.0 * dt * k4[1]; cc -= -1.0 / 6.0 * dt * k4[2];
      project_to_sphere(a, b, cc);
    } else {
      a = x[k] - dt / 6.0 * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0]);
      b = y[k] - dt / 6.0 * (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1]);
      cc = z[k] - dt / 6.0 * (k1[2] + 2 * k2[2] + 2 * k3[2] + k4[2]);
      project_to_sphere(a, b, cc);
    }
    dx[k] = a; dy[k] = b; dz[k] = cc;
  }
  return err ? ORC_E_GEOMETRY : ORC_OK;
}

// bve_time_step.cpp:38-66: new_vor = BLI[vor](dep) + 2 Omega dep_z - 2 Omega z;
// tracers[Pl][m2][T] interpolated without the Coriolis terms.
int orc_bve_departure_to_target(void* h, const double* z, const double* dx, const double* dy, const double* dz,
                                const double* vor, const double* tracers, int ntr, double omega,
                                double* new_vor, double* new_tracers) {
  Ctx& c = *static_cast<Ctx*>(h);
  int err = 0;
  size_t N = (size_t)c.Pl * c.m2;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : err)
  for (size_t k = 0; k < N; k++) {
    double xieta[2], basis[121];
    xieta_from_xyz(dx[k], dy[k], dz[k], xieta);
    int leaf = point_locate_panel(c, dx[k], dy[k], dz[k]);
    double v = 0;
    for (int t = 0; t < ntr; t++) new_tracers[k * ntr + t] = 0;
    if (leaf < 0) { err |= ORC_E_GEOMETRY; new_vor[k] = 0; continue; }
    err |= bli_in_panel(c.pn[leaf], xieta[0], xieta[1], c.md, basis);
    size_t b = (size_t)(leaf - c.offset) * c.m2;
    for (int j = 0; j < c.m2; j++) {
      v += vor[b + j] * basis[j];
      for (int t = 0; t < ntr; t++) new_tracers[k * ntr + t] += tracers[(b + j) * ntr + t] * basis[j];
    }
    v += 2 * omega * dz[k];
    new_vor[k] = v - 2 * omega * z[k];
  }
  return err ? ORC_E_GEOMETRY : ORC_OK;
}

// swe_time_step.cpp:19-74 (swe_departure_to_target): vor, div, h [Pl][m2] and
// tracers [Pl][m2][T] interpolated at the departure points, no Coriolis terms.
int orc_swe_departure_to_target(void* h, const double* dx, const double* dy, const double* dz,
                                const double* const* fields, const double* tracers, int ntr,
                                double* const* new_fields, double* new_tracers) {
  Ctx& c = *static_cast<Ctx*>(h);
  int err = 0;
  size_t N = (size_t)c.Pl * c.m2;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : err)
  for (size_t k = 0; k < N; k++) {
    double xieta[2], basis[121];
    xieta_from_xyz(dx[k], dy[k], dz[k], xieta);
    int leaf = point_locate_panel(c, dx[k], dy[k], dz[k]);
    for (int f = 0; f < 3; f++) new_fields[f][k] = 0;
    for (int t = 0; t < ntr; t++) new_tracers[k * ntr + t] = 0;
    if (leaf < 0) { err |= ORC_E_GEOMETRY; continue; }
    err |= bli_in_panel(c.pn[leaf], xieta[0], xieta[1], c.md, basis);
    size_t b = (size_t)(leaf - c.offset) * c.m2;
    for (int j = 0; j < c.m2; j++) {
      for (int f = 0; f < 3; f++) new_fields[f][k] += fields[f][b + j] * basis[j];
      for (int t = 0; t < ntr; t++) new_tracers[k * ntr + t] += tracers[(b + j) * ntr + t] * basis[j];
    }
  }
  return err ? ORC_E_GEOMETRY : ORC_OK;
}

// One FMM velocity evaluation: upward pass, interactions, downward pass
// (bve_time_step.cpp:321-344).  vel[3][Pl][m2] assigned.
static int fmm_velocity(Ctx& c, const double* area, const double* vor_eff, const double* const* disp_part,
                        const double* const* disp_proxy, double eps, double* const* vel) {
  size_t NP = (size_t)c.Pl * c.m2, NQ = (size_t)c.P * c.n;
  std::vector<double> q_part(NP), q_proxy(NQ), pp(3 * NP, 0.0), px(3 * NQ, 0.0);
  int err = orc_upward_pass(&c, area, vor_eff, q_part.data(), q_proxy.data());
  double* php[3] = {pp.data(), pp.data() + NP, pp.data() + 2 * NP};
  double* phx[3] = {px.data(), px.data() + NQ, px.data() + 2 * NQ};
  orc_bve_vel_interactions(&c, disp_part, disp_proxy, q_part.data(), q_proxy.data(), eps, php, phx);
  const double* cphp[3] = {php[0], php[1], php[2]};
  err |= orc_downward_pass(&c, 3, phx, cphp, vel);
  return err;
}

// bve_back_rk4_step, bve_time_step.cpp:311-419 (no forcing: bve_forcing is a
// copy, forcing_funcs.cpp:93-95).  vor[Pl][m2] and tracers[Pl][m2][T] are
// updated in place; vel[3][Pl][m2] receives the stage-0 velocity.
int orc_bve_back_rk4_step(void* h, const double* x, const double* y, const double* z, const double* area,
                          double* vor, double* tracers, int ntr, double dt, double omega, double eps,
                          int combine, double* const* vel) {
  Ctx& c = *static_cast<Ctx*>(h);
  size_t NP = (size_t)c.Pl * c.m2, NQ = (size_t)c.P * c.n;
  std::vector<double> dpart(3 * NP, 0.0), dprox(3 * NQ, 0.0), vs(9 * NP), dvor(NP);
  const double* dp3[3] = {dpart.data(), dpart.data() + NP, dpart.data() + 2 * NP};
  const double* dx3[3] = {dprox.data(), dprox.data() + NQ, dprox.data() + 2 * NQ};
  double* dp3w[3] = {dpart.data(), dpart.data() + NP, dpart.data() + 2 * NP};
  double* dx3w[3] = {dprox.data(), dprox.data() + NQ, dprox.data() + 2 * NQ};
  double* v1[3] = {vs.data(), vs.data() + NP, vs.data() + 2 * NP};
  double* v2[3] = {vs.data() + 3 * NP, vs.data() + 4 * NP, vs.data() + 5 * NP};
  double* v3[3] = {vs.data() + 6 * NP, vs.data() + 7 * NP, vs.data() + 8 * NP};
  int err = fmm_velocity(c, area, vor, dp3, dx3, eps, vel);
  const double* cv0[3] = {vel[0], vel[1], vel[2]};
  err |= orc_bve_displacement(&c, x, y, z, cv0, vor, 0.5 * dt, omega, dvor.data(), dp3w, dx3w);
  err |= fmm_velocity(c, area, dvor.data(), dp3, dx3, eps, v1);
  const double* cv1[3] = {v1[0], v1[1], v1[2]};
  err |= orc_bve_displacement(&c, x, y, z, cv1, vor, 0.5 * dt, omega, dvor.data(), dp3w, dx3w);
  err |= fmm_velocity(c, area, dvor.data(), dp3, dx3, eps, v2);
  const double* cv2[3] = {v2[0], v2[1], v2[2]};
  err |= orc_bve_displacement(&c, x, y, z, cv2, vor, dt, omega, dvor.data(), dp3w, dx3w);
  err |= fmm_velocity(c, area, dvor.data(), dp3, dx3, eps, v3);
  const double* all[12] = {vel[0], vel[1], vel[2], v1[0], v1[1], v1[2], v2[0], v2[1], v2[2], v3[0], v3[1], v3[2]};
  std::vector<double> dep(3 * NP), nv(NP), ntr_buf((size_t)NP * (ntr > 0 ? ntr : 1));
  err |= orc_bve_find_departure_points(&c, x, y, z, all, dt, combine, dep.data(), dep.data() + NP, dep.data() + 2 * NP);
  err |= orc_bve_departure_to_target(&c, z, dep.data(), dep.data() + NP, dep.data() + 2 * NP, vor, tracers, ntr,
                                     omega, nv.data(), ntr_buf.data());
  std::memcpy(vor, nv.data(), NP * sizeof(double));
  if (ntr > 0) std::memcpy(tracers, ntr_buf.data(), NP * ntr * sizeof(double));
  return err ? ORC_E_GEOMETRY : ORC_OK;
}

// latlon_interp_host (interp_funcs.hpp:36-59) at arbitrary points: locate,
// BLI basis on the leaf, sum over the leaf's nodes; nf fields [Pl][m2].
int orc_interp_points(void* h, int64_t n, const double* x, const double* y, const double* z, int nf,
                      const double* const* fields, double* const* out) {
  const Ctx& c = *static_cast<Ctx*>(h);
  int err = 0;
#pragma omp parallel for reduction(| : err)
  for (int64_t i = 0; i < n; i++) {
    double v[16];
    err |= interp_at(c, x[i], y[i], z[i], nf, fields, v);
    for (int f = 0; f < nf; f++) out[f][i] = v[f];
  }
  return err ? ORC_E_GEOMETRY : ORC_OK;
}

// xyz_vel_to_uv_vel (general_utils_impl.hpp:211-228).
void orc_xyz_vel_to_uv_vel(int64_t n, const double* x, const double* y, const double* z, const double* vx,
                           const double* vy, const double* vz, double* u, double* v) {
  for (int64_t i = 0; i < n; i++) {
    if (std::abs(z[i]) < 1 - 1e-16) {
      u[i] = (-y[i] * vx[i] + x[i] * vy[i]) / std::sqrt(x[i] * x[i] + y[i] * y[i]);
      v[i] = -((x[i] * vx[i] + y[i] * vy[i]) * z[i] - (x[i] * x[i] + y[i] * y[i]) * vz[i]) /
             std::sqrt(x[i] * x[i] + y[i] * y[i]);
    } else {
      u[i] = 0;
      v[i] = 0;
    }
  }
}

// cubed_sphere_2d_to_1d (initialize_cubed_sphere.cpp:290-324), the reference's
// scan: each point in order is compared with every unique point found so far
// and joins the first within 1e-15 in all three coordinates.  O(N^2): small N.
int64_t orc_dedupe(int64_t n, const double* x, const double* y, const double* z, int* two_d_to_1d, int* counts,
                   double* x1d, double* y1d, double* z1d) {
  int64_t found = 0;
  for (int64_t i = 0; i < n; i++) {
    int64_t k = 0;
    for (; k < found; k++)
      if (std::abs(x[i] - x1d[k]) < 1e-15 && std::abs(y[i] - y1d[k]) < 1e-15 && std::abs(z[i] - z1d[k]) < 1e-15) break;
    if (k == found) {
      x1d[found] = x[i];
      y1d[found] = y[i];
      z1d[found] = z[i];
      counts[found] = 0;
      found++;
    }
    counts[k] += 1;
    two_d_to_1d[i] = (int)k;
  }
  return found;
}

// unify_boundary_vals (initialize_cubed_sphere.cpp:328-370): accumulate_vals sums every entry
// into its unique point (two_d_to_1d), average_vals gives every entry the sum over its unique
// point divided by that point's multiplicity.  Sequential sums in entry order.
int orc_unify_boundary_vals(int64_t n, const int* two_d_to_1d, const int* counts, int64_t n1d, double* vals) {
  std::vector<double> sum(n1d, 0.0);
  for (int64_t i = 0; i < n; ++i) {
    const int k = two_d_to_1d[i];
    if (k < 0 || k >= n1d) return ORC_E_ARG;
    sum[k] += vals[i];
  }
  for (int64_t i = 0; i < n; ++i) vals[i] = sum[two_d_to_1d[i]] / counts[two_d_to_1d[i]];
  return ORC_OK;
}

// ---- irregular particle tree and its _ll passes -------------------------------
// initialize_cube_sphere_irreg_points (initialize_cubed_sphere.cpp:446-592),
// dual_tree_traversal_irreg (fmm_funcs.cpp:125-231), base_pots_ll /
// child_to_parent_ll / parent_to_child_3_ll / child_panel_interp_3_ll
// (fmm_funcs_impl.hpp:190-364) and swe_vel_panel_interaction_ll
// (swe_vel_impl.hpp:149-288), restated sequentially: every functor is run
// panel by panel or pair by pair in index order (no atomics, so the race of
// child_to_parent_ll, fmm_funcs_impl.hpp:258, does not arise).
struct IrrPanel {
  int face, level, parent, child[4], is_leaf, point_count;
  double min_xi, max_xi, min_eta, max_eta, radius;
};
struct IrrCtx {
  std::vector<IrrPanel> pn;
  std::vector<std::vector<int>> pts;
  std::vector<int> level_start, ilist;
  std::vector<double> x, y, z;
  int levels = 0, degree = 0, leaf_size = 0;
  double theta = 0;
};

void* orc_irreg_create(int64_t n, const double* x, const double* y, const double* z, int leaf_size, int degree,
                       double theta) {
  IrrCtx* c = new IrrCtx();
  c->x.assign(x, x + n); c->y.assign(y, y + n); c->z.assign(z, z + n);
  c->degree = degree; c->leaf_size = leaf_size; c->theta = theta;
  auto& P = c->pn;
  P.resize(6);
  c->pts.resize(6);
  for (int i = 0; i < 6; i++) {
    P[i] = IrrPanel{i, 0, -1, {-1, -1, -1, -1}, 1, 0, -M_PI / 4.0, M_PI / 4.0, -M_PI / 4.0, M_PI / 4.0, 0};
  }
  for (int64_t i = 0; i < n; i++) {
    int f = face_from_xyz(x[i], y[i], z[i]);
    P[f].point_count += 1;
    c->pts[f].push_back((int)i);
  }
  for (size_t i = 0; i < P.size(); i++) {
    if (P[i].point_count > leaf_size) {
      P[i].is_leaf = 0;
      double min_xi = P[i].min_xi, max_xi = P[i].max_xi, min_eta = P[i].min_eta, max_eta = P[i].max_eta;
      double mid_xi = 0.5 * (min_xi + max_xi), mid_eta = 0.5 * (min_eta + max_eta);
      IrrPanel s1{}, s2{}, s3{}, s4{};
      s1.min_xi = mid_xi; s1.max_xi = max_xi; s1.min_eta = mid_eta; s1.max_eta = max_eta;
      s2.min_xi = min_xi; s2.max_xi = mid_xi; s2.min_eta = mid_eta; s2.max_eta = max_eta;
      s3.min_xi = min_xi; s3.max_xi = mid_xi; s3.min_eta = min_eta; s3.max_eta = mid_eta;
      s4.min_xi = mid_xi; s4.max_xi = max_xi; s4.min_eta = min_eta; s4.max_eta = mid_eta;
      int start = (int)P.size();
      P.push_back(s1); P.push_back(s2); P.push_back(s3); P.push_back(s4);
      for (int j = 0; j < 4; j++) {
        P[start + j].level = P[i].level + 1;
        P[start + j].face = P[i].face;
        P[start + j].parent = (int)i;
        P[start + j].point_count = 0;
        P[start + j].is_leaf = 1;
        for (int k = 0; k < 4; k++) P[start + j].child[k] = -1;
        P[i].child[j] = start + j;
      }
      c->pts.resize(start + 4);
      for (int idx : c->pts[i]) {
        double xieta[2];
        xieta_from_xyz(x[idx], y[idx], z[idx], xieta);
        int which;
        if (xieta[0] < mid_xi) which = xieta[1] < mid_eta ? 2 : 1;
        else which = xieta[1] < mid_eta ? 3 : 0;
        P[start + which].point_count += 1;
        c->pts[start + which].push_back(idx);
      }
    }
  }
  for (auto& p : P) {
    double mid_xi = (p.min_xi + p.max_xi) * 0.5, mid_eta = (p.min_eta + p.max_eta) * 0.5;
    double pc[3], p1[3], p2[3], p3[3], p4[3];
    xyz_from_xieta(mid_xi, mid_eta, p.face, pc);
    xyz_from_xieta(p.max_xi, p.min_eta, p.face, p1);
    xyz_from_xieta(p.max_xi, p.max_eta, p.face, p2);
    xyz_from_xieta(p.min_xi, p.max_eta, p.face, p3);
    xyz_from_xieta(p.min_xi, p.min_eta, p.face, p4);
    double d1 = std::acos(pc[0] * p1[0] + pc[1] * p1[1] + pc[2] * p1[2]);
    double d2 = std::acos(pc[0] * p2[0] + pc[1] * p2[1] + pc[2] * p2[2]);
    double d3 = std::acos(pc[0] * p3[0] + pc[1] * p3[1] + pc[2] * p3[2]);
    double d4 = std::acos(pc[0] * p4[0] + pc[1] * p4[1] + pc[2] * p4[2]);
    p.radius = std::max(std::max(std::max(d1, d2), d3), d4);
  }
  c->levels = 0;
  for (auto& p : P) c->levels = std::max(c->levels, p.level);
  c->level_start.assign(c->levels + 2, (int)P.size());
  int target = 0;
  for (size_t i = 0; i < P.size(); i++)
    if (P[i].level == target) { c->level_start[target] = (int)i; target++; }
  // traversal
  std::deque<int> tq, sq;
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) { tq.push_back(i); sq.push_back(j); }
  while (!tq.empty()) {
    int it = tq.front(), is = sq.front();
    tq.pop_front(); sq.pop_front();
    const IrrPanel &T = P[it], &S = P[is];
    double c1[3], c2[3];
    xyz_from_xieta(0.5 * (T.min_xi + T.max_xi), 0.5 * (T.min_eta + T.max_eta), T.face, c1);
    xyz_from_xieta(0.5 * (S.min_xi + S.max_xi), 0.5 * (S.min_eta + S.max_eta), S.face, c2);
    double dist = gcdist(c1, c2), sep = 100;
    if (dist > 1e-16) sep = (T.radius + S.radius) / dist;
    auto push4 = [&](int t0, const int* ts, int s0, const int* ss) {
      for (int k = 0; k < 4; k++) { tq.push_back(ts ? ts[k] : t0); sq.push_back(ss ? ss[k] : s0); }
    };
    if (sep < theta) {
      int type = (T.level < c->levels ? 2 : 0) + (S.level < c->levels ? 1 : 0);
      c->ilist.insert(c->ilist.end(), {it, is, type});
    } else if (T.is_leaf && S.is_leaf) {
      c->ilist.insert(c->ilist.end(), {it, is, 0});
    } else if (T.is_leaf) {
      push4(it, nullptr, 0, S.child);
    } else if (S.is_leaf) {
      push4(0, T.child, is, nullptr);
    } else if (T.level <= S.level) {
      push4(0, T.child, is, nullptr);
    } else {
      push4(it, nullptr, 0, S.child);
    }
  }
  return c;
}

void orc_irreg_destroy(void* h) { delete static_cast<IrrCtx*>(h); }

void orc_irreg_sizes(void* h, int64_t* s) {
  const IrrCtx& c = *static_cast<IrrCtx*>(h);
  s[0] = (int64_t)c.pn.size(); s[1] = c.levels; s[2] = (int64_t)c.ilist.size() / 3;
}

// ints [P][9]: face, level, parent, child x4, is_leaf, point_count; dbls [P][5]: bounds, radius;
// pts: the leaf point lists concatenated in panel order.
void orc_irreg_panels(void* h, int* ints, double* dbls, int* pts) {
  const IrrCtx& c = *static_cast<IrrCtx*>(h);
  int64_t o = 0;
  for (size_t i = 0; i < c.pn.size(); i++) {
    const IrrPanel& p = c.pn[i];
    int r[9] = {p.face, p.level, p.parent, p.child[0], p.child[1], p.child[2], p.child[3], p.is_leaf, p.point_count};
    std::memcpy(ints + 9 * i, r, sizeof(r));
    double d[5] = {p.min_xi, p.max_xi, p.min_eta, p.max_eta, p.radius};
    std::memcpy(dbls + 5 * i, d, sizeof(d));
    if (p.is_leaf)
      for (int id : c.pts[i]) pts[o++] = id;
  }
}

void orc_irreg_interactions(void* h, int* out) {
  const IrrCtx& c = *static_cast<IrrCtx*>(h);
  std::memcpy(out, c.ilist.data(), c.ilist.size() * sizeof(int));
}

static void irr_node(const IrrPanel& p, int j, int k, int degree, double* xi, double* eta) {
  double xo = 0.5 * (p.min_xi + p.max_xi), xs = 0.5 * (p.max_xi - p.min_xi);
  double eo = 0.5 * (p.min_eta + p.max_eta), es = 0.5 * (p.max_eta - p.min_eta);
  *xi = std::cos(M_PI / degree * j) * xs + xo;
  *eta = std::cos(M_PI / degree * k) * es + eo;
}

// upward_pass_ll (fmm_funcs.cpp:394-404): proxy [P][n] accumulated (caller zeroes).
int orc_irreg_upward_pass(void* h, const double* area, const double* pots, double* proxy) {
  const IrrCtx& c = *static_cast<IrrCtx*>(h);
  const int d = c.degree, n = (d + 1) * (d + 1);
  int err = 0;
  for (size_t i = 0; i < c.pn.size(); i++) {  // base_pots_ll
    if (!c.pn[i].is_leaf) continue;
    for (int idx : c.pts[i]) {
      double xieta[2], basis[121];
      xieta_from_xyz(c.x[idx], c.y[idx], c.z[idx], xieta);
      err |= interp_vals_bli(basis, xieta[0], xieta[1], c.pn[i].min_xi, c.pn[i].max_xi, c.pn[i].min_eta,
                             c.pn[i].max_eta, d);
      for (int k = 0; k < n; k++) proxy[i * n + k] += basis[k] * pots[idx] * area[idx];
    }
  }
  for (int l = c.levels; l > 0; l--) {  // child_to_parent_ll
    for (int i = c.level_start[l]; i < c.level_start[l + 1]; i++) {
      const IrrPanel& ch = c.pn[i];
      const IrrPanel& par = c.pn[ch.parent];
      for (int j = 0; j <= d; j++)
        for (int k = 0; k <= d; k++) {
          double xi, eta, basis[121];
          irr_node(ch, j, k, d, &xi, &eta);
          err |= interp_vals_bli(basis, xi, eta, par.min_xi, par.max_xi, par.min_eta, par.max_eta, d);
          for (int m = 0; m < n; m++) proxy[ch.parent * n + m] += basis[m] * proxy[i * n + j * (d + 1) + k];
        }
    }
  }
  return err ? ORC_E_GEOMETRY : ORC_OK;
}

// downward_pass_3_ll (fmm_funcs.cpp:406-419): proxy rows pushed to children in place,
// then sols[f][point] += leaf interpolation.
int orc_irreg_downward_pass_3(void* h, double* const* proxy, double* const* sols) {
  const IrrCtx& c = *static_cast<IrrCtx*>(h);
  const int d = c.degree, n = (d + 1) * (d + 1);
  int err = 0;
  for (int l = 0; l < c.levels; l++) {
    for (int i = c.level_start[l]; i < c.level_start[l + 1]; i++) {
      const IrrPanel& par = c.pn[i];
      if (par.is_leaf) continue;
      for (int q = 0; q < 4; q++) {
        const int ch = par.child[q];
        for (int k = 0; k < n; k++) {
          double xi, eta, basis[121];
          irr_node(c.pn[ch], k / (d + 1), k % (d + 1), d, &xi, &eta);
          err |= interp_vals_bli(basis, xi, eta, par.min_xi, par.max_xi, par.min_eta, par.max_eta, d);
          for (int f = 0; f < 3; f++)
            for (int m = 0; m < n; m++) proxy[f][ch * n + k] += basis[m] * proxy[f][i * n + m];
        }
      }
    }
  }
  for (size_t i = 0; i < c.pn.size(); i++) {
    if (!c.pn[i].is_leaf) continue;
    for (int idx : c.pts[i]) {
      double xieta[2], basis[121];
      xieta_from_xyz(c.x[idx], c.y[idx], c.z[idx], xieta);
      err |= interp_vals_bli(basis, xieta[0], xieta[1], c.pn[i].min_xi, c.pn[i].max_xi, c.pn[i].min_eta,
                             c.pn[i].max_eta, d);
      for (int f = 0; f < 3; f++)
        for (int k = 0; k < n; k++) sols[f][idx] += basis[k] * proxy[f][i * n + k];
    }
  }
  return err ? ORC_E_GEOMETRY : ORC_OK;
}

// swe_vel_interactions_ll (swe_vel.cpp:69-100): target_pots [3][P][n] and vel [3][N] accumulated.
void orc_irreg_swe_vel_interactions(void* h, double* const* target_pots, double* const* vel, const double* proxy_vors,
                                    const double* proxy_divs, const double* vors, const double* divs,
                                    const double* area, double eps) {
  const IrrCtx& c = *static_cast<IrrCtx*>(h);
  const int d = c.degree, n = (d + 1) * (d + 1);
  const int I = (int)c.ilist.size() / 3;
  for (int e = 0; e < I; e++) {
    const int tp = c.ilist[3 * e], sp = c.ilist[3 * e + 1], it = c.ilist[3 * e + 2];
    std::vector<double> tx, ty, tz, sx, sy, sz, sv, sd;
    std::vector<int> tid;
    if (it == 0 || it == 1) {
      for (int idx : c.pts[tp]) { tx.push_back(c.x[idx]); ty.push_back(c.y[idx]); tz.push_back(c.z[idx]); tid.push_back(idx); }
    } else {
      for (int j = 0; j <= d; j++)
        for (int k = 0; k <= d; k++) {
          const IrrPanel& p = c.pn[tp];
          double xo = 0.5 * (p.min_xi + p.max_xi), xs = 0.5 * (p.max_xi - p.min_xi);
          double eo = 0.5 * (p.min_eta + p.max_eta), es = 0.5 * (p.max_eta - p.min_eta), v[3];
          xyz_from_xieta(std::cos(M_PI * j / d) * xs + xo, std::cos(M_PI * k / d) * es + eo, p.face, v);
          tx.push_back(v[0]); ty.push_back(v[1]); tz.push_back(v[2]); tid.push_back(j * (d + 1) + k);
        }
    }
    if (it == 0 || it == 2) {
      for (int idx : c.pts[sp]) {
        sx.push_back(c.x[idx]); sy.push_back(c.y[idx]); sz.push_back(c.z[idx]);
        sv.push_back(vors[idx] * area[idx]); sd.push_back(divs[idx] * area[idx]);
      }
    } else {
      for (int j = 0; j <= d; j++)
        for (int k = 0; k <= d; k++) {
          const IrrPanel& p = c.pn[sp];
          double xo = 0.5 * (p.min_xi + p.max_xi), xs = 0.5 * (p.max_xi - p.min_xi);
          double eo = 0.5 * (p.min_eta + p.max_eta), es = 0.5 * (p.max_eta - p.min_eta), v[3];
          xyz_from_xieta(std::cos(M_PI * j / d) * xs + xo, std::cos(M_PI * k / d) * es + eo, p.face, v);
          sx.push_back(v[0]); sy.push_back(v[1]); sz.push_back(v[2]);
          sv.push_back(proxy_vors[sp * n + j * (d + 1) + k]); sd.push_back(proxy_divs[sp * n + j * (d + 1) + k]);
        }
    }
    for (size_t j = 0; j < tx.size(); j++) {
      for (size_t k = 0; k < sx.size(); k++) {
        double dp = tx[j] * sx[k] + ty[j] * sy[k] + tz[j] * sz[k];
        double gf = 0.07957747154594767 / (1.0 - dp + eps);
        double gf1 = -gf * sv[k], gf2 = gf * sd[k];
        double vx = gf1 * (ty[j] * sz[k] - tz[j] * sy[k]) + gf2 * (sx[k] - tx[j] * dp);
        double vy = gf1 * (tz[j] * sx[k] - tx[j] * sz[k]) + gf2 * (sy[k] - ty[j] * dp);
        double vz = gf1 * (tx[j] * sy[k] - ty[j] * sx[k]) + gf2 * (sz[k] - tz[j] * dp);
        if (it == 0 || it == 1) {
          vel[0][tid[j]] += vx; vel[1][tid[j]] += vy; vel[2][tid[j]] += vz;
        } else {
          target_pots[0][tp * n + tid[j]] += vx; target_pots[1][tp * n + tid[j]] += vy; target_pots[2][tp * n + tid[j]] += vz;
        }
      }
    }
  }
}

// ================================================================ SWE stepper
// swe_back_rk4_step (swe_time_step.cpp:606-811) and the functions it calls.

// bli_deriv_xi / bli_deriv_eta (deriv_funcs_impl.hpp:8-92): barycentric
// differentiation along one direction of a leaf's Chebyshev grid of degree deg,
// nodes cos(pi i/deg) * half + mid (computed, not snapped), weights 1/2,
// (-1)^i, (-1)^deg/2, in the reference's difference form
//   d_i = sum_{k != i} (f_k - f_i) w_k / (x_i - x_k) / w_i.
// dir 0 differentiates along the first index (xi), dir 1 along the second;
// index a*(deg+1)+b.
static void cheb_deriv(const double* f, int deg, double lo, double hi, int dir, double* out) {
  const int m = deg + 1;
  const double half = 0.5 * (hi - lo), mid = 0.5 * (hi + lo);
  auto w = [deg](int i) {
    if (i == 0) return 0.5;
    if (i == deg) return 0.5 * ((deg & 1) ? -1.0 : 1.0);
    return (i & 1) ? -1.0 : 1.0;
  };
  for (int a = 0; a < m; a++)
    for (int b = 0; b < m; b++) {
      const int i = dir == 0 ? a : b, idx = a * m + b;
      const double xi = std::cos(kPi * i / deg) * half + mid, fi = f[idx];
      double acc = 0;
      for (int k = 0; k < m; k++) {
        if (k == i) continue;
        const int idx2 = dir == 0 ? k * m + b : a * m + k;
        const double xk = std::cos(kPi * k / deg) * half + mid;
        acc += (f[idx2] - fi) * w(k) / (xi - xk);
      }
      out[idx] = acc / w(i);
    }
}

// xyzvec_from_xietavec (cubed_sphere_transforms_impl.hpp:153-239): the
// (xi, eta) pair is mapped with the metric factors, then the 2x2 system for
// the (colatitude, longitude) components is solved (here by Cramer's rule)
// and rotated to Cartesian components.  At the pole nodes of faces 4 and 5 the
// reference returns (g_xi, g_eta, 0) (:194-198, :215-219); with pole_fix the
// surface gradient there is returned instead: (-g_eta, g_xi, 0) on face 4
// (d/dxi = +y, d/deta = -x) and (g_eta, g_xi, 0) on face 5.
static int xyzvec_from_xietavec(double& ox, double& oy, double& oz, double xd, double ed, int face, double xi,
                                double eta, bool pole_fix) {
  const double X = std::tan(xi), Y = std::tan(eta), X2 = X * X, Y2 = Y * Y;
  const double C = std::sqrt(1 + X2), D = std::sqrt(1 + Y2);
  const double gx = D * xd + X * Y / D * ed, ge = X * Y / C * xd + C * ed;
  const double isqd = 1.0 / std::sqrt(1 + X2 + Y2);
  double xyz[3];
  if (xyz_from_xieta(xi, eta, face, xyz)) return ORC_E_GEOMETRY;
  const double x = xyz[0], y = xyz[1], z = xyz[2];
  const double sqc = std::sqrt(x * x + y * y);
  double m00, m01, m10, m11;
  if (face <= 3) {
    m00 = 0; m01 = C * D * isqd; m10 = -1.0; m11 = X * Y * isqd;
  } else {
    if (std::fabs(xi) < 1e-16 && std::fabs(eta) < 1e-16) {  // at the pole
      if (!pole_fix) { ox = gx; oy = ge; }
      else if (face == 4) { ox = -ge; oy = gx; }
      else { ox = ge; oy = gx; }
      oz = 0;
      return ORC_OK;
    }
    const double isqc2 = 1.0 / std::sqrt(X2 + Y2), s = face == 4 ? 1.0 : -1.0;
    m00 = s * isqc2 * D * X; m01 = -s * isqc2 * D * Y * isqd;
    m10 = s * isqc2 * C * Y; m11 = s * isqc2 * C * X * isqd;
  }
  const double det = m00 * m11 - m01 * m10;
  const double colat = (gx * m11 - m01 * ge) / det, lon = (m00 * ge - m10 * gx) / det;
  ox = x * z / sqc * colat - y / sqc * lon;
  oy = y * z / sqc * colat + x / sqc * lon;
  oz = -sqc * colat;
  return ORC_OK;
}

// single_panel_grad (deriv_funcs_impl.hpp:272-305) of f on the grid of leaf
// panel p, the vector conversion done for face `face`.  With `metric_twice`
// (the reference) the metric combination D f_xi + XY/D f_eta, XY/C f_xi + C f_eta
// is applied to the derivatives before xyzvec_from_xietavec, which applies it
// again; without it the raw derivatives go in, which gives the surface gradient.
static int panel_grad(const Ctx& c, const Panel& p, int face, bool metric_twice, const double* f, double* gx,
                      double* gy, double* gz) {
  const int deg = c.md, m = deg + 1;
  double xo, xs, eo, es;
  panel_frame(p, xo, xs, eo, es);
  double fx[121], fe[121];
  cheb_deriv(f, deg, xo - xs, xo + xs, 0, fx);
  cheb_deriv(f, deg, eo - es, eo + es, 1, fe);
  int err = 0;
  for (int a = 0; a < m; a++) {
    const double xi = std::cos(kPi * a / deg) * xs + xo;
    for (int b = 0; b < m; b++) {
      const double eta = std::cos(kPi * b / deg) * es + eo;
      const int k = a * m + b;
      const double X = std::tan(xi), Y = std::tan(eta), C = std::sqrt(1 + X * X), D = std::sqrt(1 + Y * Y);
      const double xc = metric_twice ? D * fx[k] + X * Y / D * fe[k] : fx[k];
      const double ec = metric_twice ? X * Y / C * fx[k] + C * fe[k] : fe[k];
      err |= xyzvec_from_xietavec(gx[k], gy[k], gz[k], xc, ec, face, xi, eta, !metric_twice);
    }
  }
  return err;
}

// laplacian / panel_laplacian (deriv_funcs.cpp:27-37, deriv_funcs_impl.hpp:338-389):
// lap = delta (f_xx / C^2 + f_ee / D^2 + 2 X Y / (C^2 D^2) f_xe) on each leaf,
// f_xe = d/dxi of d/deta f.  f, lap [Pl][m2].
int orc_panel_laplacian(void* h, const double* f, double* lap) {
  Ctx& c = *static_cast<Ctx*>(h);
  const int deg = c.md, m = deg + 1;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < c.Pl; i++) {
    double xo, xs, eo, es, fx[121], fe[121], fxx[121], fxe[121], fee[121];
    panel_frame(c.pn[i + c.offset], xo, xs, eo, es);
    const double* fi = f + (size_t)i * c.m2;
    cheb_deriv(fi, deg, xo - xs, xo + xs, 0, fx);
    cheb_deriv(fi, deg, eo - es, eo + es, 1, fe);
    cheb_deriv(fx, deg, xo - xs, xo + xs, 0, fxx);
    cheb_deriv(fe, deg, xo - xs, xo + xs, 0, fxe);
    cheb_deriv(fe, deg, eo - es, eo + es, 1, fee);
    for (int a = 0; a < m; a++) {
      const double xi = std::cos(kPi * a / deg) * xs + xo;
      for (int b = 0; b < m; b++) {
        const double eta = std::cos(kPi * b / deg) * es + eo;
        const int k = a * m + b;
        const double X = std::tan(xi), Y = std::tan(eta), C2 = 1 + X * X, D2 = 1 + Y * Y, delta = 1 + X * X + Y * Y;
        lap[(size_t)i * c.m2 + k] = delta * (fxx[k] / C2 + fee[k] / D2 + 2.0 * X * Y / (C2 * D2) * fxe[k]);
      }
    }
  }
  return ORC_OK;
}

// The leaf gradient of the velocity components (swe_time_step.cpp:109-111):
// rule 0 is the reference (vectors converted with cubed_sphere_panels(i).face,
// i the leaf-local index, the metric applied twice, the pole nodes' shortcut);
// rule 1 converts with the leaf's own face from the raw derivatives and maps
// the pole nodes correctly: the surface gradient.
// grad[9][Pl][m2]: d(vel_a)/d(x_b) at 3a+b.
int orc_velocity_gradient(void* h, const double* const* vel, int rule, double* const* grad) {
  Ctx& c = *static_cast<Ctx*>(h);
  int err = 0;
#pragma omp parallel for schedule(static) reduction(| : err)
  for (int i = 0; i < c.Pl; i++) {
    const Panel& p = c.pn[i + c.offset];
    const int face = rule == 0 ? c.pn[i].face : p.face;
    const size_t b = (size_t)i * c.m2;
    for (int a = 0; a < 3; a++)
      err |= panel_grad(c, p, face, rule == 0, vel[a] + b, grad[3 * a] + b, grad[3 * a + 1] + b, grad[3 * a + 2] + b);
  }
  return err ? ORC_E_GEOMETRY : ORC_OK;
}

// swe_tendency_computation (swe_time_step.cpp:76-161), per particle:
//   vor_t = -div (vor + 2 Omega z) - 2 Omega u_z
//   div_t = 2 Omega z vor - sum_ab G_ab G_ba - gravity lap - |u|^2 - 2 Omega (-y u_x + x u_y) [x^2+y^2 > 1e-16]
// (gravity = 1 in the reference, which omits g/a^2)
//   h_t   = -h div
int orc_swe_tendency(void* h, const double* x, const double* y, const double* z, const double* const* vel,
                     const double* vor, const double* div, const double* height, const double* lap, double omega,
                     int rule, double* vor_t, double* div_t, double* h_t) {
  Ctx& c = *static_cast<Ctx*>(h);
  const size_t NP = (size_t)c.Pl * c.m2;
  std::vector<double> g(9 * NP);
  double* gp[9];
  for (int k = 0; k < 9; k++) gp[k] = g.data() + k * NP;
  int err = orc_velocity_gradient(h, vel, rule, gp);
#pragma omp parallel for schedule(static)
  for (size_t k = 0; k < NP; k++) {
    const double G11 = gp[0][k], G12 = gp[1][k], G13 = gp[2][k], G21 = gp[3][k], G22 = gp[4][k], G23 = gp[5][k],
                 G31 = gp[6][k], G32 = gp[7][k], G33 = gp[8][k];
    const double ux = vel[0][k], uy = vel[1][k], uz = vel[2][k];
    const double abs_vor = vor[k] + 2.0 * omega * z[k];
    vor_t[k] = -div[k] * abs_vor - 2.0 * omega * uz;
    double d = 2.0 * omega * z[k] * vor[k];
    const double gc = G11 * G11 + G12 * G21 + G13 * G31 + G21 * G12 + G22 * G22 + G23 * G32 + G31 * G13 +
                      G32 * G23 + G33 * G33;
    d -= gc;
    d -= c.gravity * lap[k];
    d -= ux * ux + uy * uy + uz * uz;
    if (x[k] * x[k] + y[k] * y[k] > 1e-16) d -= 2.0 * omega * (-y[k] * ux + x[k] * uy);
    div_t[k] = d;
    h_t[k] = -height[k] * div[k];
  }
  return err;
}

// cone_mountain (topography.cpp:6-45) at P(x + disp): eff = h + h0 exp(-16 r^2),
// h0 = 2000 / 6371000, r^2 = (lon - pi/2)^2 + (lat - pi/6)^2, lon in [0, 2 pi).
// topo 2 (no reference counterpart): the Gaussian mountain
// eff = h + H exp(-(r / w)^2), r the great-circle distance to the centre.
// topo 0: eff = h (apply_topography's copy, :47-55).  The displacement read at
// leaf-local i, node j is disp[i][j] of the full [P][n] array (rule 0, the
// reference: panel i, which is not leaf i; needs m2 == n) or the leaf's own
// (rule 1).
int orc_apply_topography(void* h, const double* x, const double* y, const double* z, const double* const* disp_part,
                         const double* const* disp_proxy, const double* height, int topo, int rule, double* eff) {
  Ctx& c = *static_cast<Ctx*>(h);
  const size_t NP = (size_t)c.Pl * c.m2;
  if (topo == 0) {
    std::memcpy(eff, height, NP * sizeof(double));
    return ORC_OK;
  }
  if (rule == 0 && c.m2 != c.n) return ORC_E_ARG;
  if (topo != 1 && topo != 2) return ORC_E_ARG;
  const double h0 = 2000.0 / 6371000.0, lc = 0.5 * kPi, tc = kPi / 6.0;
  // Gaussian mountain centre (topography 2)
  const double mc[3] = {std::cos(c.mtn[2]) * std::cos(c.mtn[3]), std::cos(c.mtn[2]) * std::sin(c.mtn[3]),
                        std::sin(c.mtn[2])};
#pragma omp parallel for schedule(static)
  for (int i = 0; i < c.Pl; i++)
    for (int j = 0; j < c.m2; j++) {
      const size_t k = (size_t)i * c.m2 + j;
      double d[3];
      for (int f = 0; f < 3; f++) {
        if (rule == 1 || i >= c.offset) {
          const size_t kk = rule == 1 ? k : (size_t)(i - c.offset) * c.m2 + j;
          d[f] = disp_part[f][kk];
        } else {
          d[f] = disp_proxy[f][(size_t)i * c.n + j];
        }
      }
      double a = x[k] + d[0], b = y[k] + d[1], cc = z[k] + d[2];
      project_to_sphere(a, b, cc);
      if (topo == 2) {
        // great-circle distance to the centre as atan2(|p x c|, p.c)
        const double cx = b * mc[2] - cc * mc[1], cy = cc * mc[0] - a * mc[2], cz = a * mc[1] - b * mc[0];
        const double r = std::atan2(std::sqrt(cx * cx + cy * cy + cz * cz), a * mc[0] + b * mc[1] + cc * mc[2]);
        const double q = r / c.mtn[1];
        eff[k] = height[k] + c.mtn[0] * std::exp(-q * q);
        continue;
      }
      const double lat = kPi / 2.0 - std::atan2(std::sqrt(a * a + b * b), cc);
      double lon = std::atan2(b, a);
      lon += 2 * kPi;
      lon = std::fmod(lon, 2 * kPi);
      const double r2 = (lon - lc) * (lon - lc) + (lat - tc) * (lat - tc);
      eff[k] = height[k] + h0 * std::exp(-16.0 * r2);
    }
  return ORC_OK;
}

// lap_panel_interaction (laplacian_impl.hpp:11-69) on the type-dependent point
// sets, sources at P(node + disp): for x.y > 1 - 30 eps (eps = 0.1),
//   K = [e^{-1/eps} pi^{-3/2} 2t(-1 + 3eps + t^2)
//        + (1 + erf(t/eps)) e^{(t^2-1)/eps} / (sqrt(eps) pi) (2eps^2 + 2t^2(t^2-1) + eps(7t^2-1))] / eps^{5/2},
// t = x.y; phi += K q.  phi_part [Pl][m2], phi_proxy [P][n] accumulated.
void orc_laplacian_interactions(void* h, const double* const* disp_part, const double* const* disp_proxy,
                                const double* q_part, const double* q_proxy, double* phi_part, double* phi_proxy) {
  Ctx& c = *static_cast<Ctx*>(h);
  const double eps = 0.1, c1 = std::exp(-1.0 / eps) * std::pow(kPi, -1.5), c2 = 1.0 / (std::sqrt(eps) * kPi),
               c3 = 1.0 / std::pow(eps, 2.5);
#pragma omp parallel for schedule(dynamic, 4)
  for (int t = 0; t < c.P; t++) {
    for (int e = c.tgt_start[t]; e < c.tgt_start[t + 1]; e++) {
      int ii = c.by_target[e];
      int source = c.ilist[3 * ii + 1], type = c.ilist[3 * ii + 2];
      bool tpart = (type == 0 || type == 1), spart = (type == 0 || type == 2);
      Side ts = side_of(c, t, tpart), ss = side_of(c, source, spart);
      const Panel &pt = c.pn[t], &ps = c.pn[source];
      double xot, xst, eot, est, xos, xss, eos, ess;
      panel_frame(pt, xot, xst, eot, est);
      panel_frame(ps, xos, xss, eos, ess);
      double* o = tpart ? phi_part + (size_t)(t - c.offset) * c.m2 : phi_proxy + (size_t)t * c.n;
      const size_t sb = spart ? (size_t)(source - c.offset) * c.m2 : (size_t)source * c.n;
      const double* sq = (spart ? q_part : q_proxy) + sb;
      const double* dx = (spart ? disp_part[0] : disp_proxy[0]) + sb;
      const double* dy = (spart ? disp_part[1] : disp_proxy[1]) + sb;
      const double* dz = (spart ? disp_part[2] : disp_proxy[2]) + sb;
      double xt[3], y0[3];
      for (int j = 0; j < ts.count; j++) {
        xyz_from_xieta(ts.tab[4 * j] * xst + xot, ts.tab[4 * j + 1] * est + eot, pt.face, xt);
        double acc = 0;
        for (int k = 0; k < ss.count; k++) {
          xyz_from_xieta(ss.tab[4 * k] * xss + xos, ss.tab[4 * k + 1] * ess + eos, ps.face, y0);
          double a = y0[0] + dx[k], b = y0[1] + dy[k], cc = y0[2] + dz[k];
          project_to_sphere(a, b, cc);
          const double dp = xt[0] * a + xt[1] * b + xt[2] * cc;
          if (dp > 1.0 - 30 * eps) {
            const double dp2 = dp * dp;
            const double h1 = c1 * (2.0 * dp * (-1.0 + 3.0 * eps + dp2));
            const double h2 = (1.0 + std::erf(dp / eps)) * std::exp((dp2 - 1) / eps) * c2 *
                              (2.0 * eps * eps + 2.0 * dp2 * (dp2 - 1.0) + eps * (7.0 * dp2 - 1.0));
            acc += (h1 + h2) * c3 * sq[k];
          }
        }
        o[j] += acc;
      }
    }
  }
}

// swe_apply_tendencies (swe_time_step.cpp:163-192): f' = f + dt f_t for vor,
// div, h; area' = area (1 + dt div).  area/new_area may be null (arrival values,
// swe_time_step.cpp:404-406).
void orc_swe_apply_tendencies(void* h, const double* const* f, const double* const* ft, const double* area,
                              double dt, double* const* out, double* new_area) {
  Ctx& c = *static_cast<Ctx*>(h);
  const size_t NP = (size_t)c.Pl * c.m2;
  for (size_t k = 0; k < NP; k++) {
    for (int q = 0; q < 3; q++) out[q][k] = f[q][k] + dt * ft[q][k];
    if (area && new_area) new_area[k] = area[k] * (1.0 + dt * f[1][k]);
  }
}

// swe_height_correction (swe_time_step.cpp:558-604): at P(x + dt vel) the BLI
// of h + dt h_t; lap -= 8 new_h / 0.1.
// Height of topography 2 at a unit vector (0 for other topographies).
static double topo_height(const Ctx& c, double a, double b, double cc) {
  if (c.topo_kind != 2) return 0.0;
  const double mc[3] = {std::cos(c.mtn[2]) * std::cos(c.mtn[3]), std::cos(c.mtn[2]) * std::sin(c.mtn[3]),
                        std::sin(c.mtn[2])};
  const double cx = b * mc[2] - cc * mc[1], cy = cc * mc[0] - a * mc[2], cz = a * mc[1] - b * mc[0];
  const double r = std::atan2(std::sqrt(cx * cx + cy * cy + cz * cz), a * mc[0] + b * mc[1] + cc * mc[2]);
  const double q = r / c.mtn[1];
  return c.mtn[0] * std::exp(-q * q);
}

int orc_swe_height_correction(void* h, const double* x, const double* y, const double* z, const double* const* vel,
                              const double* height, const double* h_t, double dt, double* lap) {
  Ctx& c = *static_cast<Ctx*>(h);
  const size_t NP = (size_t)c.Pl * c.m2;
  std::vector<double> hh(NP);
  for (size_t k = 0; k < NP; k++) hh[k] = height[k] + dt * h_t[k];
  const double* f1[1] = {hh.data()};
  int err = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : err)
  for (size_t k = 0; k < NP; k++) {
    double a = x[k] + dt * vel[0][k], b = y[k] + dt * vel[1][k], cc = z[k] + dt * vel[2][k];
    project_to_sphere(a, b, cc);
    double nh;
    err |= interp_at(c, a, b, cc, 1, f1, &nh);
    if (c.swe_form & 2) nh += topo_height(c, a, b, cc);
    lap[k] -= 8.0 * nh / 0.1;
  }
  return err ? ORC_E_GEOMETRY : ORC_OK;
}

// swe_displacement_upward_pass (swe_time_step.cpp:265-276): the BVE
// displacement without the vorticity update.
static int swe_displacement(Ctx& c, const double* x, const double* y, const double* z, const double* const* vel,
                            double tau, double* const* dpart, double* const* dprox) {
  std::vector<double> dummy((size_t)c.Pl * c.m2), zero((size_t)c.Pl * c.m2, 0.0);
  return orc_bve_displacement(&c, x, y, z, vel, zero.data(), tau, 0.0, dummy.data(), dpart, dprox);
}

// Upward passes of vor and div (area-weighted), swe_vel_interactions (variant 1),
// downward_pass_3 (swe_time_step.cpp:655-672).  vel[3][Pl][m2] assigned.
static int swe_fmm_velocity(Ctx& c, const double* area, const double* vor, const double* div,
                            const double* const* dpart, const double* const* dprox, double eps, double* const* vel) {
  const size_t NP = (size_t)c.Pl * c.m2, NQ = (size_t)c.P * c.n;
  std::vector<double> qzp(NP), qzx(NQ), qdp(NP), qdx(NQ), pp(3 * NP, 0.0), px(3 * NQ, 0.0);
  int err = orc_upward_pass(&c, area, vor, qzp.data(), qzx.data());
  err |= orc_upward_pass(&c, area, div, qdp.data(), qdx.data());
  double* php[3] = {pp.data(), pp.data() + NP, pp.data() + 2 * NP};
  double* phx[3] = {px.data(), px.data() + NQ, px.data() + 2 * NQ};
  orc_swe_vel_interactions(&c, 1, dpart, dprox, qzp.data(), qzx.data(), qdp.data(), qdx.data(), eps, php, phx);
  const double* cphp[3] = {php[0], php[1], php[2]};
  err |= orc_downward_pass(&c, 3, phx, cphp, vel);
  return err;
}

// Upward pass of area * f, laplacian_interactions, downward pass
// (swe_time_step.cpp:695-701).  lap [Pl][m2] assigned.
static int lap_fmm(Ctx& c, const double* area, const double* f, const double* const* dpart,
                   const double* const* dprox, double* lap) {
  const size_t NP = (size_t)c.Pl * c.m2, NQ = (size_t)c.P * c.n;
  std::vector<double> qp(NP), qx(NQ), pp(NP, 0.0), px(NQ, 0.0);
  int err = orc_upward_pass(&c, area, f, qp.data(), qx.data());
  orc_laplacian_interactions(&c, dpart, dprox, qp.data(), qx.data(), pp.data(), px.data());
  double* phx[1] = {px.data()};
  const double* cpp[1] = {pp.data()};
  double* out[1] = {lap};
  err |= orc_downward_pass(&c, 1, phx, cpp, out);
  return err;
}

// swe_back_rk4_step (swe_time_step.cpp:606-811).  vors, divs, height [Pl][m2]
// and tracers [Pl][m2][T] are updated in place; vel[3] receives the stage-1
// velocity and height_lap the stage-1 surface Laplacian.  rule 0 is the
// reference (leaf-local face index in the velocity gradient, panel-i
// displacement in the topography, stage-4 state advanced by dt/2); rule 1
// uses the leaf's own face and displacement and dt at stage 4.  The stage-4
// surface Laplacian, height correction and tendencies, and the k2..k4 part of
// swe_compute_arrival_values, do not reach any output (the arrival values are
// state + dt k1, :415-417) and are not computed.
int orc_swe_back_rk4_step(void* h, const double* x, const double* y, const double* z, const double* area,
                          double* vors, double* divs, double* height, double* tracers, int ntr, double dt,
                          double omega, double eps, int combine, int topo, int rule, double* const* vel,
                          double* height_lap) {
  Ctx& c = *static_cast<Ctx*>(h);
  const size_t NP = (size_t)c.Pl * c.m2, NQ = (size_t)c.P * c.n;
  std::vector<double> dpart(3 * NP, 0.0), dprox(3 * NQ, 0.0), eff(NP), vs(9 * NP), lap(NP), tend(9 * NP),
      st(3 * NP), da(NP);
  double* dpw[3] = {dpart.data(), dpart.data() + NP, dpart.data() + 2 * NP};
  double* dxw[3] = {dprox.data(), dprox.data() + NQ, dprox.data() + 2 * NQ};
  const double* dp3[3] = {dpw[0], dpw[1], dpw[2]};
  const double* dx3[3] = {dxw[0], dxw[1], dxw[2]};
  double* V[4][3] = {{vel[0], vel[1], vel[2]},
                     {vs.data(), vs.data() + NP, vs.data() + 2 * NP},
                     {vs.data() + 3 * NP, vs.data() + 4 * NP, vs.data() + 5 * NP},
                     {vs.data() + 6 * NP, vs.data() + 7 * NP, vs.data() + 8 * NP}};
  double* T[3][3];
  for (int s = 0; s < 3; s++)
    for (int q = 0; q < 3; q++) T[s][q] = tend.data() + (3 * s + q) * NP;
  double* S[3] = {st.data(), st.data() + NP, st.data() + 2 * NP};  // displaced vor, div, h
  const double* state[3] = {vors, divs, height};
  // stage 1
  c.topo_kind = topo;
  int err = orc_apply_topography(h, x, y, z, dp3, dx3, height, topo, rule, eff.data());
  if (c.swe_form & 1) {
    // the regularised-kernel Laplacian of stages 2-3 at the undisplaced nodes
    // less 8/0.1 times each node's own height (or free surface)
    err |= lap_fmm(c, area, eff.data(), dp3, dx3, height_lap);
    for (size_t k = 0; k < NP; k++) {
      double v = height[k];
      if (c.swe_form & 2) v += topo_height(c, x[k], y[k], z[k]);
      height_lap[k] -= 8.0 * v / 0.1;
    }
  } else {
    err |= orc_panel_laplacian(h, eff.data(), height_lap);
  }
  err |= swe_fmm_velocity(c, area, vors, divs, dp3, dx3, eps, V[0]);
  const double* cv[3] = {V[0][0], V[0][1], V[0][2]};
  err |= orc_swe_tendency(h, x, y, z, cv, vors, divs, height, height_lap, omega, rule, T[0][0], T[0][1], T[0][2]);
  const double taus[3] = {0.5 * dt, 0.5 * dt, dt};
  for (int s = 1; s <= 3; s++) {
    const double* vprev[3] = {V[s - 1][0], V[s - 1][1], V[s - 1][2]};
    err |= swe_displacement(c, x, y, z, vprev, taus[s - 1], dpw, dxw);
    const double* tp[3] = {T[s - 1][0], T[s - 1][1], T[s - 1][2]};
    const double tau_apply = (s == 3 && rule == 0) ? 0.5 * dt : taus[s - 1];
    orc_swe_apply_tendencies(h, state, tp, area, tau_apply, S, da.data());
    err |= swe_fmm_velocity(c, da.data(), S[0], S[1], dp3, dx3, eps, V[s]);
    if (s == 3) break;
    err |= orc_apply_topography(h, x, y, z, dp3, dx3, S[2], topo, rule, eff.data());
    err |= lap_fmm(c, da.data(), eff.data(), dp3, dx3, lap.data());
    err |= orc_swe_height_correction(h, x, y, z, vprev, height, T[s - 1][2], taus[s - 1], lap.data());
    const double* cvs[3] = {V[s][0], V[s][1], V[s][2]};
    err |= orc_swe_tendency(h, x, y, z, cvs, vors, divs, height, lap.data(), omega, rule, T[s][0], T[s][1], T[s][2]);
  }
  // arrival values, departure points, remap
  std::vector<double> arr(3 * NP), dep(3 * NP), nf(3 * NP), ntr_buf(NP * (ntr > 0 ? ntr : 1));
  double* A[3] = {arr.data(), arr.data() + NP, arr.data() + 2 * NP};
  const double* t1[3] = {T[0][0], T[0][1], T[0][2]};
  orc_swe_apply_tendencies(h, state, t1, nullptr, dt, A, nullptr);
  const double* all[12] = {V[0][0], V[0][1], V[0][2], V[1][0], V[1][1], V[1][2],
                           V[2][0], V[2][1], V[2][2], V[3][0], V[3][1], V[3][2]};
  err |= orc_bve_find_departure_points(h, x, y, z, all, dt, combine, dep.data(), dep.data() + NP, dep.data() + 2 * NP);
  const double* ca[3] = {A[0], A[1], A[2]};
  double* N3[3] = {nf.data(), nf.data() + NP, nf.data() + 2 * NP};
  err |= orc_swe_departure_to_target(h, dep.data(), dep.data() + NP, dep.data() + 2 * NP, ca, tracers, ntr, N3,
                                     ntr_buf.data());
  std::memcpy(vors, N3[0], NP * sizeof(double));
  std::memcpy(divs, N3[1], NP * sizeof(double));
  std::memcpy(height, N3[2], NP * sizeof(double));
  if (ntr > 0) std::memcpy(tracers, ntr_buf.data(), NP * ntr * sizeof(double));
  if (err & ORC_E_ARG) return ORC_E_ARG;
  return err ? ORC_E_GEOMETRY : ORC_OK;
}

// Shallow-water physics beyond the reference (no counterpart there): the
// coefficient of the free-surface Laplacian in the divergence tendency, the
// Gaussian mountain of topography 2 and the step form bits.
void orc_set_swe_physics(void* h, double gravity, double height, double width, double lat, double lon, int form) {
  Ctx& c = *static_cast<Ctx*>(h);
  c.swe_form = form;
  c.gravity = gravity;
  c.mtn[0] = height;
  c.mtn[1] = width;
  c.mtn[2] = lat;
  c.mtn[3] = lon;
}

// ---- Lat-lon shallow-water step (swe_time_step_ll.cpp) ----------------------
// Fields are [lat][lon] row-major on the cell-centred grid of latlon_point
// (general_utils_impl.hpp:252-276): lat_i = -90 + dx (i + 1/2), lon_j = dx (j + 1/2)
// degrees.  Sequential loops in the reference's operation order.

struct LlGrid {
  int nlat, nlon;
  double dx, dxr;  // spacing in degrees and radians (dx*pi/180 as written)
  int64_t at(int i, int j) const { return int64_t(i) * nlon + j; }
  int jp(int j) const { return j + 1 == nlon ? 0 : j + 1; }
  int jm(int j) const { return j == 0 ? nlon - 1 : j - 1; }
};

static LlGrid ll_grid(int nlat, int nlon, double dx) { return LlGrid{nlat, nlon, dx, dx * kPi / 180.0}; }

// xyz_to_latlon (general_utils_impl.hpp:232-236).
static void ll_latlon(double x, double y, double z, double& lat, double& lon) {
  lat = kPi / 2.0 - std::atan2(std::sqrt(x * x + y * y), z);
  lon = std::atan2(y, x);
}

static double ll_lat(const double* x, const double* y, const double* z, int64_t k) {
  double lat, lon;
  ll_latlon(x[k], y[k], z[k], lat, lon);
  return lat;
}

// xyz_vel_to_uv_vel's per-point body (general_utils_impl.hpp:214-226).
static void ll_uv(double x, double y, double z, double vx, double vy, double vz, double& u, double& v) {
  if (std::abs(z) < 1 - 1e-16) {
    double r = std::sqrt(x * x + y * y);
    u = (-y * vx + x * vy) / r;
    v = -((x * vx + y * vy) * z - (x * x + y * y) * vz) / r;
  } else {
    u = 0;
    v = 0;
  }
}

// swe_tendency_computation_ll (swe_time_step_ll.cpp:365-387): u, v and |vel|^2
// (xyz_vel_to_uv_vel, vel_dot_prod), then vor_tend_ll (:14-32), height_tend_ll
// (:34-95), div_tend_ll (:97-231) and filter_vals(div, 2e-12)
// (general_utils_impl.hpp:278-293).  Returns 1 for a grid under 3 x 3.
int orc_ll_tendencies(int nlat, int nlon, double dx_deg, double omega, const double* x, const double* y,
                      const double* z, const double* vor, const double* div, const double* h, const double* topo,
                      const double* vx, const double* vy, const double* vz, double* vor_t, double* div_t,
                      double* h_t) {
  if (nlat < 3 || nlon < 3) return 1;
  const LlGrid g = ll_grid(nlat, nlon, dx_deg);
  const int64_t n = int64_t(nlat) * nlon;
  std::vector<double> u(n), v(n), dp(n);
  for (int64_t k = 0; k < n; k++) {
    ll_uv(x[k], y[k], z[k], vx[k], vy[k], vz[k], u[k], v[k]);
    dp[k] = vx[k] * vx[k] + vy[k] * vy[k] + vz[k] * vz[k];
  }
  const double grav = 9.81 / 6371000.0;
  auto bern = [&](int64_t k) { return grav * (h[k] + topo[k]) + 0.5 * dp[k]; };
  for (int i = 0; i < nlat; i++) {
    for (int j = 0; j < nlon; j++) {
      const int64_t k = g.at(i, j), kp = g.at(i, g.jp(j)), km = g.at(i, g.jm(j));
      const double lat = ll_lat(x, y, z, k);
      // vorticity: -(zeta + 2 Omega z) delta - 2 Omega w
      const double avor = vor[k] + 2.0 * omega * z[k];
      vor_t[k] = -avor * div[k] - 2.0 * omega * vz[k];
      // height: -(d_lon(h u) / (2 dx in degrees) + d_lat(h v cos)) / cos
      const double d1 = (h[kp] * u[kp] - h[km] * u[km]) / (2.0 * dx_deg);
      double d2;
      if (i == 0) {
        const int64_t kn = g.at(1, j);
        d2 = (3.0 * h[k] * v[k] * std::cos(lat) + h[kn] * v[kn] * std::cos(lat + g.dxr)) / (3.0 * g.dxr);
      } else if (i == nlat - 1) {
        const int64_t ks = g.at(i - 1, j);
        d2 = (-h[ks] * v[ks] * std::cos(lat - g.dxr) - 3.0 * h[k] * v[k] * std::cos(lat)) / (3.0 * g.dxr);
      } else {
        const int64_t kn = g.at(i + 1, j), ks = g.at(i - 1, j);
        d2 = (h[kn] * v[kn] * std::cos(lat + g.dxr) - h[ks] * v[ks] * std::cos(lat - g.dxr)) / (2.0 * g.dxr);
      }
      h_t[k] = -(d1 + d2) / std::cos(lat);
      // divergence
      const double e1 = (vor[kp] + 2.0 * omega * z[kp]) * v[kp];
      const double e2 = (vor[km] + 2.0 * omega * z[km]) * v[km];
      const double b0 = bern(k);
      double curl = (e1 - e2) / (2.0 * g.dxr);
      double lap = (bern(kp) - 2.0 * b0 + bern(km)) / (g.dxr * g.dxr * std::cos(lat) * std::cos(lat));
      if (i == 0 || i == nlat - 1) {
        // one-sided rows: the pole term enters with weight 0 (the stencil's -1/2 or +1/2
        // node is the pole); note the product 2 Omega z u, not (zeta + 2 Omega z) u
        const int64_t kq = g.at(i == 0 ? 1 : i - 1, j);
        const double lq = i == 0 ? lat + g.dxr : lat - g.dxr;
        const double w1 = (vor[k] + 2.0 * omega * z[k] * u[k]) * std::cos(lat);
        const double w2 = (vor[kq] + 2.0 * omega * z[kq] * u[kq]) * std::cos(lq);
        if (i == 0)
          curl -= (3.0 * w1 + w2) / (3.0 * g.dxr);
        else
          curl -= (-3.0 * w1 - w2) / (3.0 * g.dxr);
        lap += curl / std::cos(lat);
      } else {
        const int64_t kn = g.at(i + 1, j), ks = g.at(i - 1, j);
        const double w1 = (vor[kn] + 2.0 * omega * z[kn]) * u[kn] * std::cos(lat + g.dxr);
        const double w2 = (vor[ks] + 2.0 * omega * z[ks]) * u[ks] * std::cos(lat - g.dxr);
        curl -= (w1 - w2) / (2.0 * g.dxr);
        const double bn = bern(kn), bs = bern(ks);
        lap += -std::tan(lat) * (bn - bs) / (2.0 * g.dxr) + (bn - 2.0 * b0 + bs) / (g.dxr * g.dxr);
      }
      double dt_k = curl / std::cos(lat) - lap;
      if (std::abs(lat) * 180.0 / M_PI > 60) dt_k *= std::cos(lat) * std::cos(lat);
      if (std::abs(dt_k) < 2e-12) dt_k = 0.0;
      div_t[k] = dt_k;
    }
  }
  return 0;
}

// polar_filtering (swe_time_step_ll.cpp:600-664) on rows with |lat| > 60 degrees,
// in place.  DEVIATION: every Fourier coefficient starts at zero; the reference
// zeroes only coefficient 0 inside its loop (:628-631) and accumulates into the
// others uninitialised (:644-645), which is undefined.  The clamp keeps the
// reference's rule: values above 1e-11 and below -1e-11 both become +1e-11 (:640-643).
void orc_ll_polar_filter(int nlat, int nlon, double dx_deg, const double* x, const double* y, const double* z,
                         double* f) {
  const LlGrid g = ll_grid(nlat, nlon, dx_deg);
  for (int i = 0; i < nlat; i++) {
    const double lat0 = ll_lat(x, y, z, g.at(i, 0));
    if (!(std::abs(lat0) > kPi / 3.0)) continue;
    int spec = int(std::floor(90.0 - std::abs(lat0 * 180.0 / kPi)));
    spec = std::min(spec + 1, 29);
    double cs[30] = {}, sn[30] = {};
    for (int j = 0; j < nlon; j++) {
      const int64_t k = g.at(i, j);
      const double lat = ll_lat(x, y, z, k);
      if (f[k] > 1e-11)
        f[k] = 1e-11;
      else if (f[k] < -1e-11)
        f[k] = 1e-11;
      cs[0] += f[k];
      for (int m = 1; m < spec; m++) {
        cs[m] += f[k] * std::cos(m * lat) * g.dxr;
        sn[m] += f[k] * std::sin(m * lat) * g.dxr;
      }
    }
    cs[0] /= nlon;
    for (int m = 1; m < spec; m++) {
      cs[m] /= nlon;
      sn[m] /= nlon;
    }
    for (int m = 0; m < spec; m++) {
      if (cs[m] != cs[m]) cs[m] = 0;
      if (sn[m] != sn[m]) sn[m] = 0;
    }
    for (int j = 0; j < nlon; j++) {
      const int64_t k = g.at(i, j);
      const double lat = ll_lat(x, y, z, k);
      double val = cs[0];
      for (int m = 1; m < spec; m++) val += std::cos(m * lat) * cs[m] + std::sin(m * lat) * sn[m];
      f[k] = val;
    }
  }
}

// swe_compute_arrival_values_ll (swe_time_step_ll.cpp:389-598): quadratic
// semi-Lagrangian arrival values on the lat-lon grid.  |v| < 1e-16: 1-D
// interpolation along the row at lon - u dt (:446-474); else the departure point
// x - dt vel projected to the sphere, 3 x 3 tensor Lagrange interpolation, with
// the ring mean of row 0 (row nlat-1) standing at the south (north) pole node
// (:476-593).  Indices are kept inside the grid where the reference's would leave
// it (a departure latitude of exactly 90 degrees, a longitude index rounded up to
// lon_count, or a row shift of more than 360 degrees); elsewhere they are the
// reference's.
int orc_ll_arrival_values(int nlat, int nlon, double dx, double dt, const double* x, const double* y,
                          const double* z, const double* vor, const double* div, const double* h, const double* vx,
                          const double* vy, const double* vz, const double* vor_t, const double* div_t,
                          double* a_vor, double* a_div, double* a_h) {
  if (nlat < 3 || nlon < 3) return 1;
  const LlGrid g = ll_grid(nlat, nlon, dx);
  const double vs = 1.0, ds = 1.0, hs = 1.0;  // vor_scale, div_scale, h_scale (:430-432)
  auto vval = [&](int i, int j) { int64_t k = g.at(i, j); return vor[k] + vs * dt * vor_t[k]; };
  auto hval = [&](int i, int j) { int64_t k = g.at(i, j); return h[k] - dt * hs * h[k] * div[k]; };
  auto ring = [&](int i, double out[2]) {  // pole means (:532-547, :565-580)
    double sv = 0.0, sh = 0.0;
    for (int m = 0; m < nlon; m++) {
      const int64_t k = g.at(i, m);
      sv += vor[k] + vs * dt * vor_t[k];
    }
    for (int m = 0; m < nlon; m++) {
      const int64_t k = g.at(i, m);
      sh += h[k] - hs * h[k] * div[k] * dt;
    }
    out[0] = sv / nlon;
    out[1] = sh / nlon;
  };
  double south[2], north[2];
  ring(0, south);
  ring(nlat - 1, north);
  auto lon_coeffs = [&](double lon_d, int li, double c[3]) {
    c[0] = (lon_d - dx * (li + 0.5)) * (lon_d - dx * (li + 1.5)) / (2.0 * dx * dx);
    c[1] = -(lon_d - dx * (li - 0.5)) * (lon_d - dx * (li + 1.5)) / (dx * dx);
    c[2] = (lon_d - dx * (li - 0.5)) * (lon_d - dx * (li + 0.5)) / (2.0 * dx * dx);
  };
  // a grid index from a floored coordinate, kept inside [0, hi) (NaN -> 0)
  auto index = [](double fl, int hi) {
    if (!(fl >= 0.0)) return 0;
    if (fl > double(hi - 1)) return hi - 1;
    return int(fl);
  };
  auto lon_index = [&](double& lon_d) {
    if (lon_d < 0.0) lon_d += 360.0;
    return index(std::floor(lon_d / dx), nlon);
  };
  for (int i = 0; i < nlat; i++) {
    for (int j = 0; j < nlon; j++) {
      const int64_t k = g.at(i, j);
      double u, v;
      ll_uv(x[k], y[k], z[k], vx[k], vy[k], vz[k], u, v);
      if (std::abs(v) < 1e-16) {
        double lat, lon;
        ll_latlon(x[k], y[k], z[k], lat, lon);
        lon -= u * dt;
        double lon_d = 180.0 / kPi * lon;
        lon_d = std::fmod(lon_d + 360.0, 360.0);
        const int li = lon_index(lon_d);
        const int lm = ((li - 1) + nlon) % nlon, lp = (li + 1) % nlon;
        double c[3];
        lon_coeffs(lon_d, li, c);
        const double vv[3] = {vval(i, lm), vval(i, li), vval(i, lp)};
        const double hh[3] = {hval(i, lm), hval(i, li), hval(i, lp)};
        double av = 0, ah = 0;
        for (int l = 0; l < 3; l++) {
          av += vv[l] * c[l];
          ah += hh[l] * c[l];
        }
        a_vor[k] = av;
        a_h[k] = ah;
      } else {
        double px = x[k] - dt * vx[k], py = y[k] - dt * vy[k], pz = z[k] - dt * vz[k];
        const double nrm = std::sqrt(px * px + py * py + pz * pz);
        px /= nrm;
        py /= nrm;
        pz /= nrm;
        double lat, lon;
        ll_latlon(px, py, pz, lat, lon);
        const double lat_d = 180.0 / kPi * lat;
        double lon_d = 180.0 / kPi * lon;
        lon_d = std::fmod(lon_d + 360.0, 360.0);
        int ti = index(std::floor((lat_d + 90.0) / dx), nlat);
        const int li = lon_index(lon_d);
        double a[3];
        bool done = false;
        if (ti == 0) {
          if (lat_d > -90.0 + 0.5 * dx) {
            ti += 1;
          } else {  // nodes: south pole, row 0, row 1
            a[0] = (lat_d - (-90.0 + dx * (ti + 0.5))) * (lat_d - (-90.0 + dx * (ti + 1.5))) / (dx * dx * 0.5 * 1.5);
            a[1] = -(lat_d - (-90.0)) * (lat_d - (-90.0 + dx * (ti + 1.5))) / (dx * dx * 0.5);
            a[2] = (lat_d - (-90.0)) * (lat_d - (-90.0 + dx * (ti + 0.5))) / (dx * dx * 1.5);
            done = true;
          }
        }
        if (ti == nlat - 1) {
          if (lat_d < 90.0 - 0.5 * dx) {
            ti -= 1;
          } else {  // nodes: row nlat-2, row nlat-1, north pole
            a[0] = (lat_d - (-90.0 + dx * (ti + 0.5))) * (lat_d - 90.0) / (dx * dx * 1.5);
            a[1] = -(lat_d - (-90.0 + dx * (ti - 0.5))) * (lat_d - 90.0) / (dx * dx * 0.5);
            a[2] = (lat_d - (-90.0 + dx * (ti - 0.5))) * (lat_d - (-90.0 + dx * (ti + 0.5))) / (dx * dx * 0.5 * 1.5);
            done = true;
          }
        }
        if (!done) {
          a[0] = (lat_d - (-90.0 + dx * (ti + 0.5))) * (lat_d - (-90.0 + dx * (ti + 1.5))) / (2.0 * dx * dx);
          a[1] = -(lat_d - (-90.0 + dx * (ti - 0.5))) * (lat_d - (-90.0 + dx * (ti + 1.5))) / (dx * dx);
          a[2] = (lat_d - (-90.0 + dx * (ti - 0.5))) * (lat_d - (-90.0 + dx * (ti + 0.5))) / (2.0 * dx * dx);
        }
        double c[3];
        lon_coeffs(lon_d, li, c);
        const int lm = ((li - 1) + nlon) % nlon, lp = (li + 1) % nlon;
        const int cols[3] = {lm, li, lp};
        double vv[3][3], hh[3][3];
        for (int l = 0; l < 3; l++) {
          vv[0][l] = ti == 0 ? south[0] : vval(ti - 1, cols[l]);
          hh[0][l] = ti == 0 ? south[1] : hval(ti - 1, cols[l]);
          vv[1][l] = vval(ti, cols[l]);
          hh[1][l] = hval(ti, cols[l]);
          vv[2][l] = ti == nlat - 1 ? north[0] : vval(ti + 1, cols[l]);
          hh[2][l] = ti == nlat - 1 ? north[1] : hval(ti + 1, cols[l]);
        }
        double av = 0, ah = 0;
        for (int r = 0; r < 3; r++)
          for (int l = 0; l < 3; l++) {
            av += vv[r][l] * a[r] * c[l];
            ah += hh[r][l] * a[r] * c[l];
          }
        a_vor[k] = av;
        a_h[k] = ah;
      }
      a_div[k] = div[k] + ds * dt * div_t[k];
    }
  }
  return 0;
}

// field_averaging (swe_time_step_ll.cpp:270-363): rows with |lat| > 60 degrees
// (lat of column 0) get the 5-point average with weight epsilon, times cos^2(lat);
// the other rows the same form with weight 0 (:314).  DEVIATION: the reference
// reads field(i, jp) and field(i, jm) before it sets jp and jm (:322-323, :346-347),
// which is undefined; here they are the periodic neighbours j+1 and j-1 that it
// sets just after.  The first and last rows use one latitude neighbour, weight 3.
void orc_ll_field_averaging(int nlat, int nlon, const double* x, const double* y, const double* z,
                            const double* f, double* out, double epsilon) {
  const LlGrid g = ll_grid(nlat, nlon, 1.0);
  for (int i = 0; i < nlat; i++) {
    const double lat = ll_lat(x, y, z, g.at(i, 0));
    const bool polar = std::abs(lat) > kPi / 3.0;
    const double e = polar ? epsilon : 0.00;
    for (int j = 0; j < nlon; j++) {
      const int64_t k = g.at(i, j);
      const double f1 = e * f[g.at(i, g.jp(j))], f2 = e * f[g.at(i, g.jm(j))];
      double f3, f4, coeff;
      if (i > 0 && i < nlat - 1) {
        f3 = e * f[g.at(i + 1, j)];
        f4 = e * f[g.at(i - 1, j)];
        coeff = 4;
      } else {
        f3 = e * f[g.at(i == 0 ? 1 : i - 1, j)];
        f4 = 0;
        coeff = 3;
      }
      double val = (1.0 - coeff * e) * f[k] + f1 + f2 + f3 + f4;
      if (polar) val *= std::cos(lat) * std::cos(lat);
      out[k] = val;
    }
  }
}

// laplacian_diffusion (swe_time_step_ll.cpp:233-268): interior rows get
// dt * 0.89/6371000 * (the 5-point spherical Laplacian); the first and last rows
// are copied.
void orc_ll_laplacian_diffusion(int nlat, int nlon, double dx_deg, double dt, const double* x, const double* y,
                                const double* z, const double* f, double* out) {
  const LlGrid g = ll_grid(nlat, nlon, dx_deg);
  const double coeff = 0.89 / 6371000.0;
  for (int i = 0; i < nlat; i++) {
    for (int j = 0; j < nlon; j++) {
      const int64_t k = g.at(i, j);
      out[k] = f[k];
      if (i > 0 && i < nlat - 1) {
        const double c = f[k], e = f[g.at(i, g.jp(j))], w = f[g.at(i, g.jm(j))];
        const double nn = f[g.at(i + 1, j)], ss = f[g.at(i - 1, j)];
        const double lat = ll_lat(x, y, z, k);
        const double l1 = 1.0 / (g.dxr * g.dxr * std::cos(lat) * std::cos(lat)) * (e - 2.0 * c + w);
        const double l2 = std::tan(lat) * (nn - ss) / (2.0 * g.dxr);
        const double l3 = (nn - 2.0 * c + ss) / (g.dxr * g.dxr);
        out[k] += (l1 - l2 + l3) * dt * coeff;
      }
    }
  }
}

void orc_set_threads(int n) { omp_set_num_threads(n); }
int orc_get_threads() { return omp_get_max_threads(); }

}  // extern "C"
