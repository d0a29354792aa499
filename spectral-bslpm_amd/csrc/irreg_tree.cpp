// irreg_tree.cpp -- host build of the irregular particle tree (see irreg_tree.hpp).
//
// Built with -ffp-contract=off like tree.cpp: point assignment compares label
// coordinates with panel mid-lines, so the geometry must round as written.
#include "irreg_tree.hpp"

#include <algorithm>
#include <cmath>
#include <deque>
#include <utility>

#include "geometry.hpp"

namespace sbslpm {
namespace {

// xieta_from_xyz without a face (cubed_sphere_transforms_impl.hpp:67-98): the
// face rule of face_from_xyz, then that face's formulas.
void xieta_auto(double x, double y, double z, double* xe) { xieta_from_xyz_face(x, y, z, face_from_xyz(x, y, z), xe); }

double gcdist(const double* a, const double* b) {  // general_utils.cpp:23-26
  double dp = std::max(std::min(a[0] * b[0] + a[1] * b[1] + a[2] * b[2], 1.0), -1.0);
  return std::acos(dp);
}

// initialize_cubed_sphere.cpp:446-592.
void split_panels(const double* x, const double* y, const double* z, int64_t n, IrregTree& t) {
  t.pn.assign(6, IrregPanel{});
  t.points.assign(6, {});
  for (int f = 0; f < 6; ++f) {
    IrregPanel& p = t.pn[f];
    p.face = f; p.level = 0; p.parent = -1; p.is_leaf = 1; p.point_count = 0;
    p.child[0] = p.child[1] = p.child[2] = p.child[3] = -1;
    p.min_xi = p.min_eta = -kPi / 4.0;
    p.max_xi = p.max_eta = kPi / 4.0;
  }
  for (int64_t i = 0; i < n; ++i) {
    const int f = face_from_xyz(x[i], y[i], z[i]);
    t.pn[f].point_count += 1;
    t.points[f].push_back((int)i);
  }
  for (size_t i = 0; i < t.pn.size(); ++i) {
    if (t.pn[i].point_count <= t.leaf_size) continue;
    const IrregPanel par = t.pn[i];
    const double mx = 0.5 * (par.min_xi + par.max_xi), me = 0.5 * (par.min_eta + par.max_eta);
    const double lo_x[4] = {mx, par.min_xi, par.min_xi, mx}, hi_x[4] = {par.max_xi, mx, mx, par.max_xi};
    const double lo_e[4] = {me, me, par.min_eta, par.min_eta}, hi_e[4] = {par.max_eta, par.max_eta, me, me};
    const int start = (int)t.pn.size();
    for (int c = 0; c < 4; ++c) {
      IrregPanel ch{};
      ch.face = par.face; ch.level = par.level + 1; ch.parent = (int)i; ch.is_leaf = 1; ch.point_count = 0;
      ch.child[0] = ch.child[1] = ch.child[2] = ch.child[3] = -1;
      ch.min_xi = lo_x[c]; ch.max_xi = hi_x[c]; ch.min_eta = lo_e[c]; ch.max_eta = hi_e[c];
      t.pn.push_back(ch);
      t.pn[i].child[c] = start + c;
    }
    t.pn[i].is_leaf = 0;
    t.points.resize(start + 4);
    for (int id : t.points[i]) {
      double xe[2];
      xieta_auto(x[id], y[id], z[id], xe);
      int which;
      if (xe[0] < mx)
        which = xe[1] < me ? 2 : 1;
      else
        which = xe[1] < me ? 3 : 0;
      t.pn[start + which].point_count += 1;
      t.points[start + which].push_back(id);
    }
  }
  for (auto& p : t.pn) {  // radius: largest arc from the centre to a corner
    const double mx = 0.5 * (p.min_xi + p.max_xi), me = 0.5 * (p.min_eta + p.max_eta);
    double pc[3], c[4][3];
    xyz_from_xieta(mx, me, p.face, pc);
    xyz_from_xieta(p.max_xi, p.min_eta, p.face, c[0]);
    xyz_from_xieta(p.max_xi, p.max_eta, p.face, c[1]);
    xyz_from_xieta(p.min_xi, p.max_eta, p.face, c[2]);
    xyz_from_xieta(p.min_xi, p.min_eta, p.face, c[3]);
    double r = 0;
    for (int k = 0; k < 4; ++k) r = std::max(r, std::acos(pc[0] * c[k][0] + pc[1] * c[k][1] + pc[2] * c[k][2]));
    p.radius = r;
  }
  for (size_t i = 0; i < t.pn.size(); ++i)
    if (!t.pn[i].is_leaf) t.points[i].clear();
}

// fmm_funcs.cpp:125-231.
void traverse_irreg(IrregTree& t) {
  std::deque<std::pair<int, int>> q;
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) q.emplace_back(a, b);
  t.ilist.clear();
  while (!q.empty()) {
    const auto [ti, si] = q.front();
    q.pop_front();
    const IrregPanel &T = t.pn[ti], &S = t.pn[si];
    double c1[3], c2[3];
    xyz_from_xieta(0.5 * (T.min_xi + T.max_xi), 0.5 * (T.min_eta + T.max_eta), T.face, c1);
    xyz_from_xieta(0.5 * (S.min_xi + S.max_xi), 0.5 * (S.min_eta + S.max_eta), S.face, c2);
    const double dist = gcdist(c1, c2);
    double sep = 100;
    if (dist > 1e-16) sep = (T.radius + S.radius) / dist;
    if (sep < t.theta) {
      const int type = (T.level < t.levels ? 2 : 0) + (S.level < t.levels ? 1 : 0);
      t.ilist.insert(t.ilist.end(), {ti, si, type});
    } else if (T.is_leaf && S.is_leaf) {
      t.ilist.insert(t.ilist.end(), {ti, si, 0});
    } else if (T.is_leaf) {
      for (int c = 0; c < 4; ++c) q.emplace_back(ti, S.child[c]);
    } else if (S.is_leaf) {
      for (int c = 0; c < 4; ++c) q.emplace_back(T.child[c], si);
    } else if (T.level <= S.level) {
      for (int c = 0; c < 4; ++c) q.emplace_back(T.child[c], si);
    } else {
      for (int c = 0; c < 4; ++c) q.emplace_back(ti, S.child[c]);
    }
  }
}

}  // namespace

std::string build_irreg_tree(const double* x, const double* y, const double* z, int64_t n, int leaf_size, int degree,
                             double theta, IrregTree& t) {
  if (n < 1 || n > (int64_t)1 << 30) return "point count must be in 1..2^30";
  if (leaf_size < 1) return "leaf_size must be positive";
  if (degree < 1 || degree > kMaxDeg) return "interp_degree must be in 1..10";
  if (!(theta > 0)) return "theta must be positive";
  for (int64_t i = 0; i < n; ++i)
    if (!(std::isfinite(x[i]) && std::isfinite(y[i]) && std::isfinite(z[i]))) return "points must be finite";
  t = IrregTree();
  t.leaf_size = leaf_size;
  t.degree = degree;
  t.dp1 = degree + 1;
  t.n = t.dp1 * t.dp1;
  t.theta = theta;
  t.N = n;
  split_panels(x, y, z, n, t);
  t.P = (int)t.pn.size();
  t.levels = 0;
  for (const auto& p : t.pn) t.levels = std::max(t.levels, p.level);
  t.level_start.assign(t.levels + 2, t.P);
  for (int i = t.P - 1; i >= 0; --i) t.level_start[t.pn[i].level] = i;  // panels are in breadth-first order
  traverse_irreg(t);

  // leaf-sorted points
  t.perm.clear();
  t.leaf_first.assign(t.P, -1);
  t.point_leaf.clear();
  for (int p = 0; p < t.P; ++p) {
    if (!t.pn[p].is_leaf) continue;
    t.leaf_first[p] = (int)t.perm.size();
    for (int id : t.points[p]) {
      t.perm.push_back(id);
      t.point_leaf.push_back(p);
    }
  }
  if ((int64_t)t.perm.size() != n) return "point assignment lost points";

  // Chebyshev nodes and unified coordinates (proxies as swe_vel_impl.hpp:206-218)
  t.cheb.resize(t.dp1);
  for (int i = 0; i <= degree; ++i) t.cheb[i] = std::cos(kPi * i / degree);
  t.U = n + (int64_t)t.P * t.n;
  t.px.resize(t.U); t.py.resize(t.U); t.pz.resize(t.U);
  for (int64_t k = 0; k < n; ++k) {
    t.px[k] = x[t.perm[k]]; t.py[k] = y[t.perm[k]]; t.pz[k] = z[t.perm[k]];
  }
  t.dev.resize(t.P);
  for (int p = 0; p < t.P; ++p) {
    const IrregPanel& q = t.pn[p];
    const double xo = 0.5 * (q.min_xi + q.max_xi), xs = 0.5 * (q.max_xi - q.min_xi);
    const double eo = 0.5 * (q.min_eta + q.max_eta), es = 0.5 * (q.max_eta - q.min_eta);
    for (int j = 0; j <= degree; ++j)
      for (int k = 0; k <= degree; ++k) {
        double v[3];
        xyz_from_xieta(t.cheb[j] * xs + xo, t.cheb[k] * es + eo, q.face, v);
        const int64_t u = n + (int64_t)p * t.n + j * t.dp1 + k;
        t.px[u] = v[0]; t.py[u] = v[1]; t.pz[u] = v[2];
      }
    DevPanel& d = t.dev[p];
    d.min_xi = q.min_xi; d.max_xi = q.max_xi; d.min_eta = q.min_eta; d.max_eta = q.max_eta;
    d.mid_xi = xo; d.mid_eta = eo;
    double c[3];
    xyz_from_xieta(xo, eo, q.face, c);
    d.cx = c[0]; d.cy = c[1]; d.cz = c[2]; d.pad0 = 0;
    d.face = q.face; d.is_leaf = q.is_leaf; d.child0 = q.child[0]; d.level = q.level;
  }

  // target-major tiles: per panel its proxy rows (types 2, 3), then its points (types 0, 1)
  const int64_t I = t.interaction_count();
  std::vector<std::vector<int>> by_key(2 * (size_t)t.P);
  for (int64_t e = 0; e < I; ++e) {
    const int tg = t.ilist[3 * e], type = t.ilist[3 * e + 2];
    by_key[2 * (size_t)tg + (type >= 2 ? 0 : 1)].push_back((int)e);
  }
  t.sets.clear();
  t.tiles.clear();
  std::vector<char> proxy_target(t.P, 0);
  for (size_t key = 0; key < by_key.size(); ++key) {
    const int p = (int)(key / 2);
    const bool particles = key & 1;
    const int npts = particles ? t.pn[p].point_count : t.n;
    if (by_key[key].empty() || npts == 0) continue;
    const int begin = (int)t.sets.size();
    for (int e : by_key[key]) {
      const int s = t.ilist[3 * e + 1], type = t.ilist[3 * e + 2];
      SourceSet ss{};
      ss.panel = s;
      if (type == 0 || type == 2) {
        ss.first = t.leaf_first[s];
        ss.count = t.pn[s].point_count;
      } else {
        ss.first = (int)(n + (int64_t)s * t.n);
        ss.count = t.n;
      }
      if (ss.count > 0) t.sets.push_back(ss);
    }
    const int end = (int)t.sets.size();
    if (end == begin) continue;
    if (!particles) proxy_target[p] = 1;
    const int64_t base = particles ? t.leaf_first[p] : n + (int64_t)p * t.n;
    const int chunks = (npts + 63) / 64, per = (npts + chunks - 1) / chunks;
    for (int f = 0; f < npts; f += per) {
      TargetTile tl{};
      tl.panel = p; tl.first = (int)(base + f); tl.count = std::min(per, npts - f);
      tl.src_begin = begin; tl.src_end = end;
      t.tiles.push_back(tl);
    }
  }
  t.zero_rows.clear();
  for (int p = 0; p < t.P; ++p)
    if (!proxy_target[p]) t.zero_rows.push_back(p);

  // non-leaf panels by level
  t.nonleaf.clear();
  t.level_nonleaf.assign(t.levels + 2, 0);
  for (int l = 0; l <= t.levels; ++l) {
    t.level_nonleaf[l] = (int)t.nonleaf.size();
    for (int p = t.level_start[l]; p < t.level_start[l + 1]; ++p)
      if (!t.pn[p].is_leaf) t.nonleaf.push_back(p);
  }
  t.level_nonleaf[t.levels + 1] = (int)t.nonleaf.size();

  // fixed child-to-parent matrices (children are exact halves in each direction):
  // A[a*(d+1)+k] = parent basis a at the child's node k
  std::vector<double> w(kMaxDeg + 1);
  t.m2m_hi.assign(t.dp1 * t.dp1, 0.0);
  t.m2m_lo.assign(t.dp1 * t.dp1, 0.0);
  for (int k = 0; k <= degree; ++k) {
    bli_weights_1d(0.5 * (t.cheb[k] + 1.0), -1.0, 1.0, degree, t.cheb.data(), w.data());
    for (int a = 0; a <= degree; ++a) t.m2m_hi[a * t.dp1 + k] = w[a];
    bli_weights_1d(0.5 * (t.cheb[k] - 1.0), -1.0, 1.0, degree, t.cheb.data(), w.data());
    for (int a = 0; a <= degree; ++a) t.m2m_lo[a * t.dp1 + k] = w[a];
  }
  return "";
}

}  // namespace sbslpm
