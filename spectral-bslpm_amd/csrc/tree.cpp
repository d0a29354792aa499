// tree.cpp -- host set-up: cubed-sphere quadtree, particle and proxy grids,
// dual-tree traversal, target-major work tiles, fixed transfer matrices.
//
// Built with -ffp-contract=off so the geometry rounds exactly as written,
// like the reference's host build (the proxy-to-leaf tie breaking at panel
// mid-lines depends on the last bit of atan(tan(xi))).
#include "tree.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <unordered_map>
#include <utility>

#include "geometry.hpp"

namespace sbslpm {
namespace {

double gcdist(const double* a, const double* b) {  // general_utils.cpp:23-26
  double dp = std::max(std::min(a[0] * b[0] + a[1] * b[1] + a[2] * b[2], 1.0), -1.0);
  return std::acos(dp);
}

double tri_area(const double* p1, const double* p2, const double* p3) {  // general_utils.cpp:5-21
  double e1 = 0, e2 = 0, e3 = 0;  // squared chord lengths |p2-p3|^2, |p3-p1|^2, |p1-p2|^2
  for (int k = 0; k < 3; ++k) {
    double u = p2[k] - p3[k], v = p3[k] - p1[k], w = p1[k] - p2[k];
    e1 += u * u; e2 += v * v; e3 += w * w;
  }
  double a = std::acos(1 - 0.5 * e1), b = std::acos(1 - 0.5 * e2), c = std::acos(1 - 0.5 * e3);
  double s = (a + b + c) / 2;
  double z = std::tan(s / 2) * std::tan((s - a) / 2) * std::tan((s - b) / 2) * std::tan((s - c) / 2);
  return 4 * std::atan(std::sqrt(z));
}

// Quadrature weights of interp_init (interp_funcs.cpp:42-60), literal formula.
void reference_quadrature(int deg, std::vector<double>& w) {
  w.assign(deg + 1, 0.0);
  int h = deg / 2;
  std::vector<double> vec(h + 1, 2.0);
  vec[0] = 1;
  for (int i = 1; i <= h; i++) vec[i] = 2.0 / (1 - 4 * i * i);
  for (int i = 0; i <= h; i++) {
    double sum = 0;
    for (int j = 0; j <= h; j++) sum += 2 * std::cos(2 * i * j * kPi / deg) / deg * vec[j];
    w[i] = sum;
  }
  w[0] *= 0.5;
  for (int i = 0; i < h; i++) w[deg - i] = w[i];
}

// Clenshaw-Curtis weights for nodes cos(k pi/N) (Waldvogel's closed form).
void clenshaw_curtis(int N, std::vector<double>& w) {
  w.assign(N + 1, 0.0);
  if (N == 0) { w[0] = 2; return; }
  for (int k = 0; k <= N; ++k) {
    double s = 0;
    for (int j = 1; j <= N / 2; ++j) {
      double b = (2 * j == N) ? 1.0 : 2.0;
      s += b / (4.0 * j * j - 1.0) * std::cos(2.0 * j * k * kPi / N);
    }
    double c = (k == 0 || k == N) ? 1.0 : 2.0;
    w[k] = c / N * (1.0 - s);
  }
}

// Panel frame in radians as the interaction kernels compute it
// (bve_vel_impl.hpp:38-46).
inline void frame(const HostPanel& p, double& xo, double& xs, double& eo, double& es) {
  xo = kPi / 4 * 0.5 * (p.max_xi + p.min_xi);
  xs = kPi / 4 * 0.5 * (p.max_xi - p.min_xi);
  eo = kPi / 4 * 0.5 * (p.max_eta + p.min_eta);
  es = kPi / 4 * 0.5 * (p.max_eta - p.min_eta);
}

void make_panels(Tree& t) {
  const int L = t.cfg.levels;
  t.level_start.assign(L + 1, 0);
  for (int l = 0; l < L; ++l) t.level_start[l + 1] = t.level_start[l] + 6 * (1 << (2 * l));
  t.P = t.level_start[L];
  t.Pl = 6 * (1 << (2 * (L - 1)));
  t.offset = t.P - t.Pl;
  t.pn.assign(t.P, HostPanel{});
  for (auto& p : t.pn) { p.parent = -1; p.child[0] = p.child[1] = p.child[2] = p.child[3] = -1; }
  for (int f = 0; f < 6; ++f) {  // initialize_cubed_sphere.cpp:187-196
    HostPanel& p = t.pn[f];
    p.face = f; p.level = 0;
    p.min_xi = p.min_eta = -1; p.max_xi = p.max_eta = 1;
  }
  // Quadrant refinement, children in the order (hi,hi), (lo,hi), (lo,lo),
  // (hi,lo) in (xi, eta) -- initialize_cubed_sphere.cpp:73-140.
  for (int l = 0; l + 1 < L; ++l) {
    for (int i = t.level_start[l]; i < t.level_start[l + 1]; ++i) {
      HostPanel& par = t.pn[i];
      int first = t.level_start[l + 1] + 4 * (i - t.level_start[l]);
      double mx = 0.5 * (par.max_xi + par.min_xi), me = 0.5 * (par.max_eta + par.min_eta);
      const double lo_x[4] = {mx, par.min_xi, par.min_xi, mx};
      const double hi_x[4] = {par.max_xi, mx, mx, par.max_xi};
      const double lo_e[4] = {me, me, par.min_eta, par.min_eta};
      const double hi_e[4] = {par.max_eta, par.max_eta, me, me};
      for (int c = 0; c < 4; ++c) {
        HostPanel& ch = t.pn[first + c];
        ch.face = par.face; ch.level = l + 1; ch.parent = i;
        ch.min_xi = lo_x[c]; ch.max_xi = hi_x[c]; ch.min_eta = lo_e[c]; ch.max_eta = hi_e[c];
        par.child[c] = first + c;
      }
    }
  }
  // Leaf flag, radius and area: initialize_cubed_sphere.cpp:151-182.
  for (auto& p : t.pn) {
    p.is_leaf = (p.level == L - 1);
    double ax = p.max_xi * kPi / 4.0, ix = p.min_xi * kPi / 4.0;
    double ae = p.max_eta * kPi / 4.0, ie = p.min_eta * kPi / 4.0;
    double mx = 0.5 * (ax + ix), me = 0.5 * (ae + ie);
    double c1[3], c2[3], c3[3], c4[3], cc[3];
    xyz_from_xieta(ax, ae, p.face, c1);
    xyz_from_xieta(ax, ie, p.face, c2);
    xyz_from_xieta(ix, ie, p.face, c3);
    xyz_from_xieta(ix, ae, p.face, c4);
    xyz_from_xieta(mx, me, p.face, cc);
    p.radius = std::fmax(gcdist(c1, cc), std::fmax(gcdist(c2, cc), std::fmax(gcdist(c3, cc), gcdist(c4, cc))));
    p.area = tri_area(c1, c2, c3) + tri_area(c1, c3, c4);
  }
}

// Dual-tree traversal, fmm_funcs.cpp:17-123: breadth-first over (target,
// source) pairs seeded with the 36 face pairs; a pair is accepted when
// (r_t + r_s)/dist < theta, kept as a leaf-leaf pair when both are leaves,
// and otherwise the target is split when its level is not finer.
void traverse(Tree& t) {
  const int leaf_level = t.cfg.levels - 1;
  std::deque<std::pair<int, int>> q;
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) q.emplace_back(a, b);
  t.ilist.clear();
  while (!q.empty()) {
    auto [ti, si] = q.front();
    q.pop_front();
    const HostPanel &T = t.pn[ti], &S = t.pn[si];
    double ct[3], cs[3];
    xyz_from_xieta(kPi / 4.0 * 0.5 * (T.min_xi + T.max_xi), kPi / 4.0 * 0.5 * (T.min_eta + T.max_eta), T.face, ct);
    xyz_from_xieta(kPi / 4.0 * 0.5 * (S.min_xi + S.max_xi), kPi / 4.0 * 0.5 * (S.min_eta + S.max_eta), S.face, cs);
    double dist = gcdist(ct, cs);
    double sep = dist > 1e-16 ? (T.radius + S.radius) / dist : 100.0;
    if (sep < t.cfg.theta) {
      int type = (T.level < leaf_level ? 2 : 0) + (S.level < leaf_level ? 1 : 0);
      t.ilist.insert(t.ilist.end(), {ti, si, type});
    } else if (T.is_leaf && S.is_leaf) {
      t.ilist.insert(t.ilist.end(), {ti, si, 0});
    } else if (T.is_leaf || (!S.is_leaf && T.level > S.level)) {
      for (int c = 0; c < 4; ++c) q.emplace_back(ti, S.child[c]);
    } else {
      for (int c = 0; c < 4; ++c) q.emplace_back(T.child[c], si);
    }
  }
}

}  // namespace

double sphere_tri_area(const double* p1, const double* p2, const double* p3) { return tri_area(p1, p2, p3); }

int locate_leaf(const Tree& t, double x, double y, double z) {
  double xe[2];
  return locate_uniform(t.cfg.levels, x, y, z, xe);
}

// Composite transfer matrices of the fused downward passes: the basis of an
// ancestor g levels above a panel at that panel's nodes, one 1-D matrix per
// position of the panel's interval among the ancestor's 2^g equal
// sub-intervals (counted from the lower end), g = 1 .. maxgap.  Children are
// exact halves and the interpolants are degree-d polynomials in each
// direction, so the reference's level-by-level L2L chain evaluates each
// ancestor's interpolant at the panel's nodes; these matrices do it in one
// product.  Layout [(2^g - 2 + pos)][k][a]: descendant node k, ancestor basis
// a (the transpose of m2m_hi's layout).
void composite_l2l(int degree, const double* node, const double* cheb, int maxgap, std::vector<double>& out) {
  const int dp1 = degree + 1, n = dp1 * dp1;
  std::vector<double> w(dp1);
  out.clear();
  for (int g = 1; g <= maxgap; ++g) {
    const int np = 1 << g;
    const double wd = 2.0 / np;
    out.resize((size_t)(2 * np - 2) * n, 0.0);
    for (int pos = 0; pos < np; ++pos)
      for (int k = 0; k <= degree; ++k) {
        bli_weights_1d(-1.0 + pos * wd + 0.5 * (node[k] + 1.0) * wd, -1.0, 1.0, degree, cheb, w.data());
        for (int a = 0; a <= degree; ++a) out[(size_t)(np - 2 + pos) * n + k * dp1 + a] = w[a];
      }
  }
}

std::string build_tree(const TreeConfig& cfg, Tree& t) {
  if (cfg.levels < 1 || cfg.levels > 10) return "levels must be in 1..10";
  if (cfg.degree < 1 || cfg.degree > kMaxDeg) return "interp_degree must be in 1..10";
  if (cfg.leaf_degree < 1 || cfg.leaf_degree > kMaxDeg) return "leaf_degree must be in 1..10";
  if (!(cfg.theta > 0)) return "theta must be positive";
  if (cfg.weight_rule < 0 || cfg.weight_rule > 2) return "weight_rule must be 0, 1 or 2";
  t = Tree();
  t.cfg = cfg;
  const int d = cfg.degree, md = cfg.leaf_degree;
  t.dp1 = d + 1; t.n = t.dp1 * t.dp1;
  t.m = md + 1; t.m2 = t.m * t.m;

  // Chebyshev tables (interp_funcs.cpp:33-39 and interp_funcs_impl.hpp:45).
  t.cheb_d.resize(t.dp1); t.node_d.resize(t.dp1);
  t.cheb_m.resize(t.m); t.node_m.resize(t.m);
  for (int i = 0; i <= d; ++i) {
    t.cheb_d[i] = std::cos(kPi * i / d);
    double v = std::cos(kPi / d * i);
    t.node_d[i] = std::fabs(v) < 1e-16 ? 0.0 : v;
  }
  for (int i = 0; i <= md; ++i) {
    t.cheb_m[i] = std::cos(kPi * i / md);
    double v = std::cos(kPi / md * i);
    t.node_m[i] = std::fabs(v) < 1e-16 ? 0.0 : v;
  }
  if (cfg.weight_rule == 1) clenshaw_curtis(md, t.quad_m);
  else reference_quadrature(md, t.quad_m);

  make_panels(t);
  traverse(t);
  t.npart = (int64_t)t.Pl * t.m2;
  t.U = t.npart + (int64_t)t.P * t.n;
  t.m2p = (t.m2 + 15) / 16 * 16;
  t.np16 = (t.n + 15) / 16 * 16;
  t.U16 = (int64_t)t.Pl * t.m2p + (int64_t)t.offset * t.np16;

  // Device panels.
  t.dev.resize(t.P);
  for (int i = 0; i < t.P; ++i) {
    const HostPanel& p = t.pn[i];
    DevPanel& q = t.dev[i];
    q.min_xi = kPi / 4.0 * p.min_xi; q.max_xi = kPi / 4.0 * p.max_xi;
    q.min_eta = kPi / 4.0 * p.min_eta; q.max_eta = kPi / 4.0 * p.max_eta;
    q.mid_xi = 0.5 * kPi / 4.0 * (p.max_xi + p.min_xi);
    q.mid_eta = 0.5 * kPi / 4.0 * (p.max_eta + p.min_eta);
    double c[3];
    xyz_from_xieta(kPi / 4.0 * 0.5 * (p.min_xi + p.max_xi), kPi / 4.0 * 0.5 * (p.min_eta + p.max_eta), p.face, c);
    q.cx = c[0]; q.cy = c[1]; q.cz = c[2]; q.pad0 = 0;
    q.face = p.face; q.is_leaf = p.is_leaf; q.child0 = p.child[0]; q.level = p.level;
  }

  // Unified point coordinates: leaf particle grids, then every panel's proxies,
  // node index = xi_index*(deg+1) + eta_index (interp_funcs.cpp:64-66).
  t.px.resize(t.U); t.py.resize(t.U); t.pz.resize(t.U);
  for (int i = 0; i < t.P; ++i) {
    const HostPanel& p = t.pn[i];
    double xo, xs, eo, es, v[3];
    frame(p, xo, xs, eo, es);
    if (p.is_leaf) {
      int64_t b = (int64_t)(i - t.offset) * t.m2;
      for (int a = 0; a < t.m; ++a)
        for (int c = 0; c < t.m; ++c) {
          xyz_from_xieta(t.node_m[a] * xs + xo, t.node_m[c] * es + eo, p.face, v);
          int64_t u = b + a * t.m + c;
          t.px[u] = v[0]; t.py[u] = v[1]; t.pz[u] = v[2];
        }
    }
    int64_t b = t.npart + (int64_t)i * t.n;
    for (int a = 0; a < t.dp1; ++a)
      for (int c = 0; c < t.dp1; ++c) {
        xyz_from_xieta(t.node_d[a] * xs + xo, t.node_d[c] * es + eo, p.face, v);
        int64_t u = b + a * t.dp1 + c;
        t.px[u] = v[0]; t.py[u] = v[1]; t.pz[u] = v[2];
      }
  }

  // Caller-facing particles and weights (initialize_cubed_sphere.cpp:254-277).
  t.spec_x.resize(t.npart); t.spec_y.resize(t.npart); t.spec_z.resize(t.npart); t.spec_area.resize(t.npart);
  const double equal_area = 4.0 * kPi / (double)t.npart;
  for (int i = t.offset; i < t.P; ++i) {
    const HostPanel& p = t.pn[i];
    double ax = kPi / 4.0 * p.max_xi, ix = kPi / 4.0 * p.min_xi;
    double ae = kPi / 4.0 * p.max_eta, ie = kPi / 4.0 * p.min_eta;
    double xr = 0.5 * (ax - ix), xo = 0.5 * (ax + ix), er = 0.5 * (ae - ie), eo = 0.5 * (ae + ie);
    for (int a = 0; a < t.m; ++a)
      for (int c = 0; c < t.m; ++c) {
        double v[3];
        xyz_from_xieta(xo + xr * t.node_m[a], eo + er * t.node_m[c], p.face, v);
        int64_t u = (int64_t)(i - t.offset) * t.m2 + a * t.m + c;
        t.spec_x[u] = v[0]; t.spec_y[u] = v[1]; t.spec_z[u] = v[2];
        t.spec_area[u] = cfg.weight_rule == 2 ? equal_area : (t.quad_m[a] * t.quad_m[c]) * p.area / 4.0;
      }
  }

  // Target-major tiles and source sets (interaction order kept per target).
  const int64_t I = t.interaction_count();
  std::vector<int> cnt(t.P + 1, 0);
  for (int64_t e = 0; e < I; ++e) cnt[t.ilist[3 * e] + 1]++;
  for (int p = 0; p < t.P; ++p) cnt[p + 1] += cnt[p];
  std::vector<int> order(I), fill(cnt.begin(), cnt.end() - 1);
  for (int64_t e = 0; e < I; ++e) order[fill[t.ilist[3 * e]]++] = (int)e;
  t.srcsets.resize(I);
  for (int64_t k = 0; k < I; ++k) {
    int s = t.ilist[3 * order[k] + 1], tg = t.ilist[3 * order[k]];
    SourceSet ss{};
    ss.panel = s;
    ss.m2l = (!t.pn[s].is_leaf && !t.pn[tg].is_leaf) ? 1 : 0;
    ss.ox = (float)(t.dev[tg].cx - t.dev[s].cx);
    ss.oy = (float)(t.dev[tg].cy - t.dev[s].cy);
    ss.oz = (float)(t.dev[tg].cz - t.dev[s].cz);
    if (t.pn[s].is_leaf) {
      ss.first = (int)((int64_t)(s - t.offset) * t.m2);
      ss.count = t.m2;
      ss.first16 = (int)((int64_t)(s - t.offset) * t.m2p);
    } else {
      ss.first = (int)(t.npart + (int64_t)s * t.n);
      ss.count = t.n;
      ss.first16 = (int)((int64_t)t.Pl * t.m2p + (int64_t)s * t.np16);
    }
    t.srcsets[k] = ss;
  }
  t.tiles.clear();
  for (int p = 0; p < t.P; ++p) {
    if (cnt[p + 1] == cnt[p]) continue;
    int npts = t.pn[p].is_leaf ? t.m2 : t.n;
    int64_t base = t.pn[p].is_leaf ? (int64_t)(p - t.offset) * t.m2 : t.npart + (int64_t)p * t.n;
    // balanced chunks of at most 64 points (81 -> 41 + 40, i.e. 3 + 3 MFMA tiles of 16)
    int chunks = (npts + 63) / 64, per = (npts + chunks - 1) / chunks;
    for (int f = 0; f < npts; f += per) {
      TargetTile tl{};
      tl.panel = p; tl.first = (int)(base + f); tl.count = std::min(per, npts - f);
      tl.src_begin = cnt[p]; tl.src_end = cnt[p + 1];
      t.tiles.push_back(tl);
    }
  }

  t.set_start = cnt;
  t.nonleaf_tiles = 0;
  while (t.nonleaf_tiles < (int)t.tiles.size() && !t.pn[t.tiles[t.nonleaf_tiles].panel].is_leaf) ++t.nonleaf_tiles;

  // Eight contiguous ranges of the tile list with equal estimated work, one per
  // XCD (the list is in breadth-first panel order, so a range of leaf tiles is a
  // compact patch of the sphere and its sources stay in that XCD's L2), for two
  // work models: xcd_start[0..8] for the matrix-core kernels (a tile's 16 x 16
  // pair blocks over its source sets plus one block per set for the per-set
  // flush) and xcd_start[9..17] for the vector-unit kernels, whose wave walks
  // every source of a set whatever the tile's target count (the set's sources
  // plus 16 for the per-set work).
  {
    const size_t nt = t.tiles.size();
    t.xcd_start.assign(18, (int)nt);
    for (int model = 0; model < 2; ++model) {
      std::vector<double> pre(nt + 1, 0.0);
      for (size_t k = 0; k < nt; ++k) {
        const TargetTile& tl = t.tiles[k];
        const double tb = (tl.count + 15) / 16;
        double w = 0;
        for (int e = tl.src_begin; e < tl.src_end; ++e)
          w += model == 0 ? tb * ((t.srcsets[e].count + 15) / 16) + 1.0 : t.srcsets[e].count + 16.0;
        pre[k + 1] = pre[k] + w;
      }
      int* xs = t.xcd_start.data() + 9 * model;
      xs[0] = 0;
      size_t k = 0;
      for (int x = 1; x < 8; ++x) {
        const double goal = pre[nt] * x / 8.0;
        while (k < nt && pre[k] < goal) ++k;
        xs[x] = (int)k;
      }
    }
  }

  // Proxy rows that are not interaction targets: non-leaf panels without
  // tiles, and the leaf rows when there is no L2L level to write them.
  t.zero_rows.clear();
  for (int p = 0; p < t.P; ++p)
    if ((!t.pn[p].is_leaf && cnt[p + 1] == cnt[p]) || (t.pn[p].is_leaf && t.cfg.levels == 1)) t.zero_rows.push_back(p);

  // Fixed 1-D transfer matrices.  Children are exact quadrants, so the parent
  // basis at a child node (fmm_funcs_impl.hpp:61-64, 110-113) depends only on
  // whether the child is the upper or lower half in each direction.
  std::vector<double> w(kMaxDeg + 1);
  t.m2m_hi.assign(t.dp1 * t.dp1, 0.0);
  t.m2m_lo.assign(t.dp1 * t.dp1, 0.0);
  for (int k = 0; k <= d; ++k) {
    bli_weights_1d(0.5 * (t.node_d[k] + 1.0), -1.0, 1.0, d, t.cheb_d.data(), w.data());
    for (int a = 0; a <= d; ++a) t.m2m_hi[a * t.dp1 + k] = w[a];
    bli_weights_1d(0.5 * (t.node_d[k] - 1.0), -1.0, 1.0, d, t.cheb_d.data(), w.data());
    for (int a = 0; a <= d; ++a) t.m2m_lo[a * t.dp1 + k] = w[a];
  }
  t.p2m.assign(t.dp1 * t.m, 0.0);
  for (int k = 0; k < t.m; ++k) {
    bli_weights_1d(t.node_m[k], -1.0, 1.0, d, t.cheb_d.data(), w.data());
    for (int a = 0; a <= d; ++a) t.p2m[a * t.m + k] = w[a];
  }
  // The parent basis at the particle nodes of its upper and lower half
  // children (m2m_hi p2m and m2m_lo p2m, evaluated directly), stored [k][a].
  t.l2p_hi.assign(t.dp1 * t.m, 0.0);
  t.l2p_lo.assign(t.dp1 * t.m, 0.0);
  t.l2pT_hi.assign(t.dp1 * t.m, 0.0);
  t.l2pT_lo.assign(t.dp1 * t.m, 0.0);
  for (int k = 0; k < t.m; ++k) {
    bli_weights_1d(0.5 * (t.node_m[k] + 1.0), -1.0, 1.0, d, t.cheb_d.data(), w.data());
    for (int a = 0; a <= d; ++a) t.l2p_hi[k * t.dp1 + a] = t.l2pT_hi[a * t.m + k] = w[a];
    bli_weights_1d(0.5 * (t.node_m[k] - 1.0), -1.0, 1.0, d, t.cheb_d.data(), w.data());
    for (int a = 0; a <= d; ++a) t.l2p_lo[k * t.dp1 + a] = t.l2pT_lo[a * t.m + k] = w[a];
  }
  composite_l2l(d, t.node_d.data(), t.cheb_d.data(), t.cfg.levels - 2, t.anc_l2l);

  // Leaf containing each non-leaf proxy node, located from its Cartesian
  // position exactly as bve_disp_interp does (bve_time_step.cpp:139-143).
  t.proxy_leaf.assign((size_t)t.offset * t.n, -1);
  t.proxy_xieta.assign((size_t)t.offset * t.n * 2, 0.0);
  for (int i = 0; i < t.offset; ++i) {
    double xo, xs, eo, es, v[3];
    frame(t.pn[i], xo, xs, eo, es);
    for (int j = 0; j < t.n; ++j) {
      double xi = t.node_d[j / t.dp1] * xs + xo, eta = t.node_d[j % t.dp1] * es + eo;
      xyz_from_xieta(xi, eta, t.pn[i].face, v);
      size_t o = (size_t)i * t.n + j;
      t.proxy_leaf[o] = locate_leaf(t, v[0], v[1], v[2]);
      t.proxy_xieta[2 * o] = xi;
      t.proxy_xieta[2 * o + 1] = eta;
    }
  }
  t.proxy_order.resize(t.proxy_leaf.size());
  for (size_t o = 0; o < t.proxy_order.size(); ++o) t.proxy_order[o] = (int)o;
  std::stable_sort(t.proxy_order.begin(), t.proxy_order.end(),
                   [&](int a, int b) { return t.proxy_leaf[a] < t.proxy_leaf[b]; });
  return "";
}

void build_source_major(const Tree& t, SourceMajor& sm) {
  const int64_t I = t.interaction_count();
  std::vector<int> cnt(t.P + 1, 0);
  for (int64_t e = 0; e < I; ++e) cnt[t.ilist[3 * e + 1] + 1]++;
  for (int p = 0; p < t.P; ++p) cnt[p + 1] += cnt[p];
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  sm.targets.assign(I, SourceSet{});
  for (int64_t e = 0; e < I; ++e) {
    const int tg = t.ilist[3 * e], s = t.ilist[3 * e + 1];
    SourceSet ts{};
    ts.panel = tg;
    // target-panel centre minus source-panel centre (the FP32 adjoint works in panel-centred coordinates)
    ts.ox = (float)(t.dev[tg].cx - t.dev[s].cx);
    ts.oy = (float)(t.dev[tg].cy - t.dev[s].cy);
    ts.oz = (float)(t.dev[tg].cz - t.dev[s].cz);
    if (t.pn[tg].is_leaf) {
      ts.first = (int)((int64_t)(tg - t.offset) * t.m2);
      ts.count = t.m2;
    } else {
      ts.first = (int)(t.npart + (int64_t)tg * t.n);
      ts.count = t.n;
    }
    sm.targets[fill[s]++] = ts;
  }
  sm.tiles.clear();
  for (int p = 0; p < t.P; ++p) {
    if (cnt[p + 1] == cnt[p]) continue;
    const int npts = t.pn[p].is_leaf ? t.m2 : t.n;
    const int64_t base = t.pn[p].is_leaf ? (int64_t)(p - t.offset) * t.m2 : t.npart + (int64_t)p * t.n;
    const int chunks = (npts + 63) / 64, per = (npts + chunks - 1) / chunks;
    for (int f = 0; f < npts; f += per) {
      TargetTile tl{};
      tl.panel = p; tl.first = (int)(base + f); tl.count = std::min(per, npts - f);
      tl.src_begin = cnt[p]; tl.src_end = cnt[p + 1];
      sm.tiles.push_back(tl);
    }
  }
  sm.zero_rows.clear();
  for (int p = 0; p < t.P; ++p)
    if ((!t.pn[p].is_leaf && cnt[p + 1] == cnt[p]) || (t.pn[p].is_leaf && t.cfg.levels == 1)) sm.zero_rows.push_back(p);
}

}  // namespace sbslpm

namespace sbslpm {

void dedupe_points(const Tree& t, std::vector<int>& two_d_to_1d, std::vector<int>& counts, std::vector<double>& x1d,
                   std::vector<double>& y1d, std::vector<double>& z1d) {
  constexpr double kTol = 1e-15, kCell = 1e-12;
  const int64_t N = t.npart;
  two_d_to_1d.assign(N, -1);
  counts.clear();
  x1d.clear();
  y1d.clear();
  z1d.clear();
  auto cell = [&](double v) { return (int64_t)std::floor(v / kCell); };
  auto key = [](int64_t a, int64_t b, int64_t c) {
    return (uint64_t)(a * 73856093) ^ (uint64_t)(b * 19349663) ^ (uint64_t)(c * 83492791);
  };
  std::unordered_multimap<uint64_t, int> grid;
  grid.reserve((size_t)N);
  for (int64_t u = 0; u < N; ++u) {
    const double x = t.spec_x[u], y = t.spec_y[u], z = t.spec_z[u];
    const int64_t cx = cell(x), cy = cell(y), cz = cell(z);
    int best = -1;
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = -1; dz <= 1; ++dz) {
          auto range = grid.equal_range(key(cx + dx, cy + dy, cz + dz));
          for (auto it = range.first; it != range.second; ++it) {
            const int k = it->second;
            if (std::fabs(x - x1d[k]) < kTol && std::fabs(y - y1d[k]) < kTol && std::fabs(z - z1d[k]) < kTol &&
                (best < 0 || k < best))
              best = k;
          }
        }
    if (best < 0) {
      best = (int)x1d.size();
      x1d.push_back(x);
      y1d.push_back(y);
      z1d.push_back(z);
      counts.push_back(0);
      grid.emplace(key(cx, cy, cz), best);
    }
    counts[best] += 1;
    two_d_to_1d[u] = best;
  }
}

}  // namespace sbslpm
