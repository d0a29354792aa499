// tree_api.cpp -- host-only C ABI over the tree builder (no device needed).
#include <cstring>
#include <new>
#include <vector>

#include "../../include/sbslpm.h"
#include "geometry.hpp"
#include "irreg_tree.hpp"
#include "tree.hpp"

struct sbslpm_tree {
  sbslpm::Tree t;
};

struct sbslpm_irreg_tree {
  sbslpm::IrregTree t;
};

extern "C" {

int sbslpm_tree_build(const sbslpm_params* p, sbslpm_tree** out) {
  if (!p || !out) return SBSLPM_E_ARG;
  *out = nullptr;
  sbslpm_tree* h = new (std::nothrow) sbslpm_tree();
  if (!h) return SBSLPM_E_NOMEM;
  sbslpm::TreeConfig cfg{p->levels, p->interp_degree, p->leaf_degree, p->theta, p->weight_rule};
  if (!sbslpm::build_tree(cfg, h->t).empty()) {
    delete h;
    return SBSLPM_E_ARG;
  }
  *out = h;
  return SBSLPM_OK;
}

void sbslpm_tree_destroy(sbslpm_tree* h) { delete h; }

int sbslpm_tree_sizes(const sbslpm_tree* h, int64_t* s) {
  if (!h || !s) return SBSLPM_E_ARG;
  const sbslpm::Tree& t = h->t;
  s[0] = t.P; s[1] = t.Pl; s[2] = t.n; s[3] = t.m2; s[4] = t.interaction_count(); s[5] = t.npart;
  s[6] = t.offset; s[7] = (int64_t)t.tiles.size();
  return SBSLPM_OK;
}

int sbslpm_tree_spec_points(const sbslpm_tree* h, double* x, double* y, double* z, double* area) {
  if (!h) return SBSLPM_E_ARG;
  const sbslpm::Tree& t = h->t;
  size_t bytes = (size_t)t.npart * sizeof(double);
  if (x) std::memcpy(x, t.spec_x.data(), bytes);
  if (y) std::memcpy(y, t.spec_y.data(), bytes);
  if (z) std::memcpy(z, t.spec_z.data(), bytes);
  if (area) std::memcpy(area, t.spec_area.data(), bytes);
  return SBSLPM_OK;
}

int sbslpm_tree_interaction_list(const sbslpm_tree* h, int* list) {
  if (!h || !list) return SBSLPM_E_ARG;
  std::memcpy(list, h->t.ilist.data(), h->t.ilist.size() * sizeof(int));
  return SBSLPM_OK;
}

int sbslpm_tree_panels(const sbslpm_tree* h, int* ints, double* dbls) {
  if (!h) return SBSLPM_E_ARG;
  const sbslpm::Tree& t = h->t;
  for (int i = 0; i < t.P; ++i) {
    const sbslpm::HostPanel& p = t.pn[i];
    if (ints) {
      int* r = ints + 8 * i;
      r[0] = p.face; r[1] = p.level; r[2] = p.parent;
      r[3] = p.child[0]; r[4] = p.child[1]; r[5] = p.child[2]; r[6] = p.child[3]; r[7] = p.is_leaf;
    }
    if (dbls) {
      double* q = dbls + 6 * i;
      q[0] = p.min_xi; q[1] = p.max_xi; q[2] = p.min_eta; q[3] = p.max_eta; q[4] = p.radius; q[5] = p.area;
    }
  }
  return SBSLPM_OK;
}

int sbslpm_tree_locate(const sbslpm_tree* h, double x, double y, double z) {
  if (!h) return -1;
  return sbslpm::locate_leaf(h->t, x, y, z);
}

int sbslpm_tree_transfer_matrices(const sbslpm_tree* h, double* hi, double* lo, double* p2m) {
  if (!h) return SBSLPM_E_ARG;
  const sbslpm::Tree& t = h->t;
  if (hi) std::memcpy(hi, t.m2m_hi.data(), t.m2m_hi.size() * sizeof(double));
  if (lo) std::memcpy(lo, t.m2m_lo.data(), t.m2m_lo.size() * sizeof(double));
  if (p2m) std::memcpy(p2m, t.p2m.data(), t.p2m.size() * sizeof(double));
  return SBSLPM_OK;
}

int sbslpm_tree_dedupe(const sbslpm_tree* h, int* two_d_to_1d, int* counts, double* x1d, double* y1d, double* z1d,
                       int64_t* n1d) {
  if (!h || !two_d_to_1d || !n1d) return SBSLPM_E_ARG;
  std::vector<int> map, cnt;
  std::vector<double> x, y, z;
  sbslpm::dedupe_points(h->t, map, cnt, x, y, z);
  std::memcpy(two_d_to_1d, map.data(), map.size() * sizeof(int));
  if (counts) std::memcpy(counts, cnt.data(), cnt.size() * sizeof(int));
  if (x1d) std::memcpy(x1d, x.data(), x.size() * sizeof(double));
  if (y1d) std::memcpy(y1d, y.data(), y.size() * sizeof(double));
  if (z1d) std::memcpy(z1d, z.data(), z.size() * sizeof(double));
  *n1d = (int64_t)x.size();
  return SBSLPM_OK;
}

int sbslpm_irreg_tree_build(int64_t n, const double* x, const double* y, const double* z, int leaf_size,
                            int interp_degree, double theta, sbslpm_irreg_tree** out) {
  if (!out || !x || !y || !z) return SBSLPM_E_ARG;
  *out = nullptr;
  sbslpm_irreg_tree* h = new (std::nothrow) sbslpm_irreg_tree();
  if (!h) return SBSLPM_E_NOMEM;
  if (!sbslpm::build_irreg_tree(x, y, z, n, leaf_size, interp_degree, theta, h->t).empty()) {
    delete h;
    return SBSLPM_E_ARG;
  }
  *out = h;
  return SBSLPM_OK;
}

void sbslpm_irreg_tree_destroy(sbslpm_irreg_tree* h) { delete h; }

int sbslpm_irreg_tree_sizes(const sbslpm_irreg_tree* h, int64_t* s) {
  if (!h || !s) return SBSLPM_E_ARG;
  const sbslpm::IrregTree& t = h->t;
  s[0] = t.P; s[1] = t.levels; s[2] = t.n; s[3] = t.N; s[4] = t.interaction_count(); s[5] = (int64_t)t.tiles.size();
  return SBSLPM_OK;
}

int sbslpm_irreg_tree_panels(const sbslpm_irreg_tree* h, int* ints, double* dbls) {
  if (!h) return SBSLPM_E_ARG;
  const sbslpm::IrregTree& t = h->t;
  for (int i = 0; i < t.P; ++i) {
    const sbslpm::IrregPanel& p = t.pn[i];
    if (ints) {
      int* r = ints + 10 * i;
      r[0] = p.face; r[1] = p.level; r[2] = p.parent; r[3] = p.child[0]; r[4] = p.child[1]; r[5] = p.child[2];
      r[6] = p.child[3]; r[7] = p.is_leaf; r[8] = p.point_count; r[9] = t.leaf_first[i];
    }
    if (dbls) {
      double* q = dbls + 5 * i;
      q[0] = p.min_xi; q[1] = p.max_xi; q[2] = p.min_eta; q[3] = p.max_eta; q[4] = p.radius;
    }
  }
  return SBSLPM_OK;
}

int sbslpm_irreg_tree_points(const sbslpm_irreg_tree* h, int* perm) {
  if (!h || !perm) return SBSLPM_E_ARG;
  std::memcpy(perm, h->t.perm.data(), h->t.perm.size() * sizeof(int));
  return SBSLPM_OK;
}

int sbslpm_irreg_tree_interaction_list(const sbslpm_irreg_tree* h, int* list) {
  if (!h || !list) return SBSLPM_E_ARG;
  std::memcpy(list, h->t.ilist.data(), h->t.ilist.size() * sizeof(int));
  return SBSLPM_OK;
}

// cubed_sphere_midpoints (initialize_cubed_sphere.cpp:10-62): 6 * 4^levels cell centres of
// the equiangular grid with 2^levels cells per face edge, ordered [face][xi][eta], areas
// from two spherical triangles over the cell corners.  The reference passes face i+1
// to xyz_from_xieta (:35), which aborts at face 6; here the faces are 0..5.
int sbslpm_cubed_sphere_midpoints(int levels, double* x, double* y, double* z, double* area) {
  if (levels < 0 || levels > 14 || !x || !y || !z || !area) return SBSLPM_E_ARG;
  const int side = 1 << levels;
  const double between = sbslpm::kPi / (2.0 * side), off = between / 2.0;
  for (int f = 0; f < 6; ++f)
    for (int j = 0; j < side; ++j)
      for (int k = 0; k < side; ++k) {
        const double xi = -sbslpm::kPi / 4.0 + off + between * j, eta = -sbslpm::kPi / 4.0 + off + between * k;
        double c[3], p1[3], p2[3], p3[3], p4[3];
        sbslpm::xyz_from_xieta(xi, eta, f, c);
        sbslpm::xyz_from_xieta(xi - off, eta - off, f, p1);
        sbslpm::xyz_from_xieta(xi + off, eta - off, f, p2);
        sbslpm::xyz_from_xieta(xi + off, eta + off, f, p3);
        sbslpm::xyz_from_xieta(xi - off, eta + off, f, p4);
        const int64_t o = ((int64_t)f * side + j) * side + k;
        x[o] = c[0]; y[o] = c[1]; z[o] = c[2];
        area[o] = sbslpm::sphere_tri_area(p1, p2, p3) + sbslpm::sphere_tri_area(p1, p3, p4);
      }
  return SBSLPM_OK;
}

}  // extern "C"
