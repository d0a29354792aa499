// tree.hpp -- host-side cubed-sphere quadtree, dual-tree interaction list and
// the fixed transfer operators the device passes use.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace sbslpm {

// Panel fields the hot path reads (initialize_cubed_sphere.hpp:8-29; the edge
// neighbour ids serve only the SWE derivative code and are not built).
struct HostPanel {
  int face, level, parent, child[4], is_leaf;
  double min_xi, max_xi, min_eta, max_eta;  // in [-1, 1]; radians = value * pi/4
  double radius, area;
};

// Device copy of a panel: bounds in radians, as the reference kernels convert
// them (`Kokkos::numbers::pi/4.0 * min_xi`, interp_funcs_impl.hpp:127,
// bve_time_step.cpp:45), plus the centre used by the mixed-precision kernels.
struct DevPanel {
  double min_xi, max_xi, min_eta, max_eta;
  double mid_xi, mid_eta;  // 0.5*pi/4*(max+min) as point location computes it (interp_funcs_impl.hpp:127-128)
  double cx, cy, cz, pad0;  // unit-sphere panel centre
  int face, is_leaf, child0, level;
};

struct TreeConfig {
  int levels;       // L; leaves at level L-1
  int degree;       // proxy degree d (n = (d+1)^2)
  int leaf_degree;  // particle-grid degree m-1 per leaf (== d: the reference layout)
  double theta;     // multipole acceptance parameter
  int weight_rule;  // 0 reference quadrature formula, 1 Clenshaw-Curtis, 2 equal area
};

// One unit of interaction work: up to 64 target points of one target panel
// and its source-set range.  Target points live in the unified point index
// space [0, Pl*m2) particles, [Pl*m2, Pl*m2 + P*n) proxies.
struct TargetTile {
  int panel;      // target panel id
  int first;      // first unified target point index of the tile
  int count;      // target points in the tile (1..64)
  int src_begin;  // range in the source-set list
  int src_end;
  int pad[3];
};

// A source point set: `count` consecutive records starting at `first` in the
// unified point index space, belonging to panel `panel`; `first16` is its
// first record in the 16-padded record space of the MFMA kernels.
// One interaction seen from its target: the source points [first, first+count)
// of the unified index space (first16 in the 16-padded MFMA record space) and
// the target-panel centre minus the source-panel centre, rounded to FP32 (the
// mixed-precision kernels work in panel-centred coordinates).
struct SourceSet {
  int first, count, panel, first16;
  float ox, oy, oz;
  int m2l;  // 1 when both the source and the target panel are non-leaf (proxy to proxy, type 3)
};

struct Tree {
  TreeConfig cfg{};
  int P = 0, Pl = 0, offset = 0, n = 0, dp1 = 0, m = 0, m2 = 0;
  int64_t npart = 0, U = 0;        // Pl*m2, Pl*m2 + P*n
  int m2p = 0, np16 = 0;           // m2 and n rounded up to multiples of 16
  int64_t U16 = 0;                 // 16-padded record space: Pl*m2p + offset*np16
  std::vector<int> level_start;    // [L+1]
  std::vector<HostPanel> pn;       // [P]
  std::vector<double> cheb_d, cheb_m;  // cos(pi i/deg) as interp_vals_bli evaluates it (:45)
  std::vector<double> node_d, node_m;  // interp_init nodes cos(pi/deg*i), |.|<1e-16 -> 0 (:33-39)
  std::vector<double> quad_m;          // 1-D quadrature weights of the particle grid
  std::vector<int> ilist;              // [I][3] target, source, type in the reference's BFS order

  // Device-facing structures.
  std::vector<DevPanel> dev;            // [P]
  std::vector<TargetTile> tiles;        // interaction work units
  std::vector<SourceSet> srcsets;       // per-tile source lists (CSR via tiles)
  std::vector<int> set_start;           // [P+1]: source sets of target panel p are [set_start[p], set_start[p+1])
  int nonleaf_tiles = 0;                // tiles of non-leaf target panels (they come first)
  std::vector<int> xcd_start;           // [2][9]: tiles [xcd_start[x], xcd_start[x+1]) of XCD x in the
                                        // cost-balanced contiguous launch mapping (PairShape xcd 2),
                                        // for matrix-core kernels, then for vector-unit kernels
  std::vector<double> px, py, pz;       // unified point coordinates [U] (targets, undisplaced sources)
  std::vector<double> m2m_hi, m2m_lo;   // [d+1][d+1]: parent basis a at child node k, A[a*(d+1)+k]
  std::vector<double> p2m;              // [d+1][m]: proxy basis a at particle node k
  std::vector<double> l2p_hi, l2p_lo;   // [m][d+1]: parent basis a at particle node k of an upper/lower half
                                        // child, l2p[k * (d+1) + a]
  std::vector<double> l2pT_hi, l2pT_lo; // their transposes [d+1][m]
  std::vector<double> anc_l2l;          // [2^(L-1)-2][d+1][d+1]: ancestor basis a at a descendant's node k,
                                        // [k][a] in matrix 2^g - 2 + pos for gap g and interval position pos
  std::vector<int> zero_rows;           // proxy rows no pair kernel or last-level L2L writes
  std::vector<int> proxy_leaf;          // [offset*n] leaf containing each non-leaf proxy node
  std::vector<double> proxy_xieta;      // [offset*n][2] label coordinates of those nodes (radians)
  std::vector<int> proxy_order;         // [offset*n] non-leaf proxy nodes sorted by their leaf (gather locality)
  std::vector<double> spec_x, spec_y, spec_z, spec_area;  // caller-facing particles [Pl][m2]

  int64_t interaction_count() const { return (int64_t)ilist.size() / 3; }
};

// Builds everything above on the host; returns an empty string on success or
// a message describing the invalid parameter.
std::string build_tree(const TreeConfig& cfg, Tree& t);
// Composite L2L matrices for gaps 1 .. maxgap (layout of Tree::anc_l2l); node:
// the proxy nodes in [-1, 1], cheb: the interpolation nodes.
void composite_l2l(int degree, const double* node, const double* cheb, int maxgap, std::vector<double>& out);

// The interaction list seen from its sources, for the transposed (adjoint)
// pair kernel: tiles of at most 64 source points of one source panel
// (particles of a leaf, proxies of a non-leaf panel) whose [src_begin,
// src_end) range indexes `targets`, each entry naming the target points
// [first, first+count) of one interaction (particles of a leaf target,
// proxies of a non-leaf target) and, in ox/oy/oz, the target-panel centre
// minus the source-panel centre in FP32.  `zero_rows` lists the proxy rows that are
// no interaction source (and the leaf rows when levels == 1), whose
// cotangents are zero.
struct SourceMajor {
  std::vector<TargetTile> tiles;
  std::vector<SourceSet> targets;
  std::vector<int> zero_rows;
};
void build_source_major(const Tree& t, SourceMajor& sm);

// Unique particle positions (cubed_sphere_2d_to_1d, initialize_cubed_sphere.cpp:290-324):
// particles taken in [leaf][node] order; one within 1e-15 in every coordinate of an
// earlier unique point maps to the first such point, otherwise it becomes a new
// unique point.  counts[k] = particles mapped to unique point k.  Spatial hashing
// replaces the reference's scan over all earlier points (same result).
void dedupe_points(const Tree& t, std::vector<int>& two_d_to_1d, std::vector<int>& counts, std::vector<double>& x1d,
                   std::vector<double>& y1d, std::vector<double>& z1d);

// Area of the spherical triangle p1 p2 p3 (sphere_tri_area, general_utils.cpp:5-21).
double sphere_tri_area(const double* p1, const double* p2, const double* p3);

// Host twin of the device point location (interp_funcs_impl.hpp:114-142).
int locate_leaf(const Tree& t, double x, double y, double z);

}  // namespace sbslpm
