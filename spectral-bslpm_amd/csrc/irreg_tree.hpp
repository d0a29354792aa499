// irreg_tree.hpp -- the irregular (adaptive) cubed-sphere particle tree of the
// reference's lat-lon path: face panels split into quadrants while they hold
// more than leaf_size points (initialize_cube_sphere_irreg_points,
// initialize_cubed_sphere.cpp:446-592), its dual-tree traversal
// (dual_tree_traversal_irreg, fmm_funcs.cpp:125-231) and the work tiles of
// the pair kernels over a leaf-sorted copy of the points.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "tree.hpp"

namespace sbslpm {

// Panel of the irregular tree; bounds in radians as the reference stores them.
struct IrregPanel {
  int face, level, parent, child[4], is_leaf, point_count;
  double min_xi, max_xi, min_eta, max_eta, radius;
};

struct IrregTree {
  int leaf_size = 0, degree = 0, dp1 = 0, n = 0;
  int P = 0, levels = 0;  // levels = the deepest panel level (the reference's run_config.levels)
  double theta = 0;
  int64_t N = 0, U = 0;  // points; points + P*n proxies
  std::vector<IrregPanel> pn;
  std::vector<int> level_start;             // [levels+2]: first panel of each level, then P
  std::vector<std::vector<int>> points;     // per panel: point ids in the reference's insertion order
  std::vector<int> ilist;                   // [I][3] target, source, type (reference BFS order)
  std::vector<double> cheb;                 // cos(pi i/degree), unsnapped
  // leaf-sorted points: sorted index k holds input point perm[k]; leaf panel p
  // owns sorted points [leaf_first[p], leaf_first[p] + point_count)
  std::vector<int> perm, leaf_first, point_leaf;
  std::vector<double> px, py, pz;           // unified [U]: sorted points, then proxies [P][n]
  std::vector<DevPanel> dev;                // [P] bounds in radians
  std::vector<TargetTile> tiles;            // target-major work tiles
  std::vector<SourceSet> sets;
  std::vector<int> zero_rows;               // proxy rows no interaction writes
  std::vector<int> nonleaf;                 // non-leaf panels grouped by level (level_nonleaf gives ranges)
  std::vector<int> level_nonleaf;           // [levels+2]
  std::vector<double> m2m_hi, m2m_lo;       // fixed child-to-parent matrices (exact quadrants)

  int64_t interaction_count() const { return (int64_t)ilist.size() / 3; }
};

// Builds the tree over n points (unit vectors); returns "" or an error message.
std::string build_irreg_tree(const double* x, const double* y, const double* z, int64_t n, int leaf_size, int degree,
                             double theta, IrregTree& t);

}  // namespace sbslpm
