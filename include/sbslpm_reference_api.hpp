// sbslpm_reference_api.hpp -- the reference's C++ call signatures restated over
// the C ABI of sbslpm.h, for a drop-in replacement of the Kokkos calls in
// bve_time_step.cpp / poisson_eq_fmm.cpp / poisson_eq_ds.cpp.
//
// `View2<T>` stands in for Kokkos::View<T**, LayoutRight> (device memory):
// pointer + two extents.  Function names, argument order and the
// accumulate/assign conventions are the reference's; the interpolation table,
// panels and interaction list live in the context instead of being passed.
// The uniform-tree wrappers cover the reference layout (particles are the
// leaf proxy nodes, leaf_degree == interp_degree), where the particle arrays
// are the leaf rows of the [P][n] proxy arrays, exactly as in the reference;
// the _ll wrappers take the irregular tree's handle.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>

#include "sbslpm.h"

namespace sbslpm_ref {

template <class T>
struct View2 {
  T* ptr = nullptr;
  int64_t e0 = 0, e1 = 0;
  T* data() const { return ptr; }
  int extent_int(int i) const { return (int)(i == 0 ? e0 : e1); }
  T* row(int64_t r) const { return ptr + r * e1; }
};

// Raises like the reference's std::runtime_error paths (Kokkos::abort sites
// surface from sbslpm_synchronize as SBSLPM_E_GEOMETRY).
inline void check(sbslpm_ctx* ctx, int rc, const char* what) {
  if (rc != SBSLPM_OK) throw std::runtime_error(std::string(what) + ": " + sbslpm_last_error(ctx));
}

struct Layout {
  int64_t P, Pl, n, offset;
};

inline Layout layout(sbslpm_ctx* ctx) {
  int64_t s[8];
  check(ctx, sbslpm_get_sizes(ctx, s), "get_sizes");
  if (s[2] != s[3]) throw std::runtime_error("reference-layout wrappers need leaf_degree == interp_degree");
  return Layout{s[0], s[1], s[2], s[6]};
}

// fmm_funcs.hpp:25-27
inline void upward_pass(sbslpm_ctx* ctx, View2<double>& area, View2<double>& pots,
                        View2<double>& proxy_source_pots) {
  Layout L = layout(ctx);
  check(ctx, sbslpm_upward_pass(ctx, area.data(), pots.data(), proxy_source_pots.row(L.offset),
                                proxy_source_pots.data()),
        "upward_pass");
}

// fmm_interactions/bve_vel.hpp:10-14
inline void bve_vel_interactions(sbslpm_ctx* ctx, View2<double>& disp_x, View2<double>& disp_y,
                                 View2<double>& disp_z, View2<double>& target_pots_1, View2<double>& target_pots_2,
                                 View2<double>& target_pots_3, View2<double>& source_vals) {
  Layout L = layout(ctx);
  check(ctx,
        sbslpm_bve_vel_interactions(ctx, disp_x.row(L.offset), disp_y.row(L.offset), disp_z.row(L.offset),
                                    disp_x.data(), disp_y.data(), disp_z.data(), source_vals.row(L.offset),
                                    source_vals.data(), target_pots_1.row(L.offset), target_pots_2.row(L.offset),
                                    target_pots_3.row(L.offset), target_pots_1.data(), target_pots_2.data(),
                                    target_pots_3.data()),
        "bve_vel_interactions");
}

// fmm_interactions/inverse_poisson.hpp:10-12
inline void poisson_fmm_interactions(sbslpm_ctx* ctx, View2<double>& target_pots, View2<double>& source_vals) {
  Layout L = layout(ctx);
  check(ctx,
        sbslpm_poisson_fmm_interactions(ctx, source_vals.row(L.offset), source_vals.data(),
                                        target_pots.row(L.offset), target_pots.data()),
        "poisson_fmm_interactions");
}

// fmm_interactions/swe_vel.hpp:11-23 (swe_vel_interactions, swe_vel_interactions_2)
inline void swe_vel_interactions(sbslpm_ctx* ctx, View2<double>& disp_x, View2<double>& disp_y, View2<double>& disp_z,
                                 View2<double>& target_pots_1, View2<double>& target_pots_2,
                                 View2<double>& target_pots_3, View2<double>& source_vals_vor,
                                 View2<double>& source_vals_div, int variant = 1) {
  Layout L = layout(ctx);
  auto fn = variant == 1 ? sbslpm_swe_vel_interactions : sbslpm_swe_vel_interactions_2;
  check(ctx,
        fn(ctx, disp_x.row(L.offset), disp_y.row(L.offset), disp_z.row(L.offset), disp_x.data(), disp_y.data(),
           disp_z.data(), source_vals_vor.row(L.offset), source_vals_vor.data(), source_vals_div.row(L.offset),
           source_vals_div.data(), target_pots_1.row(L.offset), target_pots_2.row(L.offset),
           target_pots_3.row(L.offset), target_pots_1.data(), target_pots_2.data(), target_pots_3.data()),
        "swe_vel_interactions");
}

// fmm_funcs.hpp:29-31 (the leaf rows already hold the leaf-target sums)
inline void downward_pass(sbslpm_ctx* ctx, View2<double>& proxy_target_pots, View2<double>& sols) {
  check(ctx, sbslpm_downward_pass(ctx, proxy_target_pots.data(), nullptr, sols.data()), "downward_pass");
}

// fmm_funcs.hpp:33-36
inline void downward_pass_3(sbslpm_ctx* ctx, View2<double>& p1, View2<double>& p2, View2<double>& p3,
                            View2<double>& s1, View2<double>& s2, View2<double>& s3) {
  check(ctx,
        sbslpm_downward_pass_3(ctx, p1.data(), p2.data(), p3.data(), nullptr, nullptr, nullptr, s1.data(), s2.data(),
                               s3.data()),
        "downward_pass_3");
}

// direct_sum.hpp:8-11, rows [lb, ub) (RunConfig::oned_lb / oned_ub)
inline void direct_sum_inv_lap(sbslpm_ctx* ctx, int64_t lb, int64_t ub, int64_t n, double* xcos, double* ycos,
                               double* zcos, double* area, double* pots, double* soln) {
  check(ctx, sbslpm_direct_sum_inv_lap(ctx, n, xcos, ycos, zcos, area, pots, soln, lb, ub), "direct_sum_inv_lap");
}

// bve_time_step.hpp:7-13 (passive_tracers: [Pl][n][tracer_count] device array)
inline void bve_back_rk4_step(sbslpm_ctx* ctx, View2<double>& xcos, View2<double>& ycos, View2<double>& zcos,
                              View2<double>& area, View2<double>& vors, View2<double>& proxy_source_vors,
                              View2<double>& vel_x, View2<double>& vel_y, View2<double>& vel_z,
                              double* passive_tracers, double time) {
  check(ctx,
        sbslpm_bve_back_rk4_step(ctx, xcos.data(), ycos.data(), zcos.data(), area.data(), vors.data(),
                                 proxy_source_vors.data(), vel_x.data(), vel_y.data(), vel_z.data(), passive_tracers,
                                 time),
        "bve_back_rk4_step");
}

// swe_back_rk4_step (swe_time_step.cpp:606-811).  interp_vals, the panels and
// the interaction list are the context's; topo_type and the indexing rule are
// the context's sbslpm_params.topography / swe_rule.
inline void swe_back_rk4_step(sbslpm_ctx* ctx, View2<double>& xcos, View2<double>& ycos, View2<double>& zcos,
                              View2<double>& area, View2<double>& vors, View2<double>& divs, View2<double>& height,
                              View2<double>& vel_x, View2<double>& vel_y, View2<double>& vel_z,
                              double* passive_tracers, View2<double>& height_lapl, double time) {
  check(ctx,
        sbslpm_swe_back_rk4_step(ctx, xcos.data(), ycos.data(), zcos.data(), area.data(), vors.data(), divs.data(),
                                 height.data(), vel_x.data(), vel_y.data(), vel_z.data(), passive_tracers,
                                 height_lapl.data(), time),
        "swe_back_rk4_step");
}

// laplacian (deriv_funcs.cpp:27-37): the spectral surface Laplacian on each leaf.
inline void laplacian(sbslpm_ctx* ctx, View2<double>& laplacian_vals, View2<double>& func_vals) {
  check(ctx, sbslpm_laplacian(ctx, func_vals.data(), laplacian_vals.data()), "laplacian");
}

// ---- the lat-lon path (irregular particle tree).  The tree, its panels and
// leaf_panel_points live in the sbslpm_irreg handle (sbslpm_irreg_create over
// the same points); the point arrays are the reference's [lat][lon] arrays.

// fmm_funcs.hpp:37-39 (proxy_source_pots is assigned; the reference accumulates
// into caller-zeroed rows)
inline void upward_pass_ll(sbslpm_ctx* ctx, sbslpm_irreg* tree, View2<double>& area, View2<double>& pots,
                           View2<double>& proxy_source_pots) {
  check(ctx, sbslpm_irreg_upward_pass(tree, area.data(), pots.data(), proxy_source_pots.data()), "upward_pass_ll");
}

// fmm_funcs.hpp:41-45
inline void downward_pass_3_ll(sbslpm_ctx* ctx, sbslpm_irreg* tree, View2<double>& proxy_target_pots_1,
                               View2<double>& proxy_target_pots_2, View2<double>& proxy_target_pots_3,
                               View2<double>& sols_1, View2<double>& sols_2, View2<double>& sols_3) {
  check(ctx,
        sbslpm_irreg_downward_pass_3(tree, proxy_target_pots_1.data(), proxy_target_pots_2.data(),
                                     proxy_target_pots_3.data(), sols_1.data(), sols_2.data(), sols_3.data()),
        "downward_pass_3_ll");
}

// fmm_interactions/swe_vel.hpp:25-29
inline void swe_vel_interactions_ll(sbslpm_ctx* ctx, sbslpm_irreg* tree, View2<double>& target_pots_1,
                                    View2<double>& target_pots_2, View2<double>& target_pots_3, View2<double>& vel_x,
                                    View2<double>& vel_y, View2<double>& vel_z, View2<double>& source_vals_vor,
                                    View2<double>& source_vals_div, View2<double>& vors, View2<double>& divs,
                                    View2<double>& area) {
  check(ctx,
        sbslpm_irreg_swe_vel_interactions(tree, target_pots_1.data(), target_pots_2.data(), target_pots_3.data(),
                                          vel_x.data(), vel_y.data(), vel_z.data(), source_vals_vor.data(),
                                          source_vals_div.data(), vors.data(), divs.data(), area.data()),
        "swe_vel_interactions_ll");
}

}  // namespace sbslpm_ref
