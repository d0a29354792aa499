/* sbslpm.h -- C ABI of the MI355X (gfx950) spectral-BSLPM hot path:
 * cubed-sphere barycentric-Lagrange-interpolation FMM for the sphere Poisson
 * and Biot-Savart problems, the direct sums, and the backward semi-Lagrangian
 * BVE step.  Library: spectral-bslpm_amd/libsbslpm.so.
 *
 * Conventions (shared with the reference operator API: see
 * fmm_funcs.hpp:16-54, bve_vel.hpp:10-14, inverse_poisson.hpp:10-12,
 * direct_sum.hpp:8-11, bve_time_step.hpp:7-13):
 *   - All array arguments are DEVICE pointers owned by the caller, FP64,
 *     row-major ("LayoutRight"): particle arrays [Pl][m*m] (leaf panel, node
 *     xi_index*m + eta_index), proxy arrays [P][n], n = (d+1)^2, panels in the
 *     reference's order (level by level, children (hi,hi),(lo,hi),(lo,lo),(hi,lo)).
 *   - "accumulate" entry points add into caller-zeroed outputs like the
 *     reference functors; "assign" ones overwrite.
 *   - Work is enqueued on the context's stream; functions return after
 *     launching.  sbslpm_synchronize() waits and reports device-side errors
 *     (the reference's Kokkos::abort sites) as SBSLPM_E_GEOMETRY.
 *   - A context is bound to one device and one stream and is not thread-safe.
 *   - Layout: each leaf carries an m x m Chebyshev (2nd kind) particle grid,
 *     m = leaf_degree + 1; proxies have degree d.  leaf_degree == d is exactly
 *     the reference layout (particles are the leaf proxy nodes).
 */
#ifndef SBSLPM_H
#define SBSLPM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sbslpm_ctx sbslpm_ctx;

enum {
  SBSLPM_OK = 0,
  SBSLPM_E_ARG = 1,      /* invalid argument (reference: std::runtime_error) */
  SBSLPM_E_GEOMETRY = 2, /* point outside its panel / location failure (reference: Kokkos::abort) */
  SBSLPM_E_HIP = 3,      /* HIP runtime error */
  SBSLPM_E_NOMEM = 4
};

enum { SBSLPM_WEIGHTS_REFERENCE = 0, SBSLPM_WEIGHTS_CLENSHAW_CURTIS = 1, SBSLPM_WEIGHTS_EQUAL_AREA = 2 };
enum { SBSLPM_RK4_REFERENCE = 0, SBSLPM_RK4_TEXTBOOK = 1 };
/* Shallow-water step indexing.  REFERENCE reproduces swe_back_rk4_step as written: the velocity
 * gradient converts vectors with cubed_sphere_panels(i).face for leaf-local i
 * (swe_time_step.cpp:109-111) and applies the metric factors twice (deriv_funcs_impl.hpp:291-292 and
 * cubed_sphere_transforms_impl.hpp:161-162), its face-4/5 pole nodes return (g_xi, g_eta, 0); the
 * topography reads the displacement of panel i (topography.cpp:33-35); the stage-4 state is advanced
 * by dt/2 (swe_time_step.cpp:761).  LEAF uses the leaf's own face and displacement, the surface
 * gradient, and dt at stage 4. */
enum { SBSLPM_SWE_RULE_REFERENCE = 0, SBSLPM_SWE_RULE_LEAF = 1 };
/* SBSLPM_MIXED: FP32 pair arithmetic on the matrix cores (Biot-Savart) / vector unit (log), FP64
 * accumulation; SBSLPM_MIXED_VALU: the same precision with the Biot-Savart pairs on the vector unit. */
enum { SBSLPM_FP64 = 0, SBSLPM_MIXED = 1, SBSLPM_MIXED_VALU = 2 };
/* SBSLPM_FLAG_GRAPH: capture steps in a hipGraph (needs a non-default stream).
 * SBSLPM_FLAG_MULTI_RHS: on a SBSLPM_MIXED context, the first RK4 stage of an ensemble step
 * (sources and targets at the same positions for every member) computes each kernel matrix once
 * and applies it to all members' charges as one multi-right-hand-side product on the matrix
 * cores.  Members then agree with the single-member step to FP32 rounding (another summation
 * order) instead of bit for bit.
 * SBSLPM_FLAG_BF16_M2L: on a SBSLPM_MIXED context, the kernel product 1 - x.y of proxy-to-proxy
 * (type 3, M2L) interactions takes bf16 operands (FP32 accumulation) in the matrix-core pair
 * kernels; every other interaction stays FP32.  An opt-in measurement of reduced-precision
 * M2L: its error is far above the FP32 form's (see the ensemble benchmark). */
enum { SBSLPM_FLAG_GRAPH = 1, SBSLPM_FLAG_MULTI_RHS = 2, SBSLPM_FLAG_BF16_M2L = 4 };

/* Run parameters: the hot-path subset of RunConfig (run_config.hpp:6-44). */
typedef struct sbslpm_params {
  int levels;         /* RunConfig::levels; leaves at level levels-1 */
  int interp_degree;  /* RunConfig::interp_degree, 1..10 */
  int leaf_degree;    /* particle grid degree per leaf (m-1), 1..10; == interp_degree for the reference layout */
  int weight_rule;    /* SBSLPM_WEIGHTS_*: particle quadrature weights (initialize_cubed_sphere.cpp:275) */
  int rk4_combine;    /* SBSLPM_RK4_*: departure-point combine (bve_time_step.cpp:288-303) */
  int precision;      /* SBSLPM_FP64, SBSLPM_MIXED or SBSLPM_MIXED_VALU pair kernels */
  int tracer_count;   /* RunConfig::tracer_count */
  int flags;          /* bit 0: capture each BVE / shallow-water step in a hipGraph and replay it */
  double theta;       /* RunConfig::fmm_theta */
  double kernel_eps;  /* RunConfig::kernel_eps */
  double omega;       /* RunConfig::omega */
  double delta_t;     /* RunConfig::delta_t */
  int topography;     /* SBSLPM_TOPO_*: RunConfig::topo_type 0 none, 1 "cone" (topography.cpp:6-56); 2 Gaussian */
  int swe_rule;       /* SBSLPM_SWE_RULE_*: the shallow-water step's gradient/topography indexing */
  /* Shallow-water physics beyond the reference (all zero = the reference's step): */
  double gravity;     /* coefficient of the free-surface Laplacian in the divergence tendency: g/a^2
                       * for heights in metres on the unit sphere; 0 = the reference's 1 (no g/a^2) */
  double mountain_height, mountain_width, mountain_lat, mountain_lon; /* SBSLPM_TOPO_GAUSSIAN:
                       * h_s = height exp(-(r/width)^2), r the great-circle distance (radians) to (lat, lon) */
  int swe_form;       /* SBSLPM_SWE_FORM_* bits, 0 = the reference */
} sbslpm_params;

/* Topography (RunConfig::topo_type, topography.cpp:47-56; 2 has no reference counterpart). */
enum { SBSLPM_TOPO_NONE = 0, SBSLPM_TOPO_CONE = 1, SBSLPM_TOPO_GAUSSIAN = 2 };
/* Shallow-water step forms (no reference counterpart; 0 is swe_back_rk4_step as written):
 * REG_LAPLACIAN: the stage-1 surface Laplacian is the regularised-kernel FMM less 8/0.1 times each
 *   node's own height (the operator of stages 2-3 at rest) instead of the spectral panel Laplacian;
 * SURFACE_CORRECTION: the height correction subtracts 8/0.1 times the free surface (height plus
 *   topography at the moved point) instead of the height alone (swe_time_step.cpp:598-600). */
enum { SBSLPM_SWE_FORM_REG_LAPLACIAN = 1, SBSLPM_SWE_FORM_SURFACE_CORRECTION = 2 };

/* ---- Host-only set-up (no device needed) ------------------------------------------- */
typedef struct sbslpm_tree sbslpm_tree;
/* cubed_sphere_panels_init (initialize_cubed_sphere.cpp:185-234), interp_init (interp_funcs.cpp:24-73),
 * cube_sphere_spec_points (initialize_cubed_sphere.cpp:280-288) and dual_tree_traversal
 * (fmm_funcs.cpp:17-123) on the host. */
int sbslpm_tree_build(const sbslpm_params* params, sbslpm_tree** out);
void sbslpm_tree_destroy(sbslpm_tree* tree);
/* Same layout as sbslpm_get_sizes. */
int sbslpm_tree_sizes(const sbslpm_tree* tree, int64_t* sizes);
int sbslpm_tree_spec_points(const sbslpm_tree* tree, double* x, double* y, double* z, double* area);
int sbslpm_tree_interaction_list(const sbslpm_tree* tree, int* list);
/* ints[P][8] = face, level, parent, child1..child4, is_leaf;
 * dbls[P][6] = min_xi, max_xi, min_eta, max_eta (in [-1,1]), radius, area (CubedSpherePanel fields). */
int sbslpm_tree_panels(const sbslpm_tree* tree, int* ints, double* dbls);
/* point_locate_panel_2 (interp_funcs.cpp:175-203): leaf panel id containing (x,y,z), -1 on failure. */
int sbslpm_tree_locate(const sbslpm_tree* tree, double x, double y, double z);
/* Fixed transfer matrices: m2m_hi/m2m_lo [(d+1)*(d+1)] (parent basis a at upper/lower-half child node k,
 * index a*(d+1)+k) and p2m [(d+1)*m] (proxy basis a at particle node k). Any pointer may be NULL. */
int sbslpm_tree_transfer_matrices(const sbslpm_tree* tree, double* m2m_hi, double* m2m_lo, double* p2m);
/* cubed_sphere_midpoints (initialize_cubed_sphere.cpp:10-62): the 6 * 4^levels cell centres of the
 * equiangular grid with 2^levels cells per face edge, [face][xi][eta], and their areas (two
 * spherical triangles over the cell corners); faces 0..5 (the reference passes face+1, :35). */
int sbslpm_cubed_sphere_midpoints(int levels, double* x, double* y, double* z, double* area);
/* Host-only view of the irregular particle tree (same construction as sbslpm_irreg_create). */
typedef struct sbslpm_irreg_tree sbslpm_irreg_tree;
int sbslpm_irreg_tree_build(int64_t n, const double* x, const double* y, const double* z, int leaf_size,
                            int interp_degree, double theta, sbslpm_irreg_tree** out);
void sbslpm_irreg_tree_destroy(sbslpm_irreg_tree* tree);
int sbslpm_irreg_tree_sizes(const sbslpm_irreg_tree* tree, int64_t* sizes);
/* ints [P][10]: face, level, parent, child x4, is_leaf, point_count, first sorted point;
 * dbls [P][5]: min_xi, max_xi, min_eta, max_eta (radians), radius */
int sbslpm_irreg_tree_panels(const sbslpm_irreg_tree* tree, int* ints, double* dbls);
/* perm [N]: input index of each leaf-sorted point (leaves in panel order, points in the
 * reference's leaf_panel_point_ids order) */
int sbslpm_irreg_tree_points(const sbslpm_irreg_tree* tree, int* perm);
int sbslpm_irreg_tree_interaction_list(const sbslpm_irreg_tree* tree, int* list);
/* Unique particle positions, cubed_sphere_2d_to_1d (initialize_cubed_sphere.cpp:290-324):
 * particles in [leaf][node] order; one within 1e-15 in every coordinate of an earlier unique
 * point maps to the first such point, otherwise it becomes the next unique point.
 * two_d_to_1d [Pl][m2] receives the map, counts/x1d/y1d/z1d (each at least Pl*m2 long, may be
 * NULL) the multiplicity and position of each unique point, *n1d their number (the reference's
 * point_count). */
int sbslpm_tree_dedupe(const sbslpm_tree* tree, int* two_d_to_1d, int* counts, double* x1d, double* y1d,
                       double* z1d, int64_t* n1d);

/* ---- Device context ----------------------------------------------------------------- */
/* Builds the cubed-sphere tree (cubed_sphere_panels_init, initialize_cubed_sphere.cpp:185-234),
 * the interpolation tables (interp_init, interp_funcs.cpp:24-73), the interaction list
 * (dual_tree_traversal, fmm_funcs.cpp:17-123) and uploads them with all scratch space. */
int sbslpm_ctx_create(const sbslpm_params* params, sbslpm_ctx** out);
void sbslpm_ctx_destroy(sbslpm_ctx* ctx);
/* stream: a hipStream_t; NULL selects the default (legacy) stream.  A new context uses a
 * stream of its own.  hipGraph capture (SBSLPM_FLAG_GRAPH) needs a non-default stream. */
int sbslpm_ctx_set_stream(sbslpm_ctx* ctx, void* stream);
int sbslpm_ctx_set_precision(sbslpm_ctx* ctx, int precision);
const char* sbslpm_last_error(const sbslpm_ctx* ctx);
/* Waits for the stream; returns SBSLPM_E_GEOMETRY once if a kernel flagged a geometry error. */
int sbslpm_synchronize(sbslpm_ctx* ctx);

/* sizes[0..7] = P (panels), Pl (leaves), n (proxies per panel), m2 (particles per leaf),
 * I (interactions, RunConfig::fmm_interaction_count), N = Pl*m2, offset = P - Pl, tiles. */
int sbslpm_get_sizes(const sbslpm_ctx* ctx, int64_t* sizes);
/* Host copies of the particle positions and weights [Pl][m2]
 * (cube_sphere_spec_points, initialize_cubed_sphere.cpp:280-288). */
int sbslpm_spec_points(const sbslpm_ctx* ctx, double* x, double* y, double* z, double* area);
/* Host copy of the interaction list [I][3] = {target, source, type}, reference order. */
int sbslpm_interaction_list(const sbslpm_ctx* ctx, int* list);

/* Kernel timing (measurement aid, no reference counterpart): while enabled, every pair-kernel
 * launch is bracketed by hipEvents on the context stream (eager calls only; not inside a captured
 * step).  sbslpm_kernel_time waits for the recorded events and returns the summed milliseconds and
 * the launch count of one kernel class since the last reset:
 * 0 Biot-Savart FP64, 1 Biot-Savart mixed, 2 log FP64, 3 log mixed, 4 shallow-water FP64,
 * 5 regularised Laplacian FP64, 6 shallow-water mixed, 7 regularised Laplacian mixed,
 * 8 Biot-Savart mixed, multi-right-hand-side form (ensemble first stage, SBSLPM_FLAG_MULTI_RHS). */
enum {
  SBSLPM_KERNEL_BVE_FP64 = 0,
  SBSLPM_KERNEL_BVE_MIXED = 1,
  SBSLPM_KERNEL_LOG_FP64 = 2,
  SBSLPM_KERNEL_LOG_MIXED = 3,
  SBSLPM_KERNEL_SWE_FP64 = 4,
  SBSLPM_KERNEL_LAP_FP64 = 5,
  SBSLPM_KERNEL_SWE_MIXED = 6,
  SBSLPM_KERNEL_LAP_MIXED = 7,
  SBSLPM_KERNEL_BVE_MRHS = 8
};
int sbslpm_ctx_enable_timing(sbslpm_ctx* ctx, int enable);
int sbslpm_kernel_time(sbslpm_ctx* ctx, int kernel, double* ms, int64_t* launches, int reset);

/* upward_pass (fmm_funcs.cpp:341-355): q_part = area*pots [Pl][m2] (base_pots),
 * q_proxy [P][n] leaf rows by P2M, parents by M2M.  Assigns every row. */
int sbslpm_upward_pass(sbslpm_ctx* ctx, const double* area, const double* pots, double* q_part, double* q_proxy);

/* bve_vel_interactions (fmm_interactions/bve_vel.cpp:9-37): Biot-Savart pair sums over the
 * interaction list with sources displaced by disp (NULL = no displacement).
 * Accumulates into phi_part_{x,y,z} [Pl][m2] (leaf targets) and phi_proxy_{x,y,z} [P][n]. */
int sbslpm_bve_vel_interactions(sbslpm_ctx* ctx, const double* disp_part_x, const double* disp_part_y,
                                const double* disp_part_z, const double* disp_proxy_x, const double* disp_proxy_y,
                                const double* disp_proxy_z, const double* q_part, const double* q_proxy,
                                double* phi_part_x, double* phi_part_y, double* phi_part_z, double* phi_proxy_x,
                                double* phi_proxy_y, double* phi_proxy_z);

/* poisson_fmm_interactions (fmm_interactions/inverse_poisson.cpp:9-33): log-kernel pair sums.
 * Accumulates into psi_part [Pl][m2] and psi_proxy [P][n]. */
int sbslpm_poisson_fmm_interactions(sbslpm_ctx* ctx, const double* q_part, const double* q_proxy, double* psi_part,
                                    double* psi_proxy);

/* swe_vel_interactions (fmm_interactions/swe_vel.cpp:9-35, pair kernel swe_vel_impl.hpp:12-76):
 * Biot-Savart term of the vorticity charges plus the irrotational term of the divergence charges,
 * u += g qz (x cross y) + g qd (y - x (x.y)), g = -1/(4 pi (1 - x.y + eps)), y = P(node + disp).
 * Displacements may be NULL.  Accumulates into phi_part [3][Pl][m2] and phi_proxy [3][P][n]. */
int sbslpm_swe_vel_interactions(sbslpm_ctx* ctx, const double* disp_part_x, const double* disp_part_y,
                                const double* disp_part_z, const double* disp_proxy_x, const double* disp_proxy_y,
                                const double* disp_proxy_z, const double* qz_part, const double* qz_proxy,
                                const double* qd_part, const double* qd_proxy, double* phi_part_x, double* phi_part_y,
                                double* phi_part_z, double* phi_proxy_x, double* phi_proxy_y, double* phi_proxy_z);
/* swe_vel_interactions_2 (swe_vel.cpp:37-67, pair kernel swe_vel_impl.hpp:78-147): the same with
 * the divergence term taken at the undisplaced node y0 and the opposite sign,
 * u += g qz (x cross y) + g0 qd (y0 - x (x.y0)), g0 = +1/(4 pi (1 - x.y0 + eps)). */
int sbslpm_swe_vel_interactions_2(sbslpm_ctx* ctx, const double* disp_part_x, const double* disp_part_y,
                                  const double* disp_part_z, const double* disp_proxy_x, const double* disp_proxy_y,
                                  const double* disp_proxy_z, const double* qz_part, const double* qz_proxy,
                                  const double* qd_part, const double* qd_proxy, double* phi_part_x,
                                  double* phi_part_y, double* phi_part_z, double* phi_proxy_x, double* phi_proxy_y,
                                  double* phi_proxy_z);

/* downward_pass (fmm_funcs.cpp:357-372): L2L accumulated into phi_proxy [P][n] top-down,
 * then sols [Pl][m2] = L2P(leaf proxies) + phi_part (phi_part may be NULL). */
int sbslpm_downward_pass(sbslpm_ctx* ctx, double* phi_proxy, const double* phi_part, double* sols);
/* downward_pass_3 (fmm_funcs.cpp:374-392): the same for three components. */
int sbslpm_downward_pass_3(sbslpm_ctx* ctx, double* phi_proxy_x, double* phi_proxy_y, double* phi_proxy_z,
                           const double* phi_part_x, const double* phi_part_y, const double* phi_part_z,
                           double* sols_x, double* sols_y, double* sols_z);

/* One Biot-Savart velocity evaluation, upward pass + interactions + downward pass
 * (the sequence at bve_time_step.cpp:321-344).  vel_{x,y,z} [Pl][m2] assigned. */
int sbslpm_bve_velocity(sbslpm_ctx* ctx, const double* area, const double* vor, const double* disp_part_x,
                        const double* disp_part_y, const double* disp_part_z, const double* disp_proxy_x,
                        const double* disp_proxy_y, const double* disp_proxy_z, double* vel_x, double* vel_y,
                        double* vel_z);
/* One shallow-water velocity evaluation (swe_time_step.cpp:640-653): upward passes of vor and div,
 * swe_vel_interactions (variant 1) or swe_vel_interactions_2 (variant 2), downward pass.
 * vel_{x,y,z} [Pl][m2] assigned; displacements may be NULL. */
int sbslpm_swe_velocity(sbslpm_ctx* ctx, int variant, const double* area, const double* vor, const double* div,
                        const double* disp_part_x, const double* disp_part_y, const double* disp_part_z,
                        const double* disp_proxy_x, const double* disp_proxy_y, const double* disp_proxy_z,
                        double* vel_x, double* vel_y, double* vel_z);
/* One sphere Poisson solve (poisson_eq_fmm.cpp:129-147): psi [Pl][m2] assigned. */
int sbslpm_poisson_solve(sbslpm_ctx* ctx, const double* area, const double* f, double* psi);

/* FMM evaluation at nt arbitrary unit-sphere targets t{x,y,z} [nt] (device arrays), in FP64.
 * Each target takes the far field of the leaf that contains it, interpolated by BLI from the
 * leaf's L2L-swept proxy potentials (child_panel_interp_3_ll, fmm_funcs_impl.hpp:318-364), plus
 * the near field over that leaf's own interactions (the particle-target types 0/1 of
 * swe_vel_impl.hpp:188-255).  Point location as in interp_funcs_impl.hpp:114-142.
 * psi [nt] assigned: Poisson solve (log kernel, poisson_eq_fmm.cpp:129-147) at the targets. */
int sbslpm_poisson_solve_targets(sbslpm_ctx* ctx, const double* area, const double* f, int64_t nt,
                                 const double* tx, const double* ty, const double* tz, double* psi);
/* The same for the Biot-Savart velocity of vorticity vor (undisplaced sources, the context's
 * eps): vel_{x,y,z} [nt] assigned. */
int sbslpm_bve_velocity_targets(sbslpm_ctx* ctx, const double* area, const double* vor, int64_t nt,
                                const double* tx, const double* ty, const double* tz, double* vel_x,
                                double* vel_y, double* vel_z);

/* direct_sum_inv_lap (direct_sum.cpp:8-15): soln[i] += sum_{j != i} -1/(4 pi) log(1 - x_i.x_j)
 * area_j pots_j for i in [lb, ub) over n points. */
int sbslpm_direct_sum_inv_lap(sbslpm_ctx* ctx, int64_t n, const double* x, const double* y, const double* z,
                              const double* area, const double* pots, double* soln, int64_t lb, int64_t ub);
/* Allocation-free sbslpm_direct_sum_inv_lap (no reference counterpart).  For callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  Same arguments, same chunking and the same result to the bit; the partial sums live
 * in a fixed 8 MiB device array (one per GPU, held until the process ends, its address looked up
 * by sbslpm_ctx_create) instead of the context's on-demand scratch.  Calls no hipMalloc/hipFree.
 * The partial sums of a slice [lb, ub) must fit 2^20 doubles, which holds for every slice of at
 * most 2^20 = 1,048,576 targets; a longer slice returns SBSLPM_E_ARG: split it.  Every call on one GPU shares the array, so calls that overlap
 * on one GPU (other streams, other contexts, other host threads) race; calls in order on one
 * stream do not. */
int sbslpm_direct_sum_inv_lap_noalloc(sbslpm_ctx* ctx, int64_t n, const double* x, const double* y,
                                      const double* z, const double* area, const double* pots, double* soln,
                                      int64_t lb, int64_t ub);
/* Log direct sum at nt separate targets over ns sources; pairs with x.y >= 1 - 1e-15 skipped
 * (the FMM kernel's guard, inverse_poisson_impl.hpp:47).  soln [nt] accumulated. */
int sbslpm_direct_sum_inv_lap_targets(sbslpm_ctx* ctx, int64_t nt, const double* tx, const double* ty,
                                      const double* tz, int64_t ns, const double* sx, const double* sy,
                                      const double* sz, const double* area, const double* pots, double* soln);
/* Biot-Savart direct sum (no reference counterpart; pair kernel of bve_vel_impl.hpp:54-66):
 * vel[i] += sum_j -1/(4 pi) area_j vor_j (x_i cross y_j)/(1 - x_i.y_j + eps), y_j = P(x_j + disp_j),
 * for i in [lb, ub); disp may be NULL.  Uses the context's eps. */
int sbslpm_direct_sum_bve_vel(sbslpm_ctx* ctx, int64_t n, const double* x, const double* y, const double* z,
                              const double* area, const double* vor, const double* disp_x, const double* disp_y,
                              const double* disp_z, double* vel_x, double* vel_y, double* vel_z, int64_t lb,
                              int64_t ub);

/* Biot-Savart direct sum at nt separate targets over ns undisplaced sources (no reference
 * counterpart; pair kernel of bve_vel_impl.hpp:63-66 with the context's eps):
 * vel[i] += sum_j -1/(4 pi) area_j vor_j (t_i cross y_j)/(1 - t_i.y_j + eps), vel_{x,y,z} [nt]. */
int sbslpm_direct_sum_bve_vel_targets(sbslpm_ctx* ctx, int64_t nt, const double* tx, const double* ty,
                                      const double* tz, int64_t ns, const double* sx, const double* sy,
                                      const double* sz, const double* area, const double* vor, double* vel_x,
                                      double* vel_y, double* vel_z);

/* bve_displacement_upward_pass (bve_time_step.cpp:162-170): new_vor = vor + 2 Omega z - 2 Omega z',
 * z' of P(x + tau v); disp_part = tau v [Pl][m2]; disp_proxy [P][n] non-leaf rows interpolated. */
int sbslpm_bve_displacement(sbslpm_ctx* ctx, const double* x, const double* y, const double* z,
                            const double* vel_x, const double* vel_y, const double* vel_z, const double* vor,
                            double tau, double* new_vor, double* disp_part_x, double* disp_part_y,
                            double* disp_part_z, double* disp_proxy_x, double* disp_proxy_y, double* disp_proxy_z);
/* bve_find_departure_points (bve_time_step.cpp:172-309); vels = 12 pointers
 * {v0x,v0y,v0z, v1x,v1y,v1z, v2x,v2y,v2z, v3x,v3y,v3z} [Pl][m2]. */
int sbslpm_bve_find_departure_points(sbslpm_ctx* ctx, const double* x, const double* y, const double* z,
                                     const double* const* vels, double dt, double* dep_x, double* dep_y,
                                     double* dep_z);
/* bve_departure_to_target (bve_time_step.cpp:15-67); tracers [Pl][m2][ntr] (may be NULL if ntr == 0). */
int sbslpm_bve_departure_to_target(sbslpm_ctx* ctx, const double* z, const double* dep_x, const double* dep_y,
                                   const double* dep_z, const double* vor, const double* tracers, int ntr,
                                   double* new_vor, double* new_tracers);
/* Ensemble form of bve_back_rk4_step: nmem members on the context's grid (shared x, y, z, area
 * [Pl][m2]), each advanced by one backward RK4 step exactly as bve_back_rk4_step would, with every
 * kernel launch covering all members (the tree, interaction list and work tiles are shared).
 * vor [nmem][Pl][m2] is updated in place; vel_{x,y,z} [nmem][Pl][m2] receive the stage-0
 * velocities.  No tracers.  Needs levels >= 2.  The workspace for nmem members is allocated on
 * the first call with that count; with SBSLPM_FLAG_GRAPH and a non-default stream the step is
 * captured once in a hipGraph and replayed. */
int sbslpm_bve_back_rk4_step_ensemble(sbslpm_ctx* ctx, int nmem, const double* x, const double* y, const double* z,
                                      const double* area, double* vor, double* vel_x, double* vel_y, double* vel_z,
                                      double time);

/* Adjoint (transpose) of the Biot-Savart velocity evaluation with undisplaced sources
 * (sbslpm_bve_velocity with NULL displacements, the first RK4 stage): for a cotangent
 * ubar_{x,y,z} [Pl][m2] of the velocity, vor_bar [Pl][m2] is assigned K^T ubar, K the FP64 FMM
 * operator of the context (tree, interaction list, upward and downward passes) -- its exact
 * transpose, which approximates
 *   vor_bar_s = -(area_s / (4 pi)) x_s . sum_t g(x_s, x_t) (ubar_t cross x_t),
 *   g = 1/(1 - x.y + eps)
 * to the FMM truncation.  On a mixed-precision context the transposed pair sums run in FP32
 * (panel-centred |x - y|^2 / 2 form, FP64 accumulation across interactions), the passes in
 * FP64.  No reference counterpart; it transposes the operator of bve_vel_interactions
 * (bve_vel.cpp:9-37, bve_vel_impl.hpp:34-70) with the passes of fmm_funcs.cpp:341-392. */
int sbslpm_bve_velocity_adjoint(sbslpm_ctx* ctx, const double* area, const double* ubar_x, const double* ubar_y,
                                const double* ubar_z, double* vor_bar);
/* Gradient of one bve_back_rk4_step (bve_time_step.cpp:311-419, no tracers) with respect to the
 * vorticity before the step: vor_bar [Pl][m2] is assigned J^T new_vor_bar, J the Jacobian of
 * vor -> vor after the step at the state vor (which is not modified).  The step's forward pass
 * is recomputed inside in the context's precision; the reverse pass is FP64 (its transposed
 * pair sums FP32 on a mixed-precision context) and differentiates
 * the four FMM velocity evaluations in the charges and the source displacements, the
 * displacement updates, the backward RK4 departure search (BLI interpolation differentiated in
 * the point, re-projections, the configured combine rule) and the remap with its Coriolis
 * terms.  Point location is piecewise constant and contributes nothing.  No reference
 * counterpart (the reference has no adjoint); for 4D-Var style gradients through many steps. */
int sbslpm_bve_back_rk4_step_adjoint(sbslpm_ctx* ctx, const double* x, const double* y, const double* z,
                                     const double* area, const double* vor, const double* new_vor_bar,
                                     double* vor_bar);
/* One bve_back_rk4_step (no tracers) that keeps what its reverse pass needs: vor_out [Pl][m2]
 * receives the vorticity after the step (must not alias vor), stages [12][Pl][m2] the four
 * stage velocities v0x v0y v0z ... v3z.  Bit-identical to sbslpm_bve_back_rk4_step's result. */
int sbslpm_bve_back_rk4_step_record(sbslpm_ctx* ctx, const double* x, const double* y, const double* z,
                                    const double* area, const double* vor, double* vor_out, double* stages);
/* sbslpm_bve_back_rk4_step_adjoint with the stage velocities of the step supplied (from
 * sbslpm_bve_back_rk4_step_record on the same state) instead of recomputed. */
int sbslpm_bve_back_rk4_step_adjoint_stages(sbslpm_ctx* ctx, const double* x, const double* y, const double* z,
                                            const double* area, const double* vor, const double* stages,
                                            const double* new_vor_bar, double* vor_bar);
/* Transpose of the remap interpolation of bve_departure_to_target (bve_time_step.cpp:38-66)
 * with respect to the field: vor_bar [Pl][m2] += B^T new_vor_bar, B the BLI weights at the
 * departure points.  Accumulated with FP64 atomic adds (summation order not fixed). */
int sbslpm_bve_departure_to_target_adjoint(sbslpm_ctx* ctx, const double* dep_x, const double* dep_y,
                                           const double* dep_z, const double* new_vor_bar, double* vor_bar);

/* Autotuner of the Biot-Savart pair kernels' launch shape (no reference counterpart): times
 * the context's pair kernel (its precision) on its own interaction list with the records of
 * (area, vor), for every candidate of waves per workgroup (1, 2, 4) x tile-to-XCD mapping
 * (0: workgroups dealt round-robin over the XCDs by the hardware; 1: a contiguous range of
 * equally many tiles per XCD; 2: a contiguous range of equal estimated work per XCD), and
 * keeps the fastest for all later evaluations, steps and ensembles
 * and the step adjoint's transposed pair sums (captured graphs are dropped and recaptured).
 * report [ncand][3] (may be NULL) receives
 * waves, xcd, ms per launch; *ncand the number of candidates timed.  Results do not depend on
 * the shape: every target is still summed by one lane in the same order. */
int sbslpm_autotune(sbslpm_ctx* ctx, const double* area, const double* vor, double* report, int* ncand);
/* Autotuner of the multi-right-hand-side pair kernel of ensemble steps (no reference
 * counterpart; mixed-precision context with SBSLPM_FLAG_MULTI_RHS, 2 <= nmem <= 65535): times
 * it on the context's interaction list with the charges of vor [nmem][Pl][m2] for every
 * candidate of waves per workgroup (1, 2, 4) x tile-to-XCD mapping (0, 1, 2 as above) x chunks
 * of four members per wave (1, 2), and keeps the fastest for later ensemble steps (a captured
 * ensemble step is dropped and recaptured).  report [ncand][4] (may be NULL) receives waves,
 * xcd, chunks per wave, ms per launch (18 rows).  Allocates the ensemble workspace for nmem
 * members.  Results do not depend on the shape: every (target, member) is summed by one lane
 * in the same order. */
int sbslpm_autotune_ensemble(sbslpm_ctx* ctx, int nmem, const double* area, const double* vor, double* report,
                             int* ncand);
/* Sets the launch shape directly (waves 1, 2 or 4; xcd 0, 1 or 2), e.g. a shape tuned earlier. */
int sbslpm_set_pair_shape(sbslpm_ctx* ctx, int waves, int xcd);
int sbslpm_get_pair_shape(const sbslpm_ctx* ctx, int* waves, int* xcd);

/* ---- Irregular particle tree (the lat-lon path) -------------------------
 * Points in any order (the reference's [lat][lon] arrays flattened, index i*lon_count + j).
 * Face panels split into quadrants while they hold more than leaf_size points
 * (initialize_cube_sphere_irreg_points, initialize_cubed_sphere.cpp:446-592; panel bounds in
 * radians); interactions from dual_tree_traversal_irreg (fmm_funcs.cpp:125-231), whose type
 * bits mark panels above the deepest level (proxies) rather than non-leaves.  The device handle
 * uses the context's stream, theta and kernel_eps; point arrays are device pointers in the
 * caller's order, proxy arrays [P][(d+1)^2] in the tree's panel order (the reference's). */
typedef struct sbslpm_irreg sbslpm_irreg;
int sbslpm_irreg_create(sbslpm_ctx* ctx, int64_t n, const double* x, const double* y, const double* z, int leaf_size,
                        int interp_degree, sbslpm_irreg** out);  /* x, y, z: host arrays */
void sbslpm_irreg_destroy(sbslpm_irreg* h);
/* sizes: P, deepest level, (d+1)^2, points, interactions, work tiles */
int sbslpm_irreg_sizes(const sbslpm_irreg* h, int64_t* sizes);
/* upward_pass_ll (fmm_funcs.cpp:394-404): proxy_source_pots assigned (the reference
 * accumulates into caller-zeroed rows); base_pots_ll by BLI anterpolation of each leaf's
 * points, child_to_parent_ll by the fixed quadrant matrices (no race). */
int sbslpm_irreg_upward_pass(sbslpm_irreg* h, const double* area, const double* pots, double* proxy_source_pots);
/* downward_pass_3_ll (fmm_funcs.cpp:406-419): proxy rows pushed to the children in place,
 * then sols += BLI interpolation of each leaf's proxies at its points. */
int sbslpm_irreg_downward_pass_3(sbslpm_irreg* h, double* proxy_1, double* proxy_2, double* proxy_3, double* sols_1,
                                 double* sols_2, double* sols_3);
/* swe_vel_interactions_ll (swe_vel.cpp:69-100, swe_vel_impl.hpp:149-288): target_pots [P][n] and
 * vel [N] accumulated; point sources carry vor*area, div*area. */
int sbslpm_irreg_swe_vel_interactions(sbslpm_irreg* h, double* target_pots_1, double* target_pots_2,
                                      double* target_pots_3, double* vel_x, double* vel_y, double* vel_z,
                                      const double* proxy_vors, const double* proxy_divs, const double* vors,
                                      const double* divs, const double* area);
/* The lat-lon shallow-water velocity sequence (upward passes of vor and div, interactions,
 * downward pass) fused; vel assigned. */
int sbslpm_irreg_swe_velocity(sbslpm_irreg* h, const double* area, const double* vors, const double* divs,
                              double* vel_x, double* vel_y, double* vel_z);
/* Biot-Savart velocity of a vorticity field on the irregular tree (the kernel of
 * bve_vel_impl.hpp:54-66 with undisplaced sources); vel assigned.  No reference counterpart
 * on this tree. */
int sbslpm_irreg_bve_velocity(sbslpm_irreg* h, const double* area, const double* vor, double* vel_x, double* vel_y,
                              double* vel_z);

/* ---- Lat-lon shallow-water step (swe_time_step_ll.cpp) ---------------------------------------
 * Fields are device arrays [lat_count][lon_count], row-major (index i*lon_count + j), on the
 * cell-centred grid of latlon_point (general_utils_impl.hpp:252-276): row i at latitude
 * -90 + grid_spacing (i + 1/2), column j at longitude grid_spacing (j + 1/2), in degrees; x, y, z
 * are its points.  lat_count and lon_count are at least 3.  Velocities are on the unit sphere per
 * second and heights in Earth radii, as in the reference.  Two rules where the reference is
 * undefined: the polar filter's Fourier coefficients start at zero (polar_filtering leaves all
 * but the first uninitialised, :628-631, :644-645), and field_averaging's longitude neighbours
 * are the periodic j+1 and j-1 (it reads them before setting them, :322-323).  Grid indices are
 * kept inside the grid where the reference's would leave it.  Outputs are assigned and may not
 * alias the inputs (the stencils read neighbouring points). */
/* swe_tendency_computation_ll (swe_time_step_ll.cpp:365-387): vor_tend_ll, height_tend_ll,
 * div_tend_ll (:14-231), then div_tend values below 2e-12 in magnitude set to zero (filter_vals,
 * general_utils_impl.hpp:278-293). */
int sbslpm_ll_tendencies(sbslpm_ctx* ctx, int lat_count, int lon_count, double grid_spacing, double omega,
                         const double* x, const double* y, const double* z, const double* vors, const double* divs,
                         const double* height, const double* topo, const double* vel_x, const double* vel_y,
                         const double* vel_z, double* vor_tend, double* div_tend, double* height_tend);
/* polar_filtering (swe_time_step_ll.cpp:600-664), in place: on each row with |lat| > 60 degrees
 * the values (clamped as the reference does: above 1e-11 and below -1e-11 both to +1e-11) are
 * replaced by their series in cos(k lat), sin(k lat), k < min(floor(90 - |lat|) + 1, 29). */
int sbslpm_ll_polar_filter(sbslpm_ctx* ctx, int lat_count, int lon_count, double grid_spacing, const double* x,
                           const double* y, const double* z, double* field);
/* swe_compute_arrival_values_ll (swe_time_step_ll.cpp:389-598): quadratic semi-Lagrangian
 * arrival values of vorticity and height (1-D along the row where |v| < 1e-16, else 3 x 3 at the
 * departure point x - delta_t vel, with the ring means of the first and last rows at the poles);
 * a_div = div + delta_t div_tend. */
int sbslpm_ll_arrival_values(sbslpm_ctx* ctx, int lat_count, int lon_count, double grid_spacing, double delta_t,
                             const double* x, const double* y, const double* z, const double* vors,
                             const double* divs, const double* height, const double* vel_x, const double* vel_y,
                             const double* vel_z, const double* vor_tend, const double* div_tend, double* a_vor,
                             double* a_div, double* a_height);
/* field_averaging (swe_time_step_ll.cpp:270-363): 5-point average with weight epsilon on rows
 * with |lat| > 60 degrees, times cos^2(lat); other rows copied. */
int sbslpm_ll_field_averaging(sbslpm_ctx* ctx, int lat_count, int lon_count, const double* x, const double* y,
                              const double* z, const double* field, double* new_field, double epsilon);
/* laplacian_diffusion (swe_time_step_ll.cpp:233-268): interior rows plus
 * dt * 0.89/6371000 * the 5-point spherical Laplacian; first and last rows copied. */
int sbslpm_ll_laplacian_diffusion(sbslpm_ctx* ctx, int lat_count, int lon_count, double grid_spacing, double dt,
                                  const double* x, const double* y, const double* z, const double* field,
                                  double* new_field);
/* swe_vel_tend_computation (swe_time_step_ll.cpp:687-735) on an irregular tree built from the
 * grid's points (x, y, z here are their device copies, in the same order): the FMM velocity
 * (sbslpm_irreg_swe_velocity), the tendencies, and the polar filter of the height tendency. */
int sbslpm_irreg_swe_vel_tend_ll(sbslpm_irreg* h, int lat_count, int lon_count, double grid_spacing, double omega,
                                 const double* x, const double* y, const double* z, const double* area,
                                 const double* vors, const double* divs, const double* height, const double* topo,
                                 double* vel_x, double* vel_y, double* vel_z, double* vor_tend, double* div_tend,
                                 double* height_tend);
/* swe_back_rk4_step_ll (swe_time_step_ll.cpp:737-774): one semi-Lagrangian step in place --
 * velocity and tendencies, arrival values, field_averaging (epsilon 0.16) and laplacian_diffusion
 * of the divergence, polar filter of the divergence.  vel receives the start state's velocity.
 * The height tendency and its filter (:728), which the step does not use, are skipped.  time is
 * unused, as in the reference. */
int sbslpm_irreg_swe_back_rk4_step_ll(sbslpm_irreg* h, int lat_count, int lon_count, double grid_spacing,
                                      double delta_t, double omega, const double* x, const double* y,
                                      const double* z, const double* area, double* vors, double* divs,
                                      double* height, const double* topo, double* vel_x, double* vel_y,
                                      double* vel_z, double time);

/* BLI interpolation of nfields (1..8) leaf fields [Pl][m2] at n arbitrary points (x, y, z on
 * the unit sphere): out[q][i] = field q interpolated at point i on the leaf that contains it.
 * The functor of latlon_interp_host / interp_to_latlon_host (interp_funcs.hpp:22-77) for any
 * point set; the particle grid's degree m-1 replaces interp_degree (equal in the reference
 * layout). */
int sbslpm_interp_points(sbslpm_ctx* ctx, int64_t n, const double* x, const double* y, const double* z, int nfields,
                         const double* const* fields, double* const* out);
/* Transpose of sbslpm_interp_points: field_bar[q] [Pl][m2] += B^T out_bar[q] (n points), with
 * the transposed interpolations summed per wave and leaf and added with FP64 atomics (order not
 * fixed).  No reference counterpart; the observation operator's adjoint for gradients. */
int sbslpm_interp_points_adjoint(sbslpm_ctx* ctx, int64_t n, const double* x, const double* y, const double* z,
                                 int nfields, const double* const* out_bar, double* const* field_bar);
/* Zonal and meridional velocity from Cartesian velocity at n points, xyz_vel_to_uv_vel
 * (general_utils_impl.hpp:195-229; zero at the poles). */
int sbslpm_xyz_vel_to_uv_vel(sbslpm_ctx* ctx, int64_t n, const double* x, const double* y, const double* z,
                             const double* vx, const double* vy, const double* vz, double* u, double* v);

/* swe_displacement_upward_pass (swe_time_step.cpp:265-276): disp_part = tau v [Pl][m2]
 * (swe_disp_update, :194-210); disp_proxy [P][n] non-leaf rows interpolated from the leaf that
 * contains each undisplaced node (swe_disp_interp, :212-263). */
int sbslpm_swe_displacement(sbslpm_ctx* ctx, const double* vel_x, const double* vel_y, const double* vel_z,
                            double tau, double* disp_part_x, double* disp_part_y, double* disp_part_z,
                            double* disp_proxy_x, double* disp_proxy_y, double* disp_proxy_z);
/* swe_find_departure_points (swe_time_step.cpp:422-556): the backward RK4 search of
 * bve_find_departure_points, same arguments. */
int sbslpm_swe_find_departure_points(sbslpm_ctx* ctx, const double* x, const double* y, const double* z,
                                     const double* const* vels, double dt, double* dep_x, double* dep_y,
                                     double* dep_z);
/* swe_departure_to_target (swe_time_step.cpp:19-74): vor, div, h [Pl][m2] and tracers [Pl][m2][ntr]
 * interpolated at the departure points (no Coriolis terms); outputs assigned. */
int sbslpm_swe_departure_to_target(sbslpm_ctx* ctx, const double* dep_x, const double* dep_y, const double* dep_z,
                                   const double* vor, const double* div, const double* h, const double* tracers,
                                   int ntr, double* new_vor, double* new_div, double* new_h, double* new_tracers);

/* ---- Shallow-water step ------------------------------------------------------------------ */
/* laplacian (deriv_funcs.cpp:27-37, panel_laplacian deriv_funcs_impl.hpp:338-389): spectral surface
 * Laplacian of f on every leaf's particle grid; lap [Pl][m2] assigned. */
int sbslpm_laplacian(sbslpm_ctx* ctx, const double* f, double* lap);
/* laplacian_interactions (fmm_interactions/laplacian.cpp:6-35, laplacian_impl.hpp:11-69): the
 * regularised-Laplacian kernel (eps = 0.1) over the interaction list, sources at P(node + disp)
 * (disp may be NULL); phi_part [Pl][m2], phi_proxy [P][n] accumulated.  On a mixed-precision
 * context the pairs are FP32 (FP64 accumulation per source set), as are the shallow-water pairs of
 * variants 1 and 3 (sbslpm_swe_vel_interactions, sbslpm_swe_velocity, the step). */
int sbslpm_laplacian_interactions(sbslpm_ctx* ctx, const double* disp_part_x, const double* disp_part_y,
                                  const double* disp_part_z, const double* disp_proxy_x, const double* disp_proxy_y,
                                  const double* disp_proxy_z, const double* q_part, const double* q_proxy,
                                  double* phi_part, double* phi_proxy);
/* apply_topography (topography.cpp:47-56): eff = height (no topography) or the cone mountain at
 * P(x + disp) (disp: [Pl][m2] particle and [P][n] proxy displacements, NULL = zero); the
 * displacement index follows the context's swe_rule. */
int sbslpm_apply_topography(sbslpm_ctx* ctx, const double* x, const double* y, const double* z,
                            const double* disp_part_x, const double* disp_part_y, const double* disp_part_z,
                            const double* disp_proxy_x, const double* disp_proxy_y, const double* disp_proxy_z,
                            const double* height, double* eff_height);
/* swe_tendency_computation (swe_time_step.cpp:76-161) with the context's omega and swe_rule:
 * vor_t, div_t, h_t [Pl][m2] assigned. */
int sbslpm_swe_tendency(sbslpm_ctx* ctx, const double* x, const double* y, const double* z, const double* vel_x,
                        const double* vel_y, const double* vel_z, const double* vor, const double* div,
                        const double* height, const double* height_lap, double* vor_t, double* div_t, double* h_t);
/* swe_height_correction (swe_time_step.cpp:558-604): height_lap -= 80 BLI[h + dt h_t](P(x + dt vel)). */
int sbslpm_swe_height_correction(sbslpm_ctx* ctx, const double* x, const double* y, const double* z,
                                 const double* vel_x, const double* vel_y, const double* vel_z, const double* height,
                                 const double* h_t, double dt, double* height_lap);
/* unify_boundary_vals (initialize_cubed_sphere.cpp:364-370, functors :328-362): vals [Pl][m2]
 * (device) is replaced, entry by entry, by the mean over all entries with the same unique index
 * two_d_to_1d [Pl][m2] (device int32, from sbslpm_tree_dedupe); counts [n1d] (device int32) are
 * the multiplicities.  Sums use FP64 atomics like the reference, so the order of the at most
 * four terms is not fixed.  Entries whose index lies outside [0, n1d) are left unchanged. */
int sbslpm_unify_boundary_vals(sbslpm_ctx* ctx, const int* two_d_to_1d, const int* counts, int64_t n1d,
                               double* vals);
/* swe_back_rk4_step (swe_time_step.cpp:606-811): vors, divs, height [Pl][m2] and tracers
 * [Pl][m2][tracer_count] updated in place; vel_{x,y,z} receive the stage-1 velocity and height_lap
 * the stage-1 surface Laplacian.  Variant-1 SWE kernel with the context's kernel_eps, omega,
 * delta_t, rk4_combine, topography and swe_rule.  The stage-4 Laplacian, height correction and
 * tendencies and the k2..k4 terms of swe_compute_arrival_values reach no output there
 * (arrival = state + dt k1, :415-417) and are not computed.  With SBSLPM_FLAG_GRAPH on a
 * non-default stream the step is captured in a hipGraph per set of pointers and replayed.  `time`
 * is accepted for signature parity (it only drives the commented-out forcing). */
int sbslpm_swe_back_rk4_step(sbslpm_ctx* ctx, const double* x, const double* y, const double* z,
                             const double* area, double* vors, double* divs, double* height, double* vel_x,
                             double* vel_y, double* vel_z, double* tracers, double* height_lap, double time);

/* Shallow-water step adjoint (no reference counterpart; it differentiates swe_back_rk4_step,
 * swe_time_step.cpp:606-811, in every form the context selects).  Needs tracer_count == 0.
 * sbslpm_swe_back_rk4_step_record: one step out of place, new_{vors,divs,height} [Pl][m2]
 * assigned, and stages [21][Pl][m2] = the stage velocities V0..V3 (x,y,z each) and the stage
 * tendencies T0..T2 (vor, div, h each) the reverse pass reads. */
int sbslpm_swe_back_rk4_step_record(sbslpm_ctx* ctx, const double* x, const double* y, const double* z,
                                    const double* area, const double* vors, const double* divs,
                                    const double* height, double* new_vors, double* new_divs, double* new_height,
                                    double* stages);
/* Gradient of one step with respect to (vors, divs, height) before it, for the cotangents of
 * (vors, divs, height) after it: {vors,divs,height}_bar [Pl][m2] assigned.  Reverse of the
 * remap and departure search, then of stages 4..1: the shallow-water and regularised-Laplacian
 * FMM evaluations transposed in their charges and displaced source positions (FP32 pair sums on
 * mixed-precision contexts, like the forward), the height correction, the topography at displaced
 * points, the tendencies with their leaf differentiation, the spectral panel Laplacian, the state
 * and area updates.  Point location is piecewise constant and contributes nothing.  The forward
 * is recomputed; the _stages form takes the record of sbslpm_swe_back_rk4_step_record instead. */
int sbslpm_swe_back_rk4_step_adjoint(sbslpm_ctx* ctx, const double* x, const double* y, const double* z,
                                     const double* area, const double* vors, const double* divs,
                                     const double* height, const double* new_vors_bar, const double* new_divs_bar,
                                     const double* new_height_bar, double* vors_bar, double* divs_bar,
                                     double* height_bar);
int sbslpm_swe_back_rk4_step_adjoint_stages(sbslpm_ctx* ctx, const double* x, const double* y, const double* z,
                                            const double* area, const double* vors, const double* divs,
                                            const double* height, const double* stages, const double* new_vors_bar,
                                            const double* new_divs_bar, const double* new_height_bar,
                                            double* vors_bar, double* divs_bar, double* height_bar);

/* bve_back_rk4_step (bve_time_step.cpp:311-419), no forcing (bve_forcing copy,
 * forcing_funcs.cpp:93-95).  vor [Pl][m2] and tracers [Pl][m2][tracer_count] updated in place;
 * vel_{x,y,z} receive the stage-0 velocity; proxy_source_vors [P][n] is scratch (may be NULL).
 * `time` is accepted for signature parity (it only drives forcing). */
int sbslpm_bve_back_rk4_step(sbslpm_ctx* ctx, const double* x, const double* y, const double* z,
                             const double* area, double* vor, double* proxy_source_vors, double* vel_x,
                             double* vel_y, double* vel_z, double* tracers, double time);

#ifdef __cplusplus
}
#endif

#endif /* SBSLPM_H */
