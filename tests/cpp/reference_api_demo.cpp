// Drop-in demo: the reference's call sequence of bve_back_rk4_step stage 0
// (bve_time_step.cpp:321-344) written against sbslpm_reference_api.hpp, checked
// against the fused entry point.  "host" mode needs no GPU.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "sbslpm_reference_api.hpp"

using sbslpm_ref::View2;

static int host_mode() {
  sbslpm_params p{};
  p.levels = 4; p.interp_degree = 8; p.leaf_degree = 7; p.weight_rule = SBSLPM_WEIGHTS_CLENSHAW_CURTIS;
  p.theta = 0.7; p.kernel_eps = 1e-4; p.omega = 7.292e-5; p.delta_t = 3600;
  sbslpm_tree* t = nullptr;
  if (sbslpm_tree_build(&p, &t) != SBSLPM_OK) return 1;
  int64_t s[8];
  sbslpm_tree_sizes(t, s);
  std::printf("panels %lld leaves %lld particles %lld interactions %lld\n", (long long)s[0], (long long)s[1],
              (long long)s[5], (long long)s[4]);
  sbslpm_tree_destroy(t);
  p.levels = 0;
  return sbslpm_tree_build(&p, &t) == SBSLPM_E_ARG ? 0 : 2;
}

static double* dalloc(size_t n) {
  double* d = nullptr;
  if (hipMalloc(&d, n * sizeof(double)) != hipSuccess) return nullptr;
  hipMemset(d, 0, n * sizeof(double));
  return d;
}

static int gpu_mode() {
  sbslpm_params p{};
  p.levels = 3; p.interp_degree = 8; p.leaf_degree = 8; p.weight_rule = SBSLPM_WEIGHTS_REFERENCE;
  p.theta = 0.7; p.kernel_eps = 1e-4; p.omega = 7.292e-5; p.delta_t = 3600;
  sbslpm_ctx* ctx = nullptr;
  if (sbslpm_ctx_create(&p, &ctx) != SBSLPM_OK) return 1;
  int64_t s[8];
  sbslpm_get_sizes(ctx, s);
  const int64_t P = s[0], Pl = s[1], n = s[2], N = s[5];
  std::vector<double> x(N), y(N), z(N), a(N), w(N);
  sbslpm_spec_points(ctx, x.data(), y.data(), z.data(), a.data());
  for (int64_t i = 0; i < N; ++i) w[i] = 1e-5 * (z[i] + 0.3 * x[i] * y[i]);
  double *dx = dalloc(N), *dy = dalloc(N), *dz = dalloc(N), *da = dalloc(N), *dw = dalloc(N);
  hipMemcpy(dx, x.data(), N * 8, hipMemcpyHostToDevice);
  hipMemcpy(dy, y.data(), N * 8, hipMemcpyHostToDevice);
  hipMemcpy(dz, z.data(), N * 8, hipMemcpyHostToDevice);
  hipMemcpy(da, a.data(), N * 8, hipMemcpyHostToDevice);
  hipMemcpy(dw, w.data(), N * 8, hipMemcpyHostToDevice);
  // reference-shaped arrays
  View2<double> xcos{dx, Pl, n}, area{da, Pl, n}, vors{dw, Pl, n};
  View2<double> q{dalloc(P * n), P, n}, dispx{dalloc(P * n), P, n}, dispy{dalloc(P * n), P, n},
      dispz{dalloc(P * n), P, n}, t1{dalloc(P * n), P, n}, t2{dalloc(P * n), P, n}, t3{dalloc(P * n), P, n};
  View2<double> v1{dalloc(N), Pl, n}, v2{dalloc(N), Pl, n}, v3{dalloc(N), Pl, n};
  try {
    sbslpm_ref::upward_pass(ctx, area, vors, q);
    sbslpm_ref::bve_vel_interactions(ctx, dispx, dispy, dispz, t1, t2, t3, q);
    sbslpm_ref::downward_pass_3(ctx, t1, t2, t3, v1, v2, v3);
  } catch (const std::exception& e) {
    std::printf("error: %s\n", e.what());
    return 3;
  }
  double *f1 = dalloc(N), *f2 = dalloc(N), *f3 = dalloc(N);
  sbslpm_bve_velocity(ctx, da, dw, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, f1, f2, f3);
  if (sbslpm_synchronize(ctx) != SBSLPM_OK) return 4;
  std::vector<double> r(N), g(N);
  double num = 0, den = 0;
  double* refs[3] = {v1.data(), v2.data(), v3.data()};
  double* fused[3] = {f1, f2, f3};
  for (int c = 0; c < 3; ++c) {
    hipMemcpy(r.data(), refs[c], N * 8, hipMemcpyDeviceToHost);
    hipMemcpy(g.data(), fused[c], N * 8, hipMemcpyDeviceToHost);
    for (int64_t i = 0; i < N; ++i) { num += (r[i] - g[i]) * (r[i] - g[i]); den += g[i] * g[i]; }
  }
  const double rel = std::sqrt(num / den);
  std::printf("reference-API sequence vs fused velocity: rel-L2 %.3e\n", rel);
  // one full step through the reference signature
  View2<double> ycos{dy, Pl, n}, zcos{dz, Pl, n};
  sbslpm_ref::bve_back_rk4_step(ctx, xcos, ycos, zcos, area, vors, q, v1, v2, v3, nullptr, 0.0);
  if (sbslpm_synchronize(ctx) != SBSLPM_OK) return 5;
  sbslpm_ctx_destroy(ctx);
  return rel < 1e-13 ? 0 : 6;
}

int main(int argc, char** argv) {
  if (argc > 1 && std::strcmp(argv[1], "gpu") == 0) return gpu_mode();
  return host_mode();
}
