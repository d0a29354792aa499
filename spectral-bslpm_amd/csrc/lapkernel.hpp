// lapkernel.hpp -- the regularised-Laplacian kernel of the shallow-water step
// (lap_panel_interaction, laplacian_impl.hpp:46-66) with its constants folded
// into short polynomials, for the forward pair kernels (k_lap64, k_lap_mixed)
// and the transposed ones (k_lap_adj64, k_lap_adj32).
//
// For t = x.y > 1 - 30 eps:
//   K(t)  = c1 2t(t^2 - 1 + 3 eps) + (1 + erf(t/eps)) E c2 Q(t),
//   K'(t) = c1 (6t^2 - 2 + 6 eps)
//           + c2 [(2/(sqrt(pi) eps)) e^{-t^2/eps^2} E Q + (1 + erf(t/eps)) E ((2t/eps) Q + Q')],
//   E = e^{(t^2-1)/eps},  Q = 2 eps^2 + 2t^2(t^2 - 1) + eps(7t^2 - 1),
//   c1 = e^{-1/eps} pi^{-3/2},  c2 = 1/(sqrt(eps) pi).
// erf(s) rounds to exactly +-1 in FP64 for |s| >= 6 (erfc(6) = 2.2e-17 is
// below half an ulp of 1) and the e^{-s^2} term is below an ulp of the others,
// so for t >= 6 eps (over 99% of the pair work at the tuned sizes)
//   K = A + E B,  K' = A' + E D  with
//   A  = c1 2t(t^2 - 1 + 3 eps),     B = 2 c2 Q,
//   A' = c1 (6t^2 - 2 + 6 eps),      D = 2 c2 ((2t/eps) Q + Q')
//      = 2 c2 t ((4/eps) t^4 + (22 - 4/eps) t^2 + 18 eps - 6).
// Elsewhere, with f = (1 + erf(t/eps))/2 and g = (2/(sqrt(pi) eps)) e^{-t^2/eps^2}:
//   K = A + f E B,  K' = A' + E (f D + g B/2).
//
// LapCoef64 holds these in powers of t (every term times `scale`); LapCoef32 in
// powers of u = (t^2 - 1) log2(e)/eps, the base-2 exponent of E, which the FP32
// kernels form without cancellation from the squared chord
// h2 = |x - y|^2 = 2(1 - t) as t^2 - 1 = h2 (h2/4 - 1).
#pragma once
#include <cmath>

namespace sbslpm {

struct LapCoef64 {
  double a1, a0;      // A  = t (a1 t^2 + a0)
  double b2, b1, b0;  // B  = (b2 t^2 + b1) t^2 + b0
  double e1, e0;      // A' = e1 t^2 + e0
  double d2, d1, d0;  // D  = t ((d2 t^2 + d1) t^2 + d0)
  double inv_eps;     // 1/eps: the exponent of E is (t^2 - 1)/eps
  double cut;         // 1 - 30 eps: pairs at or below it contribute nothing
  double t_sat;       // t >= t_sat gives t/eps >= 6 after rounding
};

struct LapCoef32 {
  float a1, a0;      // A  = t (a1 u + a0)
  float b2, b1, b0;  // B  = (b2 u + b1) u + b0
  float e1, e0;      // A' = e1 u + e0
  float d2, d1, d0;  // D  = t ((d2 u + d1) u + d0)
  float k;           // log2(e)/eps: u = k (t^2 - 1), E = 2^u
  float inv_eps, cut, t_sat;
};

inline void lap_constants(double eps, double& c1, double& c2) {
  const double pi = 3.14159265358979323846;
  c1 = std::exp(-1.0 / eps) * std::pow(pi, -1.5);
  c2 = 1.0 / (std::sqrt(eps) * pi);
}

inline LapCoef64 lap_coef64(double eps, double scale) {
  double c1, c2;
  lap_constants(eps, c1, c2);
  LapCoef64 k{};
  k.a1 = scale * 2.0 * c1;
  k.a0 = scale * 2.0 * c1 * (3.0 * eps - 1.0);
  k.b2 = scale * 4.0 * c2;
  k.b1 = scale * 2.0 * c2 * (7.0 * eps - 2.0);
  k.b0 = scale * 2.0 * c2 * (2.0 * eps * eps - eps);
  k.e1 = scale * 6.0 * c1;
  k.e0 = scale * c1 * (6.0 * eps - 2.0);
  k.d2 = scale * 2.0 * c2 * (4.0 / eps);
  k.d1 = scale * 2.0 * c2 * (22.0 - 4.0 / eps);
  k.d0 = scale * 2.0 * c2 * (18.0 * eps - 6.0);
  k.inv_eps = 1.0 / eps;
  k.cut = 1.0 - 30.0 * eps;
  k.t_sat = 6.0 * eps * (1.0 + 1e-12);
  return k;
}

// In u = k (t^2 - 1) with t^2 - 1 = m:  Q = 2m^2 + (2 + 7 eps) m + 2 eps^2 + 6 eps,
// A = t (2 c1 m + 6 c1 eps), A' = c1 (6m + 4 + 6 eps),
// D = 2 c2 t ((4/eps) m^2 + (4/eps + 22) m + 16 + 18 eps).
inline LapCoef32 lap_coef32(double eps, double scale) {
  double c1, c2;
  lap_constants(eps, c1, c2);
  const double kk = 1.44269504088896340736 / eps, ik = 1.0 / kk;
  LapCoef32 k{};
  k.a1 = (float)(scale * 2.0 * c1 * ik);
  k.a0 = (float)(scale * 6.0 * c1 * eps);
  k.b2 = (float)(scale * 2.0 * c2 * 2.0 * ik * ik);
  k.b1 = (float)(scale * 2.0 * c2 * (2.0 + 7.0 * eps) * ik);
  k.b0 = (float)(scale * 2.0 * c2 * (2.0 * eps * eps + 6.0 * eps));
  k.e1 = (float)(scale * 6.0 * c1 * ik);
  k.e0 = (float)(scale * c1 * (4.0 + 6.0 * eps));
  k.d2 = (float)(scale * 2.0 * c2 * (4.0 / eps) * ik * ik);
  k.d1 = (float)(scale * 2.0 * c2 * (4.0 / eps + 22.0) * ik);
  k.d0 = (float)(scale * 2.0 * c2 * (16.0 + 18.0 * eps));
  k.k = (float)kk;
  k.inv_eps = (float)(1.0 / eps);
  k.cut = (float)(1.0 - 30.0 * eps);
  k.t_sat = (float)(6.0 * eps * (1.0 + 1e-6));
  return k;
}

}  // namespace sbslpm
