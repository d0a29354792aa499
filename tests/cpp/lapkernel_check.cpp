// Host check of the folded regularised-Laplacian kernel
// (spectral-bslpm_amd/csrc/lapkernel.hpp) against the kernel as the reference
// writes it (laplacian_impl.hpp:46-66, eps = 0.1), evaluated in long double:
//   K(t) = c1 2t(-1 + 3 eps + t^2)
//          + (1 + erf(t/eps)) e^{(t^2-1)/eps} c2 (2 eps^2 + 2t^2(t^2-1) + eps(7t^2-1)),
// and its derivative by central differences of that expression.  The FP64
// coefficients (powers of t) are checked in both forms (t >= 6 eps: A + E B and
// A' + E D; everywhere: A + f E B and A' + E (f D + g B/2)); the FP32 ones
// (powers of u = (t^2-1) log2(e)/eps, rounded to float) in double arithmetic.
// Prints the worst relative errors and fails above 1e-12 for the FP64 K, 1e-8 for
// the FP64 K' (the central difference's own error is ~1e-12), 1e-6 for FP32.
#include <cmath>
#include <cstdio>

#include "lapkernel.hpp"

using sbslpm::LapCoef32;
using sbslpm::LapCoef64;

static const long double kEps = 0.1L;

static long double kref(long double t) {
  const long double pi = 3.14159265358979323846264338327950288L;
  const long double c1 = expl(-1.0L / kEps) * powl(pi, -1.5L), c2 = 1.0L / (sqrtl(kEps) * pi);
  return c1 * 2.0L * t * (-1.0L + 3.0L * kEps + t * t) +
         (1.0L + erfl(t / kEps)) * expl((t * t - 1.0L) / kEps) * c2 *
             (2.0L * kEps * kEps + 2.0L * t * t * (t * t - 1.0L) + kEps * (7.0L * t * t - 1.0L));
}
static long double kdref(long double t) {
  const long double h = 1e-7L;
  return (kref(t + h) - kref(t - h)) / (2.0L * h);
}

static void eval64(const LapCoef64& k, double t, bool fast, double& K, double& Kd) {
  const double t2 = t * t;
  const double A = t * (k.a1 * t2 + k.a0), B = (k.b2 * t2 + k.b1) * t2 + k.b0;
  const double Ad = k.e1 * t2 + k.e0, D = t * ((k.d2 * t2 + k.d1) * t2 + k.d0);
  const double E = std::exp((t2 - 1.0) * k.inv_eps);
  if (fast) {
    K = A + E * B;
    Kd = Ad + E * D;
    return;
  }
  const double s = t * k.inv_eps;
  const double f = 0.5 * (1.0 + std::erf(s));
  const double g = 2.0 / (std::sqrt(3.14159265358979323846) * 0.1) * std::exp(-s * s);
  K = A + f * E * B;
  Kd = Ad + E * (f * D + 0.5 * g * B);
}

static void eval32(const LapCoef32& k, double t, bool fast, double& K, double& Kd) {
  const double u = (double)k.k * (t * t - 1.0);
  const double A = t * ((double)k.a1 * u + k.a0), B = ((double)k.b2 * u + k.b1) * u + k.b0;
  const double Ad = (double)k.e1 * u + k.e0, D = t * (((double)k.d2 * u + k.d1) * u + k.d0);
  const double E = std::exp2(u);
  if (fast) {
    K = A + E * B;
    Kd = Ad + E * D;
    return;
  }
  const double s = t * (double)k.inv_eps;
  const double f = 0.5 * (1.0 + std::erf(s));
  const double g = 2.0 / (std::sqrt(3.14159265358979323846) * 0.1) * std::exp(-s * s);
  K = A + f * E * B;
  Kd = Ad + E * (f * D + 0.5 * g * B);
}

int main() {
  const LapCoef64 c64 = sbslpm::lap_coef64(0.1, 1.0);
  const LapCoef32 c32 = sbslpm::lap_coef32(0.1, 1.0);
  // scales: the largest |K| and |K'| over [-1, 1]
  long double kmax = 0, kdmax = 0;
  for (int i = -2000; i <= 2000; ++i) {
    const long double t = i / 2000.0L;
    kmax = fmaxl(kmax, fabsl(kref(t)));
    kdmax = fmaxl(kdmax, fabsl(kdref(t)));
  }
  double w64 = 0, w64d = 0, w32 = 0, w32d = 0;
  for (int i = -20000; i <= 20000; ++i) {
    const double t = i / 20000.0;
    const long double K = kref(t), Kd = kdref(t);
    double a, b;
    const bool sat = t >= c64.t_sat;
    for (int form = 0; form < 2; ++form) {
      if (form == 1 && !sat) continue;  // the fast form holds only where erf has saturated
      eval64(c64, t, form == 1, a, b);
      w64 = std::fmax(w64, (double)(fabsl(a - K) / kmax));
      w64d = std::fmax(w64d, (double)(fabsl(b - Kd) / kdmax));
      eval32(c32, t, form == 1, a, b);
      w32 = std::fmax(w32, (double)(fabsl(a - K) / kmax));
      w32d = std::fmax(w32d, (double)(fabsl(b - Kd) / kdmax));
    }
  }
  std::printf("FP64 form: K %.2e, K' %.2e; FP32 form: K %.2e, K' %.2e (relative to max |K|, |K'|)\n", w64, w64d,
              w32, w32d);
  return (w64 <= 1e-12 && w64d <= 1e-8 && w32 <= 1e-6 && w32d <= 1e-6) ? 0 : 1;
}
