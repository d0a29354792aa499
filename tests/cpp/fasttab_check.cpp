// Host check of log_tab, exp_tab and exp_neg
// (spectral-bslpm_amd/csrc/fasttab.hpp) against the long-double logarithm and
// exponential: log over [1e-15, 2] (the range of 1 - x.y in the log pair
// kernel), near 1 and near 2, at every cell boundary and over [1e-300, 1e300];
// exp_tab over [-10, 0] (the range of (t^2 - 1)/eps in the regularised-
// Laplacian kernels) and [-40, 40]; exp_neg over [-10.375, 0] and at every
// table point.  Prints the worst errors in units in the last place and fails
// above 2.
#include <cstdio>
#include <cmath>
#include <random>
#include <vector>

#include "fasttab.hpp"

int main() {
  std::vector<sbslpm::LogTabEntry> lt(sbslpm::kLogTabSize);
  std::vector<double> et(sbslpm::kExpTabSize), nt(sbslpm::kExpNegTabSize);
  sbslpm::build_log_table(lt.data());
  sbslpm::build_exp_table(et.data());
  sbslpm::build_exp_neg_table(nt.data());
  std::mt19937_64 g(1);
  double wlog = 0, wexp = 0, wneg = 0;
  long n = 0;
  auto ulp_err = [](double got, long double ref) {
    double a = std::fabs((double)ref);
    double ulp = std::nextafter(a, INFINITY) - a;
    return std::fabs((double)((long double)got - ref)) / ulp;
  };
  auto clog = [&](double x) {
    double e = ulp_err(sbslpm::log_tab(x, lt.data()), logl((long double)x));
    if (e > wlog) wlog = e;
    ++n;
  };
  auto cexp = [&](double t) {
    double e = ulp_err(sbslpm::exp_tab(t, et.data()), expl((long double)t));
    if (e > wexp) wexp = e;
    ++n;
  };
  auto cneg = [&](double t) {
    double e = ulp_err(sbslpm::exp_neg(t, nt.data()), expl((long double)t));
    if (e > wneg) wneg = e;
    ++n;
  };
  std::uniform_real_distribution<double> u(-15.0, 0.30103);
  for (int i = 0; i < 3000000; ++i) clog(std::pow(10.0, u(g)));
  for (int i = 1; i < 200000; ++i) { clog(1.0 + i * 1e-12); clog(1.0 - i * 1e-12); clog(2.0 - i * 1e-15); }
  for (int i = 0; i < 100000; ++i) { clog(1.0 + i * 4e-8); clog(1.0 - i * 4e-8); }
  clog(1e-300); clog(2.0); clog(0.7071067811865476); clog(0.7071067811865475); clog(1.0);
  // both sides of every cell boundary of m in [sqrt(1/2), sqrt(2)), at several exponents
  for (int i = 1; i < sbslpm::kLogTabSize; ++i) {
    const int half = sbslpm::kLogTabSize / 2;
    const double m = i < half ? 0.5 + i * (0.5 / half) : 1.0 + (i - half) * (1.0 / half);
    if (m < 0.70710678118654746 || m > 1.4142135623730951) continue;
    for (double sc : {1.0, 0.5, 2.0, 0x1p-20, 0x1p-45})
      for (double v : {m, std::nextafter(m, 0.0), std::nextafter(m, 4.0)}) clog(v * sc);
  }
  std::uniform_real_distribution<double> wide(-300.0, 300.0);
  for (int i = 0; i < 200000; ++i) clog(std::pow(10.0, wide(g)));
  std::uniform_real_distribution<double> v(-10.0, 0.0), w(-40.0, 40.0);
  for (int i = 0; i < 3000000; ++i) cexp(v(g));
  for (int i = 0; i < 1000000; ++i) cexp(w(g));
  for (int i = -100000; i <= 100000; ++i) cexp(i * 1e-9);
  cexp(0.0); cexp(-10.0);
  std::uniform_real_distribution<double> vn(sbslpm::kExpNegMin, 0.0);
  for (int i = 0; i < 3000000; ++i) cneg(vn(g));
  for (int i = 0; i <= 100000; ++i) cneg(-i * 1e-9);
  for (int j = 0; j < sbslpm::kExpNegTabSize; ++j) {
    const double t = (j - (sbslpm::kExpNegTabSize - 1)) * 0.69314718055994531 / 256.0;
    if (t < sbslpm::kExpNegMin) continue;
    for (double d : {0.0, 1.3e-3, -1.3e-3}) if (t + d <= 0.0 && t + d >= sbslpm::kExpNegMin) cneg(t + d);
  }
  cneg(0.0); cneg(-10.0); cneg(sbslpm::kExpNegMin);
  std::printf("%ld points, log worst %.3f ulp, exp_tab worst %.3f ulp, exp_neg worst %.3f ulp\n", n, wlog, wexp,
              wneg);
  return (wlog <= 2.0 && wexp <= 2.0 && wneg <= 2.0) ? 0 : 1;
}
