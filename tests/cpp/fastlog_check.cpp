// Host check of log_pos (spectral-bslpm_amd/csrc/fastlog.hpp) against the
// long-double logarithm: 2.6 million points over [1e-15, 2] (the range of
// 1 - x.y in the log pair kernel), near 1 and near 2; prints the worst error
// in units in the last place and fails above 2.5.
#include <cstdio>
#include <cmath>
#include <random>

#include "fastlog.hpp"
int main() {
  std::mt19937_64 g(1);
  std::uniform_real_distribution<double> u(-15.0, 0.30103);
  double worst = 0; long n = 0;
  auto rcp = [](double d) { return 1.0 / d; };
  auto check = [&](double x) {
    long double ref = logl((long double)x);
    double got = sbslpm::log_pos(x, rcp);
    double ulp = std::nextafter(std::fabs((double)ref), INFINITY) - std::fabs((double)ref);
    double err = std::fabs((double)((long double)got - ref)) / ulp;
    if (err > worst) worst = err;
    ++n;
  };
  for (int i = 0; i < 2000000; ++i) check(std::pow(10.0, u(g)));
  for (int i = 1; i < 200000; ++i) { check(1.0 + i * 1e-12); check(1.0 - i * 1e-12); check(2.0 - i * 1e-15); }
  check(1e-300); check(2.0); check(0.7071067811865476); check(0.7071067811865475);
  std::printf("%ld points, worst %.3f ulp\n", n, worst);
  return worst <= 2.5 ? 0 : 1;
}
