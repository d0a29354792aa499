// fastlog.hpp -- natural logarithm of a positive normal double, for the FP64
// log pair kernel (about half the instructions of the library routine).
//
// x = 2^e m with m in [sqrt(1/2), sqrt(2)).  log m = 2 atanh(s) with
// s = (m - 1)/(m + 1), |s| <= 3 - 2 sqrt(2) = 0.1716, summed by the odd series
//   2 (s + s^3/3 + s^5/5 + ... + s^21/21),
// whose first omitted term is below 1e-18 of the result.  e ln 2 is added with
// ln 2 split in two (the Cody-Waite reduction): ln2_hi holds 32 significant bits,
// so e * ln2_hi is exact for every exponent of a double.  m - 1 is exact; the
// quotient comes from the reciprocal the caller supplies (the hardware one on
// the device), two Newton steps and one residual correction.
#pragma once
#include <cmath>

#ifndef SBSLPM_HD
#define SBSLPM_HD __host__ __device__
#endif

namespace sbslpm {

constexpr double kLn2Hi = 0x1.62e42fee00000p-1;  // ln 2 rounded to 32 significant bits
constexpr double kLn2Lo = 0x1.a39ef35793c76p-33;  // ln 2 - kLn2Hi

template <class Rcp>
SBSLPM_HD inline double log_pos(double x, Rcp rcp) {
  int e;
  double m = std::frexp(x, &e);  // [0.5, 1)
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    e -= 1;
  }
  const double num = m - 1.0, den = m + 1.0;
  double r = rcp(den);
  r = std::fma(r, std::fma(-den, r, 1.0), r);
  r = std::fma(r, std::fma(-den, r, 1.0), r);
  double s = num * r;
  s = std::fma(r, std::fma(-s, den, num), s);
  const double z = s * s;
  double p = 1.0 / 21;
  p = std::fma(p, z, 1.0 / 19);
  p = std::fma(p, z, 1.0 / 17);
  p = std::fma(p, z, 1.0 / 15);
  p = std::fma(p, z, 1.0 / 13);
  p = std::fma(p, z, 1.0 / 11);
  p = std::fma(p, z, 1.0 / 9);
  p = std::fma(p, z, 1.0 / 7);
  p = std::fma(p, z, 1.0 / 5);
  p = std::fma(p, z, 1.0 / 3);
  const double s2 = 2.0 * s;
  const double lm = std::fma(s2 * z, p, s2);
  const double ed = (double)e;
  return std::fma(ed, kLn2Hi, std::fma(ed, kLn2Lo, lm));
}

}  // namespace sbslpm
