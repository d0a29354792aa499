// fasttab.hpp -- table-driven FP64 logarithm and exponential for the FP64 pair
// kernels (log pairs of the Poisson solve, exponentials of the regularised
// Laplacian), each read from a table the kernel stages in LDS.
//
// Logarithm (Tang's method with accurate table values): x = 2^k m with
// m in [sqrt(1/2), sqrt(2)), taken from the bits of x (an add carries m past
// sqrt(2) into the exponent).  Bits 20..11 of m's high word (the lowest
// exponent bit and 9 mantissa bits) pick one of 1024 cells, of width 1/1024
// below 1 and 1/512 above.  Each cell stores a double c_inv near the
// reciprocal of its centre, chosen so that -log(c_inv) lies within 2^-59 of a
// double, and that double, logc (the two cells next to 1 store c_inv = 1,
// logc = 0, so that log is accurate in relative terms near 1).  Then
//   r = m c_inv - 1 (one fma, |r| < 2^-9),
//   log x = k ln2 + logc + log1p(r),  log1p(r) = r + r^2 q(r),
// q the degree-4 Taylor polynomial (first omitted term r^7/7 < 2^-65), with
// k ln2 split as k ln2_hi (exact) + k ln2_lo, and k converted to a double by
// placing it in the low bits of 2^52 (one subtraction).
//
// Exponential (general, exp_tab): n = rint(t 128/ln2), t = n ln2/128 + r with
// |r| <= ln2/256 (Cody-Waite split of ln2/128),
// exp t = 2^(n>>7) 2^((n&127)/128) (1 + p(r)), p(r) = r + r^2 (1/2 + r/6 +
// r^2/24 + r^3/120) (first omitted term r^6/720 < 2^-60), 2^(j/128) from a
// 128-entry table, the power of two by ldexp.  For |t| < 700.
//
// Exponential on [-10.375, 0] (exp_neg): n = round(t 256/ln2) by the 1.5 2^52
// shifter (n lands in the low word of the sum, and the sum minus the shifter
// is n as a double), t = n ln2/256 + r, |r| <= ln2/512, and
// exp t = 2^(n/256) (1 + p(r)) with 2^(n/256) read directly from a table over
// every n in [-3839, 0] (no ldexp, no masking) and p(r) = r + r^2 (1/2 + r/6 +
// r^2/24) (first omitted term r^5/120 < 2^-56).  The argument is clamped to
// the table's range (a NaN becomes its lower end).
//
// The tables are built on the host in long double (build_log_table,
// build_exp_table, build_exp_neg_table) and uploaded with the context's
// constant tables.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

#ifndef SBSLPM_HD
#define SBSLPM_HD __host__ __device__
#endif

namespace sbslpm {

constexpr int kLogTabSize = 1024;
constexpr int kExpTabSize = 128;
constexpr int kExpNegTabSize = 3840;  // 2^(n/256), n = -3839..0
constexpr double kExpNegMin = -10.375;  // exp_neg clamps its argument here (n >= -3832)

struct LogTabEntry {
  double inv_c, logc;
};

constexpr double kTabLn2Hi = 0x1.62e42fee00000p-1;  // ln 2 to 32 significant bits
constexpr double kTabLn2Lo = 0x1.a39ef35793c76p-33;  // ln 2 - kTabLn2Hi
constexpr double kExpInvL = 0x1.71547652b82fep+7;     // 128 / ln 2
constexpr double kExpLHi = 0x1.62e42fee00000p-8;      // kTabLn2Hi / 128 (n kExpLHi exact for |n| < 2^21)
constexpr double kExpLLo = 0x1.a39ef35793c76p-40;     // kTabLn2Lo / 128
constexpr double kExpNegInvL = 0x1.71547652b82fep+8;  // 256 / ln 2
constexpr double kExpNegLHi = 0x1.62e42fee00000p-9;   // kTabLn2Hi / 256 (n kExpNegLHi exact for |n| < 2^21)
constexpr double kExpNegLLo = 0x1.a39ef35793c76p-41;  // kTabLn2Lo / 256
constexpr double kShifter = 0x1.8p52;                 // 1.5 2^52: adding it rounds to an integer in the low word

SBSLPM_HD inline double bits_to_double(uint64_t b) {
  double d;
  __builtin_memcpy(&d, &b, sizeof d);
  return d;
}
SBSLPM_HD inline uint64_t double_to_bits(double d) {
  uint64_t b;
  __builtin_memcpy(&b, &d, sizeof b);
  return b;
}

// log x for a positive normal double x.  c6 is -1/6, the leading coefficient
// of q: a kernel passes it from a vector register, so that the first Horner
// step is one three-operand fma (with both of its constants in scalar
// registers the instruction set allows only one, and the compiler adds a copy
// per evaluation).
SBSLPM_HD inline double log_tab(double x, const LogTabEntry* __restrict__ tab, double c6 = -1.0 / 6.0) {
  const uint64_t b = double_to_bits(x);
  const uint32_t h = (uint32_t)(b >> 32) + (0x3ff00000u - 0x3fe6a09eu);
  // k + 1023 in the low bits of 2^52, minus 2^52 + 1023: k as a double
  const double kd = bits_to_double(0x4330000000000000ull | (h >> 20)) - (0x1p52 + 1023.0);
  const uint32_t hm = (h & 0x000fffffu) + 0x3fe6a09eu;
  const double m = bits_to_double(((uint64_t)hm << 32) | (b & 0xffffffffu));
  const LogTabEntry e = tab[(hm >> 11) & (kLogTabSize - 1)];
  const double r = fma(m, e.inv_c, -1.0);
  // q(r) = -1/2 + r/3 - r^2/4 + r^3/5 - r^4/6
  double q = fma(r, c6, 1.0 / 5.0);
  q = fma(q, r, -1.0 / 4.0);
  q = fma(q, r, 1.0 / 3.0);
  q = fma(q, r, -1.0 / 2.0);
  const double y = fma(r * r, q, r);
  return fma(kd, kTabLn2Hi, e.logc) + fma(kd, kTabLn2Lo, y);
}

// exp t for |t| < 700.
SBSLPM_HD inline double exp_tab(double t, const double* __restrict__ tab) {
  const double n = rint(t * kExpInvL);
  const int k = (int)n;
  double r = fma(-n, kExpLHi, t);
  r = fma(-n, kExpLLo, r);
  double q = fma(r, 1.0 / 120.0, 1.0 / 24.0);
  q = fma(q, r, 1.0 / 6.0);
  q = fma(q, r, 0.5);
  const double p = fma(r * r, q, r);
  const double s = tab[k & (kExpTabSize - 1)];
  return ldexp(fma(s, p, s), k >> 7);
}

// exp t for t in [kExpNegMin, 2^-40] (arguments below are clamped to
// kExpNegMin).  c24 is 1/24, the leading coefficient of p (passed as c6 is to
// log_tab).
SBSLPM_HD inline double exp_neg(double t, const double* __restrict__ tab, double c24 = 1.0 / 24.0) {
  t = fmax(t, kExpNegMin);
  const double z = fma(t, kExpNegInvL, kShifter);
  const double n = z - kShifter;
  const double* __restrict__ top = tab + (kExpNegTabSize - 1);  // 2^0; n <= 0 indexes below it
  double r = fma(-n, kExpNegLHi, t);
  r = fma(-n, kExpNegLLo, r);
  double q = fma(r, c24, 1.0 / 6.0);
  q = fma(q, r, 0.5);
  const double p = fma(r * r, q, r);
  const double s = top[(int)(uint32_t)double_to_bits(z)];
  return fma(s, p, s);
}

// Host: the tables.
inline void build_log_table(LogTabEntry* tab) {
  constexpr int half = kLogTabSize / 2;
  for (int i = 0; i < kLogTabSize; ++i) {
    // the cell's range of m: exponent 0x3fe (i < half, m in [0.5, 1)) or 0x3ff
    const long double w = i < half ? 0.5L / half : 1.0L / half;
    const long double lo = i < half ? 0.5L + (i % half) * w : 1.0L + (i % half) * w;
    if (i == half - 1 || i == half) {
      tab[i] = LogTabEntry{1.0, 0.0};
      continue;
    }
    const double base = (double)(1.0L / (lo + 0.5L * w));
    double best = base, best_err = 1e300;
    uint64_t bb = double_to_bits(base);
    for (int j = 0; j < 4096; ++j) {
      const int64_t off = (j & 1) ? -(int64_t)((j + 1) / 2) : (int64_t)(j / 2);
      const double c = bits_to_double(bb + (uint64_t)off);
      const long double L = -logl((long double)c);
      const double d = (double)L;
      const long double ulp = (long double)(std::nextafter(std::fabs(d), INFINITY) - std::fabs(d));
      const long double err = fabsl(L - (long double)d) / ulp;
      if (err < best_err) {
        best_err = (double)err;
        best = c;
      }
      if (err < 1.0L / 64) break;
    }
    tab[i] = LogTabEntry{best, (double)(-logl((long double)best))};
  }
}

inline void build_exp_table(double* tab) {
  for (int j = 0; j < kExpTabSize; ++j) tab[j] = (double)exp2l(j / 128.0L);
}

// tab[j] = 2^((j - (kExpNegTabSize - 1)) / 256)
inline void build_exp_neg_table(double* tab) {
  for (int j = 0; j < kExpNegTabSize; ++j) tab[j] = (double)exp2l((j - (kExpNegTabSize - 1)) / 256.0L);
}

}  // namespace sbslpm
