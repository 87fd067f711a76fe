// Correctly rounded float64 power (the ideal-gas equation of state) and exponential (the
// exponential material coupling).
//
// pyfds evaluates p = rho0 c^2/gamma * (((rho0 + rho')/rho0)**gamma - 1) with numpy's power
// (pyfds/nonlinear_acoustics.py:82-87).  The subtraction turns a one-ulp difference in the
// power into a relative error of about 1/(gamma |rho'/rho0|) in p, so the power has to be the
// exact one: pow_cr returns x**y rounded to nearest-even, whatever libm a host would use.
//
// Method: x**y = exp(y ln x) in double-double arithmetic (Dekker 1971; error-free sums after
// Knuth, error-free products by fused multiply-add), with table-driven argument reduction for
// both functions after Tang ("Table-driven implementation of the exponential function", ACM
// TOMS 15, 1989; "... of the logarithm function", ACM TOMS 16, 1990):
//   * ln x = e ln2 - ln c_j + log1p(r),  x = 2^e m, m in [0.75, 1.5), c_j ~ 1/m from a
//     128-entry table, r = m c_j - 1 formed exactly as a double-double (|r| < 2^-7.9),
//     log1p(r) by its Taylor series to degree 14;
//   * exp z = 2^(k/64) exp(t),  k = round(64 z/ln2), t = z - k ln2/64 with ln2/64 in three
//     parts (exact products and an exact leading difference, |t| < 2^-7.5), exp(t) by its
//     Taylor series to degree 11;
//   * series terms whose size is below 2^-55 of the leading term are summed in double, the
//     others in double-double; the tables (fds_crmath_tables.h, gen_crmath_tables.py) hold
//     every constant to 2^-107.
// The double-double value of x**y carries a relative error below 2^-100, and its rounding to
// the nearest double (the last normalising sum) is the correctly rounded result unless x**y
// lies within 2^-100 relative of a midpoint between two doubles: about one input in 2^47, and
// for the non-integer exponents of this use never exactly (an exact midpoint needs a y with
// few significant bits).  Results below 2^-1022 (subnormal) are rounded a second time by the
// final scaling.  IEEE special cases (zeros, infinities, NaN, negative bases with integer y)
// follow C99 Annex F.
//
// The header also compiles as plain C++ (no HIP), so the CPU test suite can check the same
// source against an independent correctly rounded power on millions of inputs.
#pragma once

#include <stdint.h>
#include <math.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define FDS_CRM_FN __device__ __forceinline__
#define FDS_CRM_CONST static __constant__ const
#else
#define FDS_CRM_FN static inline
#define FDS_CRM_CONST static const
#endif

namespace fds {
namespace crm {

#include "fds_crmath_tables.h"

struct dd {
  double hi, lo;
};

// error-free transformations
FDS_CRM_FN dd two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  const double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}
FDS_CRM_FN dd fast_two_sum(double a, double b) {  // |a| >= |b| or a == 0
  const double s = a + b;
  const double e = b - (s - a);
  return {s, e};
}
FDS_CRM_FN dd two_prod(double a, double b) {
  const double p = a * b;
  return {p, fma(a, b, -p)};
}

// double-double arithmetic (relative errors about 2^-104)
FDS_CRM_FN dd add(dd a, dd b) {
  dd s = two_sum(a.hi, b.hi);
  const dd t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return fast_two_sum(s.hi, s.lo);
}
FDS_CRM_FN dd mul(dd a, dd b) {
  dd p = two_prod(a.hi, b.hi);
  p.lo += a.hi * b.lo + a.lo * b.hi;
  return fast_two_sum(p.hi, p.lo);
}
FDS_CRM_FN dd mul_d(dd a, double b) {
  dd p = two_prod(a.hi, b);
  p.lo += a.lo * b;
  return fast_two_sum(p.hi, p.lo);
}

FDS_CRM_FN uint64_t bits_of(double x) {
  uint64_t u;
  __builtin_memcpy(&u, &x, sizeof u);
  return u;
}
FDS_CRM_FN double from_bits(uint64_t u) {
  double x;
  __builtin_memcpy(&x, &u, sizeof x);
  return x;
}

// ln x for finite x > 0, as a double-double
FDS_CRM_FN dd log_dd(double x) {
  int e = 0;
  uint64_t ix = bits_of(x);
  if ((ix >> 52) == 0) {  // subnormal: scale into the normal range
    ix = bits_of(x * 0x1p54);
    e = -54;
  }
  e += (int)(ix >> 52) - 1023;
  double m = from_bits((ix & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);  // [1, 2)
  if (m >= 1.5) {
    m *= 0.5;
    e += 1;
  }
  // bucket of m in [0.75, 1.5); a bucket edge may go either way, the bound on r allows it
  int j = (int)((m - 0.75) * (kLogBuckets / 0.75));
  j = j < 0 ? 0 : (j >= kLogBuckets ? kLogBuckets - 1 : j);
  const double c = kLogC[j];
  dd r;
  if (j == kLogOneBucket) {
    r = {m - 1.0, 0.0};  // exact (m within a factor 2 of 1)
  } else {
    const dd p = two_prod(m, c);  // m c exactly
    r = two_sum(p.hi - 1.0, p.lo);  // p.hi - 1 exact: p.hi in [0.5, 2]
  }
  // log1p(r) = sum_k (-1)^(k+1) r^k / k: degrees 8..14 in double, 1..7 in double-double
  double s = kLog1pHi[kLog1pDegree - 1];
  for (int k = kLog1pDegree - 1; k >= 8; --k) s = s * r.hi + kLog1pHi[k - 1];
  dd q = {s, 0.0};
  for (int k = 7; k >= 1; --k) q = add(mul(q, r), dd{kLog1pHi[k - 1], kLog1pLo[k - 1]});
  q = mul(q, r);
  // e ln2 + (-ln c) + log1p(r)
  dd el = two_prod((double)e, kLn2Hi);
  el.lo += (double)e * kLn2Lo;
  el = fast_two_sum(el.hi, el.lo);
  return add(add(el, dd{kLogHi[j], kLogLo[j]}), q);
}

// exp z for a double-double z, rounded to the nearest double
FDS_CRM_FN double exp_dd(dd z) {
  if (z.hi > 710.0) return HUGE_VAL;
  if (z.hi < -746.0) return 0.0;
  const double kf = rint(z.hi * kInvL64);
  const int k = (int)kf;
  // t = z - k ln2/64: k kL64a is exact and the difference with z.hi is exact (Sterbenz)
  const double a = z.hi - kf * kL64a;
  const dd p2 = two_prod(kf, kL64b);
  dd t = two_sum(a, -p2.hi);
  t.lo += (z.lo - p2.lo) - kf * kL64c;
  t = fast_two_sum(t.hi, t.lo);
  // exp(t) = sum_k t^k / k!: degrees 7..11 in double, 0..6 in double-double
  double s = kExpCHi[kExpDegree];
  for (int i = kExpDegree - 1; i >= 7; --i) s = s * t.hi + kExpCHi[i];
  dd q = {s, 0.0};
  for (int i = 6; i >= 0; --i) q = add(mul(q, t), dd{kExpCHi[i], kExpCLo[i]});
  const int j = k & (kExpTable - 1);
  const int n = (k - j) / kExpTable;
  const dd r = mul(dd{kExpHi[j], kExpLo[j]}, q);  // r.hi = r.hi + r.lo rounded to nearest
  return ldexp(r.hi, n);
}

// e**z rounded to nearest (the same exp; infinities and NaN follow C99 Annex F)
FDS_CRM_FN double exp_cr(double z) {
  if (z != z) return z;
  return exp_dd(dd{z, 0.0});
}

// x**y rounded to nearest (see the header comment for the method and its bound)
FDS_CRM_FN double pow_cr(double x, double y) {
  if (y == 0.0 || x == 1.0) return 1.0;
  if (x != x || y != y) return x + y;
  const double ax = fabs(x), ay = fabs(y);
  if (ay == HUGE_VAL) {
    if (ax == 1.0) return 1.0;
    return ((ax < 1.0) == (y < 0.0)) ? HUGE_VAL : 0.0;
  }
  // y finite and nonzero: 0 = not an integer, 1 = odd integer, 2 = even integer
  int yint = 0;
  if (floor(y) == y) yint = (ay < 0x1p53 && fmod(ay, 2.0) == 1.0) ? 1 : 2;
  const bool negx = signbit(x);
  if (negx && yint == 0 && ax != HUGE_VAL && ax != 0.0) return (x - x) / (x - x);  // NaN
  const double sign = (negx && yint == 1) ? -1.0 : 1.0;
  if (ax == 0.0) return y < 0.0 ? sign * HUGE_VAL : sign * 0.0;
  if (ax == HUGE_VAL) return y < 0.0 ? sign * 0.0 : sign * HUGE_VAL;
  return sign * exp_dd(mul_d(log_dd(ax), y));
}

}  // namespace crm
}  // namespace fds
