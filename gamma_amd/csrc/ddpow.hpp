// ddpow.hpp -- x^y, tanh and log10 from IEEE-exact operations only, identical
// on host and device.
//
// The hot path raises to non-integer powers in the synchrotron update, the
// shock detector, the Blandford-McKee inflow and the BM regrid rule.  A
// library pow is not the same function on the CPU and on the GPU: the C
// library rounds correctly in nearly all cases, the device library to within
// one unit in the last place, and they disagree on about a quarter of all
// arguments.  Kernels call gh_pow instead, which evaluates exp(y log x) in
// double-double arithmetic (Dekker 1971; Knuth TAOCP vol. 2, 4.2.2) with only
// +, -, *, /, fma, frexp, ldexp and rint -- operations IEEE 754 defines
// exactly -- so the host build and the device build return the same bits.
// The double-double result carries about 2^-80 relative error before the
// final rounding, so gh_pow is the correctly rounded x^y except when x^y lies
// within that distance of a rounding boundary.  Table-driven (Tang 1989,
// 1990), tables of 80-digit values in ddtables.h:
//
//   log x: x = m 2^e with m in [1, 2); log m = log c + log1p((m - c)/c) for
//          the nearest c = 1 + i/128, series to the 10th power;
//   exp z: z = (64 k + j) log(2)/64 + r, |r| <= log(2)/128; exp z =
//          2^k 2^(j/64) exp(r), exp(r) by its Taylor series to r^9/9!.
#pragma once
#include <math.h>

#include "ddtables.h"

#ifndef GH_HD
#define GH_HD __host__ __device__ __forceinline__
#endif

namespace gh {
namespace dd {

struct D2 { double hi, lo; };

GH_HD D2 two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}
GH_HD D2 fast_two_sum(double a, double b) {
  double s = a + b;
  return {s, b - (s - a)};
}
GH_HD D2 two_prod(double a, double b) {
  double p = a * b;
  return {p, fma(a, b, -p)};
}
GH_HD D2 add(D2 a, D2 b) {
  D2 s = two_sum(a.hi, b.hi);
  double e = s.lo + (a.lo + b.lo);
  return fast_two_sum(s.hi, e);
}
GH_HD D2 mul(D2 a, D2 b) {
  D2 p = two_prod(a.hi, b.hi);
  double e = p.lo + (a.hi * b.lo + a.lo * b.hi);
  return fast_two_sum(p.hi, e);
}
GH_HD D2 mul_d(D2 a, double b) {
  D2 p = two_prod(a.hi, b);
  double e = p.lo + a.lo * b;
  return fast_two_sum(p.hi, e);
}
GH_HD D2 div(D2 a, D2 b) {
  double q1 = a.hi / b.hi;
  D2 r = add(a, mul_d({-b.hi, -b.lo}, q1));
  double q2 = r.hi / b.hi;
  return fast_two_sum(q1, q2);
}

// natural logarithm of a positive double, as a double-double: x = m 2^e
// with m in [1, 2); c = 1 + i/128 the table point nearest m (|m - c| <= 2^-8,
// m - c exact); log x = e log 2 + log c + log1p(z), z = (m - c) (1/c),
// |z| <= 2^-8, by the series to z^10 with the terms through z^3 in
// double-double.  For x just below 1 (e = -1, c = 2) the first two terms
// cancel exactly.
GH_HD D2 log_dd(double x) {
  int e;
  double m = frexp(x, &e);                       // x = m 2^e, m in [0.5, 1)
  m *= 2.;
  e -= 1;
  int i = (int)rint((m - 1.) * 128.);            // 0 .. 128
  double c = (double)(128 + i) * 0x1p-7;
  const double* T = gh_logc[i];
  D2 z = mul_d({T[2], T[3]}, m - c);
  double zh = z.hi;
  double q = -1. / 10.;                          // -1/4 + z/5 - ... - z^6/10
  q = fma(q, zh, 1. / 9.); q = fma(q, zh, -1. / 8.); q = fma(q, zh, 1. / 7.);
  q = fma(q, zh, -1. / 6.); q = fma(q, zh, 1. / 5.); q = fma(q, zh, -1. / 4.);
  D2 z2 = mul(z, z);
  D2 z3 = mul(z2, z);
  D2 s = add(mul(z3, {0x1.5555555555555p-2, 0x1.5555555555555p-56}), {(z2.hi * z2.hi) * q, 0.});
  s = add({-0.5 * z2.hi, -0.5 * z2.lo}, s);
  s = add(z, s);
  return add(add(mul_d({gh_ln2[0], gh_ln2[1]}, (double)e), {T[0], T[1]}), s);
}

// exp of a double-double as E * 2^k (E a double-double in [1, 2.02)), for |z|
// within the range of doubles: z = (64 k + j) log(2)/64 + r, j in 0..63,
// |r| <= log(2)/128; exp z = 2^k 2^(j/64) exp(r), exp(r) by its Taylor series
// to r^9 with the terms through r^3 in double-double.
GH_HD D2 exp_dd2(D2 z, int* k2) {
  double kk = rint(z.hi * 0x1.71547652b82fep+6);   // 64 / log 2
  int n = (int)kk;
  int j = n & 63;
  D2 r = add(z, mul_d({-gh_ln2[0] * 0x1p-6, -gh_ln2[1] * 0x1p-6}, kk));
  double rh = r.hi;
  double q = 1. / 362880.;                           // 1/24 + r/120 + ... + r^5/9!
  q = fma(q, rh, 1. / 40320.); q = fma(q, rh, 1. / 5040.); q = fma(q, rh, 1. / 720.);
  q = fma(q, rh, 1. / 120.); q = fma(q, rh, 1. / 24.);
  D2 r2 = mul(r, r);
  D2 r3 = mul(r2, r);
  D2 s = add(mul(r3, {0x1.5555555555555p-3, 0x1.5555555555555p-57}), {(r2.hi * r2.hi) * q, 0.});
  s = add({0.5 * r2.hi, 0.5 * r2.lo}, s);
  s = add(r, s);
  s = add({1., 0.}, s);
  *k2 = (n - j) / 64;
  return mul({gh_exp2j[j][0], gh_exp2j[j][1]}, s);
}

// exp of a double-double, rounded to double
GH_HD double exp_dd(D2 z) {
  if (z.hi > 709.8) return HUGE_VAL;
  if (z.hi < -745.2) return 0.;
  int k;
  D2 E = exp_dd2(z, &k);
  return ldexp(E.hi, k);
}

// x^(3/4) and x^(3/2) for x in [2^-300, 2^300] -- the powers of the shock
// detector (1/G, G = 4/3) and of the BM regrid rule: a first value y0 from
// square roots (a few units in the last place off), then one Newton step on
// y^4 = x^3 (y^2 = x^3) with the residual formed in double-double.  y0 minus
// the step is within about 2^-100 of the power, so its rounding is the
// correctly rounded power except within that distance of a rounding
// boundary; about 30 operations against 150 for exp(y log x).
GH_HD double pow34(double x) {
  double s = sqrt(x), t = sqrt(s);
  double y0 = s * t;
  D2 y2 = two_prod(y0, y0);
  D2 y4 = mul(y2, y2);
  D2 x3 = mul_d(two_prod(x, x), x);
  D2 r = add(y4, {-x3.hi, -x3.lo});
  return y0 - r.hi / (4. * y0 * y2.hi);
}
GH_HD double pow32(double x) {
  double y0 = x * sqrt(x);
  D2 y2 = two_prod(y0, y0);
  D2 x3 = mul_d(two_prod(x, x), x);
  D2 r = add(y2, {-x3.hi, -x3.lo});
  return y0 - r.hi / (2. * y0);
}

}  // namespace dd

// x^y for the physics: x > 0 finite; x = 0 gives 0 (y > 0) or +inf (y < 0),
// y = 0 gives 1, x < 0 or NaN gives NaN (the C library's answers for
// non-integer y).
GH_HD double gh_pow(double x, double y) {
  if (y == 0.) return 1.;
  if (!(x >= 0.)) return (x - x) / (x - x);      // NaN
  if (x == 0.) return y > 0. ? 0. : HUGE_VAL;
  if (x == HUGE_VAL) return y > 0. ? HUGE_VAL : 0.;
  if (x >= 0x1p-300 && x <= 0x1p+300) {
    if (y == 0.75) return dd::pow34(x);
    if (y == 1.5) return dd::pow32(x);
  }
  return dd::exp_dd(dd::mul_d(dd::log_dd(x), y));
}

// tanh y = (e^{2|y|} - 1) / (e^{2|y|} + 1) with the sign of y, in double-double
// (no cancellation for |y| above 1e-8, below which tanh y rounds to y); the
// shock detector's tanh (radiation_sph.cpp:90, :29)
GH_HD double gh_tanh(double y) {
  if (!(y == y)) return y;
  double a = fabs(y);
  if (a < 1e-8) return y;
  if (a > 22.) return y > 0. ? 1. : -1.;     // 1 - tanh(22) < 2^-62
  int k;
  dd::D2 E = dd::exp_dd2({2. * a, 0.}, &k);
  dd::D2 Ek = {ldexp(E.hi, k), ldexp(E.lo, k)};
  dd::D2 t = dd::div(dd::add(Ek, {-1., 0.}), dd::add(Ek, {1., 0.}));
  return y > 0. ? t.hi : -t.hi;
}

// log10 x = log x / log 10 in double-double, for positive normal x (0 gives
// -inf, negative or NaN gives NaN, +inf gives +inf)
GH_HD double gh_log10(double x) {
  if (!(x > 0.)) return x == 0. ? -HUGE_VAL : (x - x) / (x - x);
  if (x == HUGE_VAL) return x;
  dd::D2 l = dd::log_dd(x);
  dd::D2 r = dd::div(l, {0x1.26bb1bbb55516p+1, -0x1.f48ad494ea3e9p-53});   // log 10
  return r.hi;
}

}  // namespace gh
