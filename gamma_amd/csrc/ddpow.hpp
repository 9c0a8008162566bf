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
// The double-double result carries about 2^-76 relative error before the
// final rounding, so gh_pow is the correctly rounded x^y except when x^y lies
// within that distance of a rounding boundary.
//
//   log x: x = m 2^e with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(t),
//          t = (m-1)/(m+1), |t| <= 0.172, series in u = t^2 to u^15
//          (first four terms in double-double, the rest in double);
//   exp z: z = k log 2 + r, |r| <= log(2)/2; exp(r) = exp(r/256)^256, the
//          inner exponential by its Taylor series to r^9/9!.
#pragma once
#include <math.h>

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

// natural logarithm of a positive normal double, as a double-double
GH_HD D2 log_dd(double x) {
  int e;
  double m = frexp(x, &e);                       // x = m 2^e, m in [0.5, 1)
  if (m < 0x1.6a09e667f3bcdp-1) { m *= 2.; e -= 1; }
  D2 t = div({m - 1., 0.}, two_sum(m, 1.));      // m - 1 is exact (Sterbenz)
  D2 u = mul(t, t);
  double uh = u.hi;
  double r = 1. / 31.;
  r = 1. / 29. + uh * r; r = 1. / 27. + uh * r; r = 1. / 25. + uh * r; r = 1. / 23. + uh * r;
  r = 1. / 21. + uh * r; r = 1. / 19. + uh * r; r = 1. / 17. + uh * r; r = 1. / 15. + uh * r;
  r = 1. / 13. + uh * r; r = 1. / 11. + uh * r; r = 1. / 9. + uh * r;
  D2 s = add({0x1.2492492492492p-3, 0x1.2492492492492p-57}, mul_d(u, r));     // 1/7 + u r
  s = add({0x1.999999999999ap-3, -0x1.999999999999ap-57}, mul(u, s));          // 1/5 + u s
  s = add({0x1.5555555555555p-2, 0x1.5555555555555p-56}, mul(u, s));           // 1/3 + u s
  s = add({1., 0.}, mul(u, s));
  D2 lm = mul(mul_d(t, 2.), s);
  D2 ln2 = {0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56};
  return add(mul_d(ln2, (double)e), lm);
}

// exp of a double-double as E * 2^k (E a double-double in [0.7, 1.5)),
// for |z| within the range of doubles
GH_HD D2 exp_dd2(D2 z, int* k2) {
  D2 ln2 = {0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56};
  double k = rint(z.hi * 0x1.71547652b82fep+0);
  D2 r = add(z, mul_d({-ln2.hi, -ln2.lo}, k));
  r.hi *= 0x1p-8; r.lo *= 0x1p-8;
  double rh = r.hi;
  double q = 1. / 40320.;
  q = 1. / 5040. + rh * q; q = 1. / 720. + rh * q; q = 1. / 120. + rh * q; q = 1. / 24. + rh * q;
  D2 E = add({0x1.5555555555555p-3, 0x1.5555555555555p-57}, mul_d(r, q));      // 1/6 + r q
  E = add({0.5, 0.}, mul(r, E));
  E = add({1., 0.}, mul(r, E));
  E = add({1., 0.}, mul(r, E));
  for (int n = 0; n < 8; ++n) E = mul(E, E);
  *k2 = (int)k;
  return E;
}

// exp of a double-double, rounded to double
GH_HD double exp_dd(D2 z) {
  if (z.hi > 709.8) return HUGE_VAL;
  if (z.hi < -745.2) return 0.;
  int k;
  D2 E = exp_dd2(z, &k);
  return ldexp(E.hi, k);
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
