/* og_pow.h -- x^y used by the oracle in place of the C library's pow.
 *
 * TEST INFRASTRUCTURE.  The reference calls pow() (glibc).  The device has no
 * glibc: its pow rounds differently in about a quarter of all cases.  Both
 * sides therefore evaluate powers with the same double-double algorithm
 * written with IEEE-exact operations only (+, -, *, /, fma, frexp, ldexp,
 * rint): x^y = exp(y log x), log and exp table-driven (Tang 1989, 1990) on
 * the 80-digit double-double tables of og_tables.h (Dekker 1971 double-double
 * arithmetic).  This file restates that algorithm in C,
 * operation for operation; tests/test_oracle_pins.py checks it against the C
 * library's pow (equal in all but about 1 case in 1200, never more than one
 * unit in the last place apart) and against correctly rounded values.
 */
#ifndef OG_POW_H
#define OG_POW_H
#include <math.h>
#include "og_tables.h"

typedef struct { double hi, lo; } og_d2;

static inline og_d2 og_two_sum(double a, double b) {
  og_d2 r; double s = a + b; double bb = s - a;
  r.hi = s; r.lo = (a - (s - bb)) + (b - bb); return r;
}
static inline og_d2 og_fast_two_sum(double a, double b) {
  og_d2 r; double s = a + b; r.hi = s; r.lo = b - (s - a); return r;
}
static inline og_d2 og_two_prod(double a, double b) {
  og_d2 r; double p = a * b; r.hi = p; r.lo = fma(a, b, -p); return r;
}
static inline og_d2 og_dd_add(og_d2 a, og_d2 b) {
  og_d2 s = og_two_sum(a.hi, b.hi);
  double e = s.lo + (a.lo + b.lo);
  return og_fast_two_sum(s.hi, e);
}
static inline og_d2 og_dd_mul(og_d2 a, og_d2 b) {
  og_d2 p = og_two_prod(a.hi, b.hi);
  double e = p.lo + (a.hi * b.lo + a.lo * b.hi);
  return og_fast_two_sum(p.hi, e);
}
static inline og_d2 og_dd_mul_d(og_d2 a, double b) {
  og_d2 p = og_two_prod(a.hi, b);
  double e = p.lo + a.lo * b;
  return og_fast_two_sum(p.hi, e);
}
static inline og_d2 og_mk(double hi, double lo) { og_d2 r; r.hi = hi; r.lo = lo; return r; }
static inline og_d2 og_dd_div(og_d2 a, og_d2 b) {
  double q1 = a.hi / b.hi;
  og_d2 r = og_dd_add(a, og_dd_mul_d(og_mk(-b.hi, -b.lo), q1));
  double q2 = r.hi / b.hi;
  return og_fast_two_sum(q1, q2);
}
/* log x = e log 2 + log c + log1p(z): x = m 2^e, m in [1, 2), c = 1 + i/128
 * nearest m, z = (m - c)(1/c), series to z^10 (terms through z^3 in
 * double-double); tables og_tables.h */
static inline og_d2 og_log_dd(double x) {
  int e;
  double m = frexp(x, &e);
  m *= 2.;
  e -= 1;
  int i = (int)rint((m - 1.) * 128.);
  double c = (double)(128 + i) * 0x1p-7;
  const double *T = og_logc[i];
  og_d2 z = og_dd_mul_d(og_mk(T[2], T[3]), m - c);
  double zh = z.hi;
  double q = -1. / 10.;
  q = fma(q, zh, 1. / 9.); q = fma(q, zh, -1. / 8.); q = fma(q, zh, 1. / 7.);
  q = fma(q, zh, -1. / 6.); q = fma(q, zh, 1. / 5.); q = fma(q, zh, -1. / 4.);
  og_d2 z2 = og_dd_mul(z, z);
  og_d2 z3 = og_dd_mul(z2, z);
  og_d2 s = og_dd_add(og_dd_mul(z3, og_mk(0x1.5555555555555p-2, 0x1.5555555555555p-56)),
                      og_mk((z2.hi * z2.hi) * q, 0.));
  s = og_dd_add(og_mk(-0.5 * z2.hi, -0.5 * z2.lo), s);
  s = og_dd_add(z, s);
  return og_dd_add(og_dd_add(og_dd_mul_d(og_mk(og_ln2[0], og_ln2[1]), (double)e), og_mk(T[0], T[1])), s);
}
/* exp z = 2^k 2^(j/64) exp(r): z = (64 k + j) log(2)/64 + r, |r| <= log(2)/128,
 * Taylor series to r^9 (terms through r^3 in double-double); E in [1, 2.02) */
static inline og_d2 og_exp_dd2(og_d2 z, int *k2) {
  double kk = rint(z.hi * 0x1.71547652b82fep+6);
  int n = (int)kk;
  int j = n & 63;
  og_d2 r = og_dd_add(z, og_dd_mul_d(og_mk(-og_ln2[0] * 0x1p-6, -og_ln2[1] * 0x1p-6), kk));
  double rh = r.hi;
  double q = 1. / 362880.;
  q = fma(q, rh, 1. / 40320.); q = fma(q, rh, 1. / 5040.); q = fma(q, rh, 1. / 720.);
  q = fma(q, rh, 1. / 120.); q = fma(q, rh, 1. / 24.);
  og_d2 r2 = og_dd_mul(r, r);
  og_d2 r3 = og_dd_mul(r2, r);
  og_d2 s = og_dd_add(og_dd_mul(r3, og_mk(0x1.5555555555555p-3, 0x1.5555555555555p-57)),
                      og_mk((r2.hi * r2.hi) * q, 0.));
  s = og_dd_add(og_mk(0.5 * r2.hi, 0.5 * r2.lo), s);
  s = og_dd_add(r, s);
  s = og_dd_add(og_mk(1., 0.), s);
  *k2 = (n - j) / 64;
  return og_dd_mul(og_mk(og_exp2j[j][0], og_exp2j[j][1]), s);
}
static inline double og_exp_dd(og_d2 z) {
  if (z.hi > 709.8) return HUGE_VAL;
  if (z.hi < -745.2) return 0.;
  int k;
  og_d2 E = og_exp_dd2(z, &k);
  return ldexp(E.hi, k);
}
/* x^(3/4) and x^(3/2) for x in [2^-300, 2^300]: square roots for a first
 * value y0, then one Newton step on y^4 = x^3 (y^2 = x^3) with the residual
 * in double-double; y0 minus the step is within about 2^-100 of the power */
static inline double og_pow34(double x) {
  double s = sqrt(x), t = sqrt(s);
  double y0 = s * t;
  og_d2 y2 = og_two_prod(y0, y0);
  og_d2 y4 = og_dd_mul(y2, y2);
  og_d2 x3 = og_dd_mul_d(og_two_prod(x, x), x);
  og_d2 r = og_dd_add(y4, og_mk(-x3.hi, -x3.lo));
  return y0 - r.hi / (4. * y0 * y2.hi);
}
static inline double og_pow32(double x) {
  double y0 = x * sqrt(x);
  og_d2 y2 = og_two_prod(y0, y0);
  og_d2 x3 = og_dd_mul_d(og_two_prod(x, x), x);
  og_d2 r = og_dd_add(y2, og_mk(-x3.hi, -x3.lo));
  return y0 - r.hi / (2. * y0);
}
static inline double og_pow(double x, double y) {
  if (y == 0.) return 1.;
  if (!(x >= 0.)) return (x - x) / (x - x);
  if (x == 0.) return y > 0. ? 0. : HUGE_VAL;
  if (x == HUGE_VAL) return y > 0. ? HUGE_VAL : 0.;
  if (x >= 0x1p-300 && x <= 0x1p+300) {
    if (y == 0.75) return og_pow34(x);
    if (y == 1.5) return og_pow32(x);
  }
  return og_exp_dd(og_dd_mul_d(og_log_dd(x), y));
}
/* tanh and log10 the same way (the reference calls the C library's; these
 * agree with it to one ulp and are what the kernels evaluate) */
static inline double og_tanh(double y) {
  if (!(y == y)) return y;
  double a = fabs(y);
  if (a < 1e-8) return y;
  if (a > 22.) return y > 0. ? 1. : -1.;
  int k;
  og_d2 E = og_exp_dd2(og_mk(2. * a, 0.), &k);
  og_d2 Ek = og_mk(ldexp(E.hi, k), ldexp(E.lo, k));
  og_d2 t = og_dd_div(og_dd_add(Ek, og_mk(-1., 0.)), og_dd_add(Ek, og_mk(1., 0.)));
  return y > 0. ? t.hi : -t.hi;
}
static inline double og_log10(double x) {
  if (!(x > 0.)) return x == 0. ? -HUGE_VAL : (x - x) / (x - x);
  if (x == HUGE_VAL) return x;
  og_d2 l = og_log_dd(x);
  og_d2 r = og_dd_div(l, og_mk(0x1.26bb1bbb55516p+1, -0x1.f48ad494ea3e9p-53));
  return r.hi;
}
#endif
