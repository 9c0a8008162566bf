/* og_pow.h -- x^y used by the oracle in place of the C library's pow.
 *
 * TEST INFRASTRUCTURE.  The reference calls pow() (glibc).  The device has no
 * glibc: its pow rounds differently in about a quarter of all cases.  Both
 * sides therefore evaluate powers with the same double-double algorithm
 * written with IEEE-exact operations only (+, -, *, /, fma, frexp, ldexp,
 * rint): x^y = exp(y log x), log by 2 atanh((m-1)/(m+1)) on m in
 * [sqrt(1/2), sqrt(2)), exp by exp(r/256)^256 with |r| <= log(2)/2 (Dekker
 * 1971 double-double arithmetic).  This file restates that algorithm in C,
 * operation for operation; tests/test_oracle_pins.py checks it against the C
 * library's pow (equal in all but about 1 case in 1200, never more than one
 * unit in the last place apart) and against correctly rounded values.
 */
#ifndef OG_POW_H
#define OG_POW_H
#include <math.h>

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
static inline og_d2 og_log_dd(double x) {
  int e;
  double m = frexp(x, &e);
  if (m < 0x1.6a09e667f3bcdp-1) { m *= 2.; e -= 1; }
  og_d2 t = og_dd_div(og_mk(m - 1., 0.), og_two_sum(m, 1.));
  og_d2 u = og_dd_mul(t, t);
  double uh = u.hi;
  double r = 1. / 31.;
  r = 1. / 29. + uh * r; r = 1. / 27. + uh * r; r = 1. / 25. + uh * r; r = 1. / 23. + uh * r;
  r = 1. / 21. + uh * r; r = 1. / 19. + uh * r; r = 1. / 17. + uh * r; r = 1. / 15. + uh * r;
  r = 1. / 13. + uh * r; r = 1. / 11. + uh * r; r = 1. / 9. + uh * r;
  og_d2 s = og_dd_add(og_mk(0x1.2492492492492p-3, 0x1.2492492492492p-57), og_dd_mul_d(u, r));
  s = og_dd_add(og_mk(0x1.999999999999ap-3, -0x1.999999999999ap-57), og_dd_mul(u, s));
  s = og_dd_add(og_mk(0x1.5555555555555p-2, 0x1.5555555555555p-56), og_dd_mul(u, s));
  s = og_dd_add(og_mk(1., 0.), og_dd_mul(u, s));
  og_d2 lm = og_dd_mul(og_dd_mul_d(t, 2.), s);
  og_d2 ln2 = og_mk(0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56);
  return og_dd_add(og_dd_mul_d(ln2, (double)e), lm);
}
static inline og_d2 og_exp_dd2(og_d2 z, int *k2) {
  og_d2 ln2 = og_mk(0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56);
  double k = rint(z.hi * 0x1.71547652b82fep+0);
  og_d2 r = og_dd_add(z, og_dd_mul_d(og_mk(-ln2.hi, -ln2.lo), k));
  r.hi *= 0x1p-8; r.lo *= 0x1p-8;
  double rh = r.hi;
  double q = 1. / 40320.;
  q = 1. / 5040. + rh * q; q = 1. / 720. + rh * q; q = 1. / 120. + rh * q; q = 1. / 24. + rh * q;
  og_d2 E = og_dd_add(og_mk(0x1.5555555555555p-3, 0x1.5555555555555p-57), og_dd_mul_d(r, q));
  E = og_dd_add(og_mk(0.5, 0.), og_dd_mul(r, E));
  E = og_dd_add(og_mk(1., 0.), og_dd_mul(r, E));
  E = og_dd_add(og_mk(1., 0.), og_dd_mul(r, E));
  for (int n = 0; n < 8; ++n) E = og_dd_mul(E, E);
  *k2 = (int)k;
  return E;
}
static inline double og_exp_dd(og_d2 z) {
  if (z.hi > 709.8) return HUGE_VAL;
  if (z.hi < -745.2) return 0.;
  int k;
  og_d2 E = og_exp_dd2(z, &k);
  return ldexp(E.hi, k);
}
static inline double og_pow(double x, double y) {
  if (y == 0.) return 1.;
  if (!(x >= 0.)) return (x - x) / (x - x);
  if (x == 0.) return y > 0. ? 0. : HUGE_VAL;
  if (x == HUGE_VAL) return y > 0. ? HUGE_VAL : 0.;
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
