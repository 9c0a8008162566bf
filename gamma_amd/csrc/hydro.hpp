// hydro.hpp -- pointwise FP64 physics of GAMMA's hot path for gfx950 kernels.
//
// Register-resident state vectors (prim/cons/flux of one state), templated on
// the track dimensionality ND (1 or 2) and the number of advected variables NQ
// (6, or 10 with the synchrotron tracers).  Geometry is a run-time value that
// is uniform across a track (cartesian: rel_cart + cartesian modules;
// spherical: rel_sph + spherical[1D] modules), so branches on it never
// diverge inside a wavefront.
//
// Formulas follow the reference (eliotayache-nv/GAMMA, src/) term for term and
// in the same evaluation order, so that with contraction disabled the results
// are bit-identical to the CPU restatement; each function cites its source.
#pragma once
#ifdef GH_EMU
// CPU emulation build (test tool): kernel bodies compile as plain C++.
#define __host__
#define __device__
#define __forceinline__ inline
#else
#include <hip/hip_runtime.h>
#endif
#include <math.h>
#include "ddpow.hpp"

#define GH_HD __host__ __device__ __forceinline__

// the user setup's hooks (GAMMAHIP_USER_HOOKS names another file at build time)
#define GH_USER_FN static inline __host__ __device__
#ifdef GH_USER_HOOKS
#include GH_USER_HOOKS
#else
#include "user_hooks.h"
#endif

namespace gh {

enum { RHO = 0, PPP = 1, UU1 = 2, UU2 = 3 };
enum { DEN = 0, TAU = 1, SS1 = 2, SS2 = 3 };
enum { TR1 = 4, GMN = 6, GMX = 7, GMD = 8, PSN = 9 };
enum { CART = 0, SPH = 1 };

constexpr double kPi = 3.14159265358979323846;
constexpr double kPFloor = 1.e-10;       // constants.h:6 P_FLOOR_
constexpr double kPspec = 2.22;          // constants.h:24 p_ (VARIABLE_PSPEC_ disabled)
constexpr double kEpsE = 0.1;            // constants.h:25
constexpr double kEpsB = 0.01;           // constants.h:26
constexpr double kZeta = 1.;             // constants.h:27
constexpr double kBig = 1.e15;           // dt_loc reset value (cell.h:32)

struct Phys {          // run-time constants shared by all kernels
  double gamma, cfl, vi, shock_thr, gamma_max_init;
  double Nalpha, Nmp, Nme;
  int synge;           // EOS_ == SYNGE_EOS_ (spherical module)
};

// FluidState::gamma  rel_sph.cpp:37-52: the adiabatic index; with the Synge
// equation of state (spherical module only) an effective index from p/rho,
// between G for a cold gas and (G + 1)/2 for a hot one
GH_HD double eos_gamma(double rho, double p, double G, bool synge) {
  if (!synge) return G;
  double a = p / rho / (G - 1.);
  double e_ratio = a + sqrt(a * a + 1.);
  return G - (G - 1.) / 2. * (1. - 1. / (e_ratio * e_ratio));
}

template <int NQ>
struct State {
  double prim[NQ];
  double cons[NQ];
};

// FluidState::lfac  fluid.cpp:10-23 (sqrt of the norm, then squared again)
template <int ND, int NQ>
GH_HD double lfac(const double* prim) {
  double u = 0;
  for (int i = 0; i < ND; ++i) u += prim[UU1 + i] * prim[UU1 + i];
  u = sqrt(u);
  return sqrt(1 + u * u);
}

// FluidState::prim2cons  rel_sph.cpp:55-82 / rel_cart.cpp:37-64
template <int ND, int NQ>
GH_HD void prim2cons(const double* prim, double* cons, double r, int geom, double G, bool synge = false) {
  double rho = prim[RHO], p = prim[PPP];
  double u = 0;
  for (int i = 0; i < ND; ++i) u += prim[UU1 + i] * prim[UU1 + i];
  u = sqrt(u);
  double lf = sqrt(1 + u * u);
  double gma = eos_gamma(rho, p, G, synge && geom == SPH);
  double h = 1. + p * gma / (gma - 1.) / rho;
  double D = lf * rho;
  cons[DEN] = D;
  cons[TAU] = D * h * lf - p - D;
  for (int i = 0; i < ND; ++i) cons[SS1 + i] = D * h * prim[UU1 + i];
  if (ND == 1) cons[SS2] = 0.;               // unused slot in one dimension
  else if (geom == SPH) cons[SS2] *= r;
  for (int t = TR1; t < NQ; ++t) cons[t] = prim[t] * D;
}

// FluidState::state2flux along direction n  rel_sph.cpp:85-113 / rel_cart.cpp:67-95
template <int ND, int NQ>
GH_HD void state2flux(const double* prim, const double* cons, double* flux, int n,
                      double r, int geom) {
  double p = prim[PPP];
  double u = 0;
  for (int i = 0; i < ND; ++i) u += prim[UU1 + i] * prim[UU1 + i];
  u = sqrt(u);
  double lf = sqrt(1 + u * u);
  double D = cons[DEN];
  double ss[2] = {cons[SS1], ND == 2 ? cons[SS2] : 0.};
  if (ND == 2 && geom == SPH) ss[1] /= r;
  double un = prim[UU1 + n];
  flux[DEN] = D * un / lf;
  for (int i = 0; i < ND; ++i) {
    if (i == n) flux[SS1 + i] = ss[i] * un / lf + p;
    else flux[SS1 + i] = ss[i] * un / lf;
  }
  if (ND == 1) flux[SS2] = 0.;
  else if (geom == SPH) flux[SS2] *= r;
  flux[TAU] = ss[n] - D * un / lf;
  for (int t = TR1; t < NQ; ++t) flux[t] = cons[t] * un / lf;
}

// cons2prim residual  rel_sph.cpp:191-210 (fabs) / rel_cart.cpp:108-123
GH_HD double c2p_f(double p, double D, double S, double E, double G, int geom, bool synge = false) {
  double w = 1. - (S * S) / ((E + p) * (E + p));
  double lf = 1. / sqrt(geom == SPH ? fabs(w) : w);
  double gma = eos_gamma(D / lf, p, G, synge && geom == SPH);   // rel_sph.cpp:199-203
  return (E + p - D * lf - (gma * p * lf * lf) / (gma - 1.));
}

// Newton's residual in double-double arithmetic, scaled by a = 1/W^2 > 0
// (same sign as f, same root).  Where E = D + tau holds, f is the same
// function of p as c2p_f, written without forming E - D:
//   f = tau + p - D (W - 1) - G/(G-1) p W^2,
//   a f = a (tau + p) - D r (1 - a) / (1 + r) - G/(G-1) p,   r = sqrt(a) = 1/W,
// with a = w = 1 - S^2/(E+p)^2 (|w| in spherical geometry, as c2p_f).
// Evaluated in doubles, any form of f rounds at eps times the energy of the
// state, which for a cold gas in motion is the kinetic energy, many orders
// above the pressure's own scale: the root is then only known to
// eps (kinetic / thermal) relative, and a root finder stops at a point of that
// band that depends on where it started.  In double-double (the Dekker/Knuth
// operations of ddpow.hpp, exact on host and device alike) the band shrinks
// to eps^2 times the same ratio, below one unit in the last place of p for
// every state the solver meets, so the sign of f at a double is reliable and
// the root rounds to one double whatever the start.  gg = G/(G-1) in
// double-double; *a_out = a rounded to a double.
GH_HD double c2p_afdd(double p, double D, double S, double tau, dd::D2 gg, int geom, double* a_out) {
  using namespace dd;
  D2 tp = two_sum(tau, p);
  D2 Ep = add(two_sum(D, tau), {p, 0.});
  D2 v2 = div(two_prod(S, S), mul(Ep, Ep));
  D2 a = add({1., 0.}, {-v2.hi, -v2.lo});
  D2 oma = v2;                                      // 1 - a
  if (geom == SPH && a.hi < 0.) { a = {-a.hi, -a.lo}; oma = add({2., 0.}, {-v2.hi, -v2.lo}); }
  double s = sqrt(a.hi);                            // r = sqrt(a): one Newton correction
  D2 s2 = two_prod(s, s);
  D2 r = fast_two_sum(s, ((a.hi - s2.hi) - s2.lo + a.lo) / (2. * s));
  D2 t1 = mul_d(div(mul(r, oma), add({1., 0.}, r)), D);
  D2 g = add(mul(a, tp), {-t1.hi, -t1.lo});
  D2 t2 = mul_d(gg, p);
  g = add(g, {-t2.hi, -t2.lo});
  *a_out = a.hi;
  return g.hi;
}
GH_HD dd::D2 c2p_gg(double G) { return dd::div({G, 0.}, {G - 1., 0.}); }

// d f / d p for the Newton steps (in doubles: the derivative only sets the
// step), the residual in doubles, the magnitude of its terms (*fscale), a
// bound on its rounding (*fnoise: the terms holding W weighted by 1 + W^2, as
// w, formed by cancellation, carries a relative error of about W^2 eps) and
// a bound on the derivative's own rounding (*dferr).
GH_HD double c2p_df(double p, double D, double S, double tau, double G, int geom, double* fval,
                    double* fscale, double* fnoise = nullptr, double* dferr = nullptr) {
  double E = D + tau;
  double Ep = E + p;
  double v2 = (S * S) / (Ep * Ep);
  double w = 1. - v2;
  double aw = geom == SPH ? fabs(w) : w;
  double W2 = 1. / aw;
  double lf = sqrt(W2);
  double om = w >= 0. ? v2 : 2. - v2;      // 1 - |w|
  double gg = G / (G - 1.);
  double t1 = D * (om * W2 / (lf + 1.));
  double t2 = gg * p * W2;
  *fval = tau + p - t1 - t2;
  double tw = fabs(t1) + fabs(t2);
  *fscale = fabs(tau) + fabs(p) + tw;
  if (fnoise) *fnoise = *fscale + W2 * tw;
  // d W / dp = -W^3 S^2 / (E+p)^3 (sign of w folded in for |w|)
  double dw = 2. * (S * S) / (Ep * Ep * Ep);
  double dlf = -0.5 * lf * lf * lf * (geom == SPH && w < 0 ? -dw : dw);
  double a1 = D * dlf, a2 = gg * W2, a3 = gg * 2. * p * lf * dlf;
  if (dferr) *dferr = 16. * 2.2204460492503131e-16 * (1. + W2) * (1. + fabs(a1) + fabs(a2) + fabs(a3));
  return 1. - a1 - (a2 + a3);
}

enum { C2P_BRENT = 0, C2P_NEWTON = 1 };

// iters: root-finder iterations; fail: 1 no bracket, 2 max iter, 4 p_hi cap;
// evals: residual evaluations outside the root finder (bracketing, judging
// a second start); used: calls that started from the second start
struct C2PStat { int iters; int fail; int evals; int used; };

// Brent's method (Brent 1973, ch. 4) in the iteration form of GSL's brent
// fsolver (roots/brent.c: the interpolation test min(3 m q - |tol q|, |e q|),
// the bisection fallback and the bracket bookkeeping, step for step), with
// the GSL fsolver contract used by rel_sph.cpp:273-285: root = best iterate b, bracket = [min(b,c), max(b,c)],
// stop on |hi-lo| < 1e-13 * min(|lo|,|hi|) (0 if the bracket straddles 0).
// flo, fhi: the residual at lo and hi, which the caller has already evaluated
// while bracketing (the fsolver's set call evaluates them again; the residual
// is a pure function, so the values are the same bits)
GH_HD double brent(double D, double S, double E, double G, int geom, double lo, double hi,
                   double flo, double fhi, int maxit, C2PStat* st, bool synge = false) {
  double a = lo, b = hi;
  double fa = flo, fb = fhi;
  double c = b, fc = fb, d = b - a, e = b - a;
  double root = 0.5 * (lo + hi), xl = lo, xu = hi;
  int it = 0;
  for (;;) {
    it++;
    bool ac_equal = false;
    if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) { ac_equal = true; c = a; fc = fa; d = b - a; e = b - a; }
    if (fabs(fc) < fabs(fb)) {
      ac_equal = true;
      a = b; b = c; c = a;
      fa = fb; fb = fc; fc = fa;
    }
    double tol = 0.5 * 2.2204460492503131e-16 * fabs(b);
    double m = 0.5 * (c - b);
    bool done = false;
    if (fb == 0) { root = b; xl = b; xu = b; done = true; }
    else if (fabs(m) <= tol) {
      root = b;
      if (b < c) { xl = b; xu = c; } else { xl = c; xu = b; }
      done = true;
    }
    if (!done) {
      if (fabs(e) < tol || fabs(fa) <= fabs(fb)) { d = m; e = m; }
      else {
        double pp, q, rr, s = fb / fa;
        if (ac_equal) { pp = 2 * m * s; q = 1 - s; }
        else {
          q = fa / fc; rr = fb / fc;
          pp = s * (2 * m * q * (q - rr) - (b - a) * (rr - 1));
          q = (q - 1) * (rr - 1) * (s - 1);
        }
        if (pp > 0) q = -q; else pp = -pp;
        double l1 = 3 * m * q - fabs(tol * q), l2 = fabs(e * q);
        if (2 * pp < (l1 < l2 ? l1 : l2)) { e = d; d = pp / q; }
        else { d = m; e = m; }
      }
      a = b; fa = fb;
      if (fabs(d) > tol) b += d; else b += (m > 0 ? +tol : -tol);
      fb = c2p_f(b, D, S, E, G, geom, synge);
      root = b;
      if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) c = a;
      if (b < c) { xl = b; xu = c; } else { xl = c; xu = b; }
    }
    double mn = 0.;
    if ((xl > 0.0 && xu > 0.0) || (xl < 0.0 && xu < 0.0)) mn = fabs(xl) < fabs(xu) ? fabs(xl) : fabs(xu);
    if (fabs(xu - xl) < 1.e-13 * mn) break;
    if (it >= maxit) { st->fail |= 2; break; }
  }
  st->iters += it;
  return root;
}

// Newton cons2prim: the largest double p with f(p) > 0 (f decreases through
// the root), i.e. the root rounded down, whatever the start.
//   1. Newton on the residual in doubles (c2p_df), safeguarded by the bracket
//      [lo, hi] the iterates establish (lo from p_lo, the reference's lower
//      bound, hi open until f < 0 is seen), until the step falls inside that
//      residual's rounding band; where that band is far narrower than f' times
//      one unit, the decision of step 3 is taken on the doubles and the
//      solve ends here;
//   2. Newton on the double-double residual until the step is within two
//      units in the last place;
//   3. the decision: from f and f' at x, the sign of f one unit above (or
//      below) x is clear whenever f' is known well enough and f's rounding is
//      far below f' times one unit; otherwise f is evaluated there, stepping one
//      unit at a time (at most 8 steps each way).  Where even the
//      double-double residual is noisy at that scale (see below) the last
//      iterate is the answer.
// *fail gets 1 when f(p_lo) < 0 (no root above the lower bound, the
// reference's test), 2 when the iteration or the steps run out.
GH_HD double newton(double D, double S, double tau, double G, int geom, double p_lo, double x0, int maxit,
                    C2PStat* st) {
  const double eps = 2.2204460492503131e-16;
  const dd::D2 gg = c2p_gg(G);
  double lo = p_lo, hi = HUGE_VAL;
  double x = x0 > p_lo && x0 < HUGE_VAL ? x0 : (p_lo > 0. ? 2. * p_lo : fmax((G - 1.) * tau, 1e-300));
  int it = 0;
  bool ok = false;
  // 1. doubles
  for (;;) {
    it++;
    double f, fs, fn;
    double df = c2p_df(x, D, S, tau, G, geom, &f, &fs, &fn);
    if (!(f == f)) break;
    double step = f / df;
    // inside the rounding band the sign of f means nothing: stop before the
    // bracket test (a zero step would sit on the bracket's end)
    if (fabs(f) <= 4. * eps * fn || fabs(step) <= 1e-15 * fabs(x)) {
      if (x - step > p_lo) x -= step;
      ok = true;
      break;
    }
    if (f > 0) lo = x; else hi = x;
    double xn = x - step;
    if (!(xn > lo && xn < hi)) xn = hi < HUGE_VAL ? 0.5 * (lo + hi) : 10. * x;
    x = xn;
    if (hi - lo <= 1e-15 * lo) { ok = true; break; }
    if (it >= maxit) break;
  }
  bool found = false;
  // 1b. where the double residual's rounding is far below f' times one unit
  // (hot or static gas: the thermal energy is not small next to the kinetic
  // one), the doubles already decide the sign of f one unit either side of x
  bool dd_needed = true;
  if (ok) {
    it++;
    double f1, fs1, fn1, dfe1;
    double df1 = c2p_df(x, D, S, tau, G, geom, &f1, &fs1, &fn1, &dfe1);
    double up = nextafter(x, HUGE_VAL) - x, dn = x - nextafter(x, 0.);
    double m = 0.25 * fabs(df1) * fmin(up, dn), nb = 4. * eps * fn1;
    if (df1 < 0. && dfe1 < 0.1 * fabs(df1) && nb < 0.05 * m) {
      if (f1 > nb && f1 + df1 * up < -(m + nb)) { found = true; dd_needed = false; }
      else if (f1 < -nb && f1 - df1 * dn > m + nb) { x -= dn; found = true; dd_needed = false; }
    }
  }
  // 2. double-double
  double g = 0., a = 1., f = 0., df = -1., dfe = 0., fn = 0.;
  if (ok && dd_needed) {
    // the double residual's signs inside its rounding band are not to be
    // trusted: start the bracket afresh
    ok = false;
    lo = p_lo;
    hi = HUGE_VAL;
    for (int k = 0; k < 8; ++k) {
      it++;
      g = c2p_afdd(x, D, S, tau, gg, geom, &a);
      double fd, fs;
      df = c2p_df(x, D, S, tau, G, geom, &fd, &fs, &fn, &dfe);
      f = g / a;
      double step = f / df;
      double up = nextafter(x, HUGE_VAL) - x;
      if (fabs(step) <= 2. * up || fabs(f) <= 4. * eps * eps * fn) { ok = true; break; }
      if (g > 0) lo = fmax(lo, x); else hi = fmin(hi, x);
      double xn = x - step;
      if (!(xn > lo && xn < hi)) xn = hi < HUGE_VAL ? 0.5 * (lo + hi) : 10. * x;
      x = xn;
    }
  }
  // 3. the root rounded down
  double x2 = x;
  if (ok && dd_needed) {
    double up = nextafter(x, HUGE_VAL) - x, dn = x - nextafter(x, 0.);
    double m = 0.25 * fabs(df) * fmin(up, dn);
    bool sure = df < 0. && dfe < 0.1 * fabs(df) && 4. * eps * eps * fn < 0.05 * m;
    if (sure && g > 0. && f + df * up < -m) found = true;                     // f(x + up) < 0
    else if (sure && !(g > 0.) && f - df * dn > m) { x -= dn; found = true; }  // f(x - dn) > 0
    else if (g > 0.) {
      for (int k = 0; k < 8; ++k) {
        double xu = nextafter(x, HUGE_VAL), au;
        it++;
        if (!(c2p_afdd(xu, D, S, tau, gg, geom, &au) > 0.)) { found = true; break; }
        x = xu;
      }
    } else {
      for (int k = 0; k < 8 && x > 0.; ++k) {
        double ad;
        x = nextafter(x, 0.);
        it++;
        if (c2p_afdd(x, D, S, tau, gg, geom, &ad) > 0.) { found = true; break; }
      }
    }
  }
  // Near W ~ 1e3 with p/rho ~ 1e-10 the double-double residual itself rounds
  // at eps^2 W^2 times its terms, which spans a few units in the last place
  // of p: no double is the rounded root, and the point where f reached that
  // floor is the answer (it depends on the start by those few units).
  if (!found && ok && dd_needed && fabs(f) <= 8. * eps * eps * fn) { x = x2; found = true; }
  // the reference's bracket test: no root above p_lo
  if (!found || x < 2. * p_lo) {
    double al;
    it++;
    if (c2p_afdd(p_lo, D, S, tau, gg, geom, &al) < 0.) st->fail |= 1;
    else if (!found) st->fail |= 2;
  }
  st->iters += it;
  return x;
}

// (D, |S|, tau) of a conserved state as cons2prim below forms them (the
// spherical 2D angular momentum is divided by r); inputs of the learned guess
template <int ND, int NQ>
GH_HD void c2p_inputs(const double* cons, double r, int geom, double* D, double* S, double* tau) {
  double ss[2] = {cons[SS1], ND == 2 ? cons[SS2] : 0.};
  if (ND == 2 && geom == SPH) ss[1] /= r;
  double S2 = 0.;
  for (int d = 0; d < ND; ++d) S2 += ss[d] * ss[d];
  *D = cons[DEN];
  *S = sqrt(S2);
  *tau = cons[TAU];
}

// FluidState::cons2prim  rel_sph.cpp:213-320 / rel_cart.cpp:125-229.
// prim[PPP] on entry is the previous pressure (the reference's guess).
// guess_override > 0 (Newton only) is a second start, e.g. the learned guess:
// the solve starts from whichever of the two has the smaller residual |f|,
// so the previous pressure stays the fallback and the arbiter.  Newton's answer
// does not depend on its start (see newton).
// cons on exit = prim2cons(prim) (rel_sph.cpp:318).
template <int ND, int NQ>
GH_HD void cons2prim(double* prim, double* cons, double r, int geom, double G, int solver,
                     double guess_override, int maxit, C2PStat* st, bool synge = false,
                     const double* upar = nullptr) {
  double D = cons[DEN], tau = cons[TAU], E = D + tau;
  double ss[2] = {cons[SS1], ND == 2 ? cons[SS2] : 0.};
  if (ND == 2 && geom == SPH) ss[1] /= r;
  double S2 = 0.;
  for (int d = 0; d < ND; ++d) S2 += ss[d] * ss[d];
  double S = sqrt(S2);
  double pprev = geom == SPH ? fmax(1.e-13, prim[PPP]) : prim[PPP];
  double pCand = pprev;
  if (solver == C2P_NEWTON) {
    double start = pCand;
    if (guess_override > 0.) {
      // judge the two starts by the residual in doubles (the answer does not
      // depend on the start, only the iteration count does)
      double f1, f2, s1, s2;
      st->evals += 2;
      (void)c2p_df(guess_override, D, S, tau, G, geom, &f1, &s1);
      (void)c2p_df(pCand, D, S, tau, G, geom, &f2, &s2);
      if (fabs(f1) < fabs(f2)) { start = guess_override; st->used++; }
    }
    double p_lo = fmax(0., (1. + 1e-13) * S - E);
    int f0 = st->fail;
    st->fail = 0;
    pCand = newton(D, S, tau, G, geom, p_lo, start, 100, st);
    if (st->fail & 1) pCand = pprev;                 // no bracket: keep the previous pressure
    st->fail |= f0;
  } else {
    double f_init = c2p_f(pCand, D, S, E, G, geom, synge);
    if (fabs(f_init) > 0.) {
      double p_lo = fmax(0., (1. + 1e-13) * S - E), p_hi;
      bool find_p = true;
      st->evals += 2;
      double f_lo = c2p_f(p_lo, D, S, E, G, geom, synge);
      if (f_lo < 0) find_p = false;
      // p_hi grows by decades from the previous pressure until f(p_hi) <= 0;
      // f at the previous pressure is f_init
      p_hi = pCand;
      double f_hi = f_init;
      int guard = 0;
      while (f_hi > 0) {
        p_hi *= 10;
        if (++guard > 400) { find_p = false; st->fail |= 4; break; }
        st->evals++;
        f_hi = c2p_f(p_hi, D, S, E, G, geom, synge);
      }
      if (p_hi < p_lo) find_p = false;
      if (find_p) pCand = brent(D, S, E, G, geom, p_lo, p_hi, f_lo, f_hi, maxit, st, synge);
      else {
        st->fail |= 1;
        pCand = pprev;
      }
    }
  }
  double p = pCand;
  double Ep2 = (E + p) * (E + p);
  double v2 = S2 / Ep2;
  double lf = 1. / sqrt(fabs(1. - v2));
  double rho = D / lf;
  double uu[ND];
  if (fabs(S) == 0.) {
    for (int d = 0; d < ND; ++d) uu[d] = 0;
  } else {
    double v = sqrt(1. - 1. / (lf * lf));
    for (int d = 0; d < ND; ++d) uu[d] = lf * v * (ss[d] / S);
  }
  // cons2prim_user (rel_sph.cpp:304): the user setup's, when the context asks
  if (upar) gh_user_c2p(upar, ND, &rho, &p, uu);
  for (int d = 0; d < ND; ++d) prim[UU1 + d] = uu[d];
  if (geom == SPH) p = fmax(p, kPFloor);
  prim[PPP] = p;
  prim[RHO] = rho;
  for (int t = TR1; t < NQ; ++t) prim[t] = cons[t] / D;
  prim2cons<ND, NQ>(prim, cons, r, geom, G, synge);
}

// ---- HLLC (Mignone & Bodo 2006) through a moving face ----
struct Waves { double lL, lR, lS; };

// Interface::wavespeedEstimates + computeLambda  rel_sph.cpp:322-394 / rel_cart.cpp:232-302
template <int ND, int NQ>
GH_HD Waves lambdas(const double* pL, const double* cL, const double* pR, const double* cR,
                    int dim, double rface, int geom, double G) {
  int u = UU1 + dim, s = SS1 + dim;
  double lfL = lfac<ND, NQ>(pL), lfR = lfac<ND, NQ>(pR);
  double vL = pL[u] / lfL, vR = pR[u] / lfR;
  double cSL = sqrt(G * pL[PPP] / (pL[RHO] + pL[PPP] * G / (G - 1.)));
  double cSR = sqrt(G * pR[PPP] / (pR[RHO] + pR[PPP] * G / (G - 1.)));
  double sgSL = cSL * cSL / (lfL * lfL * (1. - cSL * cSL));
  double sgSR = cSR * cSR / (lfR * lfR * (1. - cSR * cSR));
  double l1, l2;
  Waves w;
  l1 = (vR - sqrt(sgSR * (1. - vR * vR + sgSR))) / (1. + sgSR);
  l2 = (vL - sqrt(sgSL * (1. - vL * vL + sgSL))) / (1. + sgSL);
  w.lL = fmin(l1, l2);
  l1 = (vR + sqrt(sgSR * (1. - vR * vR + sgSR))) / (1. + sgSR);
  l2 = (vL + sqrt(sgSL * (1. - vL * vL + sgSL))) / (1. + sgSL);
  w.lR = fmax(l1, l2);
  bool sph = geom == SPH;
  double d = (sph && dim == 1) ? rface : 1.;
  double mL = sph ? cL[s] / d : cL[s];
  double mR = sph ? cR[s] / d : cR[s];
  double EL = cL[TAU] + cL[DEN], ER = cR[TAU] + cR[DEN];
  double lL = w.lL, lR = w.lR;
  double AL = lL * EL - mL, AR = lR * ER - mR;
  double BL = mL * (lL - vL) - pL[PPP], BR = mR * (lR - vR) - pR[PPP];
  double FhllE = (lL * AR - lR * AL) / (lR - lL);
  double Ehll = (AR - AL) / (lR - lL);
  double Fhllm = (lL * BR - lR * BL) / (lR - lL);
  double mhll = (BR - BL) / (lR - lL);
  bool degenerate = sph ? (fabs(FhllE) == 0) : (fabs(FhllE) < 1.e-15);
  if (degenerate) { w.lS = mhll / (Ehll + Fhllm); return w; }
  double ef = Ehll + Fhllm;
  double delta = ef * ef - 4. * FhllE * mhll;
  w.lS = ((Ehll + Fhllm) - sqrt(delta)) / (2. * FhllE);
  return w;
}

// Interface::starState  rel_sph.cpp:397-450 / rel_cart.cpp:305-355 -> star cons
template <int ND, int NQ>
GH_HD void star_state(const double* pin, const double* cin, double lbda, double lS, int dim,
                      double rface, int geom, double* out) {
  int u = UU1 + dim, s = SS1 + dim;
  bool sph = geom == SPH;
  double mm[2] = {cin[SS1], ND == 2 ? cin[SS2] : 0.};
  if (ND == 2 && sph) mm[1] /= rface;
  double lf = lfac<ND, NQ>(pin);
  double v = pin[u] / lf;
  double m = mm[dim];
  double p = pin[PPP];
  double D = cin[DEN];
  double E = cin[TAU] + cin[DEN];
  double A = lbda * E - m;
  double B = m * (lbda - v) - p;
  double pS = (A * lS - B) / (1. - lbda * lS);
  double k = (lbda - v) / (lbda - lS);
  for (int q = 0; q < NQ; ++q) out[q] = 0.;
  out[DEN] = D * k;
  out[s] = (m * (lbda - v) + pS - p) / (lbda - lS);
  if (ND == 2) { int o = 1 - dim; out[SS1 + o] = mm[o] * k; }
  if (ND == 2 && sph) out[SS2] *= rface;
  out[TAU] = (E * (lbda - v) + pS * lS - p * v) / (lbda - lS) - D * k;
  for (int t = TR1; t < NQ; ++t) out[t] = cin[t] * k;
}

// Interface::computeFlux  hllc.cpp:12-39: sequential region tests against the
// face speed v (last matching one wins), F - v U*, times dA.  fL/fR are the
// physical fluxes of the two states along dim.  NaN wave speeds leave the
// flux NaN (flagged by the caller).
template <int ND, int NQ>
GH_HD void hllc_flux(const double* pL, const double* cL, const double* fL,
                     const double* pR, const double* cR, const double* fR,
                     const Waves& w, double v, double dA, int dim, double rface, int geom,
                     double* flux) {
  int region = -1;
  if (v < w.lL) region = 0;
  if (w.lL <= v && v < w.lS) region = 1;
  if (w.lS <= v && v < w.lR) region = 2;
  if (w.lR <= v) region = 3;
  if (region == 0 || region == 3) {
    const double* c = region == 0 ? cL : cR;
    const double* f = region == 0 ? fL : fR;
    for (int q = 0; q < NQ; ++q) flux[q] = f[q] - v * c[q];
  } else if (region == 1 || region == 2) {
    const double* pK = region == 1 ? pL : pR;
    const double* cK = region == 1 ? cL : cR;
    const double* fK = region == 1 ? fL : fR;
    double lK = region == 1 ? w.lL : w.lR;
    double st[NQ];
    star_state<ND, NQ>(pK, cK, lK, w.lS, dim, rface, geom, st);
    for (int q = 0; q < NQ; ++q) flux[q] = fK[q] + lK * (st[q] - cK[q]) - v * st[q];
  } else {
    for (int q = 0; q < NQ; ++q) flux[q] = __builtin_nan("");
  }
  for (int q = 0; q < NQ; ++q) flux[q] *= dA;
}

// The same flux with the physical flux formed only for the side the face's
// wave region uses: side(right, f) writes the flux of the left (false) or
// right (true) state.  Same arithmetic as hllc_flux; one state2flux instead of
// two and six fewer live values per state in the face kernels.  Where no
// region matches (a NaN wave speed: a reconstructed state without a sound
// speed) the flux is `none`: NaN on MV faces (counted by the caller), 0 on F1
// faces -- the reference forms each F1 face in a fresh local Interface
// (2d.cpp:923) whose flux the constructor leaves unset (interface.cpp:10-17)
// and computeFlux then leaves untouched; the zero-initialised face is the
// reading taken.
template <int ND, int NQ, class Side>
GH_HD void hllc_flux_lazy(const double* pL, const double* cL, const double* pR, const double* cR,
                          const Waves& w, double v, double dA, int dim, double rface, int geom, Side side,
                          double* flux, double none = __builtin_nan("")) {
  int region = -1;
  if (v < w.lL) region = 0;
  if (w.lL <= v && v < w.lS) region = 1;
  if (w.lS <= v && v < w.lR) region = 2;
  if (w.lR <= v) region = 3;
  if (region < 0) {
    for (int q = 0; q < NQ; ++q) flux[q] = none;
    return;
  }
  bool right = region >= 2;
  const double* pK = right ? pR : pL;
  const double* cK = right ? cR : cL;
  double fK[NQ];
  side(right, fK);
  if (region == 0 || region == 3) {
    for (int q = 0; q < NQ; ++q) flux[q] = fK[q] - v * cK[q];
  } else {
    double lK = region == 1 ? w.lL : w.lR;
    double st[NQ];
    star_state<ND, NQ>(pK, cK, lK, w.lS, dim, rface, geom, st);
    for (int q = 0; q < NQ; ++q) flux[q] = fK[q] + lK * (st[q] - cK[q]) - v * st[q];
  }
  for (int q = 0; q < NQ; ++q) flux[q] *= dA;
}

// ---- shock detector: radiation_sph.cpp:37-158 (Rezzolla & Zanotti 2013) ----
// Returns -1 without touching *pspec when no shock is possible (:56-57).
// spectral index of the electrons accelerated at a shock of four-velocity
// fvelu: spectral_p, radiation_sph.cpp:18-35 (VARIABLE_PSPEC_ or the constant)
GH_HD double spectral_p(double fvelu, bool varp) {
  if (!varp) return kPspec;
  return 2.11 + 0.11 * gh_tanh((gh_log10(fabs(fvelu)) - .5));
}

template <int ND, int NQ>
GH_HD double shock_sd(const double* S1, const double* S2, int dim, bool reverse, double G,
                      double* pspec, bool varp = false) {
  int uux = UU1 + dim, uut = UU2 - dim;
  const int n_evals = 10;
  double chi = 0.5;
  double p1 = S1[PPP], p2 = S2[PPP];
  if (fabs((p1 - p2) / p1) < 1.e-10) return (-1);
  if (p1 < p2) return (-1);
  double delta_p = p2 - p1;
  double dp = delta_p / (double)n_evals;
  double rho1 = S1[RHO], rho2 = S2[RHO];
  double h1 = 1 + p1 * G / (G - 1.) / rho1;
  double h2 = 1 + p2 * G / (G - 1.) / rho2;
  double lfac1 = lfac<ND, NQ>(S1), lfac2 = lfac<ND, NQ>(S2);
  double vx1 = S1[uux] / lfac1, vx2 = S2[uux] / lfac2;
  if (reverse) { vx1 *= -1; vx2 *= -1; }
  double ut1 = 0;
  if (ND > 1) ut1 = S1[uut];
  double v12 = (vx1 - vx2) / (1 - vx1 * vx2);
  double A1 = h1 * ut1;
  double I = 0;
  for (int ip = 0; ip < n_evals; ++ip) {
    double p = p1 + (ip + .5) * dp;
    double rho = rho1 * gh_pow(p / p1, 1. / G);
    double h = 1 + p * G / (G - 1.) / rho;
    double cs = sqrt(G * p / (rho + p * G / (G - 1.)));
    double dI = sqrt(h * h + A1 * A1 * (1. - cs * cs)) / ((h * h + A1 * A1) * rho * cs) * dp;
    I += dI;
  }
  double vSR = gh_tanh(I);
  double rho22 = rho2 * rho2, lfac22 = lfac2 * lfac2, vx22 = vx2 * vx2;
  double g = G, gm1 = (g - 1.), gm12 = gm1 * gm1;
  double Da = 4. * g * p1 * ((g - 1) * p2 + p1) / (gm12 * (p1 - p2) * (p1 - p2));
  double Db = h2 * (p2 - p1) / rho2 - h2 * h2;
  double Dd = 1. - Da * Db;
  double h3 = (sqrt(Dd) - 1) * gm1 * (p1 - p2) / (2 * (gm1 * p2 + p1));
  double J232a = -g / gm1 * (p1 - p2);
  double J232b = h3 * (h3 - 1) / p1 - h2 * (h2 - 1) / p2;
  double J232 = J232a / J232b;
  double J23 = sqrt(J232);
  double Vsa = rho22 * lfac22 * vx2 + fabs(J23) * sqrt(J232 + rho22 * lfac22 * (1. - vx22));
  double Vsb = rho22 * lfac22 + J232;
  double Vs = Vsa / Vsb;
  double v2Sa = (p1 - p2) * (1. - vx2 * Vs);
  double v2Sb = (Vs - vx2) * (h2 * rho2 * lfac22 * (1. - vx22) + p1 - p2);
  double v2S = v2Sa / v2Sb;
  double vlim = vSR + chi * (v2S - vSR);
  // shock Lorentz factor setup (:146-155): the shock runs into gas at rest
  double betau = Vs;
  double lfacu = 1. / sqrt(1. - betau * betau);
  *pspec = spectral_p(lfacu * betau, varp);
  return v12 - vlim;
}

// ---- geometry (Geometry/*.cpp) ----
struct Geo { double dV, dl0, dl1, cen0, cen1; };

// Sines and cosines of a track's angular position y and width dy (2D
// spherical geometry): the only transcendental functions of the geometry, and
// functions of the track alone.  They are evaluated once per track on the
// host (the C library's sin/cos, as the reference uses) and tabulated, so the
// device never depends on its own math library for them.
struct Trig { double s_hdy, s_y, c_y, s_1, s_2, c_1, c_2; };
inline Trig trig_of(double y, double dy) {
  Trig t;
  double th1 = y - dy / 2., th2 = y + dy / 2.;
  t.s_hdy = sin(.5 * dy); t.s_y = sin(y); t.c_y = cos(y);
  t.s_1 = sin(th1); t.s_2 = sin(th2); t.c_1 = cos(th1); t.c_2 = cos(th2);
  return t;
}

template <int ND>
GH_HD Geo cell_geo(double x, double dx, double y, double dy, int geom, const Trig& T) {
  Geo g;
  if (geom == CART) {
    g.dl0 = dx; g.dl1 = dy;
    double dV = 1.;
    dV *= dx;
    if (ND == 2) dV *= dy;
    g.dV = dV;
    g.cen0 = x; g.cen1 = y;
  } else if (ND == 1) {                        // spherical1D.cpp:12-48
    g.dl0 = dx; g.dl1 = 0.;
    g.dV = 4. * kPi * x * x * dx;
    double r2 = x * x, dr2 = dx * dx;
    g.cen0 = 3. * x * (r2 + dr2) / (3 * r2 + dr2);
    g.cen1 = y;
  } else {                                     // spherical.cpp:12-69
    double r = x, dr = dx, th = y, dth = dy;
    g.dl0 = dr; g.dl1 = r * dth;
    double r1 = r - dr / 2., r2 = r + dr / 2.;
    double rsq = r1 * r1 + r1 * r2 + r2 * r2;
    g.dV = 4. / 3. * kPi * rsq * dr * fabs(T.s_hdy * T.s_y);
    double rr = r * r, dr2 = dr * dr;
    double th1 = th - dth / 2., th2 = th + dth / 2.;
    g.cen0 = 3. * r * (rr + dr2) / (3 * rr + dr2);
    g.cen1 = (th1 * T.c_1 - th2 * T.c_2 + fabs(T.s_2) - fabs(T.s_1)) / (T.c_1 - T.c_2);
  }
  return g;
}

// Interface::computedA for a face normal to dim at (x, y) with transverse
// extent dxf (cartesian.cpp:135-143, spherical1D.cpp:60-76, spherical.cpp:81-112).
// 2D spherical: sy = sin(y) and shd = sin(dxf / 2) (dim 0 only), from the
// track's Trig table.
template <int ND>
GH_HD double face_area(int dim, double x, double dxf, int geom, double sy, double shd) {
  if (geom == CART) return ND == 1 ? 1. : dxf;
  if (ND == 1) return 4. * kPi * x * x;
  if (dim == 0) return 4. * kPi * x * x * fabs(shd * sy);
  double dr = dxf, r1 = x - dr / 2., r2 = x + dr / 2.;
  double rsq = (r1 + r2) * dr;
  return kPi * fabs(sy) * rsq;
}

GH_HD double minmod(double a, double b) {      // 1d.cpp:250-260
  if (fabs(a) < fabs(b) && a * b > 0) return (a);
  if (fabs(b) < fabs(a) && a * b > 0) return (b);
  return (0);
}

// Cell::sourceTerms  rel_sph.cpp:452-496 (cartesian: none)
template <int ND, int NQ>
GH_HD void sources(const double* prim, double* U, double x, double dx, const Trig& T, double dV,
                   double dt, int geom, double G, bool synge = false) {
  if (geom != SPH) return;
  double r = x, dr = dx;
  double r1 = r - dr / 2., r2 = r + dr / 2.;
  double rsq = (r2 * r2 + r1 * r1 + r2 * r1) / 3.;
  if (ND == 1) {
    U[SS1] += (2. * prim[PPP]) * (r / rsq) * dV * dt;
  } else {
    double rho = prim[RHO], ut = prim[UU2], p = prim[PPP];
    double gma = eos_gamma(rho, p, G, synge);
    double h = 1. + p * gma / (gma - 1.) / rho;
    U[SS1] += (rho * h * ut * ut + 2. * p) * (r / rsq) * dV * dt;
    U[SS2] += p * T.c_y / fabs(T.s_y) * dV * dt;
  }
}

// Cell::radiativeSourceTerms  radiation_sph.cpp:285-300
GH_HD void radiative(const double* prim, double* U, double dV, double dt, const Phys& P) {
  double rho = prim[RHO], p = prim[PPP], gma = eos_gamma(rho, p, P.gamma, P.synge);
  double h = 1 + p * gma / (gma - 1.) / rho;
  double eps = rho * (h - 1.) / gma;
  double eB = kEpsB * eps;
  double B = sqrt(8. * kPi * eB);
  double dgma = P.Nalpha * gh_pow(rho, 4. / 3.) * B * B * dV * dt;
  U[GMX] += dgma;
  U[GMN] += dgma;
}

// Cell::radiation_injectParticles  radiation_sph.cpp:196-283
template <int ND, int NQ>
GH_HD void inject(double* prim, bool shocked, double pspec, const Phys& P) {
  double lf = lfac<ND, NQ>(prim), rho = prim[RHO];
  double r43 = gh_pow(rho, 4. / 3.);
  double lim = lf * r43 / (lf * rho);
  if (shocked) {
    if (pspec > prim[PSN] || isnan(prim[PSN])) prim[PSN] = pspec;
    double gmax = P.gamma_max_init;
    prim[GMX] = (gmax != 0. ? lf * r43 / gmax : 0.) / (lf * rho);
    double p = prim[PPP], psyn = prim[PSN], gma = eos_gamma(rho, p, P.gamma, P.synge);
    double h = 1. + p * gma / (gma - 1.) / rho;
    double eps = rho * (h - 1.) / gma;
    double ee = kEpsE * eps;
    double ne = kZeta * rho / P.Nmp;
    double lfac_av = ee / (ne * P.Nme);
    double gmin = (psyn - 2.) / (psyn - 1.) * lfac_av;
    prim[GMN] = (gmin != 0. ? lf * r43 / gmin : 0.) / (lf * rho);
  }
  if (isnan(prim[GMX])) prim[GMX] = lim;
  if (isnan(prim[GMN])) prim[GMN] = lim;
}

// Cell::radiation_initGammas  radiation_sph.cpp:246-260 (prim side)
template <int ND, int NQ>
GH_HD void init_gammas(double* prim) {
  double lf = lfac<ND, NQ>(prim), rho = prim[RHO];
  double lim = lf * gh_pow(rho, 4. / 3.);
  if (prim[GMX] > lim / (lf * rho) || prim[GMX] <= 0.) prim[GMX] = lim / (lf * rho);
  if (prim[GMN] > lim / (lf * rho) || prim[GMN] <= 0.) prim[GMN] = lim / (lf * rho);
}

// BM analytic profile in code units (BM1D.cpp:62-86, c = 1, k = 0):
// bm = {lfacShock2(t0), t0, rho_ambient, eta}
GH_HD void bm_profile(const double* bm, double r, double t, double G, double* rho, double* u,
                      double* p) {
  double L0 = bm[0], t0 = bm[1], rhoa = bm[2], eta = bm[3];
  double tr = t0 / t;
  double L = L0 * tr * tr * tr;
  double pa = rhoa * eta;
  double ha = 1. + pa * G / (G - 1.) / rhoa;
  double pf = 2. / 3. * L * ha * rhoa;
  double lfacf = sqrt(fmax(1., .5 * L));
  double Df = 2. * L * rhoa;
  double chi = (1. + 8. * L) * (1. - r / t);
  *p = pf * gh_pow(chi, -17. / 12.);
  double lf = sqrt(lfacf * lfacf / chi + 1);
  double Dn = Df * gh_pow(chi, -7. / 4.);
  *rho = Dn / lf;
  *u = sqrt(1. - 1. / (lf * lf)) * lf;
}

}  // namespace gh
