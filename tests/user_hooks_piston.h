/* An alternative user hooks file (test input): a piston.  The walls at the
 * inner end of each track move at the constant speed par[0], the outer walls
 * stay, the faces between move at the solver's speeds; ghost cells copy the
 * edge cell with the normal four-velocity mirrored about the wall's
 * (u -> 2 u_wall - u, u_wall = par[0] / sqrt(1 - par[0]^2) at the inner end,
 * 0 at the outer); no regrid.  Per cell: a uniform force par[1] along the
 * track (momentum -par[1] rho dV dt, energy -par[1] rho v dV dt) and a density
 * floor par[2] after each cons2prim. */
#ifndef GH_USER_HOOKS_H
#define GH_USER_HOOKS_H
#ifndef GH_USER_FN
#define GH_USER_FN static inline
#endif
#include <math.h>

enum { GH_USER_NPAR = 16 };

typedef struct gh_user_track {
  int j, ntrack, ngst, geom;
  double t, y, dy;
  const double* par;
} gh_user_track;

GH_USER_FN double gh_user_face_speed(const gh_user_track* T, int i, double x, double v) {
  (void)x;
  if (i < T->ngst) return T->par[0];
  if (i > T->ntrack - 2 - T->ngst) return 0.;
  return v;
}

GH_USER_FN void gh_user_ghost(const gh_user_track* T, int i, double x, double dx, double* prim) {
  (void)x; (void)dx;
  double vw = i < T->ngst ? T->par[0] : 0.;
  double uw = vw / sqrt(1. - vw * vw);
  prim[2] = 2. * uw - prim[2];
}

GH_USER_FN int gh_user_regrid(const gh_user_track* T, int i, double x, double dx, const double* prim) {
  (void)T; (void)i; (void)x; (void)dx; (void)prim;
  return 0;
}

GH_USER_FN void gh_user_sources(const gh_user_track* T, int i, double x, double dx, double dV, double dt,
                                const double* prim, double* cons) {
  (void)i; (void)x; (void)dx;
  double g = T->par[1], rho = prim[0], u = prim[2];
  double v = u / sqrt(1. + u * u);
  cons[2] -= g * rho * dV * dt;
  cons[1] -= g * rho * v * dV * dt;
}

GH_USER_FN void gh_user_c2p(const double* par, int nd, double* rho, double* p, double* u) {
  (void)nd; (void)p; (void)u;
  if (*rho < par[2]) *rho = par[2];
}

#endif
