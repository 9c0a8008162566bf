/* user_hooks.h -- the hooks of a user setup, for tracks whose hook kinds are
 * GAMMAHIP_KIN_USER, GAMMAHIP_BND_USER or GAMMAHIP_RG_USER.
 *
 * In the reference a setup is a source file in src/Initial/Tests/ that
 * defines Grid::userKinematics, Grid::userBoundaries and
 * Grid::checkCellForRegrid, selected at link time (Makefile:82-92).  Here the
 * built-in kinds cover the shipped setups (fixed boundary-face speeds, the
 * Blandford-McKee inflow, the aspect-ratio and BM1D rules), and this file is
 * the link-time slot for any other: edit the three functions below, or build
 * with GAMMAHIP_USER_HOOKS=<file> naming another file that defines them, and
 * rebuild.  The kernels call them per face or per cell on the GPU, and the CPU
 * oracle calls the same functions, so a user setup is checked against it like
 * any other.
 *
 * The functions must be pure functions of their arguments: they run one
 * thread per face or cell, in any order, and in plain C as well as in HIP
 * (the oracle is C).  prim[] holds the reference's slots (rho, p, u1, u2,
 * tracers, synchrotron tracers; u2 reads 0 in 1D and is not kept there),
 * velocities as four-velocity components as the solver stores them.
 *
 * Two more hooks run per cell when the context's GAMMAHIP_OPT_USER_CELL bits
 * ask for them (the reference calls them in every run; its shipped setups
 * leave them empty): userSourceTerms (bit 0) and cons2prim_user (bit 1).
 *
 * The shipped functions define a reflecting box on a slowly moving mesh:
 *   userKinematics      the ngst faces at each end of a track stay still (the
 *                       walls); the others move at par[0] times the speed the
 *                       solver gives them (0: a fixed mesh, 1: the usual
 *                       moving mesh);
 *   userBoundaries      ghost cells copy the edge cell (the outflow copy made
 *                       before the hook) with the normal velocity reversed;
 *   checkCellForRegrid  split a cell wider than par[1] whose density exceeds
 *                       par[3], merge a cell narrower than par[2] (no rule
 *                       where par[1] or par[2] is 0);
 *   userSourceTerms,    none (empty, as in the reference's setups).
 *   cons2prim_user
 * par[] are the context's user parameters (gammahip_set_user_params). */
#ifndef GH_USER_HOOKS_H
#define GH_USER_HOOKS_H

#ifndef GH_USER_FN
#define GH_USER_FN static inline
#endif

enum { GH_USER_NPAR = 16 };

/* the track a hook is evaluated on */
typedef struct gh_user_track {
  int j;              /* track index (ghost tracks included in 2D) */
  int ntrack;         /* cells on the track, ghost cells included */
  int ngst;           /* ghost cells at each end */
  int geom;           /* 0 cartesian, 1 spherical */
  double t;           /* the track's time */
  double y, dy;       /* F1 centre and width (2D; 0 in 1D) */
  const double* par;  /* GH_USER_NPAR user parameters */
} gh_user_track;

/* userKinematics: the speed of face i (0 .. ntrack-2, between cells i and
 * i+1) at position x, given v = VI * lambda* from the solver */
GH_USER_FN double gh_user_face_speed(const gh_user_track* T, int i, double x, double v) {
  (void)x;
  if (i < T->ngst || i > T->ntrack - 2 - T->ngst) return 0.;
  return T->par[0] * v;
}

/* userBoundaries: ghost cell i (i < ngst: inner end, else outer end), centre
 * x and width dx, whose primitives prim[] hold the outflow copy of the edge
 * cell; the hook writes the ghost's primitives */
GH_USER_FN void gh_user_ghost(const gh_user_track* T, int i, double x, double dx, double* prim) {
  (void)T; (void)i; (void)x; (void)dx;
  prim[2] = -prim[2];
}

/* checkCellForRegrid: 0 no action, 1 merge, 2 split, for active cell i with
 * centre x, width dx and primitives prim[] */
GH_USER_FN int gh_user_regrid(const gh_user_track* T, int i, double x, double dx, const double* prim) {
  (void)i; (void)x;
  const double* p = T->par;
  if (p[1] > 0. && dx > p[1] && prim[0] > p[3]) return 2;
  if (p[2] > 0. && dx < p[2]) return 1;
  return 0;
}

/* userSourceTerms (cell.cpp:112, after the geometric source terms): add the
 * sources of active cell i over dt to cons[] (conserved variables times the
 * cell volume, as Cell::update holds them there), given the stage's starting
 * primitives prim[], the cell's centre x, width dx and volume dV */
GH_USER_FN void gh_user_sources(const gh_user_track* T, int i, double x, double dx, double dV, double dt,
                                const double* prim, double* cons) {
  (void)T; (void)i; (void)x; (void)dx; (void)dV; (void)dt; (void)prim; (void)cons;
}

/* cons2prim_user (rel_sph.cpp:304, rel_cart.cpp:216): adjust the recovered
 * density, pressure and four-velocity u[0..nd) before the pressure floor */
GH_USER_FN void gh_user_c2p(const double* par, int nd, double* rho, double* p, double* u) {
  (void)par; (void)nd; (void)rho; (void)p; (void)u;
}

#endif
