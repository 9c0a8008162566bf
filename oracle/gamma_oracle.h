/*
 * gamma_oracle.h -- CPU oracle for GAMMA's per-step hydro hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (PLM reconstruction, HLLC on moving faces, Brent cons2prim, RK3
 * with moving faces, CFL dt, shock detection, local synchrotron update, host
 * AMR split/merge).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the reference of the parity
 * checks or the CPU baseline.
 * The product (gamma_amd/) never links or calls it.
 *
 * Every function in gamma_oracle.c cites the reference file:line it follows.
 */
#ifndef GAMMA_ORACLE_H_
#define GAMMA_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

#define OG_QMAX 10
#define OG_DMAX 2

/* variable indices (reference environment.h:33-35, 70-82) */
enum { OG_RHO = 0, OG_PPP = 1, OG_UU1 = 2, OG_UU2 = 3 };
enum { OG_DEN = 0, OG_TAU = 1, OG_SS1 = 2, OG_SS2 = 3 };
enum { OG_TR1 = 4, OG_GMN = 6, OG_GMX = 7, OG_GMD = 8, OG_PSN = 9 };

enum { OG_GEOM_CART = 0, OG_GEOM_SPH = 1 };
/* hook kinds; the USER kinds call the user setup's functions
   (gamma_amd/csrc/user_hooks.h, or the file named by OG_USER_HOOKS) */
enum { OG_KIN_NONE = 0, OG_KIN_FIXED = 1, OG_KIN_USER = 2 };   /* userKinematics kinds   */
enum { OG_BND_NONE = 0, OG_BND_BM = 1, OG_BND_USER = 2 };      /* userBoundaries kinds   */
enum { OG_RG_NONE = 0, OG_RG_AR2D = 1, OG_RG_BM1D = 2, OG_RG_USER = 3 }; /* checkCellForRegrid */

typedef struct {
  int    nd;        /* NUM_D: 1 or 2 */
  int    geom;      /* OG_GEOM_CART (rel_cart + cartesian) or OG_GEOM_SPH (rel_sph + spherical[1D]) */
  int    nq;        /* NUM_Q = 4 + NUM_TR(2) + NUM_S(4 if synchrotron) */
  int    shock;     /* SHOCK_DETECTION_ enabled */
  int    sync;      /* LOCAL_SYNCHROTRON_ enabled */
  int    ngst;      /* ghost cells per side (2) */
  int    nmax;      /* max cells per track in MV (allocation: nmax + 2 ngst) */
  int    ncell_mv;  /* initial active cells per track */
  int    ncell_f1;  /* active tracks (2D); 1 in 1D */
  double gamma;     /* GAMMA_ */
  double cfl;       /* CFL_ */
  double vi;        /* VI */
  double shock_thr; /* DETECT_SHOCK_THRESHOLD_ */
  double gamma_max_init; /* GAMMA_MAX_INIT_ */
  double Nalpha, Nmp, Nme;  /* normalised constants (constants.cpp:12-20) */
  /* hooks */
  int    kin_mode;  double v_in, v_out;        /* userKinematics */
  int    bnd_mode;  double bm[6];              /* userBoundaries: BM params */
  int    rg_mode;   double rg[6];              /* checkCellForRegrid params */
  int    c2p_maxiter;                          /* 1000 in the reference */
  int    timestep;    /* TIMESTEP_: 0 RK3, 1 Euler (euler.cpp:12-19) */
  int    recon;       /* SPATIAL_RECONSTRUCTION_: 0 piecewise linear, 1 piecewise constant */
  int    vpspec;      /* VARIABLE_PSPEC_ (radiation_sph.cpp:18-35) */
  int    eos;         /* EOS_: 0 ideal, 1 Synge (rel_sph.cpp:42-50) */
  int    circ_regrid; /* CIRC_REGRID_ (1d.cpp:532-551, 2d.cpp:658-681): every split
                         paired with a merge of the track's smallest cell, every
                         merge with a split of its biggest */
  double c2p_tol;     /* Brent interval tolerance, 1e-13 in the reference (rel_sph.cpp:281);
                         other values only measure how much the reference's own
                         tolerance moves its trajectories */
  int    c2p_form;    /* 0: the reference's residual E + p - D W - G/(G-1) p W^2;
                         1: the same function written tau + p - D (W - 1) - G/(G-1) p W^2,
                         which rounds differently;
                         2: the reference's residual in long double (E = D + tau unrounded),
                         with c2p_tol 1e-16 the root to the last place of a double:
                         1 and 2 only measure how far rounding inside the residual
                         moves the reference's trajectories */
  double dt_fixed;    /* > 0: every step advances by this dt instead of CFL_ * collect_dt
                         (collect_dt still runs and resets dt_loc); pins the time
                         integrator's order at a fixed step.  0: the reference's rule */
  double upar[16];    /* parameters of the user setup's hooks */
  int    user_cell;   /* the user setup's per-cell hooks: bit 0 userSourceTerms,
                         bit 1 cons2prim_user (the reference always calls them;
                         its shipped setups leave them empty) */
} og_config;

typedef struct og_grid og_grid;

/* lifecycle (reference simu.cpp:24-38, 41-78) */
og_grid *og_create(const og_config *cfg);
void     og_destroy(og_grid *g);
/* Load the initial active-cell data (what Grid::initialGeometry/initialValues
 * write into Cinit): x,dx [ncell_f1][ncell_mv] in MV, y,dy [ncell_f1] in F1,
 * prim [nq][ncell_f1][ncell_mv] with 3-velocities in the UU slots. */
int      og_set_track_cells(og_grid *g, const int *ncells);
int      og_set_track_bm(og_grid *g, const double *bm);
void     og_pow_batch(const double *x, const double *y, double *out, long n);
void     og_fn_batch(int fn, const double *x, double *out, long n);   /* 0 tanh, 1 log10 */
int      og_load(og_grid *g, const double *x, const double *dx,
                 const double *y, const double *dy, const double *prim);
/* prepForRun + radiation_initGammas + the first updateGhosts of Simu::run */
int      og_init(og_grid *g, double t0);
/* nsteps iterations of the Simu::run loop body; returns 0 or an error code */
int      og_run(og_grid *g, int nsteps);
/* threads the track loops run on (OpenMP), and setting them (results do not
 * depend on the count) */
int      og_threads(void);
void     og_set_threads(int n);
/* grids of fewer cells (tracks x slots) than this run their loops on one thread */
void     og_set_par_min(long cells);

/* state access */
int      og_ntracks(const og_grid *g);     /* nde_nax[F1] (1 in 1D) */
int      og_nax(const og_grid *g);         /* nde_nax[MV] */
void     og_track_info(const og_grid *g, int *ntrack, int *nact, int *iL, int *iR);
double   og_time(const og_grid *g);
long     og_iter(const og_grid *g);
double   og_last_dt(const og_grid *g);
/* field ids for og_get_cells: 0..nq-1 prim, 100+q cons, then: */
enum { OG_F_X = 200, OG_F_DX, OG_F_Y, OG_F_DY, OG_F_DV, OG_F_CEN, OG_F_DTLOC,
       OG_F_SD, OG_F_SHOCKED, OG_F_PSPEC, OG_F_GRADMV0 = 300, OG_F_GRADF10 = 400 };
void     og_get_cells(const og_grid *g, int field, double *out /* [nj][nax] */);
enum { OG_I_X = 0, OG_I_V, OG_I_LL, OG_I_LR, OG_I_LS, OG_I_DA, OG_I_FLUX0 = 10 };
void     og_get_faces(const og_grid *g, int field, double *out /* [nj][nax-1] */);
/* counters: [0] c2p calls, [1] Brent iterations, [2] no-bracket, [3] max-iter,
 * [4] splits, [5] merges, [6] p_hi search cap hit, [7] tangled faces */
void     og_counters(const og_grid *g, long *out /* [8] */);
/* per-track split/merge counts [nj] */
void     og_regrid_counts(const og_grid *g, long *splits, long *merges);

/* ---- pieces of a step, for the oracle's own pin tests ----
 * og_prep: the step's first half as Simu::run runs it before collect_dt
 * (updateGhosts, prepForUpdate: neighbours, MV faces, kinematics, fluxes).
 * og_update_ghosts: Grid::updateGhosts alone.  Both return 0 or an error code. */
int      og_prep(og_grid *g);
int      og_update_ghosts(og_grid *g);
/* F1 neighbour list of cell (j, i): out = {j_down, first, count, j_up, first, count};
 * the list is the cells first .. first+count-1 of track j_down / j_up */
void     og_get_neighbors(const og_grid *g, int j, int i, int *out /* [6] */);
/* Grid::split / Grid::merge of cell (j, i) as applyRegrid calls them, and the
 * setup's checkCellForRegrid decision for (j, i): 0 skip, 1 merge, 2 split */
void     og_split(og_grid *g, int j, int i);
void     og_merge(og_grid *g, int j, int i);
int      og_regrid_decision(og_grid *g, int j, int i);
/* Record, after each prepForUpdate of a step (stage 1, 2, 3), the faces'
 * v, lL, lR and the cells' dl[MV]: og_get_stage(g, stage, field, out) with
 * field 0 v, 1 lL, 2 lR ([nj][nax-1]), 3 dl ([nj][nax]). */
int      og_record_stages(og_grid *g, int on);
int      og_get_stage(const og_grid *g, int stage, int field, double *out);

/* ---- pointwise entry points used by the oracle's pin tests ---- */
void og_prim2cons(const og_config *c, const double *prim, double r, double *cons);
void og_state2flux(const og_config *c, const double *prim, const double *cons,
                   double r, double *flux /* [nd][nq] */);
/* cons2prim of one state: prim_io holds the guess in, result out; cons_io is
 * round-tripped.  Returns Brent iterations (>=0) or -1 no bracket. */
int  og_cons2prim(const og_config *c, double *prim_io, double *cons_io, double r);
double og_c2p_residual(const og_config *c, double p, double D, double S, double E);
/* One face: states SL/SR given as prims (cons/flux computed inside with r=xf),
 * dim 0/1, face position xf (r for sph), face speed v, area dA.  Outputs
 * lL,lR,lS and flux[nq]. */
void og_face_flux(const og_config *c, int dim, const double *primL,
                  const double *primR, double xf, double v, double dA,
                  double *lam /* [3] */, double *flux);
double og_compute_sd(const og_config *c, const double *prim1,
                     const double *prim2, int dim, int reverse, double *pspec);
/* geometry of one 1D/2D cell: out = {dV, dl0, dl1, cen0, cen1} */
void og_cell_geom(const og_config *c, double x, double dx, double y, double dy,
                  double *out);
double og_face_area(const og_config *c, int dim, double x, double y, double dxf);
/* radiativeSourceTerms increment of cons[GMX] and cons[GMN] for one cell */
double og_sync_source(const og_config *c, const double *prim, double dV, double dt);
/* radiation_injectParticles on one cell (prim in/out) */
void og_inject(const og_config *c, double *prim, int shocked, double pspec);

#ifdef __cplusplus
}
#endif
#endif
