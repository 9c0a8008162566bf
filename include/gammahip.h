/*
 * gammahip.h -- C-ABI of the MI355X (gfx950) backend for GAMMA's per-step
 * hydro hot path.
 *
 * The reference (eliotayache-nv/GAMMA) binds its hot path at link time: the
 * driver loop Simu::run (src/simu.cpp:41-78) calls Grid methods declared in
 * src/grid.h:61-112 and defined by the Dimensions/Timestep/Hydro/Solver/
 * Geometry/Radiation module chosen in the Makefile (Makefile:1-8, 82-92).
 * This header is the C boundary those Grid methods call when the HIP backend
 * is selected: plain pointers and sizes, an opaque context, int status codes
 * (0 ok, <0 error; message from gammahip_last_error), no exceptions across
 * the ABI.  One host thread per context.  Calls are stream-ordered on the
 * context's stream; the ones documented "synchronous" wait for it.
 *
 * Data layout of host buffers: structure of arrays over tracks.  A "cell
 * field" buffer is [ntracks][nax] doubles (nax = nmax + 2*ngst, ntracks =
 * nde_nax[F1] including ghost tracks in 2D, the number of independent runs in
 * 1D), a "face field" buffer [ntracks][nax] (entry nax-1 unused), matching the
 * reference's Ctot[j][i] / Itot[j][i] indexing (src/Dimensions/2d.cpp:90-94).
 */
#ifndef GAMMAHIP_H_
#define GAMMAHIP_H_

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gammahip_ctx gammahip_ctx;

/* Compile-time environment of the reference (src/environment.h:46-66) plus
 * the setup's hook selection (src/Initial/Tests/ setups) and backend options. */
typedef struct gammahip_config {
  int    nd;          /* NUM_D: 1 (independent tracks) or 2 (coupled tracks)          */
  int    nq;          /* NUM_Q: 6, or 10 with LOCAL_SYNCHROTRON_                      */
  int    geom;        /* 0 cartesian (rel_cart + cartesian), 1 spherical (rel_sph +   */
                      /* spherical[1D]), -1 per track (1D ensembles, see tgeom)       */
  int    shock;       /* SHOCK_DETECTION_                                             */
  int    sync;        /* LOCAL_SYNCHROTRON_ (requires nq == 10)                       */
  int    ngst;        /* ghost cells per side (2)                                     */
  int    nmax;        /* max active cells per track (Grid::initialise, 1d.cpp:60)     */
  int    ncell_mv;    /* initial active cells per track                               */
  int    ntracks_active; /* active tracks (2D: ncell[F1]; 1D: independent runs)       */
  double gamma, cfl, vi, shock_thr, gamma_max_init;
  double Nalpha, Nmp, Nme;               /* normalizeConstants (constants.cpp:12-20)  */
  int    rg_mode;     /* checkCellForRegrid: 0 none, 1 dx/dy aspect ratio (2DST.cpp:  */
                      /* 304-322), 2 Blandford-McKee rule (BM1D.cpp:277-319), 3 the  */
                      /* user setup's; the rule of tracks whose hooks.rg_mode is -1  */
  double rg[6];       /* rule parameters {split, merge, dth, eta, target, -}          */
  int    c2p_solver;  /* 0 Brent (reference algorithm), 1 safeguarded Newton          */
  int    c2p_guess;   /* 0 previous pressure (reference), 1 learned MLP guess         */
  int    c2p_maxiter; /* 1000 in the reference                                        */
  int    timestep;    /* TIMESTEP_: 0 RK3 (rk3.cpp:12-195), 1 Euler (euler.cpp:12-19) */
  int    recon;       /* SPATIAL_RECONSTRUCTION_: 0 piecewise linear, 1 piecewise      */
                      /* constant (1d.cpp:378-381, 2d.cpp:439-447)                    */
  int    vpspec;      /* VARIABLE_PSPEC_: spectral index from the shock's four-velocity */
                      /* (radiation_sph.cpp:18-35, 146-155); 0: the constant p_       */
  int    eos;         /* EOS_: 0 ideal (GAMMA_), 1 Synge (rel_sph.cpp:42-50, spherical */
                      /* module; Brent cons2prim only)                                */
  int    circ_regrid; /* CIRC_REGRID_ (1d.cpp:532-551, 2d.cpp:658-681): each split     */
                      /* pairs with a merge of the track's smallest cell, each merge  */
                      /* with a split of its biggest (Grid::targetRegridVictims,      */
                      /* 1d.cpp:498-520); a victim equal to the acting cell, the      */
                      /* reference's exit(10), is counted in gammahip_stats.          */
                      /* regrid_clash and ends that track's regrid                    */
} gammahip_config;

/* Per-track hook parameters (userKinematics / userBoundaries /
 * checkCellForRegrid of the setup).  Per-track rules and initial lengths let
 * one context advance an ensemble of independent 1D runs of different setups
 * and sizes, or start a 2D grid from uneven tracks. */
typedef struct gammahip_track_hooks {
  int    geom;        /* used when config.geom == -1                                  */
  int    kin_mode;    /* 0 none, 1 fixed boundary-face speeds (BM1D.cpp:239-244),
                         2 the user setup's gh_user_face_speed                        */
  double v_in, v_out;
  int    bnd_mode;    /* 0 outflow only, 1 Blandford-McKee inflow (BM1D.cpp:254-274),
                         2 the user setup's gh_user_ghost                             */
  double bm[4];       /* {lfacShock2(t0), t0, rho_ambient, eta}                       */
  int    rg_mode;     /* this track's checkCellForRegrid rule (0 none, 1 aspect ratio
                         2DST.cpp:304-322, 2 BM1D.cpp:277-319, 3 the user setup's
                         gh_user_regrid); -1: config.rg_mode                          */
  int    ncell;       /* initial active cells of this track (2..ncell_mv); 0: ncell_mv */
} gammahip_track_hooks;
/* hook kinds (kin_mode, bnd_mode, rg_mode) */
enum { GAMMAHIP_KIN_NONE = 0, GAMMAHIP_KIN_FIXED = 1, GAMMAHIP_KIN_USER = 2 };
enum { GAMMAHIP_BND_OUTFLOW = 0, GAMMAHIP_BND_BM = 1, GAMMAHIP_BND_USER = 2 };
enum { GAMMAHIP_RG_NONE = 0, GAMMAHIP_RG_AR2D = 1, GAMMAHIP_RG_BM1D = 2, GAMMAHIP_RG_USER = 3 };

typedef struct gammahip_stats {
  unsigned long long c2p_calls, c2p_iters, c2p_nobracket, c2p_maxiter, c2p_phicap;
  unsigned long long negative_dt, nan_flux, f1_overflow, splits, merges, mlp_used;
  unsigned long long regrid_clash;   /* CIRC_REGRID_ victim equal to the acting cell */
  unsigned long long reserved[4];
} gammahip_stats;

/* Grid::Grid + Grid::initialise (src/Dimensions/1d.cpp:53-81, 2d.cpp:59-96):
 * allocates device storage for nmax cells per track. */
int  gammahip_create(const gammahip_config* cfg, gammahip_ctx** out);
/* Grid::destruct (1d.cpp:870-875) */
int  gammahip_destroy(gammahip_ctx* ctx);
const char* gammahip_last_error(gammahip_ctx* ctx);

/* Grid::initialGeometry/initialValues output + mpi_distribute + Grid::prepForRun
 * (grid.cpp:20-26) + apply(radiation_initGammas) (simu.cpp:34-36) + the first
 * Grid::updateGhosts of Simu::run (simu.cpp:43).  x,dx: [ntracks_active][ncell_mv];
 * y,dy: [ntracks_active] (F1 centre and width, 2D); prim: [nq][ntracks_active]
 * [ncell_mv] with 3-velocities in the UU slots (track a read up to its
 * hooks[a].ncell); hooks: [ntracks_active] in 1D and 2D; t0: start time.
 * Synchronous. */
int  gammahip_load(gammahip_ctx* ctx, const double* x, const double* dx, const double* y,
                   const double* dy, const double* prim, const gammahip_track_hooks* hooks,
                   double t0);

/* One or more iterations of the body of Simu::run (simu.cpp:47-77): regrid,
 * resetShock, updateGhosts, prepForUpdate, detectShock, dt = CFL*collect_dt,
 * evolve (rk3.cpp:12-195), t += dt, radiation_injectParticles.  dt stays on
 * the device.  Asynchronous. */
int  gammahip_step(gammahip_ctx* ctx, int nsteps);

/* The same loop body as separate Grid methods (src/grid.h:61-112). */
int  gammahip_regrid(gammahip_ctx* ctx);                       /* Grid::regrid        */
int  gammahip_update_ghosts(gammahip_ctx* ctx);                /* Grid::updateGhosts  */
int  gammahip_prep_for_update(gammahip_ctx* ctx);              /* Grid::prepForUpdate */
int  gammahip_collect_dt(gammahip_ctx* ctx, double* dt_out);   /* Grid::collect_dt * CFL_, synchronous */
int  gammahip_evolve(gammahip_ctx* ctx);                       /* Grid::evolve with the collected dt */
/* Grid::collect_dt as the reference returns it (src/Dimensions/2d.cpp:954-971,
 * 1d.cpp:779-795): the minimum of the cells' local time steps, before CFL_.
 * The device still sets its own dt = CFL * minimum, so Simu::run's
 * dt = CFL_ * collect_dt() (simu.cpp:61) gives the same bits.  One run per
 * context (2D, or 1D with one track).  Synchronous. */
int  gammahip_collect_dt_min(gammahip_ctx* ctx, double* min_out);
/* Grid::evolve(it, t, dt) with the caller's dt (src/Timestep/rk3.cpp:13,
 * 99, euler.cpp:12): dt replaces the device's dt of every track (a caller
 * that clamps dt is followed), then the RK3 (or Euler) update runs.  One run
 * per context; dt > 0 and finite.  Waits for the context's stream before
 * copying dt. */
int  gammahip_evolve_dt(gammahip_ctx* ctx, double dt);
enum { GAMMAHIP_RESET_SHOCK = 0, GAMMAHIP_DETECT_SHOCK = 1, GAMMAHIP_INJECT = 2,
       GAMMAHIP_INIT_GAMMAS = 3, GAMMAHIP_ADVANCE_TIME = 4 };
int  gammahip_apply(gammahip_ctx* ctx, int what);              /* Grid::apply(&Cell::...) */

/* Synchronous downloads (Grid::printCols reads the same state).  field:
 * 0..nq-1 primitive q, 100+q conserved q (prim2cons of the stored state),
 * 200 x, 201 dx, 202 y, 203 dy, 204 dV, 205 centroid, 206 dt_loc, 207 Sd,
 * 208 isShocked, 209 pspec, 210 F1 faces above the cell (2D), 300+q grad_mv,
 * 400+q grad_f1;
 * face fields: 1000 x, 1001 v, 1002 lL, 1003 lR, 1010+q flux. */
int  gammahip_download(gammahip_ctx* ctx, int field, double* out /* [ntracks][nax] */);
int  gammahip_track_info(gammahip_ctx* ctx, int* ntrack, int* nact, int* iLbnd, int* iRbnd);

/* Exact state between two steps, for resuming a run bit for bit: the time
 * and dt of every track, the track lengths, the primitive, position and width
 * planes of all cells, the face positions and velocities, the cells' local
 * time steps (the next step's CFL minimum still sees the last stage's), with
 * shock detection Sd, pspec and isShocked, and in 2D the gradients, F1
 * neighbour lists and F1 face values (the face pass reads the upper track's
 * transverse gradient of the previous stage, ghost tracks keep the lists of
 * earlier stages).
 * The reference resumes from its text snapshots (Simu::reinitialise,
 * src/IO/text1d.cpp:103-134, src/IO/text2d.cpp:108-145), which round the state
 * to the printed digits; this is the exact form of the same step.  Restore
 * into a context created with the same config and loaded with the same setup
 * and hooks (they carry the per-track boundary and regrid parameters); the
 * buffer's dimensions are checked against the context's.  Synchronous. */
int  gammahip_state_size(gammahip_ctx* ctx, unsigned long long* nbytes);
int  gammahip_state_save(gammahip_ctx* ctx, void* buf, unsigned long long nbytes);
int  gammahip_state_restore(gammahip_ctx* ctx, const void* buf, unsigned long long nbytes);
int  gammahip_dims(gammahip_ctx* ctx, int* ntracks, int* nax);
int  gammahip_time(gammahip_ctx* ctx, double* t /* [ntracks] */, double* dt /* [ntracks] */);
int  gammahip_stats_get(gammahip_ctx* ctx, gammahip_stats* st);
int  gammahip_regrid_counts(gammahip_ctx* ctx, long long* splits, long long* merges);
/* Average duration (ms) of the dominant kernel launches since the last call,
 * measured with HIP events on the context's stream: entry 0 the update, 1 the
 * MV face pass, 2 the F1 face pass, 3 the transverse-gradient pass, 4 the F1
 * flux sums, 5 the fused gradient-and-face pass. */
int  gammahip_kernel_times(gammahip_ctx* ctx, double* ms_per_launch /* [16] */,
                           long long* launches /* [16] */);
int  gammahip_sync(gammahip_ctx* ctx);

/* Launch options (no counterpart in the reference; chosen by the autotuner in
 * gamma_amd/autotune.py): threads per workgroup of the per-cell kernels (64,
 * 128, 256; default 128), the order of the (track, chunk) tiles (REMAP 0
 * hardware order; 1 XCD-aware bands of tracks; 2 track groups, the default:
 * each workgroup takes one 64-cell chunk of block/64 consecutive tracks, one
 * track per wavefront), replay of
 * two-step hipGraphs by gammahip_step, and n for the learned guess's trigger:
 * a wavefront runs the network when one of its cells' previous pressures
 * leaves a residual |f| above 10^-n times the magnitude of f's terms
 * (default 3), and the packed launch: n > 0 runs the per-cell kernels as
 * persistent grids of at most n workgroups per CU (fewer where a kernel's
 * registers or LDS let a CU hold fewer) walking a list of (track, chunk)
 * tiles that covers only each track's live cells (uneven tracks, ensembles),
 * 0 launches one workgroup per chunk of the nmax-wide track slots; OCC bit 0
 * runs the F1 face kernel in a build capped at 128 registers (4 wavefronts
 * per SIMD) with six fields, at 168 (3) with ten, bit 1 the update kernel in
 * one capped at 96 (5 wavefronts; the default build of 2D Cartesian six-field
 * grids), spilling what does not fit; F1FUSE 1 sums each cell's F1 face fluxes inside
 * the update kernel instead of in a pass of its own (fewer bytes, a longer
 * dependent-load chain in the update); SOLVER switches cons2prim between
 * Brent (0, the reference's algorithm) and Newton (1; required while the
 * learned guess is loaded); F1CELL 1 computes each cell's transverse gradient
 * and then the fluxes of its own F1 faces in one pass, the owner's state kept
 * in registers, instead of a gradient pass and a one-thread-per-face pass. */
enum { GAMMAHIP_OPT_BLOCK = 0, GAMMAHIP_OPT_REMAP = 1, GAMMAHIP_OPT_GRAPH = 2,
       GAMMAHIP_OPT_MLP_THR = 3, GAMMAHIP_OPT_PERSIST = 4, GAMMAHIP_OPT_OCC = 5,
       GAMMAHIP_OPT_F1FUSE = 6, GAMMAHIP_OPT_SOLVER = 7, GAMMAHIP_OPT_F1CELL = 8,
       GAMMAHIP_OPT_FUSE = 9, GAMMAHIP_OPT_USER_CELL = 10 };
/* GAMMAHIP_OPT_FUSE (bit set, 2D grids): bit 0 runs the F1 neighbour search
 * inside the transverse-gradient pass (one kernel instead of two); bit 1 has
 * the F1 passes recompute the MV gradients from the primitives instead of
 * reading them; bit 2 forms the F1 faces with one thread per owner cell
 * instead of one per face slot.  Every combination gives the same bits.
 * GAMMAHIP_OPT_USER_CELL (bit set): the user setup's per-cell hooks, called
 * where the reference calls them -- bit 0 userSourceTerms (cell.cpp:112, in
 * every cell update), bit 1 cons2prim_user (rel_sph.cpp:304,
 * rel_cart.cpp:216, in every cons2prim of the updates and merges); 0 (the
 * default) leaves them out, as the reference's shipped setups leave them
 * empty. */
int  gammahip_set_option(gammahip_ctx* ctx, int key, long long value);

/* Parameters of the user setup's hooks (gamma_amd/csrc/user_hooks.h, or the
 * file named by GAMMAHIP_USER_HOOKS at build time): the link-time setup file
 * of the reference (a file of src/Initial/Tests/ defining userKinematics,
 * userBoundaries and checkCellForRegrid, selected by Makefile:82-92), for
 * tracks whose kin_mode, bnd_mode or rg_mode selects the user kind.  par[0..n)
 * with n <= 16; the rest read 0.  Synchronous; keeps its values until the
 * next call (0 until the first). */
int  gammahip_set_user_params(gammahip_ctx* ctx, const double* par, int n);

/* ---- learned cons2prim guess (no counterpart in the reference) ----
 * A 3 -> 64 -> 64 -> 1 MLP predicting log p of a conserved state from
 * (log D, |S|/D, tau/D); parameters are one float array of
 * gammahip_mlp_nparams() entries (weights, then the fixed feature/target
 * normalisation).  states: [n][4] = (D, |S|, tau, p) for training,
 * [n][3] = (D, |S|, tau) for inference.  Training runs FP32 forward/backward
 * kernels with Adam on minibatches drawn with replacement; inference runs the
 * hidden layer as BF16 MFMA with FP32 accumulation (mfma = 1) or FP32 (0). */
int  gammahip_mlp_nparams(void);
int  gammahip_mlp_train(const double* states, long long n, float* params, int steps, int batch,
                        float lr, unsigned long long seed, float* loss_hist /* [steps] */);
int  gammahip_mlp_grad(const float* params, const double* states, int n, float* grad, float* loss);
int  gammahip_mlp_infer(const float* params, const double* states, long long n, double* p_out,
                        int mfma);
/* gammahip_mlp_infer without any allocation (no counterpart in the reference;
 * for callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops).  params [gammahip_mlp_nparams()], states
 * [n][3] and p_out [n] are DEVICE pointers; the call is asynchronous on
 * stream (a hipStream_t; null = the default stream) and returns the bits of
 * gammahip_mlp_infer.  The packed BF16 hidden-layer operands go into a fixed
 * 16 KiB device array of the library (one per device, held until the process
 * ends; 8 KiB used, whatever n), fully rewritten by every call: calls that
 * overlap on one device race on it.  Rejects null pointers, mfma not 0/1 and
 * n outside [1, 64 * INT_MAX] with -1 and a message on stderr. */
int  gammahip_mlp_infer_noalloc(const float* params, const double* states, long long n, double* p_out,
                                int mfma, void* stream);
int  gammahip_mfma_probe(const float* A /* 32x16 */, const float* B /* 16x32 */, float* D /* 32x32 */);
/* Cartesian cons2prim of raw states [n][3] = (D, |S|, tau) started from
 * guess[n] (FluidState::cons2prim, src/Hydro/rel_cart.cpp:125-229, with the
 * previous pressure replaced by the given start).  guess2 (nullable): a second
 * start; each state then starts from the one with the smaller |f| and
 * used[s] = 1 when guess[s] won.  iters[s] counts residual evaluations
 * (root-finder iterations plus bracketing), fail[s] the failure bits
 * (1 no bracket, 2 iteration cap, 4 bracket-growth cap). */
/* x^y exactly as the kernels evaluate it (double-double exp(y log x) from
 * IEEE-exact operations, identical bits on host and device; no counterpart
 * in the reference, which calls the C library's pow), for numerics checks. */
int  gammahip_pow_batch(const double* x, const double* y, double* out, long long n);
/* tanh (fn 0) and log10 (fn 1) as the kernels evaluate them (the shock
 * detector's, radiation_sph.cpp:18-35, 90): double-double, identical on host
 * and device.  No counterpart in the reference; for numerics checks. */
int  gammahip_fn_batch(int fn, const double* x, double* out, long long n);
/* Seed the context's Newton cons2prim with the network (params = null
 * switches it off).  The previous pressure stays the arbiter: each solve starts
 * from whichever start has the smaller residual; stats.mlp_used counts the
 * solves that started from the network. */
int  gammahip_mlp_load(gammahip_ctx* ctx, const float* params);
int  gammahip_c2p_batch(const double* states, const double* guess, const double* guess2, long long n,
                        double gamma, int solver, int maxit, double* p_out, int* iters, int* fail,
                        int* used);

#ifdef __cplusplus
}
#endif
#endif
