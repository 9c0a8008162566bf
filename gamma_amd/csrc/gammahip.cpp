// gammahip.cpp -- MI355X (gfx950) backend of GAMMA's per-step hydro hot path:
// device storage, kernels, stage/step driver and the C-ABI of include/gammahip.h.
//
// Layout in HBM: structure of arrays over tracks.  A cell field is a
// [ntracks][nax] plane (nax = nmax + 2*ngst), a multi-variable field is
// [nvar][ntracks][nax]; consecutive cells of a track are consecutive doubles,
// so every per-cell kernel reads and writes fully coalesced 64-bit lanes.
// In one dimension the unused momentum slot (SS2/UU2) is not stored.
//
// Kernel bodies are __host__ __device__ functions of (track j, index i); the
// __global__ wrappers map blockIdx.y -> track and blockIdx.x*blockDim.x +
// threadIdx.x -> cell.  Built with -DGH_EMU the same bodies run in plain loops
// on the CPU (test tool for index checks); the product library is the HIP build.
#include "hydro.hpp"
#include "mlp.hpp"
#include "../../include/gammahip.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <climits>
#include <string>
#include <vector>
#include <algorithm>
#include <map>
#include <mutex>
#include <utility>

#ifdef GH_EMU
#define __global__
#define __shared__ static
#define GH_ATOMIC_ADD(p, v) (*(p) += (v))
#define GH_ATOMIC_MIN(p, v) (*(p) = (*(p) < (v) ? *(p) : (v)))
#define GH_ATOMIC_MAX(p, v) (*(p) = (*(p) > (v) ? *(p) : (v)))
#else
#define GH_ATOMIC_ADD(p, v) atomicAdd((p), (v))
#define GH_ATOMIC_MIN(p, v) atomicMin((p), (v))
#define GH_ATOMIC_MAX(p, v) atomicMax((p), (v))
#endif

using namespace gh;

namespace {

enum { CNT_C2P_CALLS = 0, CNT_C2P_ITERS, CNT_NOBRACKET, CNT_MAXITER, CNT_PHICAP, CNT_NEGDT,
       CNT_NANFLUX, CNT_F1OVF, CNT_SPLITS, CNT_MERGES, CNT_MLP, CNT_RGCLASH, CNT_N = 16 };
// Counters live in CNT_PARTS copies of one 128-byte line each; a workgroup adds
// into copy (linear workgroup id mod CNT_PARTS) and the host sums the copies.
// Device-wide atomics to one address serialise at the memory side (the XCDs
// have separate L2s), so a single copy would cost every kernel that counts on
// each wavefront (the update counts every cons2prim) a queue of ~10^5
// same-address atomics per launch.
constexpr int CNT_PARTS = 1024;
enum { KIN_NONE = 0, KIN_FIXED = 1, KIN_USER = 2 };
enum { BND_NONE = 0, BND_BM = 1, BND_USER = 2 };
enum { RG_NONE = 0, RG_AR2D = 1, RG_BM1D = 2, RG_USER = 3 };

constexpr int DT_CH = 1024;   // cells per workgroup of the per-track dt minimum

// ---------------------------------------------------------------------------
// device view (passed by value to every kernel)
struct Dev {
  int nd, nq, nqs, ngst, nj, nax, jL, jR, fcap, shock, sync, solver, guess, maxit;
  int euler;               // TIMESTEP_ Euler (one update per step) instead of RK3
  int pc;                  // piecewise-constant reconstruction instead of PLM
  int vpspec;              // VARIABLE_PSPEC_: spectral index from the shock's four-velocity
  int rg_mode;             // bit set of the tracks' regrid rules (1 << RG_*)
  int circ;                // CIRC_REGRID_: splits and merges in pairs with the track's victims
  long long cs;            // nj * nax
  long long fs1;           // nj * fcap
  Phys P;
  double rg[6];
  // per track
  int *ntrack, *tgeom, *tkin, *tbnd, *tfirst, *tiS, *trg;
  int *tlast;              // last cell of each track the regrid screen flagged
  unsigned char* rgflag;   // [nj][nax] the regrid screen's decision per active cell (1: an admissible action)
  double *tvin, *tvout, *tbm, *ty, *tdy, *tdt, *tt, *tmin;
  double *tpart;           // [nj][dtch] dt minimum of each chunk of DT_CH cells of each track
  int dtch;                // chunks per track: ceil(nax / DT_CH)
  // cells
  double *prim, *cx, *cdx, *dtloc, *U0, *dV0, *gmv, *gf1, *gf1n, *sd, *pspec;
  int *shk, *nbdj, *nbdlo, *nbdcnt, *nbuj, *nbulo, *nbucnt, *f1off, *f1tot;
  int *f1own;              // [nj][fcap] cell owning each F1 face slot
  // faces
  double *fx, *fxn, *fx0, *fv, *flL, *flR, *fflux;
  double *ff1, *ff1sd;
  double *f1net;           // [NQ][nj][nax] net F1 flux (dn - up) of each cell, this stage
  int f1fused;             // 1: the update kernel sums the F1 fluxes itself (no k_f1sum pass)
  int gmvre;               // 1: the F1 passes recompute MV gradients from the primitives (gmv_at)
  unsigned long long* cnt;
  long long *nsplit, *nmerge;
  // learned cons2prim guess: FP32 parameters and the packed BF16 hidden-layer
  // operands (null when the network is not loaded)
  const float* mw;
  const void* mw2b;
  double mlp_thr;          // relative residual of the previous pressure that asks for the network
  // packed launch: (track, chunk) tiles covering only the live cells of every
  // track, walked by a persistent grid (0: one workgroup per chunk of nax)
  int packed;
  int* tiles;              // [cap][2] = (track, chunk)
  int* ntiles;             // [1]
  double* ttrig;           // [nj][7] Trig of each track (2D spherical geometry)
  double* upar;            // [GH_USER_NPAR] parameters of the user hooks
  int ucell;               // per-cell user hooks: bit 0 userSourceTerms, bit 1 cons2prim_user
};

GH_HD long long CI(const Dev& d, int j, int i) { return (long long)j * d.nax + i; }

// the track as the user hooks see it
GH_HD gh_user_track user_track(const Dev& d, int j) {
  gh_user_track T;
  T.j = j; T.ntrack = d.ntrack[j]; T.ngst = d.ngst; T.geom = d.tgeom[j]; T.t = d.tt[j];
  T.y = d.ty[j]; T.dy = d.tdy[j]; T.par = d.upar;
  return T;
}

// this workgroup's copy of counter id
GH_HD unsigned long long* cnt_at(const Dev& d, int id) {
#ifdef GH_EMU
  return d.cnt + id;
#else
  int part = (int)((blockIdx.x + gridDim.x * blockIdx.y) & (CNT_PARTS - 1));
  return d.cnt + (long long)part * CNT_N + id;
#endif
}

template <int ND>
GH_HD int sq(int q) { return (ND == 1 && q > 3) ? q - 1 : q; }

template <int ND, int NQ>
GH_HD void load_v(const Dev& d, const double* a, long long k, double* v) {
  for (int q = 0; q < NQ; ++q) v[q] = (ND == 1 && q == 3) ? 0. : a[sq<ND>(q) * d.cs + k];
}
template <int ND, int NQ>
GH_HD void store_v(const Dev& d, double* a, long long k, const double* v) {
  for (int q = 0; q < NQ; ++q)
    if (!(ND == 1 && q == 3)) a[sq<ND>(q) * d.cs + k] = v[q];
}

// the track's Trig (zeros where the geometry does not use it)
template <int ND>
GH_HD Trig trig_at(const Dev& d, int j) {
  Trig t = {0., 0., 0., 0., 0., 0., 0.};
  if (ND == 2 && d.tgeom[j] == SPH) {
    const double* a = d.ttrig + 7 * (long long)j;
    t.s_hdy = a[0]; t.s_y = a[1]; t.c_y = a[2]; t.s_1 = a[3]; t.s_2 = a[4]; t.c_1 = a[5]; t.c_2 = a[6];
  }
  return t;
}
// the same with the geometry given (a template constant where the grid has
// one geometry: 2D grids), so that the kernel carries only that geometry's code
template <int ND>
GH_HD Trig trig_at_g(const Dev& d, int j, int geom) {
  Trig t = {0., 0., 0., 0., 0., 0., 0.};
  if (ND == 2 && geom == SPH) {
    const double* a = d.ttrig + 7 * (long long)j;
    t.s_hdy = a[0]; t.s_y = a[1]; t.c_y = a[2]; t.s_1 = a[3]; t.s_2 = a[4]; t.c_1 = a[5]; t.c_2 = a[6];
  }
  return t;
}
template <int ND>
GH_HD Geo geo_of_g(const Dev& d, int j, long long k, int geom) {
  return cell_geo<ND>(d.cx[k], d.cdx[k], d.ty[j], d.tdy[j], geom, trig_at_g<ND>(d, j, geom));
}
template <int ND>
GH_HD Geo geo_of(const Dev& d, int j, long long k) {
  return cell_geo<ND>(d.cx[k], d.cdx[k], d.ty[j], d.tdy[j], d.tgeom[j], trig_at<ND>(d, j));
}
// r argument of the grid sweeps: centroid in 1D (1d.cpp:915), centre in 2D (2d.cpp:1127)
template <int ND>
GH_HD double sweep_r(const Dev& d, long long k, const Geo& g) { return ND == 1 ? g.cen0 : d.cx[k]; }

// prim and cons of a cell as the prepForUpdate sweeps leave them; returns the radius its
// physical flux takes
template <int ND, int NQ>
GH_HD double cell_pc(const Dev& d, int j, int i, double* p, double* c) {
  long long k = CI(d, j, i);
  load_v<ND, NQ>(d, d.prim, k, p);
  Geo g = geo_of<ND>(d, j, k);
  double r = sweep_r<ND>(d, k, g);
  prim2cons<ND, NQ>(p, c, r, d.tgeom[j], d.P.gamma, d.P.synge);
  return r;
}

template <int ND, int NQ>
GH_HD double cell_pc_g(const Dev& d, int j, int i, double* p, double* c, int geom) {
  long long k = CI(d, j, i);
  load_v<ND, NQ>(d, d.prim, k, p);
  Geo g = geo_of_g<ND>(d, j, k, geom);
  double r = sweep_r<ND>(d, k, g);
  prim2cons<ND, NQ>(p, c, r, geom, d.P.gamma, d.P.synge);
  return r;
}

GH_HD bool active_track(const Dev& d, int j) { return d.nd == 1 || (j > d.jL && j < d.jR); }

// ---------------------------------------------------------------------------
// Grid::updateGhosts (1d.cpp:217-246, 2d.cpp:241-321)

// copy every persistent per-cell field of (js,is) into (j,i)
template <int ND>
GH_HD void copy_cell(const Dev& d, long long kd, long long ks) {
  for (int s = 0; s < d.nqs; ++s) d.prim[s * d.cs + kd] = d.prim[s * d.cs + ks];
  d.cx[kd] = d.cx[ks];
  d.cdx[kd] = d.cdx[ks];
  d.dtloc[kd] = d.dtloc[ks];
  if (ND == 2) {
    for (int s = 0; s < d.nqs; ++s) {
      d.gmv[s * d.cs + kd] = d.gmv[s * d.cs + ks];
      d.gf1[s * d.cs + kd] = d.gf1[s * d.cs + ks];
    }
    d.nbdj[kd] = d.nbdj[ks]; d.nbdlo[kd] = d.nbdlo[ks]; d.nbdcnt[kd] = d.nbdcnt[ks];
    d.nbuj[kd] = d.nbuj[ks]; d.nbulo[kd] = d.nbulo[ks]; d.nbucnt[kd] = d.nbucnt[ks];
  }
  if (d.shock) { d.sd[kd] = d.sd[ks]; d.pspec[kd] = d.pspec[ks]; d.shk[kd] = d.shk[ks]; }
}

// ghost tracks: whole-track copy of the edge active track (2d.cpp:246-289)
GH_HD void body_ghost_tracks(const Dev& d, int g, int i) {
  int j = g < d.ngst ? g : d.nj - d.ngst + (g - d.ngst);
  int js = j <= d.jL ? d.jL + 1 : d.jR - 1;
  int nts = d.ntrack[js];
  if (i == 0) d.ntrack[j] = nts;
  if (i < nts) copy_cell<2>(d, CI(d, j, i), CI(d, js, i));
  if (i < nts - 1) d.fx[CI(d, j, i)] = d.fx[CI(d, js, i)];
}
// the same copy without the ghost cells and faces along the track, which
// body_ghost_mv_src then forms from the edge track directly (one pass for
// both: the two write disjoint cells and read only the edge track's active
// cells)
GH_HD void body_ghost_track_inner(const Dev& d, int g, int i) {
  int j = g < d.ngst ? g : d.nj - d.ngst + (g - d.ngst);
  int js = j <= d.jL ? d.jL + 1 : d.jR - 1;
  int nts = d.ntrack[js];
  if (i == 0) d.ntrack[j] = nts;
  if (i >= d.ngst && i < nts - d.ngst) copy_cell<2>(d, CI(d, j, i), CI(d, js, i));
  if (i >= d.ngst - 1 && i <= nts - d.ngst - 1) d.fx[CI(d, j, i)] = d.fx[CI(d, js, i)];
}

// userBoundaries of a user setup on ghost cell (j, i): its primitives after
// the outflow copy, in the reference's slots, through gh_user_ghost
template <int ND, int NQ>
GH_HD void user_ghost(const Dev& d, int j, int nt, int i, long long kc) {
  double p[NQ];
  load_v<ND, NQ>(d, d.prim, kc, p);
  gh_user_track T = user_track(d, j);
  T.ntrack = nt;
  gh_user_ghost(&T, i, d.cx[kc], d.cdx[kc], p);
  store_v<ND, NQ>(d, d.prim, kc, p);
}

// ghost cells and faces along track j + userBoundaries (BM1D.cpp:254-274),
// formed from the edge cells of track js: j itself, or for a ghost track the
// edge track it copies (whose active cells it holds)
template <int ND, int NQ>
GH_HD void body_ghost_mv_src(const Dev& d, int j, int kk, int js) {
  if (kk >= 2 * d.ngst) return;
  int nt = d.ntrack[js], iL = d.ngst - 1, iR = nt - d.ngst;
  if (kk < d.ngst) {
    int i = kk;
    long long ke = CI(d, js, iL + 1), kc = CI(d, j, i);
    double dxe = d.cdx[ke];
    copy_cell<ND>(d, kc, ke);
    d.cx[kc] = d.cx[ke] - (iL - i + 1) * dxe;
    if (i <= iL - 1) d.fx[kc] = d.fx[ke] - (iL - i + 1) * dxe;
    if (d.tbnd[j] == BND_BM) {
      double rho, u, p;
      bm_profile(d.tbm + 4 * j, d.cx[kc], d.tt[j], d.P.gamma, &rho, &u, &p);
      d.prim[RHO * d.cs + kc] = rho;
      d.prim[UU1 * d.cs + kc] = u;
      if (ND == 2) d.prim[UU2 * d.cs + kc] = 0;
      d.prim[PPP * d.cs + kc] = p;
      d.prim[sq<ND>(TR1) * d.cs + kc] = 2.;
    }
    if (d.tbnd[j] == BND_USER) user_ghost<ND, NQ>(d, j, nt, i, kc);
  } else {
    int i = iR + (kk - d.ngst);
    long long ke = CI(d, js, iR - 1), kc = CI(d, j, i);
    double dxe = d.cdx[ke];
    copy_cell<ND>(d, kc, ke);
    d.cx[kc] = d.cx[ke] + (i - iR + 1) * dxe;
    if (i <= nt - 2) d.fx[kc] = d.fx[ke] + (i - iR) * dxe;   // (i-iRbnd): 1d.cpp:241
    if (d.tbnd[j] == BND_USER) user_ghost<ND, NQ>(d, j, nt, i, kc);
  }
}
template <int ND, int NQ>
GH_HD void body_ghost_mv(const Dev& d, int j, int kk) { body_ghost_mv_src<ND, NQ>(d, j, kk, j); }

// ---------------------------------------------------------------------------
// Grid::movDir_ComputeLambda + updateKinematics + computeFluxes, MV faces
// (1d.cpp:370-461, 731-776; 2d.cpp:451-490, 835-888; hllc.cpp:12-39)
template <int ND, int NQ, bool SHK>
GH_HD void body_faces_s(const Dev& d, int j, int i) {
  int nt = d.ntrack[j];
  if (i > nt - 2) return;
  int geom = d.tgeom[j];
  double G = d.P.gamma;
  long long k = CI(d, j, i);
  double xf = d.fx[k];
  double pL[NQ], cL[NQ], pR[NQ], cR[NQ];
  // side fluxes: at the track ends the cells' own states (sweep radius), inside
  // the reconstructed states at the face
  bool edge = d.pc || i == 0 || i == nt - 2;     // piecewise constant: 1d.cpp:378-381
  double rL = xf, rR = xf;
  if (edge) {
    rL = cell_pc<ND, NQ>(d, j, i, pL, cL);
    rR = cell_pc<ND, NQ>(d, j, i + 1, pR, cR);
  } else {
    double p0[NQ], p3[NQ];
    load_v<ND, NQ>(d, d.prim, k - 1, p0);
    load_v<ND, NQ>(d, d.prim, k, pL);
    load_v<ND, NQ>(d, d.prim, k + 1, pR);
    load_v<ND, NQ>(d, d.prim, k + 2, p3);
    double x0 = geo_of<ND>(d, j, k - 1).cen0, x1 = geo_of<ND>(d, j, k).cen0;
    double x2 = geo_of<ND>(d, j, k + 1).cen0, x3 = geo_of<ND>(d, j, k + 2).cen0;
    double gLs[NQ], gRs[NQ];
    for (int q = 0; q < NQ; ++q) {
      double gl = (pL[q] - p0[q]) / (x1 - x0);
      double gn = (pR[q] - pL[q]) / (x2 - x1);
      double gr = (p3[q] - pR[q]) / (x3 - x2);
      gLs[q] = minmod(gl, gn);
      gRs[q] = minmod(gr, gn);
    }
    if (ND == 2) {
      store_v<ND, NQ>(d, d.gmv, k, gLs);
      if (i == nt - 3) store_v<ND, NQ>(d, d.gmv, k + 1, gRs);
    }
    for (int q = 0; q < NQ; ++q) {
      pL[q] = pL[q] + gLs[q] * (xf - x1);
      pR[q] = pR[q] + gRs[q] * (xf - x2);
    }
    prim2cons<ND, NQ>(pL, cL, xf, geom, G, d.P.synge);
    prim2cons<ND, NQ>(pR, cR, xf, geom, G, d.P.synge);
  }
  Waves w = lambdas<ND, NQ>(pL, cL, pR, cR, 0, xf, geom, G);
  double v = d.P.vi * w.lS;
  if (d.tkin[j] == KIN_FIXED) {
    for (int n = 0; n < d.ngst; ++n) {
      if (i == n) v = d.tvin[j];
      if (i == nt - 2 - n) v = d.tvout[j];
    }
  } else if (d.tkin[j] == KIN_USER) {
    gh_user_track T = user_track(d, j);
    v = gh_user_face_speed(&T, i, xf, v);
  }
  Trig T = trig_at<ND>(d, j);
  double dA = face_area<ND>(0, xf, d.tdy[j], geom, T.s_y, T.s_hdy);
  double fl[NQ];
  hllc_flux_lazy<ND, NQ>(pL, cL, pR, cR, w, v, dA, 0, xf, geom, [&](bool right, double* f) {
    if (right) state2flux<ND, NQ>(pR, cR, f, 0, rR, geom);
    else state2flux<ND, NQ>(pL, cL, f, 0, rL, geom);
  }, fl);
  if (fl[0] != fl[0]) GH_ATOMIC_ADD(cnt_at(d, CNT_NANFLUX), 1ull);
  d.fv[k] = v;
  d.flL[k] = w.lL;
  d.flR[k] = w.lR;
  store_v<ND, NQ>(d, d.fflux, k, fl);
  if (SHK) {
    double ps = -1.e300;
    double Sd = shock_sd<ND, NQ>(pL, pR, 0, false, G, &ps, d.vpspec);
    d.sd[k] = fmax(d.sd[k], Sd);
    if (ps > d.pspec[k]) d.pspec[k] = ps;
  }
}

template <int ND, int NQ>
GH_HD void body_faces(const Dev& d, int j, int i) {
  if (d.shock) body_faces_s<ND, NQ, true>(d, j, i);
  else body_faces_s<ND, NQ, false>(d, j, i);
}

// Cell::update_dt, MV candidates (cell.cpp:47-99)
// the smaller of the dt candidates cell k's two MV faces give it
GH_HD double dt_cand(const Dev& d, long long k) {
  double a = d.cdx[k];
  double l, v, dtL, dtR;
  l = d.flR[k - 1]; v = d.fv[k];
  if (l > v) dtL = a / (l - v); else dtL = 1.e15;
  l = d.flL[k]; v = d.fv[k - 1];
  if (l < v) dtR = a / (v - l); else dtR = 1.e15;
  return fmin(dtL, dtR);
}
GH_HD void body_dt(const Dev& d, int j, int i) {
  int nt = d.ntrack[j];
  if (i < 1 || i > nt - 2) return;
  long long k = CI(d, j, i);
  double cand = dt_cand(d, k);
  if (cand < 0) GH_ATOMIC_ADD(cnt_at(d, CNT_NEGDT), 1ull);
  d.dtloc[k] = fmin(d.dtloc[k], cand);
}

// ---------------------------------------------------------------------------
// 2D transverse (F1) coupling (2d.cpp:366-430, 491-527, 534-606, 890-950)

// First t in [0, n) with X(t) >= xr (n if none) for non-decreasing X(t),
// read for t in [0, n) only.  Galloping search from the hint h (adjacent
// tracks have nearly the same structure, so the answer is usually within a
// few cells of the cell's own index), then bisection inside the bracket
// found: the same lower bound as a full binary search.  On the faces of a
// track (X = face positions, n = nt - 1): the first face index k in
// [0, nt-2] with x >= xr, nt-1 if none.
template <class XF, class T>
GH_HD int lb_x(int n, const XF& X, T xr, int h) {
  if (h < 0) h = 0;
  if (h > n) h = n;
  int lo, hi;
  if (h < n && X(h) < xr) {                  // answer right of h
    lo = h + 1;
    int step = 1;
    hi = lo;
    while (hi < n && X(hi) < xr) { lo = hi + 1; hi += step; step <<= 1; }
    if (hi > n) hi = n;
  } else {                                   // answer at or left of h
    hi = h;
    int step = 1;
    lo = h - 1;
    while (lo >= 0 && X(lo) >= xr) { hi = lo; lo -= step; step <<= 1; }
    lo = lo < 0 ? 0 : lo + 1;
  }
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (X(mid) >= xr) hi = mid; else lo = mid + 1;
  }
  return lo;
}
GH_HD int lb_face(const Dev& d, int ja, double xr, int h) {
  const double* X = d.fx + (long long)ja * d.nax;
  return lb_x(d.ntrack[ja] - 1, [X](int t) { return X[t]; }, xr, h);  // searchable faces [0, nt-1)
}

// Grid::computeNeighbors: the two-pointer sweep of 2d.cpp:551-576 evaluated
// per cell.  With face positions non-decreasing along each track, the sweep
// pointer after cell i is min(lb(xR_i), n-2); the list of cell i starts one
// step back from the previous pointer when that face lies right of xL_i.
GH_HD void nb_side(const Dev& d, int j, int i, int ja, int* lo, int* cnt) {
  int nta = d.ntrack[ja];
  long long k = CI(d, j, i);
  double xjL = d.cx[k] - d.cdx[k] / 2., xjR = d.cx[k] + d.cdx[k] / 2.;
  int fin = 0;
  if (i > 1) {
    double xpR = d.cx[k - 1] + d.cdx[k - 1] / 2.;
    int l = lb_face(d, ja, xpR, i - 1);
    fin = l < nta - 2 ? l : nta - 2;
  }
  int s = fin;
  if (d.fx[CI(d, ja, s)] > xjL && s > 0) s--;
  int L = lb_face(d, ja, xjR, i);
  if (L < s) L = s;
  *lo = s + 1;
  *cnt = L - s;
}

GH_HD void body_nb(const Dev& d, int j, int i) {
  int nt = d.ntrack[j];
  if (i < 1 || i > nt - 2) return;
  long long k = CI(d, j, i);
  int lo = 0, cnt = 0;
  if (j != 0) nb_side(d, j, i, j - 1, &lo, &cnt);
  d.nbdj[k] = j - 1; d.nbdlo[k] = lo; d.nbdcnt[k] = cnt;
  lo = 0; cnt = 0;
  if (j != d.nj - 1) nb_side(d, j, i, j + 1, &lo, &cnt);
  d.nbuj[k] = j + 1; d.nbulo[k] = lo; d.nbucnt[k] = cnt;
}

// MV gradient of cell (j, m) as this stage's MV face pass left it (2D): face m
// stores gLs for cell m (1 <= m <= nt-3), face nt-3 stores gRs for cell nt-2
// -- the same minmod of the two one-sided slopes, whose arguments commute.
// With d.gmvre the F1 passes recompute it from the primitives they read
// anyway instead of loading the stored plane; cells 0 and nt-1 (and every
// cell under piecewise-constant reconstruction or on tracks shorter than 4)
// keep the value their ghost copy carried, which is read.
template <int NQ, int GEO = -1>
GH_HD void gmv_at(const Dev& d, int j, int m, double* g) {
  long long k = CI(d, j, m);
  int nt = d.ntrack[j];
  if (!d.gmvre || d.pc || nt < 4 || m < 1 || m > nt - 2) {
    load_v<2, NQ>(d, d.gmv, k, g);
    return;
  }
  int geom = GEO >= 0 ? GEO : d.tgeom[j];
  double pa[NQ], pb[NQ], pc[NQ];
  load_v<2, NQ>(d, d.prim, k - 1, pa);
  load_v<2, NQ>(d, d.prim, k, pb);
  load_v<2, NQ>(d, d.prim, k + 1, pc);
  double xa = geo_of_g<2>(d, j, k - 1, geom).cen0, xb = geo_of_g<2>(d, j, k, geom).cen0;
  double xc = geo_of_g<2>(d, j, k + 1, geom).cen0;
  for (int q = 0; q < NQ; ++q) {
    double gl = (pb[q] - pa[q]) / (xb - xa);
    double gn = (pc[q] - pb[q]) / (xc - xb);
    g[q] = minmod(gl, gn);
  }
}

// Grid::computeTransGradients (2d.cpp:366-430) of cell (j,i) into gf.  The
// second pass (minmod against every neighbour's gradient) reuses the
// gradients of the first KC neighbours from cache (KC x NQ doubles with
// stride cs, thread-private) and recomputes the rest, so that a neighbour's
// state is read once instead of twice.
// GEO: the grid's geometry when fixed at compile time (-1: the track's)
template <int NQ, int KC, int GEO = -1>
GH_HD void trans_grad(const Dev& d, int j, long long k, const double* p0, const double* m0,
                      const Geo& g0, double* gf, double* cache, int cs) {
  int geom = GEO >= 0 ? GEO : d.tgeom[j];
  Trig Tj = trig_at_g<2>(d, j, geom);
  double x0 = d.cx[k];
  double xL0 = x0 - d.cdx[k] / 2., xR0 = x0 + d.cdx[k] / 2.;
  double gavg[NQ];
  for (int q = 0; q < NQ; ++q) gavg[q] = 0.;
  double dAtot = 0;
  for (int pass = 0; pass < 2; ++pass) {
    int idx = 0;
    for (int side = 0; side < 2; ++side) {
      int tj = side == 0 ? d.nbdj[k] : d.nbuj[k];
      int lo = side == 0 ? d.nbdlo[k] : d.nbulo[k];
      int cnt = side == 0 ? d.nbdcnt[k] : d.nbucnt[k];
      for (int n = 0; n < cnt; ++n, ++idx) {
        if (pass == 1 && idx < KC) {
          for (int q = 0; q < NQ; ++q) gf[q] = minmod(gf[q], cache[(idx * NQ + q) * cs]);
          continue;
        }
        long long kn = CI(d, tj, lo + n);
        double xn = d.cx[kn];
        double xLn = xn - d.cdx[kn] / 2., xRn = xn + d.cdx[kn] / 2.;
        // face at y_j - dy_j/2 on the lower side; the upper side keeps y_j (the
        // second branch of 2d.cpp:386-391 tests d == 0 again)
        double xm = (fmax(xL0, xLn) + fmin(xR0, xRn)) / 2.;
        double ov = fmax(0, fmin(xR0, xRn) - fmax(xL0, xLn));
        double dA = face_area<2>(1, xm, ov, geom, side == 0 ? Tj.s_1 : Tj.s_y, 0.);
        Geo gn = geo_of_g<2>(d, tj, kn, geom);     // 2D: one geometry for all tracks
        double dxmL = xm - g0.cen0, dxmR = xm - gn.cen0;
        double pn[NQ], mn[NQ];
        load_v<2, NQ>(d, d.prim, kn, pn);
        gmv_at<NQ, GEO>(d, tj, lo + n, mn);
        if (pass == 0) dAtot += dA;
        for (int q = 0; q < NQ; ++q) {
          double qL = p0[q] + m0[q] * dxmL;
          double qR = pn[q] + mn[q] * dxmR;
          double gr = (qR - qL) / (gn.cen1 - g0.cen1);
          if (pass == 0) {
            gavg[q] += gr * dA;
            if (idx < KC) cache[(idx * NQ + q) * cs] = gr;
          } else {
            gf[q] = minmod(gf[q], gr);
          }
        }
      }
    }
    if (pass == 0)
      for (int q = 0; q < NQ; ++q) { gavg[q] /= dAtot; gf[q] = gavg[q]; }
  }
}

// The F1 loop of Grid::computeFluxes (2d.cpp:901-950) in two passes.
// Cell pass: each cell's transverse gradient (computeTransGradients), stored
// as the new value.
// neighbours per cell whose gradients the cell pass caches (LDS budget:
// KC x NQ doubles per thread)
template <int NQ> constexpr int f1_kc() { return NQ <= 6 ? 4 : 2; }

template <int NQ, int GEO = -1>
GH_HD void body_f1grad_c(const Dev& d, int j, int i, double* cache, int cs) {
  if (j > d.nj - 2) return;
  int nt = d.ntrack[j];
  if (i >= nt) return;
  long long k = CI(d, j, i);
  Geo g0 = geo_of_g<2>(d, j, k, GEO >= 0 ? GEO : d.tgeom[j]);
  double pc[NQ], mc[NQ], fc[NQ];
  load_v<2, NQ>(d, d.prim, k, pc);
  gmv_at<NQ, GEO>(d, j, i, mc);
  trans_grad<NQ, f1_kc<NQ>(), GEO>(d, j, k, pc, mc, g0, fc, cache, cs);
  store_v<2, NQ>(d, d.gf1n, k, fc);
}
template <int NQ>
GH_HD void body_f1grad(const Dev& d, int j, int i) {
  double cache[f1_kc<NQ>() * NQ];
  body_f1grad_c<NQ>(d, j, i, cache, 1);
}

// Face n of cell (j, i)'s own F1 list (slot s = f1off + n), between the
// cell and its upper neighbour n, given the owner's primitive state pc, MV
// gradient mc, new transverse gradient fc and geometry g0 -- from memory
// (face pass) or registers (fused cell pass).  The owner's gradient is the
// one just computed; the upper track's is the previous stage's (2d.cpp:901
// loops j upward, so track j+1 has not been recomputed yet); stale lists of
// the top ghost track pointing into their own track read the previous values
// too (their faces only feed ghost cells).
template <int NQ, bool SHK, int GEO = -1>
GH_HD void f1face_at(const Dev& d, int j, int i, int s, int tj, int in, const double* pc, const double* mc,
                     const double* fc, const Geo& g0) {
  int geom = GEO >= 0 ? GEO : d.tgeom[j];
  Trig Tj = trig_at_g<2>(d, j, geom);
  double G = d.P.gamma;
  long long k = CI(d, j, i);
  double x0 = d.cx[k];
  double xL0 = x0 - d.cdx[k] / 2., xR0 = x0 + d.cdx[k] / 2.;
  double jpos = d.ty[j] + d.tdy[j] / 2.;
  bool pc0 = (d.pc || j == 0 || j == d.nj - 2 || d.nbdcnt[k] == 0);
  long long kn = CI(d, tj, in);
  double xn = d.cx[kn];
  double xLn = xn - d.cdx[kn] / 2., xRn = xn + d.cdx[kn] / 2.;
  double xm = (fmax(xL0, xLn) + fmin(xR0, xRn)) / 2.;
  double ov = fmax(0, fmin(xR0, xRn) - fmax(xL0, xLn));
  double dA = face_area<2>(1, xm, ov, geom, Tj.s_2, 0.);
  double pL[NQ], cL[NQ], pR[NQ], cR[NQ];
  double rL = xm, rR = xm;
  if (pc0 || d.nbucnt[kn] == 0) {
    rL = cell_pc_g<2, NQ>(d, j, i, pL, cL, geom);
    rR = cell_pc_g<2, NQ>(d, tj, in, pR, cR, geom);
  } else {
    Geo gn = geo_of_g<2>(d, tj, kn, geom);
    double dxmL = xm - g0.cen0, dxmR = xm - gn.cen0;
    double dxfL = jpos - g0.cen1, dxfR = jpos - gn.cen1;
    double pn[NQ], mn[NQ], fn[NQ];
    load_v<2, NQ>(d, d.prim, kn, pn);
    gmv_at<NQ, GEO>(d, tj, in, mn);
    load_v<2, NQ>(d, d.gf1, kn, fn);
    for (int q = 0; q < NQ; ++q) {
      double qL = pc[q] + mc[q] * dxmL;
      double qR = pn[q] + mn[q] * dxmR;
      pL[q] = qL + fc[q] * dxfL;
      pR[q] = qR + fn[q] * dxfR;
    }
    prim2cons<2, NQ>(pL, cL, xm, geom, G, d.P.synge);
    prim2cons<2, NQ>(pR, cR, xm, geom, G, d.P.synge);
  }
  Waves w = lambdas<2, NQ>(pL, cL, pR, cR, 1, xm, geom, G);
  double fl[NQ];
  hllc_flux_lazy<2, NQ>(pL, cL, pR, cR, w, 0., dA, 1, xm, geom, [&](bool right, double* f) {
    if (right) state2flux<2, NQ>(pR, cR, f, 1, rR, geom);
    else state2flux<2, NQ>(pL, cL, f, 1, rL, geom);
  }, fl, 0.);
  long long fk = (long long)j * d.fcap + s;
  for (int q = 0; q < NQ; ++q) d.ff1[q * d.fs1 + fk] = fl[q];
  if (SHK) {
    // Forward (pL over pR) and reverse (pR over pL) shocks: with finite
    // positive pressures at most one of the two passes the detector's entry
    // tests (higher pressure first, a relative jump of 1e-10 or more); the
    // other returns -1 and leaves its spectral index unset.  Each lane then
    // evaluates only the candidate direction, so a wavefront whose lanes face
    // both ways runs the detector's powers once instead of twice.  Other
    // pressures (zero, infinite, NaN: the entry tests can pass both ways) run
    // both, as written.
    double psf = -1.e300, psr = -1.e300, sdf = -1., sdr = -1.;
    double pl = pL[PPP], pr = pR[PPP];
    if (pl > 0. && pr > 0. && pl < HUGE_VAL && pr < HUGE_VAL) {
      bool fw = pl >= pr;
      double S1[NQ], S2[NQ], ps = -1.e300;
      for (int q = 0; q < NQ; ++q) { S1[q] = fw ? pL[q] : pR[q]; S2[q] = fw ? pR[q] : pL[q]; }
      double sd = shock_sd<2, NQ>(S1, S2, 1, !fw, G, &ps, d.vpspec);
      if (fw) { sdf = sd; psf = ps; } else { sdr = sd; psr = ps; }
    } else {
      sdf = shock_sd<2, NQ>(pL, pR, 1, false, G, &psf, d.vpspec);
      sdr = shock_sd<2, NQ>(pR, pL, 1, true, G, &psr, d.vpspec);
    }
    d.ff1sd[0 * d.fs1 + fk] = sdf; d.ff1sd[1 * d.fs1 + fk] = psf;
    d.ff1sd[2 * d.fs1 + fk] = sdr; d.ff1sd[3 * d.fs1 + fk] = psr;
  }
}

// Face pass: one F1 face slot s of track j.
template <int NQ, bool SHK, int GEO = -1>
GH_HD void body_f1face_s(const Dev& d, int j, int s) {
  if (j > d.nj - 2) return;
  int tot = d.f1tot[j];
  if (s >= tot || s >= d.fcap) return;
  int i = d.f1own[(long long)j * d.fcap + s];
  long long k = CI(d, j, i);
  double pc[NQ], mc[NQ], fc[NQ];
  load_v<2, NQ>(d, d.prim, k, pc);
  gmv_at<NQ, GEO>(d, j, i, mc);
  load_v<2, NQ>(d, d.gf1n, k, fc);
  f1face_at<NQ, SHK, GEO>(d, j, i, s, d.nbuj[k], d.nbulo[k] + (s - d.f1off[k]), pc, mc, fc,
                          geo_of_g<2>(d, j, k, GEO >= 0 ? GEO : d.tgeom[j]));
}
template <int NQ>
GH_HD void body_f1face(const Dev& d, int j, int s) {
  if (d.shock) body_f1face_s<NQ, true>(d, j, s);
  else body_f1face_s<NQ, false>(d, j, s);
}

// Face pass, one thread per owner cell (launch option FUSE bit 2): the cell's
// state is loaded once and the faces of its own list formed in turn (where
// the tracks are aligned a cell has one upper neighbour).  The slot pass finds
// its owner through the slot table, one dependent load more.  Same faces,
// same arithmetic: the slots of every cell with a list, stale lists of the
// track-end cells included.
template <int NQ, bool SHK, int GEO = -1>
GH_HD void body_f1own(const Dev& d, int j, int i) {
  if (j > d.nj - 2) return;
  int nt = d.ntrack[j];
  if (i >= nt) return;
  long long k = CI(d, j, i);
  int cnt = d.nbucnt[k], off = d.f1off[k];
  if (cnt <= 0) return;
  double pc[NQ], mc[NQ], fc[NQ];
  load_v<2, NQ>(d, d.prim, k, pc);
  gmv_at<NQ, GEO>(d, j, i, mc);
  load_v<2, NQ>(d, d.gf1n, k, fc);
  Geo g0 = geo_of_g<2>(d, j, k, GEO >= 0 ? GEO : d.tgeom[j]);
  int tj = d.nbuj[k], lo = d.nbulo[k];
  for (int n = 0; n < cnt && off + n < d.fcap; ++n)
    f1face_at<NQ, SHK, GEO>(d, j, i, off + n, tj, lo + n, pc, mc, fc, g0);
}

// Fused cell pass: the cell's transverse gradient (as body_f1grad_c), then
// the faces of its own F1 list with the owner's state kept in registers.
// The same arithmetic as the two passes; the gradient is still stored, as the
// next stage's face pass reads it for the track above.
template <int NQ>
GH_HD void body_f1cell_c(const Dev& d, int j, int i, double* cache, int cs) {
  if (j > d.nj - 2) return;
  int nt = d.ntrack[j];
  if (i >= nt) return;
  long long k = CI(d, j, i);
  Geo g0 = geo_of<2>(d, j, k);
  double pc[NQ], mc[NQ], fc[NQ];
  load_v<2, NQ>(d, d.prim, k, pc);
  gmv_at<NQ>(d, j, i, mc);
  trans_grad<NQ, f1_kc<NQ>()>(d, j, k, pc, mc, g0, fc, cache, cs);
  store_v<2, NQ>(d, d.gf1n, k, fc);
  int cnt = d.nbucnt[k], off = d.f1off[k], tj = d.nbuj[k], lo = d.nbulo[k];
  for (int n = 0; n < cnt && off + n < d.fcap; ++n) {
    if (d.shock) f1face_at<NQ, true>(d, j, i, off + n, tj, lo + n, pc, mc, fc, g0);
    else f1face_at<NQ, false>(d, j, i, off + n, tj, lo + n, pc, mc, fc, g0);
  }
}
template <int NQ>
GH_HD void body_f1cell(const Dev& d, int j, int i) {
  double cache[f1_kc<NQ>() * NQ];
  body_f1cell_c<NQ>(d, j, i, cache, 1);
}

// Visit, in the reference's accumulation order (track j' ascending, cell i'
// ascending, 2d.cpp:901-906), every F1 face whose upper cell is (J, i):
// fn(fk) with fk the face's slot in the lower cell's list (ff1 / ff1sd).
// The fresh lists of track J-1 (cells 1..nt-2, this stage's neighbour pass)
// are ranges [lo, lo + cnt) with lo and lo + cnt non-decreasing along the
// track, so the cells whose list holds i are consecutive: from the first
// whose list ends past i to the last that starts at or before i, found by
// two searches from the hint of (J, i)'s own lower list.  The stale lists
// that ghost copies carry on cells 0 and nt-1 are checked one by one.
template <class F>
GH_HD void for_incoming_f1(const Dev& d, int J, int i, F&& fn) {
  auto src = [&](int jp, long long kp) {
    long long slot = d.f1off[kp] + (i - d.nbulo[kp]);
    if (slot < d.fcap) fn((long long)jp * d.fcap + slot);
  };
  auto stale = [&](int jp, int ip) {
    long long kp = CI(d, jp, ip);
    int c = d.nbucnt[kp], l = d.nbulo[kp];
    if (d.nbuj[kp] == J && c > 0 && l <= i && i < l + c) src(jp, kp);
  };
  // ghost tracks below the domain carry stale lists copied from track jL+1,
  // which point two tracks up: they are the only sources besides track J-1
  int jstale = J - 2 < d.jL ? J - 2 : d.jL;
  for (int jp = 0; jp <= jstale; ++jp) {
    stale(jp, 0);
    stale(jp, d.ntrack[jp] - 1);
  }
  {
    int jp = J - 1;
    if (jp < 0 || jp > d.nj - 2) return;
    int ntp = d.ntrack[jp];
    stale(jp, 0);
    int m = ntp - 2;                               // fresh cells 1..m
    if (m >= 1) {
      const int* LO = d.nbulo + (long long)jp * d.nax + 1;
      const int* CN = d.nbucnt + (long long)jp * d.nax + 1;
      long long k = CI(d, J, i);
      int h = d.nbdlo[k] - 1;
      int a = lb_x(m, [LO, CN](int t) { return LO[t] + CN[t]; }, i + 1, h);
      int b = a < m ? a + lb_x(m - a, [LO, a](int t) { return LO[a + t]; }, i + 1, 0) : a;
      for (int t = a; t < b; ++t) src(jp, CI(d, jp, t + 1));
    }
    stale(jp, ntp - 1);
  }
}

// measureShock of F1 faces onto their cells (2d.cpp:936-938), after the face pass
GH_HD void body_f1_shock(const Dev& d, int j, int i) {
  int nt = d.ntrack[j];
  if (i >= nt) return;
  long long k = CI(d, j, i);
  double Sd = d.sd[k], ps = d.pspec[k];
  if (j <= d.nj - 2) {
    int cnt = d.nbucnt[k], off = d.f1off[k];
    for (int n = 0; n < cnt && off + n < d.fcap; ++n) {
      long long fk = (long long)j * d.fcap + off + n;
      Sd = fmax(Sd, d.ff1sd[fk]);
      if (d.ff1sd[d.fs1 + fk] > ps) ps = d.ff1sd[d.fs1 + fk];
    }
  }
  for_incoming_f1(d, j, i, [&](long long fk) {
    double sd = d.ff1sd[2 * d.fs1 + fk];
    double pv = d.ff1sd[3 * d.fs1 + fk];
    Sd = fmax(Sd, sd);
    if (pv > ps) ps = pv;
  });
  d.sd[k] = Sd;
  d.pspec[k] = ps;
}

// Net F1 flux of a cell that the update will advance (2d.cpp:974-1036: the
// faces of its own list, then the incoming ones in the reference's order),
// as dn - up, ahead of the update kernel: the gathers over other cells' slots
// run here, in a light kernel, not in the update's register-heavy one.
template <int NQ>
GH_HD void f1_net(const Dev& d, int j, int i, long long k, double* net) {
  double up[NQ], dn[NQ];
  for (int q = 0; q < NQ; ++q) { up[q] = 0.; dn[q] = 0.; }
  int cnt = d.nbucnt[k], off = d.f1off[k];
  for (int n = 0; n < cnt && off + n < d.fcap; ++n) {
    long long fk = (long long)j * d.fcap + off + n;
    for (int q = 0; q < NQ; ++q) up[q] += d.ff1[q * d.fs1 + fk];
  }
  for_incoming_f1(d, j, i, [&](long long fk) {
    for (int q = 0; q < NQ; ++q) dn[q] += d.ff1[q * d.fs1 + fk];
  });
  for (int q = 0; q < NQ; ++q) net[q] = dn[q] - up[q];
}
template <int NQ>
GH_HD void body_f1sum(const Dev& d, int j, int i) {
  if (!active_track(d, j)) return;
  int nt = d.ntrack[j];
  int iL = d.ngst - 1, iR = nt - d.ngst;
  if (i < iL + 1 || i > iR - 1) return;
  long long k = CI(d, j, i);
  double net[NQ];
  f1_net<NQ>(d, j, i, k, net);
  for (int q = 0; q < NQ; ++q) d.f1net[q * d.cs + k] = net[q];
}

// ---------------------------------------------------------------------------
// Starting-pressure provider of the update's cons2prim calls: NoGuess keeps
// the previous pressure alone (the reference's rule).  A provider is called by
// every thread of the kernel (want = false for threads without a cell), so
// providers may use wavefront-wide operations.
struct NoGuess {
  GH_HD double operator()(bool, double, double, double, double, int) const { return -1.; }
};

#ifdef GH_EMU
// CPU build only (a test tool): records the update's cons2prim solves as rows
// (D, |S|, tau, previous pressure, root, shocked, stage) for every stride-th
// solve, until cap rows are taken (gammahip_emu_capture).
struct C2PCapture { double* buf; long long cap, n, stride, seen; };
C2PCapture g_c2p_capture = {nullptr, 0, 0, 1, 0};
inline void capture_c2p(double D, double S, double tau, double pprev, double p, int shocked, int stage) {
  C2PCapture& c = g_c2p_capture;
  if (!c.buf || c.n >= c.cap) return;
  if (c.seen++ % c.stride) return;
  double* r = c.buf + 7 * c.n++;
  r[0] = D; r[1] = S; r[2] = tau; r[3] = pprev; r[4] = p; r[5] = shocked; r[6] = stage;
}
#endif

// GEO: the grid's geometry when fixed at compile time (2D grids have one;
// -1: the track's, as 1D ensembles mix them)
template <int ND, int NQ, int SV, class GuessT, int GEO = -1>
GH_HD void body_update_g(const Dev& d, int j, int i, int stage, bool valid, const GuessT& guess) {
  bool run = valid && active_track(d, j);
  int nt = run ? d.ntrack[j] : 0;
  run = run && i < nt;
  int iL = d.ngst - 1, iR = nt - d.ngst;
  int geom = GEO >= 0 ? GEO : run ? d.tgeom[j] : 0;
  double G = d.P.gamma;
  double dt = run ? d.tdt[j] : 0.;
  double w0 = stage == 2 ? 3. / 4. : 1. / 3.;
  double w1 = stage == 2 ? 1. / 4. : 2. / 3.;
  long long k = run ? CI(d, j, i) : 0;
  bool cell = false;
  double p[NQ], U[NQ];
  for (int q = 0; q < NQ; ++q) { p[q] = 0.; U[q] = 0.; }
  double xnL = 0., xnR = 0., cxn = 0., cdxn = 0., r1 = 0.;
  Geo g2 = {};
  if (run) {
    if (i <= nt - 2) {
      double xo = d.fx[k];
      double xn = xo + dt * d.fv[k];
      if (stage == 1) d.fx0[k] = xo;
      double xa = xn;
      if (stage > 1 && i >= iL && i <= iR - 1) xa = w0 * d.fx0[k] + w1 * xn;
      d.fxn[k] = xa;
    }
    if (stage > 1 && (i == iL || i == iR)) {
      // CellGeomFromInterfacePos covers cells 1..ntrack-2 (1d.cpp:834), so the
      // ghost cells next to the active range follow their faces too (split at
      // the first / last active cell reads them)
      double xoL = d.fx[k - 1] + dt * d.fv[k - 1];
      double xoR = d.fx[k] + dt * d.fv[k];
      if (i - 1 >= iL && i - 1 <= iR - 1) xoL = w0 * d.fx0[k - 1] + w1 * xoL;
      if (i >= iL && i <= iR - 1) xoR = w0 * d.fx0[k] + w1 * xoR;
      d.cx[k] = (xoR + xoL) / 2.;
      d.cdx[k] = (xoR - xoL);
    } else if (i >= iL + 1 && i <= iR - 1) {
      cell = true;
    }
  }
  if (cell) {
    xnL = d.fx[k - 1] + dt * d.fv[k - 1];
    xnR = d.fx[k] + dt * d.fv[k];
    load_v<ND, NQ>(d, d.prim, k, p);
    Geo g = geo_of_g<ND>(d, j, k, geom);
    prim2cons<ND, NQ>(p, U, sweep_r<ND>(d, k, g), geom, G, d.P.synge);
    if (stage == 1) {
      store_v<ND, NQ>(d, d.U0, k, U);
      d.dV0[k] = g.dV;
    }
    for (int q = 0; q < NQ; ++q) U[q] *= g.dV;
    {
      double fa[NQ], fb[NQ];
      load_v<ND, NQ>(d, d.fflux, k - 1, fa);
      load_v<ND, NQ>(d, d.fflux, k, fb);
      for (int q = 0; q < NQ; ++q) U[q] += (fa[q] - fb[q]) * dt;
    }
    if (ND == 2) {   // net F1 flux: summed here, or read from body_f1sum's pass
      double f1[NQ];
      if (d.f1fused) f1_net<NQ>(d, j, i, k, f1);
      else for (int q = 0; q < NQ; ++q) f1[q] = d.f1net[q * d.cs + k];
      for (int q = 0; q < NQ; ++q) U[q] += f1[q] * dt;
    }
    sources<ND, NQ>(p, U, d.cx[k], d.cdx[k], trig_at_g<ND>(d, j, geom), g.dV, dt, geom, G, d.P.synge);
    if (d.ucell & 1) {       // userSourceTerms (cell.cpp:112)
      gh_user_track T = user_track(d, j);
      gh_user_sources(&T, i, d.cx[k], d.cdx[k], g.dV, dt, p, U);
    }
    if (NQ == 10 && d.sync) radiative(p, U, g.dV, dt, d.P);
    cxn = (xnR + xnL) / 2.;
    cdxn = (xnR - xnL);
    g2 = cell_geo<ND>(cxn, cdxn, d.ty[j], d.tdy[j], geom, trig_at_g<ND>(d, j, geom));
    for (int q = 0; q < NQ; ++q) U[q] /= g2.dV;
    r1 = g2.cen0;
  }
  C2PStat st = {0, 0, 0, 0};
  {
    double D = 1., S = 0., tau = 1.;
    if (cell) c2p_inputs<ND, NQ>(U, r1, geom, &D, &S, &tau);
    double gs = guess(cell, D, S, tau, p[PPP], geom);
#ifdef GH_EMU
    double pprev = p[PPP];
#endif
    if (cell) cons2prim<ND, NQ>(p, U, r1, geom, G, SV, gs, d.maxit, &st, d.P.synge, (d.ucell & 2) ? d.upar : nullptr);
#ifdef GH_EMU
    if (cell) capture_c2p(D, S, tau, pprev, p[PPP], d.shock ? d.shk[k] : 0, stage);
#endif
  }
  int ncalls = 1;
  if (stage > 1) {
    if (cell) {
      double U0[NQ];
      load_v<ND, NQ>(d, d.U0, k, U0);
      double dV0 = d.dV0[k];
      for (int q = 0; q < NQ; ++q) U[q] = w0 * U0[q] * dV0 + w1 * U[q] * g2.dV;
      double xaL = w0 * d.fx0[k - 1] + w1 * xnL;
      double xaR = w0 * d.fx0[k] + w1 * xnR;
      cxn = (xaR + xaL) / 2.;
      cdxn = (xaR - xaL);
      Geo g3 = cell_geo<ND>(cxn, cdxn, d.ty[j], d.tdy[j], geom, trig_at_g<ND>(d, j, geom));
      for (int q = 0; q < NQ; ++q) U[q] /= g3.dV;
    }
    double D = 1., S = 0., tau = 1.;
    if (cell) c2p_inputs<ND, NQ>(U, cxn, geom, &D, &S, &tau);
    double gs = guess(cell, D, S, tau, p[PPP], geom);
#ifdef GH_EMU
    double pprev = p[PPP];
#endif
    if (cell) cons2prim<ND, NQ>(p, U, cxn, geom, G, SV, gs, d.maxit, &st, d.P.synge, (d.ucell & 2) ? d.upar : nullptr);
#ifdef GH_EMU
    if (cell) capture_c2p(D, S, tau, pprev, p[PPP], d.shock ? d.shk[k] : 0, stage);
#endif
    ncalls = 2;
  }
  if (!cell) return;
  store_v<ND, NQ>(d, d.prim, k, p);
  d.cx[k] = cxn;
  d.cdx[k] = cdxn;
  GH_ATOMIC_ADD(cnt_at(d, CNT_C2P_CALLS), (unsigned long long)ncalls);
  if (st.iters) GH_ATOMIC_ADD(cnt_at(d, CNT_C2P_ITERS), (unsigned long long)st.iters);
  if (st.used) GH_ATOMIC_ADD(cnt_at(d, CNT_MLP), (unsigned long long)st.used);
  if (st.fail & 1) GH_ATOMIC_ADD(cnt_at(d, CNT_NOBRACKET), 1ull);
  if (st.fail & 2) GH_ATOMIC_ADD(cnt_at(d, CNT_MAXITER), 1ull);
  if (st.fail & 4) GH_ATOMIC_ADD(cnt_at(d, CNT_PHICAP), 1ull);
}

// Grid::update + Cell::update + RK3 averages (1d.cpp:798-841, 2d.cpp:974-1036,
// cell.cpp:102-127, rk3.cpp:12-195).  One thread owns face i and cell i.
// The solver is a template argument so that each kernel holds one root
// finder only (smaller code, fewer instruction-cache misses).
template <int ND, int NQ, int SV, int GEO = -1>
GH_HD void body_update_s(const Dev& d, int j, int i, int stage) {
  body_update_g<ND, NQ, SV, NoGuess, GEO>(d, j, i, stage, true, NoGuess{});
}
template <int ND, int NQ>
GH_HD void body_update(const Dev& d, int j, int i, int stage) {
  if (stage > 1) body_dt(d, j, i);         // the faces' dt candidates (prep with dt = false)
  if (d.solver == C2P_NEWTON) body_update_s<ND, NQ, C2P_NEWTON>(d, j, i, stage);
  else body_update_s<ND, NQ, C2P_BRENT>(d, j, i, stage);
}

// ---------------------------------------------------------------------------
// apply() sweeps (radiation_sph.cpp:177-283) and time advance
GH_HD void body_shock_reset(const Dev& d, int j, int i) {
  if (i >= d.nax) return;
  long long k = CI(d, j, i);
  d.sd[k] = -1; d.pspec[k] = 0; d.shk[k] = 0;
}
GH_HD void body_shock_detect(const Dev& d, int j, int i) {
  if (i >= d.nax) return;
  long long k = CI(d, j, i);
  if (d.sd[k] > d.P.shock_thr) d.shk[k] = 1;
}
template <int ND, int NQ>
GH_HD void body_inject(const Dev& d, int j, int i) {
  if (i >= d.ntrack[j]) return;
  long long k = CI(d, j, i);
  double p[NQ];
  load_v<ND, NQ>(d, d.prim, k, p);
  inject<ND, NQ>(p, d.shk[k] != 0, d.pspec[k], d.P);
  store_v<ND, NQ>(d, d.prim, k, p);
}
template <int ND, int NQ>
GH_HD void body_init_gammas(const Dev& d, int j, int i) {
  if (i >= d.nax) return;
  long long k = CI(d, j, i);
  double p[NQ];
  load_v<ND, NQ>(d, d.prim, k, p);
  init_gammas<ND, NQ>(p);
  store_v<ND, NQ>(d, d.prim, k, p);
}

// ---------------------------------------------------------------------------
// AMR: Grid::regrid / applyRegrid / split / merge (1d.cpp:484-728,
// 2d.cpp:609-832) with the checkCellForRegrid hooks of the setups.
enum { RG_SKIP = 0, RG_MERGE = 1, RG_SPLIT = 2 };

template <int ND, int NQ>
GH_HD int rg_check(const Dev& d, int j, int i, double iS) {
  long long k = CI(d, j, i);
  int mode = d.trg[j];
  if (mode == RG_AR2D) {
    double ar = d.cdx[k] / d.tdy[j];
    if (ar > d.rg[0]) return RG_SPLIT;
    if (ar < d.rg[1]) return RG_MERGE;
    return RG_SKIP;
  }
  if (mode == RG_BM1D) {
    int nt = d.ntrack[j], iR = nt - d.ngst;
    double dth = d.rg[2], target = d.rg[4];
    double rOut = d.cx[CI(d, j, iR - 1)];
    double ar = d.cdx[k] / (rOut * dth);
    double dist = i - iS;
    if (0 < dist && dist < 10) ar *= 10;
    double p[NQ];
    load_v<ND, NQ>(d, d.prim, k, p);
    double lf = lfac<ND, NQ>(p);
    if (ar > d.rg[0] * target / gh_pow(lf, 3. / 2.)) return RG_SPLIT;
    if (ar < d.rg[1] * target / gh_pow(lf, 3. / 2.)) return RG_MERGE;
    return RG_SKIP;
  }
  if (mode == RG_USER) {
    double p[NQ];
    load_v<ND, NQ>(d, d.prim, k, p);
    gh_user_track T = user_track(d, j);
    int a = gh_user_regrid(&T, i, d.cx[k], d.cdx[k], p);
    return a == 2 ? RG_SPLIT : a == 1 ? RG_MERGE : RG_SKIP;
  }
  return RG_SKIP;
}

// Cell::regridVal with the hook set's user_regridVal: the geometry's value
// (cartesian.cpp:58-63 dx; spherical1D.cpp:79-86 dr/r; spherical.cpp:115-123
// dr/(r dth), dth the track's width), replaced under the BM1D hooks by
// dr/(r dth) lfac^(3/2) with the rule's dth (BM1D.cpp:322-331); the 2DST hooks
// keep the geometry's value (2DST.cpp:171-178)
template <int ND, int NQ>
GH_HD double rg_val(const Dev& d, int j, long long k) {
  double r = d.cx[k], dr = d.cdx[k];
  if (d.trg[j] == RG_BM1D) {
    double p[NQ];
    load_v<ND, NQ>(d, d.prim, k, p);
    return dr / (r * d.rg[2]) * gh_pow(lfac<ND, NQ>(p), 3. / 2.);
  }
  if (d.tgeom[j] != SPH) return dr;
  return ND == 1 ? dr / r : dr / (r * d.tdy[j]);
}
// Grid::targetRegridVictims (1d.cpp:498-520, 2d.cpp:625-646): first cell of the
// smallest value below 1e20 and first of the largest above 0 among cells
// iLbnd+3 .. iRbnd-3, negative values skipped; -1 where none qualifies
template <int ND, int NQ>
GH_HD void rg_victims(const Dev& d, int j, int* is, int* ib) {
  int nt = d.ntrack[j], iL = d.ngst - 1, iR = nt - d.ngst;
  double mn = 1.e20, mx = 0.;
  *is = -1;
  *ib = -1;
  for (int i = iL + 3; i <= iR - 3; ++i) {
    double v = rg_val<ND, NQ>(d, j, CI(d, j, i));
    if (v < 0) continue;
    if (v < mn) { mn = v; *is = i; }
    if (v > mx) { mx = v; *ib = i; }
  }
}

// shock position of BM1D.cpp:292-301: last active cell with p > 1.5 eta, else iLbnd+1
GH_HD double rg_shock_index(const Dev& d, int j) {
  int nt = d.ntrack[j], iL = d.ngst - 1, iR = nt - d.ngst;
  double eta = d.rg[3];
  double iS = iL + 1;
  for (int ii = iR - 1; ii > iL; --ii) {
    iS = ii;
    if (d.prim[PPP * d.cs + CI(d, j, ii)] > 1.5 * eta) break;
  }
  return iS;
}
// rg_shock_index after a split (delta = +1) or merge (delta = -1) that wrote
// cells up to last_mod and moved every cell after them by delta: when the old
// index shifted lies past last_mod, its cell is an unchanged candidate and the
// cells after it are unchanged non-candidates, so it is the new index (the
// scan's default, iL + 1, never passes this test); otherwise scan again
GH_HD bool rg_shock_shifts(double iS, int delta, int last_mod) { return iS + delta > last_mod; }
#ifdef GH_EMU
// CPU build only: every shifted shock index checked against a new scan --
// [0] shifts taken, [1] shifts that differ from the scan, [2] actions where
// the shifted index lands on last_mod and the scan differs from it (the test
// above must then scan)
extern "C" long long gammahip_emu_shock_shift_checks[3];
long long gammahip_emu_shock_shift_checks[3] = {0, 0, 0};
#endif

template <int ND>
GH_HD void rg_move_cell(const Dev& d, long long kd, long long ks) {
  for (int s = 0; s < d.nqs; ++s) d.prim[s * d.cs + kd] = d.prim[s * d.cs + ks];
  d.cx[kd] = d.cx[ks];
  d.cdx[kd] = d.cdx[ks];
  d.dtloc[kd] = d.dtloc[ks];
  if (ND == 2) {
    for (int s = 0; s < d.nqs; ++s) {
      d.gmv[s * d.cs + kd] = d.gmv[s * d.cs + ks];
      d.gf1[s * d.cs + kd] = d.gf1[s * d.cs + ks];
    }
    d.nbdj[kd] = d.nbdj[ks]; d.nbdlo[kd] = d.nbdlo[ks]; d.nbdcnt[kd] = d.nbdcnt[ks];
    d.nbuj[kd] = d.nbuj[ks]; d.nbulo[kd] = d.nbulo[ks]; d.nbucnt[kd] = d.nbucnt[ks];
  }
}

// Grid::split (1d.cpp:560-662): cells i+1.. and faces i.. move up one
// slot, then the two halves of cell i are filled (split_fill).
template <int ND, int NQ>
GH_HD void split_fill(const Dev& d, int j, int i);
template <int ND, int NQ>
GH_HD void rg_split(const Dev& d, int j, int i) {
  int nt = d.ntrack[j];
  long long b = CI(d, j, 0);
  for (int m = nt - 1; m >= i + 1; --m) rg_move_cell<ND>(d, b + m + 1, b + m);
  for (int m = nt - 2; m >= i; --m) d.fx[b + m + 1] = d.fx[b + m];
  d.ntrack[j] = nt + 1;
  split_fill<ND, NQ>(d, j, i);
}
template <int ND, int NQ>
GH_HD void split_fill(const Dev& d, int j, int i) {
  long long b = CI(d, j, 0);
  int geom = d.tgeom[j];
  double pLc[NQ], pc[NQ], pRo[NQ], gC[NQ];
  load_v<ND, NQ>(d, d.prim, b + i - 1, pLc);
  load_v<ND, NQ>(d, d.prim, b + i, pc);
  load_v<ND, NQ>(d, d.prim, b + i + 2, pRo);
  double xLc = geo_of<ND>(d, j, b + i - 1).cen0, xc = geo_of<ND>(d, j, b + i).cen0;
  double xRo = geo_of<ND>(d, j, b + i + 2).cen0;
  for (int q = 0; q < NQ; ++q) {
    double gL = (pc[q] - pLc[q]) / (xc - xLc);
    double gR = (pRo[q] - pc[q]) / (xRo - xc);
    if (ND == 1) {
      double gN = (pRo[q] - pLc[q]) / (xRo - xLc);
      gL = minmod(gL, gN);
      gR = minmod(gR, gN);
    }
    gC[q] = minmod(gL, gR);
  }
  double x_old = d.cx[b + i], dl_old = d.cdx[b + i];
  d.cx[b + i] = x_old - dl_old / 4.;
  d.cx[b + i + 1] = x_old + dl_old / 4.;
  d.cdx[b + i] = dl_old / 2.;
  d.cdx[b + i + 1] = dl_old / 2.;
  Trig T = trig_at<ND>(d, j);
  Geo gc = cell_geo<ND>(d.cx[b + i], d.cdx[b + i], d.ty[j], d.tdy[j], geom, T);
  Geo gr = cell_geo<ND>(d.cx[b + i + 1], d.cdx[b + i + 1], d.ty[j], d.tdy[j], geom, T);
  double xL = gc.cen0, xR = gr.cen0;
  double xI = d.cx[b + i] + d.cdx[b + i] / 2.;
  double pn[NQ], pr[NQ];
  for (int q = 0; q < NQ; ++q) {
    double Sc = pc[q];
    pn[q] = Sc + gC[q] * (xL - xI);
    pr[q] = Sc + gC[q] * (xR - xI);
  }
  store_v<ND, NQ>(d, d.prim, b + i, pn);
  store_v<ND, NQ>(d, d.prim, b + i + 1, pr);
  d.fx[b + i] = x_old;
}

// Grid::merge (1d.cpp:665-728): cell i absorbs its smaller neighbour
// v (merge_fill, returns whether v = i-1), then the cells after the pair move
// down one slot.
template <int ND, int NQ>
GH_HD bool merge_fill(const Dev& d, int j, int i);
template <int ND, int NQ>
GH_HD void rg_merge(const Dev& d, int j, int i) {
  int nt = d.ntrack[j];
  long long b = CI(d, j, 0);
  bool left = merge_fill<ND, NQ>(d, j, i);
  if (left) {
    for (int m = i; m < nt; ++m) rg_move_cell<ND>(d, b + m - 1, b + m);
    for (int m = i; m < nt - 1; ++m) d.fx[b + m - 1] = d.fx[b + m];
  } else {
    for (int m = i + 2; m < nt; ++m) rg_move_cell<ND>(d, b + m - 1, b + m);
    for (int m = i + 1; m < nt - 1; ++m) d.fx[b + m - 1] = d.fx[b + m];
  }
  d.ntrack[j] = nt - 1;
}
template <int ND, int NQ>
GH_HD bool merge_fill(const Dev& d, int j, int i) {
  int nt = d.ntrack[j];
  int nact = nt - 2 * d.ngst;
  long long b = CI(d, j, 0);
  int geom = d.tgeom[j];
  double G = d.P.gamma;
  bool left = (d.cdx[b + i - 1] < d.cdx[b + i + 1] && i != d.ngst) || i == nact + d.ngst - 1;
  int v = left ? i - 1 : i + 1;
  double pc[NQ], pv[NQ], cc[NQ], cv[NQ];
  load_v<ND, NQ>(d, d.prim, b + i, pc);
  load_v<ND, NQ>(d, d.prim, b + v, pv);
  Geo gc = geo_of<ND>(d, j, b + i), gv = geo_of<ND>(d, j, b + v);
  prim2cons<ND, NQ>(pc, cc, sweep_r<ND>(d, b + i, gc), geom, G, d.P.synge);
  prim2cons<ND, NQ>(pv, cv, sweep_r<ND>(d, b + v, gv), geom, G, d.P.synge);
  double dVl = gc.dV, dVv = gv.dV;
  for (int q = 0; q < NQ; ++q) cc[q] = (cc[q] * dVl + cv[q] * dVv) / (dVl + dVv);
  double xc = d.cx[b + i], dxc = d.cdx[b + i], xv = d.cx[b + v], dxv = d.cdx[b + v];
  double xnew = left ? ((xv - dxv / 2.) + (xc + dxc / 2.)) / 2. : ((xv + dxv / 2.) + (xc - dxc / 2.)) / 2.;
  double dxnew = dxc + dxv;
  d.cx[b + i] = xnew;
  d.cdx[b + i] = dxnew;
  Geo g2 = cell_geo<ND>(xnew, dxnew, d.ty[j], d.tdy[j], geom, trig_at<ND>(d, j));
  C2PStat st = {0, 0, 0, 0};
  cons2prim<ND, NQ>(pc, cc, ND == 1 ? g2.cen0 : xnew, geom, G, d.solver, -1., d.maxit, &st, d.P.synge,
                    (d.ucell & 2) ? d.upar : nullptr);
  store_v<ND, NQ>(d, d.prim, b + i, pc);
  return left;
}

// parallel screen: first active cell of each track whose rule asks for an
// admissible action (cells before it would be skipped by the sequential loop)
GH_HD void body_rg_init(const Dev& d, int j, int i) {
  if (i != 0) return;
  d.tfirst[j] = INT_MAX;
  d.tlast[j] = -1;
  d.tiS[j] = d.ngst;        // iLbnd + 1
}
// In the plain launch (a thread for every slot of the track) the threads take
// the cells from the far end of the track (cell nax-1-i), so that the
// wavefronts dispatched first hold the largest indices; a thread whose
// candidate does not exceed the maximum already stored (it only grows; a
// stale read is lower) leaves out its atomic: the same maximum with far fewer
// same-address atomics per track.  The packed launch walks the live cells
// only, in their own order.
GH_HD void body_rg_shockpos(const Dev& d, int j, int i) {
  if (!active_track(d, j) || d.trg[j] != RG_BM1D) return;
  if (!d.packed) i = d.nax - 1 - i;
  int nt = d.ntrack[j], iL = d.ngst - 1, iR = nt - d.ngst;
  if (i < iL + 1 || i > iR - 1) return;
  if (d.prim[PPP * d.cs + CI(d, j, i)] > 1.5 * d.rg[3] && i > d.tiS[j]) GH_ATOMIC_MAX(&d.tiS[j], i);
}
template <int ND, int NQ>
GH_HD void body_rg_screen(const Dev& d, int j, int i) {
  if (!active_track(d, j) || d.trg[j] == RG_NONE) return;
  int nt = d.ntrack[j], iL = d.ngst - 1, iR = nt - d.ngst;
  if (i < iL + 1 || i > iR - 1) return;
  int a = rg_check<ND, NQ>(d, j, i, (double)d.tiS[j]);
  bool ok = (a == RG_SPLIT && nt < d.nax - 1) || (a == RG_MERGE && nt > 2 * d.ngst + 2);
  d.rgflag[CI(d, j, i)] = ok ? 1 : 0;
  if (ok) {
    GH_ATOMIC_MIN(&d.tfirst[j], i);
    GH_ATOMIC_MAX(&d.tlast[j], i);
  }
}
// The sequential loop may stop at cell i when no cell from i on can act in
// this pass: every cell from i on is an unmodified cell of the screened track
// (i past the cells the last split or merge wrote), shifted by the net count
// dn of actions so far; its rule reads, besides the cell, only the shock
// index relative to the cell and the outer radius (BM1D) or nothing (aspect
// ratio), and the guards read the track length; when these are what the
// screen saw, the cell decides as the screen found, and the screen flagged
// none past tlast.  (CIRC_REGRID_ and the user rule read more of the track:
// never stopped early.)
struct RgStop { int tl, nt0; double iS0, rOut0; bool on; };
GH_HD RgStop rg_stop_init(const Dev& d, int j, double iS) {
  RgStop s;
  s.tl = d.tlast[j];
  s.nt0 = d.ntrack[j];
  s.iS0 = iS;
  s.rOut0 = d.cx[CI(d, j, s.nt0 - d.ngst - 1)];
  s.on = !d.circ && (d.trg[j] == RG_AR2D || d.trg[j] == RG_BM1D) && s.tl >= 0;
  return s;
}
// whether the rule inputs besides the cell (guards, shock index relative to
// the cell, outer radius) are what the screen saw, shifted by dn
GH_HD bool rg_inputs_same(const Dev& d, int j, const RgStop& s, double iS) {
  int nt = d.ntrack[j], dn = nt - s.nt0;
  if ((s.nt0 < d.nax - 1) != (nt < d.nax - 1) || (s.nt0 > 2 * d.ngst + 2) != (nt > 2 * d.ngst + 2)) return false;
  if (d.trg[j] == RG_BM1D && (iS != s.iS0 + dn || d.cx[CI(d, j, nt - d.ngst - 1)] != s.rOut0)) return false;
  return true;
}
GH_HD bool rg_stop(const Dev& d, int j, const RgStop& s, int i, int last_mod, double iS) {
  if (!s.on || i <= last_mod) return false;
  int dn = d.ntrack[j] - s.nt0;
  if (i <= s.tl + dn) return false;
  return rg_inputs_same(d, j, s, iS);
}
// By the same argument the loop may skip from cell i to the next cell the
// screen flagged: cell m >= i is the screened cell m - dn, and decides as the
// screen found (rgflag).  Returns where to continue: i itself when the
// argument does not hold, the next flagged cell, or -1 when none is left.
GH_HD int rg_next(const Dev& d, int j, const RgStop& s, int i, int last_mod, double iS) {
  if (!s.on || i <= last_mod || !rg_inputs_same(d, j, s, iS)) return i;
  int dn = d.ntrack[j] - s.nt0;
  long long b = CI(d, j, 0);
  for (int o = i - dn > 0 ? i - dn : 0; o <= s.tl; ++o)
    if (d.rgflag[b + o]) return o + dn;
  return -1;
}
// sequential applyRegrid loop of one track from its first flagged cell (the
// cells before it take no action, so the victims found at the start of the
// track's loop are the same); with CIRC_REGRID_ a split of i first merges the
// smallest cell is (then splits i-1 if is < i), a merge of i is followed by a
// split of the biggest cell ib (ib-1 if ib > i), and a victim equal to i -- the
// reference's exit(10) -- is counted and ends the track's regrid
template <int ND, int NQ>
GH_HD void body_rg_apply(const Dev& d, int j, int i0) {
  if (i0 != 0 || !active_track(d, j)) return;
  int i = d.tfirst[j];
  if (i == INT_MAX) return;
  bool bm = d.trg[j] == RG_BM1D;
  // the pass's shock index as body_rg_shockpos left it (the same maximum as
  // rg_shock_index: no cell has changed since)
  double iS = bm ? (double)d.tiS[j] : 0.;
  int is = -1, ib = -1;
  if (d.circ) rg_victims<ND, NQ>(d, j, &is, &ib);
  RgStop stop = rg_stop_init(d, j, iS);
  int last_mod = -1;                       // the last cell a split or merge wrote
  for (; i <= d.ntrack[j] - d.ngst - 1; ++i) {
    if (rg_stop(d, j, stop, i, last_mod, iS)) break;
    i = rg_next(d, j, stop, i, last_mod, iS);
    if (i < 0) break;
    int a = rg_check<ND, NQ>(d, j, i, iS);
    int nt = d.ntrack[j];
    bool acted = false;
    int delta = 0;
    if (a == RG_SPLIT && nt < d.nax - 1) {
      int isplit = i;
      if (d.circ) {
        if (is < 0 || is == i) { GH_ATOMIC_ADD(cnt_at(d, CNT_RGCLASH), 1ull); return; }
        rg_merge<ND, NQ>(d, j, is);
        d.nmerge[j]++;
        if (i > is) isplit--;
      }
      rg_split<ND, NQ>(d, j, isplit);
      d.nsplit[j]++;
      acted = true;
      delta = 1;
      last_mod = i + 1;
    }
    if (a == RG_MERGE && nt > 2 * d.ngst + 2) {
      if (d.circ && (ib < 0 || ib == i)) { GH_ATOMIC_ADD(cnt_at(d, CNT_RGCLASH), 1ull); return; }
      rg_merge<ND, NQ>(d, j, i);
      d.nmerge[j]++;
      if (d.circ) {
        rg_split<ND, NQ>(d, j, ib > i ? ib - 1 : ib);
        d.nsplit[j]++;
      }
      acted = true;
      delta = -1;
      last_mod = i;
    }
    if (acted) {
#ifdef GH_EMU
      if (bm && !d.circ) {
        double fresh = rg_shock_index(d, j);
        long long* ck = gammahip_emu_shock_shift_checks;
        if (rg_shock_shifts(iS, delta, last_mod)) { ck[0]++; if (iS + delta != fresh) ck[1]++; }
        else if (iS + delta == last_mod && fresh != iS + delta) ck[2]++;
      }
#endif
      if (bm) iS = !d.circ && rg_shock_shifts(iS, delta, last_mod) ? iS + delta : rg_shock_index(d, j);
      if (d.circ) rg_victims<ND, NQ>(d, j, &is, &ib);
    }
  }
}

GH_HD void body_advance_time(const Dev& d, int j, int i) {
  if (i == 0) d.tt[j] += d.tdt[j];
}

}  // namespace

// ===========================================================================
// launch layer
#ifndef GH_EMU
#define KB 256
// Block -> (track, cell-chunk) mapping.  remap == 0: blockIdx.y is the track,
// blockIdx.x the chunk.  remap == 1 (XCD-aware): workgroups are dispatched
// round-robin over the 8 XCDs by linear id, so linear id L runs on XCD L % 8;
// L is renumbered so that each XCD walks a contiguous band of tiles (whole
// tracks in order), keeping adjacent tracks -- the F1 neighbours -- in one L2.
// On the uniform 2048^2 grid this cuts the transverse-gradient pass's HBM
// traffic from 1.87 to 0.81 GB per launch; on tracks of uneven length the
// bands are uneven work, so it is a tuned option, off by default.
// remap 2 (track groups; grid of ceil(ni/64) x ceil(ntr/W) workgroups of W
// wavefronts, plain_launch): workgroup (x, y) takes the 64-cell chunk x of
// tracks yW .. yW+W-1, one per wavefront, so the F1 neighbours of its inner
// tracks are read by the workgroup itself, on one CU, at the same time.
// Track indices past the launch's last track fail the kernels' j < nj test
// (launches over fewer tracks than the grid holds use this order only when
// their count is a multiple of W).  j is the same on every lane of a
// wavefront; readfirstlane says so to the compiler, which otherwise treats it
// as varying per lane: per-track values would then be loaded per lane and
// same-address atomics (the shock-index maximum) issued per lane instead of
// once per wavefront (0.025 -> 0.7 ms per launch on the 1D ensemble).
__device__ __forceinline__ void tile_of(int remap, int& j, int& i) {
  if (remap == 2) {
    j = __builtin_amdgcn_readfirstlane((int)blockIdx.y * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6));
    i = (int)blockIdx.x * 64 + (int)(threadIdx.x & 63);
    return;
  }
  int gx = gridDim.x, gy = gridDim.y;
  int L = blockIdx.x + gx * blockIdx.y;
  if (remap) {
    // XCD x runs blocks x, x+8, x+16, ...: n_x = N/8 + (x < N%8) of them.  Give
    // it the logical range [x*(N/8) + min(x, N%8), +n_x): a bijection on [0, N).
    int N = gx * gy;
    int x = L % 8, r = L / 8;
    int q = N / 8, m = N % 8;
    L = x * q + (x < m ? x : m) + r;
  }
  j = L / gx;
  i = (L % gx) * blockDim.x + threadIdx.x;
}
#define GH_KERNEL2(NAME, BODY)                                                  \
  __global__ void __launch_bounds__(KB) NAME(Dev d, int ni, int remap) {        \
    int i, j;                                                                    \
    tile_of(remap, j, i);                                                        \
    if (i < ni && j < d.nj) BODY(d, j, i);                                       \
  }
// Packed form: a persistent grid walks the tile list built by k_tiles.  With
// remap the list is cut into 8 contiguous bands, band x walked by the
// workgroups that the dispatcher places on XCD x (linear id % 8), so adjacent
// tracks -- F1 neighbours -- share an L2.  Grid sizes are multiples of 8.
// The trip count depends on blockIdx only, so every lane of a wavefront runs
// the same iterations.  PK (a template argument of every kernel built on this
// loop) selects the form at compile time: the plain form (one workgroup per
// tile) is then straight-line code holding none of the loop's state in
// registers (in the update kernel those scalars spilled into vector lanes).
#define GH_TILE_LOOP(PK, ...)                                                    \
  if constexpr (!(PK)) {                                                         \
    int i, j;                                                                    \
    tile_of(remap, j, i);                                                        \
    __VA_ARGS__;                                                                 \
  } else {                                                                       \
    int nt_ = *d.ntiles, G_ = gridDim.x;                                         \
    int t_ = blockIdx.x, end_ = nt_, st_ = G_;                                   \
    if (remap) {                                                                 \
      int x_ = blockIdx.x & 7;                                                   \
      t_ = (int)((long long)nt_ * x_ / 8) + (int)(blockIdx.x >> 3);              \
      end_ = (int)((long long)nt_ * (x_ + 1) / 8);                               \
      st_ = G_ >> 3;                                                             \
    }                                                                            \
    for (; t_ < end_; t_ += st_) {                                               \
      int j = d.tiles[2 * t_];                                                   \
      int i = d.tiles[2 * t_ + 1] * (int)blockDim.x + (int)threadIdx.x;          \
      __VA_ARGS__;                                                               \
    }                                                                            \
  }
// TPL: the kernel's other template parameters, each preceded by a comma, in
// parentheses: () or (, int ND, int NQ)
#define GH_UNPAREN(...) __VA_ARGS__
#define GH_KERNEL2P(TPL, NAME, BODY)                                            \
  template <bool PK GH_UNPAREN TPL>                                              \
  __global__ void __launch_bounds__(KB) NAME(Dev d, int ni, int remap) {        \
    GH_TILE_LOOP(PK, if (i < ni && j < d.nj) BODY(d, j, i))                      \
  }
namespace {
GH_KERNEL2P((), k_dt, body_dt)
// body_nb on a wavefront of 64 consecutive cells of one track: the lower
// bound of the previous cell's right face (first search of nb_side) is the
// previous lane's own bound of its right face (same face, same track, and the
// search returns the bound whatever its starting hint), so each lane searches
// once per side and takes the other from its neighbour; lane 0 searches both.
// Every lane runs the shuffles and the barrier (valid = false for lanes
// without a tile cell).  The faces of each neighbour track around the
// wavefront's cells, [i0 - NBW_LO, i0 - NBW_LO + NBW) for the first cell i0,
// are staged in LDS with one coalesced load: the searches start at the cell's
// own index and mostly stay inside that window, so their chains of dependent
// probes read LDS instead of memory (probes outside it read memory, so the
// answer is the same).
constexpr int NBW = 96, NBW_LO = 16;
__device__ void body_nb_lanes(const Dev& d, int j, int i, bool valid) {
  __shared__ double nbwin[KB / 64][2][NBW];
  int lane = (int)(threadIdx.x & 63);
  double* win0 = nbwin[threadIdx.x >> 6][0];
  int w0 = i - lane - NBW_LO;
  if (w0 < 0) w0 = 0;
  for (int side = 0; side < 2; ++side) {
    int ja = side == 0 ? j - 1 : j + 1;
    if (ja >= 0 && ja < d.nj) {
      const double* X = d.fx + (long long)ja * d.nax;
      for (int t = lane; t < NBW; t += 64)
        if (w0 + t < d.nax) win0[side * NBW + t] = X[w0 + t];
    }
  }
  __syncthreads();
  int nt = valid ? d.ntrack[j] : 0;
  bool cell = valid && i >= 1 && i <= nt - 2;
  long long k = cell ? CI(d, j, i) : 0;
  double xjL = 0., xjR = 0.;
  if (cell) { xjL = d.cx[k] - d.cdx[k] / 2.; xjR = d.cx[k] + d.cdx[k] / 2.; }
  int lo2[2] = {0, 0}, cnt2[2] = {0, 0};
  for (int side = 0; side < 2; ++side) {
    int ja = side == 0 ? j - 1 : j + 1;
    bool has = cell && (side == 0 ? j != 0 : j != d.nj - 1);
    const double* X = d.fx + (long long)ja * d.nax;
    const double* win = win0 + side * NBW;
    // the window covers [w0, w0 + NBW) but never reaches past the plane's row
    auto XF = [X, win, w0](int t) { unsigned u = (unsigned)(t - w0); return u < (unsigned)NBW ? win[u] : X[t]; };
    int n = has ? d.ntrack[ja] - 1 : 0;
    int L = has ? lb_x(n, XF, xjR, i) : 0;
    int lp = __shfl(L, lane == 0 ? 0 : lane - 1, 64);
    if (has) {
      int nta = n + 1;
      int fin = 0;
      if (i > 1) {
        if (lane == 0) lp = lb_x(n, XF, d.cx[k - 1] + d.cdx[k - 1] / 2., i - 1);
        fin = lp < nta - 2 ? lp : nta - 2;
      }
      int s0 = fin;
      if (XF(s0) > xjL && s0 > 0) s0--;
      if (L < s0) L = s0;
      lo2[side] = s0 + 1;
      cnt2[side] = L - s0;
    }
  }
  if (!cell) return;
  d.nbdj[k] = j - 1; d.nbdlo[k] = lo2[0]; d.nbdcnt[k] = cnt2[0];
  d.nbuj[k] = j + 1; d.nbulo[k] = lo2[1]; d.nbucnt[k] = cnt2[1];
}
template <bool PK>
__global__ void __launch_bounds__(KB) k_nb(Dev d, int ni, int remap) {
  GH_TILE_LOOP(PK, body_nb_lanes(d, j, i, i < ni && j < d.nj))
}
GH_KERNEL2P((), k_f1_shock, body_f1_shock)
GH_KERNEL2(k_shock_reset, body_shock_reset)
GH_KERNEL2(k_shock_detect, body_shock_detect)
GH_KERNEL2(k_rg_init, body_rg_init)
GH_KERNEL2P((), k_rg_shockpos, body_rg_shockpos)
GH_KERNEL2(k_advance_time, body_advance_time)
template <int ND, int NQ> GH_KERNEL2(k_ghost_mv, (body_ghost_mv<ND, NQ>))
// updateGhosts of a 2D grid in one launch: workgroups [0, ncopy) copy the
// ghost tracks' active cells and inner faces from the edge tracks (128 cells
// each), the rest form the ghost cells along every track, two tracks per
// workgroup (one per wavefront, 2 ngst lanes busy), ghost tracks from their
// edge tracks' cells -- the state the copy then the ghost-cell pass leave
template <int NQ>
__global__ void __launch_bounds__(128) k_ghosts2d(Dev d, int ncopy) {
  int b = (int)blockIdx.x, t = (int)threadIdx.x;
  if (b < ncopy) {
    long long gi = (long long)b * 128 + t;
    int g = (int)(gi / d.nax), i = (int)(gi % d.nax);
    if (g < 2 * d.ngst) body_ghost_track_inner(d, g, i);
    return;
  }
  int j = __builtin_amdgcn_readfirstlane((b - ncopy) * 2 + (t >> 6)), kk = t & 63;
  if (j < d.nj && kk < 2 * d.ngst) {
    int js = j <= d.jL ? d.jL + 1 : j >= d.jR ? d.jR - 1 : j;
    body_ghost_mv_src<2, NQ>(d, j, kk, js);
  }
}
// MV face pass: built for 5 wavefronts per SIMD with six fields and no shock
// detector (the body then fits 96 registers without spilling; left to itself
// the compiler hoists loads into more and runs 4)
template <bool PK, int ND, int NQ, bool SHK>
__global__ void __launch_bounds__(KB, (SHK || NQ > 6) ? 1 : 5) k_faces(Dev d, int ni, int remap) {
  GH_TILE_LOOP(PK, if (i < ni && j < d.nj) body_faces_s<ND, NQ, SHK>(d, j, i))
}
// cell pass with the gradient cache in LDS: thread t's entry m at
// cache[m * blockDim + t] (consecutive lanes, consecutive banks); the launch
// passes blockDim * f1_kc * NQ doubles of dynamic LDS.  The pass waits on
// chains of dependent neighbour loads: scheduled for 3 wavefronts per SIMD it
// keeps more of them in flight per wavefront, which beats a fourth wavefront
// (0.336 against 0.363 ms on the 2048^2 grid)
template <bool PK, int NQ, int GEO>
__global__ void __launch_bounds__(KB) __attribute__((amdgpu_waves_per_eu(1, 3)))
k_f1grad(Dev d, int ni, int remap) {
  extern __shared__ double f1cache[];
  double* cache = f1cache + threadIdx.x;
  GH_TILE_LOOP(PK, if (i < ni && j < d.nj) body_f1grad_c<NQ, GEO>(d, j, i, cache, (int)blockDim.x))
}
// Neighbour search and transverse gradient in one pass (launch option FUSE
// bit 0): a cell's gradient (2d.cpp:366-430) reads only its own neighbour
// lists (2d.cpp:534-606), so the two sweeps run back to back in one thread;
// the search's dependent loads overlap the gradient's arithmetic in other
// wavefronts, and the lists are read back by the thread that wrote them.
// Every lane runs the search (its shuffles need the whole wavefront).
template <bool PK, int NQ, int GEO>
__global__ void __launch_bounds__(KB) __attribute__((amdgpu_waves_per_eu(1, 3)))
k_nbgrad(Dev d, int ni, int remap) {
  extern __shared__ double f1cache[];
  double* cache = f1cache + threadIdx.x;
  GH_TILE_LOOP(PK, {
    bool valid = i < ni && j < d.nj;
    body_nb_lanes(d, j, i, valid);
    if (valid) body_f1grad_c<NQ, GEO>(d, j, i, cache, (int)blockDim.x);
  })
}
// fused gradient + own-face pass (same LDS cache as k_f1grad)
template <bool PK, int NQ>
__global__ void __launch_bounds__(KB) k_f1cell(Dev d, int ni, int remap) {
  extern __shared__ double f1cache[];
  double* cache = f1cache + threadIdx.x;
  GH_TILE_LOOP(PK, if (i < ni && j < d.nj) body_f1cell_c<NQ>(d, j, i, cache, (int)blockDim.x))
}
GH_KERNEL2P((, int NQ), k_f1sum, body_f1sum<NQ>)
// face pass over the F1 slots of each track: one thread per slot of the
// track's capacity (4 nax); threads past the track's slot count leave at
// once.  (A grid-stride loop over the slots kept the shock detector's
// constants live across it: 153 scalars spilled and 203 VGPRs in the ten-field
// spherical build, against 4 and 169 without the loop.)
// W = minimum wavefronts per SIMD the compiler must allow (registers capped
// accordingly, spilling if it must): 1 leaves the choice to the compiler, 4
// trades spills for latency hiding.  The launch option picks one.
// (XCD-aware with remap, as the cell passes: the face reads the upper track)
template <int NQ, int W, bool SHK, int GEO>
__global__ void __launch_bounds__(KB, W) k_f1faces(Dev d, int remap) {
  int j, s;
  tile_of(remap, j, s);
  if (j >= d.nj) return;
  int tot = min(d.f1tot[j], d.fcap);
  if (s < tot) body_f1face_s<NQ, SHK, GEO>(d, j, s);
}
// the face pass with one thread per owner cell (body_f1own)
template <int NQ, int W, bool SHK, int GEO>
__global__ void __launch_bounds__(KB, W) k_f1own(Dev d, int ni, int remap) {
  int i, j;
  tile_of(remap, j, i);
  if (i < ni && j < d.nj) body_f1own<NQ, SHK, GEO>(d, j, i);
}
// W: minimum wavefronts per SIMD (1 = the compiler's choice; 5 caps the
// registers at 96 and spills the rest), picked by the launch option; GEO as
// body_update_g (-1 in 1D).  The 2D Cartesian six-field body fits 96
// registers but for 4 spilled dwords, so its default build runs 5 wavefronts
// too (0.358 against 0.370 ms on the 2048^2 grid).
template <int ND, int NQ, int W, int GEO>
constexpr int update_waves() { return (W == 1 && ND == 2 && NQ == 6 && GEO == CART) ? 5 : W; }
template <bool PK, int ND, int NQ, int SV, int W, int GEO>
__global__ void __launch_bounds__(KB, (update_waves<ND, NQ, W, GEO>())) k_update(Dev d, int ni, int remap, int stage) {
  GH_TILE_LOOP(PK, if (i < ni && j < d.nj) {
    if (stage > 1) body_dt(d, j, i);
    body_update_s<ND, NQ, SV, GEO>(d, j, i, stage);
  })
}
GH_KERNEL2P((, int ND, int NQ), k_inject, (body_inject<ND, NQ>))
template <int ND, int NQ> GH_KERNEL2(k_init_gammas, (body_init_gammas<ND, NQ>))
GH_KERNEL2P((, int ND, int NQ), k_rg_screen, (body_rg_screen<ND, NQ>))

// Tile list of the packed launch: ceil(ntrack / bs) chunks of every track, in
// track order (one workgroup; block-wide scan of the per-thread counts).
__global__ void __launch_bounds__(1024) k_tiles(Dev d, int bs) {
  __shared__ int part[1024];
  int tid = threadIdx.x;
  int per = (d.nj + 1023) / 1024;
  int j0 = tid * per, j1 = min(j0 + per, d.nj);
  int sum = 0;
  for (int j = j0; j < j1; ++j) sum += (d.ntrack[j] + bs - 1) / bs;
  part[tid] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int t = part[tid] - sum;
  for (int j = j0; j < j1; ++j) {
    int n = (d.ntrack[j] + bs - 1) / bs;
    for (int c = 0; c < n; ++c, ++t) {
      d.tiles[2 * t] = j;
      d.tiles[2 * t + 1] = c;
    }
  }
  if (tid == 1023) d.ntiles[0] = part[1023];
}

// per-track minimum of dt_loc with reset (1d.cpp:779-795, 2d.cpp:954-971)
// FORM: the stage-1 candidates of the faces too (body_dt's, folded into the
// minimum instead of stored: the step's collect_dt reads dtloc right after).
// One workgroup per chunk of DT_CH cells of a track (workgroup j * dtch + ch),
// so that a long track spreads over many CUs; its minimum goes to tpart and
// k_dt_finish takes the track's minimum over its chunks (a minimum does not
// depend on the order it is taken in).
template <bool FORM>
__global__ void __launch_bounds__(KB) k_dt_track_min(Dev d) {
  __shared__ double sm[KB];
  int j = (int)blockIdx.x / d.dtch, ch = (int)blockIdx.x - j * d.dtch;
  int nt = d.ntrack[j];
  int i0 = ch * DT_CH, i1 = min(nt, i0 + DT_CH);
  if (i0 >= nt) {
    if (threadIdx.x == 0) d.tpart[blockIdx.x] = 1.e15;
    return;
  }
  double m = 1.e15;
  for (int i = i0 + (int)threadIdx.x; i < i1; i += KB) {
    long long k = CI(d, j, i);
    double v = d.dtloc[k];
    if (FORM && i >= 1 && i <= nt - 2) {
      double cand = dt_cand(d, k);
      if (cand < 0) GH_ATOMIC_ADD(cnt_at(d, CNT_NEGDT), 1ull);
      v = fmin(v, cand);
    }
    m = fmin(m, v);
    d.dtloc[k] = 1.e15;
  }
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = KB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] = fmin(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) d.tpart[blockIdx.x] = sm[0];
}
// each track's minimum over its chunks; dt = CFL * min: one global value for
// coupled tracks, per track for independent runs
__global__ void __launch_bounds__(KB) k_dt_finish(Dev d) {
  __shared__ double sm[KB];
  double m = 1.e15;
  for (int j = threadIdx.x; j < d.nj; j += KB) {
    double mj = 1.e15;
    for (int ch = 0; ch < d.dtch; ++ch) mj = fmin(mj, d.tpart[(long long)j * d.dtch + ch]);
    d.tmin[j] = mj;
    m = fmin(m, mj);
  }
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = KB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] = fmin(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  for (int j = threadIdx.x; j < d.nj; j += KB) d.tdt[j] = d.P.cfl * (d.nd == 2 ? sm[0] : d.tmin[j]);
}
// exclusive prefix sum of upper-neighbour counts along each track (F1 face slots)
// inclusive prefix sum of one value per thread over the workgroup: shuffles
// within each wavefront, the wavefront totals through LDS; *total = the sum
// over the workgroup
__device__ int wg_scan_incl(int v, int* wsum, int* total) {
  int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  if (lane == 63) wsum[w] = v;
  __syncthreads();
  int add = 0, tot = 0;
  for (int q = 0; q < KB / 64; ++q) {
    if (q < w) add += wsum[q];
    tot += wsum[q];
  }
  __syncthreads();
  *total = tot;
  return v + add;
}
// Per track: exclusive scans of the F1 face counts (slot offsets, owner of
// each slot, the track's total) and of the fresh-source counts (down-slot
// offsets), KB cells at a time; the next chunk's counts are loaded while the
// current chunk is scanned.
__global__ void __launch_bounds__(KB) k_f1_scan(Dev d) {
  __shared__ int wsum[KB / 64];
  int j = blockIdx.x;
  int nt = d.ntrack[j];
  const int* cnt = d.nbucnt;
  int base = 0;
  int vnext = (int)threadIdx.x < nt ? cnt[CI(d, j, threadIdx.x)] : 0;
  for (int i0 = 0; i0 < nt; i0 += KB) {
    int i = i0 + (int)threadIdx.x;
    int v = vnext < 0 ? 0 : vnext;
    vnext = i + KB < nt ? cnt[CI(d, j, i + KB)] : 0;
    int tot;
    int off = base + wg_scan_incl(v, wsum, &tot) - v;
    if (i < nt) {
      d.f1off[CI(d, j, i)] = off;
      for (int n = 0; n < v && off + n < d.fcap; ++n) d.f1own[(long long)j * d.fcap + off + n] = i;
    }
    base += tot;
  }
  if (threadIdx.x == 0) {
    d.f1tot[j] = base;
    if (base > d.fcap) atomicAdd(cnt_at(d, CNT_F1OVF), 1ull);
  }
}
// The applyRegrid loop of one track on one workgroup, in the sequential
// order.  The rule is evaluated for RGB cells at once (nothing it reads
// changes before the next action), the arithmetic of each split/merge runs on
// one lane, and the O(ntrack) work per action -- moving the cells after it
// one slot and locating the shock -- is spread over the workgroup.  Moves go
// in chunks of RGB cells, each chunk read into registers before any of it is
// written, chunks taken from the far end towards the action, which is the
// order in which the sequential loops copy.
constexpr int RGB = 512;
template <int ND, int NQ>
struct CellRec {
  static constexpr int NS = ND == 1 ? NQ - 1 : NQ;
  double p[NS], gm[ND == 2 ? NS : 1], gf[ND == 2 ? NS : 1];
  double x, dx, dt;
  int nb[6];
};
template <int ND, int NQ>
__device__ __forceinline__ void rec_load(const Dev& d, long long k, CellRec<ND, NQ>& r) {
  constexpr int NS = CellRec<ND, NQ>::NS;
  for (int s = 0; s < NS; ++s) r.p[s] = d.prim[s * d.cs + k];
  r.x = d.cx[k]; r.dx = d.cdx[k]; r.dt = d.dtloc[k];
  if (ND == 2) {
    for (int s = 0; s < NS; ++s) { r.gm[s] = d.gmv[s * d.cs + k]; r.gf[s] = d.gf1[s * d.cs + k]; }
    r.nb[0] = d.nbdj[k]; r.nb[1] = d.nbdlo[k]; r.nb[2] = d.nbdcnt[k];
    r.nb[3] = d.nbuj[k]; r.nb[4] = d.nbulo[k]; r.nb[5] = d.nbucnt[k];
  }
}
template <int ND, int NQ>
__device__ __forceinline__ void rec_store(const Dev& d, long long k, const CellRec<ND, NQ>& r) {
  constexpr int NS = CellRec<ND, NQ>::NS;
  for (int s = 0; s < NS; ++s) d.prim[s * d.cs + k] = r.p[s];
  d.cx[k] = r.x; d.cdx[k] = r.dx; d.dtloc[k] = r.dt;
  if (ND == 2) {
    for (int s = 0; s < NS; ++s) { d.gmv[s * d.cs + k] = r.gm[s]; d.gf1[s * d.cs + k] = r.gf[s]; }
    d.nbdj[k] = r.nb[0]; d.nbdlo[k] = r.nb[1]; d.nbdcnt[k] = r.nb[2];
    d.nbuj[k] = r.nb[3]; d.nbulo[k] = r.nb[4]; d.nbucnt[k] = r.nb[5];
  }
}
// cells [lo, hi) of the track at b move by dir (+1 or -1), faces [flo, fhi) too
template <int ND, int NQ>
__device__ void wg_shift(const Dev& d, long long b, int lo, int hi, int flo, int fhi, int dir) {
  int t = threadIdx.x;
  int n = hi - lo, nf = fhi - flo;
  int nmax = n > nf ? n : nf;
  for (int c0 = 0; c0 < nmax; c0 += RGB) {
    // up: serve the top first; down: serve the bottom first
    int o = c0 + t;
    int m = dir > 0 ? hi - 1 - o : lo + o;
    int mf = dir > 0 ? fhi - 1 - o : flo + o;
    bool act = o < n, actf = o < nf;
    CellRec<ND, NQ> r;
    double fx = 0.;
    if (act) rec_load<ND, NQ>(d, b + m, r);
    if (actf) fx = d.fx[b + mf];
    __syncthreads();
    if (act) rec_store<ND, NQ>(d, b + m + dir, r);
    if (actf) d.fx[b + mf + dir] = fx;
    __syncthreads();
  }
}
// minimum (MAX = false) or maximum of one int per thread over the workgroup,
// returned on every thread: shuffles within each wavefront, then the
// wavefronts' results through red[0 .. RGB/64); the barrier in front lets a
// previous use of red finish its reads
template <bool MAX>
__device__ int wg_reduce(int v, int* red) {
  for (int o = 32; o > 0; o >>= 1) {
    int u = __shfl_xor(v, o, 64);
    v = MAX ? (u > v ? u : v) : (u < v ? u : v);
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  int r = red[0];
  for (int w = 1; w < RGB / 64; ++w) r = MAX ? (red[w] > r ? red[w] : r) : (red[w] < r ? red[w] : r);
  return r;
}
// rg_shock_index over the workgroup: the last active cell with p > 1.5 eta
__device__ double wg_shock_index(const Dev& d, int j, int* red) {
  int nt = d.ntrack[j], iL = d.ngst - 1, iR = nt - d.ngst;
  double thr = 1.5 * d.rg[3];
  int best = iL + 1;
  for (int ii = iL + 1 + (int)threadIdx.x; ii <= iR - 1; ii += RGB)
    if (d.prim[PPP * d.cs + CI(d, j, ii)] > thr && ii > best) best = ii;
  double r = wg_reduce<true>(best, red);
  __syncthreads();
  return r;
}
// rg_victims over the workgroup: each lane keeps the first minimum and the
// first maximum of its cells (taken in increasing order), and the reduction
// prefers the smaller value, then the smaller index -- the first cell of the
// extreme value, as the sequential scan finds it
template <int ND, int NQ>
__device__ void wg_victims(const Dev& d, int j, double* rv, int* ri, int* is, int* ib) {
  int nt = d.ntrack[j], iL = d.ngst - 1, iR = nt - d.ngst;
  int t = threadIdx.x;
  double mn = 1.e20, mx = 0.;
  int imn = -1, imx = -1;
  for (int ii = iL + 3 + t; ii <= iR - 3; ii += RGB) {
    double v = rg_val<ND, NQ>(d, j, CI(d, j, ii));
    if (v < 0) continue;
    if (v < mn) { mn = v; imn = ii; }
    if (v > mx) { mx = v; imx = ii; }
  }
  for (int side = 0; side < 2; ++side) {
    rv[t] = side == 0 ? mn : mx;
    ri[t] = side == 0 ? imn : imx;
    __syncthreads();
    for (int w = RGB / 2; w > 0; w >>= 1) {
      if (t < w) {
        double o = rv[t + w];
        int oi = ri[t + w];
        bool better = side == 0 ? o < rv[t] : o > rv[t];
        if (oi >= 0 && (ri[t] < 0 || better || (o == rv[t] && oi < ri[t]))) { rv[t] = o; ri[t] = oi; }
      }
      __syncthreads();
    }
    if (side == 0) *is = ri[0]; else *ib = ri[0];
    __syncthreads();
  }
}
// Grid::split and Grid::merge of cell i over the workgroup: the cells after
// the action move one slot (wg_shift), its arithmetic runs on lane 0
template <int ND, int NQ>
__device__ void wg_split(const Dev& d, int j, long long b, int i) {
  int nt = d.ntrack[j];
  __syncthreads();
  wg_shift<ND, NQ>(d, b, i + 1, nt, i, nt - 1, +1);
  if (threadIdx.x == 0) {
    d.ntrack[j] = nt + 1;
    split_fill<ND, NQ>(d, j, i);
    d.nsplit[j]++;
  }
  __syncthreads();
}
template <int ND, int NQ>
__device__ void wg_merge(const Dev& d, int j, long long b, int i, int* s_left) {
  int nt = d.ntrack[j];
  __syncthreads();
  if (threadIdx.x == 0) *s_left = merge_fill<ND, NQ>(d, j, i) ? 1 : 0;
  __syncthreads();
  if (*s_left) wg_shift<ND, NQ>(d, b, i, nt, i, nt - 1, -1);
  else wg_shift<ND, NQ>(d, b, i + 2, nt, i + 1, nt - 1, -1);
  if (threadIdx.x == 0) {
    d.ntrack[j] = nt - 1;
    d.nmerge[j]++;
  }
  __syncthreads();
}
template <int ND, int NQ>
__global__ void __launch_bounds__(RGB) k_rg_apply(Dev d) {
  __shared__ int red[RGB];
  __shared__ double rv[RGB];
  __shared__ int s_act;
  int j = blockIdx.x;
  if (!active_track(d, j)) return;
  int i = d.tfirst[j];
  if (i == INT_MAX) return;
  bool bm = d.trg[j] == RG_BM1D;
  long long b = CI(d, j, 0);
  double iS = bm ? (double)d.tiS[j] : 0.;    // as body_rg_apply
  int is = -1, ib = -1;
  if (d.circ) wg_victims<ND, NQ>(d, j, rv, red, &is, &ib);
  RgStop stop = rg_stop_init(d, j, iS);
  int last_mod = -1;                       // the last cell a split or merge wrote
  for (;;) {
    int nt = d.ntrack[j];
    int last = nt - d.ngst - 1;
    if (i > last) break;
    if (rg_stop(d, j, stop, i, last_mod, iS)) break;
    // rg_next over the workgroup: the next cell the screen flagged, 4 RGB
    // cells a round
    if (stop.on && i > last_mod && rg_inputs_same(d, j, stop, iS)) {
      int dn = nt - stop.nt0, nx = INT_MAX;
      for (int o = i - dn > 0 ? i - dn : 0; o <= stop.tl; o += 4 * RGB) {
        int f = INT_MAX;
        for (int u = 0; u < 4; ++u) {
          int mm = o + 4 * (int)threadIdx.x + u;
          if (mm <= stop.tl && d.rgflag[b + mm]) { f = mm; break; }
        }
        nx = wg_reduce<false>(f, red);
        if (nx != INT_MAX) break;
      }
      if (nx == INT_MAX) break;
      i = nx + dn;
    }
    // The rule of a cell reads that cell, the track's last active cell and the
    // shock index, none of which change until an action is taken: evaluate
    // the next RGB cells at once and take the first action among them, which
    // is what the sequential loop would reach next.
    int m = i + (int)threadIdx.x;
    int a = 0;
    if (m <= last) {
      int r = rg_check<ND, NQ>(d, j, m, iS);
      a = (r == RG_SPLIT && nt < d.nax - 1) ? 1 : (r == RG_MERGE && nt > 2 * d.ngst + 2) ? 2 : 0;
    }
    int first = wg_reduce<false>(a ? (int)threadIdx.x : RGB, red);
    if (first < RGB && (int)threadIdx.x == first) s_act = a;
    __syncthreads();
    if (first == RGB) { i += RGB; continue; }
    i += first;
    int act = s_act, delta = 0;
    __syncthreads();
    // is, ib and i are the same on every lane: a clash returns the whole workgroup
    if (act == 1) {
      int isplit = i;
      if (d.circ) {
        if (is < 0 || is == i) {
          if (threadIdx.x == 0) GH_ATOMIC_ADD(cnt_at(d, CNT_RGCLASH), 1ull);
          return;
        }
        wg_merge<ND, NQ>(d, j, b, is, &s_act);
        if (i > is) isplit--;
      }
      wg_split<ND, NQ>(d, j, b, isplit);
      last_mod = i + 1;
      delta = 1;
    } else {
      if (d.circ && (ib < 0 || ib == i)) {
        if (threadIdx.x == 0) GH_ATOMIC_ADD(cnt_at(d, CNT_RGCLASH), 1ull);
        return;
      }
      wg_merge<ND, NQ>(d, j, b, i, &s_act);
      if (d.circ) wg_split<ND, NQ>(d, j, b, ib > i ? ib - 1 : ib);
      last_mod = i;
      delta = -1;
    }
    if (bm) iS = !d.circ && rg_shock_shifts(iS, delta, last_mod) ? iS + delta : wg_shock_index(d, j, red);
    if (d.circ) wg_victims<ND, NQ>(d, j, rv, red, &is, &ib);
    ++i;
  }
}
}  // namespace
#endif


// ===========================================================================
// learned cons2prim guess: training and batch inference kernels
// pointwise cons2prim of raw states (D, |S|, tau) from given starting
// pressures; with guess2 != null each state starts from whichever of the two
// guesses has the smaller residual |f| (used[s] = 1 when guess[s] won)
GH_HD void body_c2p_batch(long long s, const double* st, const double* guess, const double* guess2, double G,
                          int solver, int maxit, double* p_out, int* iters, int* fail, int* used) {
  double D = st[3 * s], S = st[3 * s + 1], tau = st[3 * s + 2];
  double prim[6] = {0., 0., 0., 0., 0., 0.}, cons[6] = {0., 0., 0., 0., 0., 0.};
  // with a second start, guess2 plays the previous pressure and guess the challenger
  prim[PPP] = guess2 ? guess2[s] : guess[s];
  cons[DEN] = D; cons[SS1] = S; cons[TAU] = tau;
  C2PStat cs = {0, 0, 0, 0};
  cons2prim<1, 6>(prim, cons, 1., CART, G, solver, guess2 ? guess[s] : -1., maxit, &cs);
  p_out[s] = prim[PPP];
  iters[s] = cs.iters + cs.evals;
  fail[s] = cs.fail;
  if (used) used[s] = guess2 ? cs.used : 1;
}

#ifndef GH_EMU
namespace {
using namespace gh::mlp;

// normalised features and targets of raw states (D, |S|, tau, p)
__global__ void k_mlp_prepare(const float* __restrict__ w, const double* __restrict__ st, long long n,
                              float* __restrict__ X, float* __restrict__ T) {
  long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  double D = st[4 * s], S = st[4 * s + 1], tau = st[4 * s + 2], p = st[4 * s + 3];
  features(w, D, S, tau, X + 3 * s);
  double lpd = log(p / (D > 1e-300 ? D : 1e-300));
  T[s] = (float)((lpd - w[ONORM + 6]) * w[ONORM + 7]);
}

// gradient of the mean squared error over the samples idx[0..nb) (FP32
// forward/backward, one wavefront of 64 samples per block), accumulated into g
__global__ void __launch_bounds__(64) k_mlp_grad(const float* __restrict__ w, const float* __restrict__ X,
                                                 const float* __restrict__ T, const long long* __restrict__ idx,
                                                 int nb, float* __restrict__ g, float* __restrict__ loss) {
  __shared__ float h1s[64][H + 1], zs[64][H + 1], dys[64], xs[64][3];
  int t = threadIdx.x;
  int s = blockIdx.x * 64 + t;
  bool valid = s < nb;
  float x[3] = {0.f, 0.f, 0.f}, h1[H], h2[H], dy = 0.f;
  if (valid) {
    long long k = idx[s];
    x[0] = X[3 * k]; x[1] = X[3 * k + 1]; x[2] = X[3 * k + 2];
    float y = forward_f32(w, x, h1, h2);
    float e = y - T[k];
    dy = 2.f * e / nb;
    atomicAdd(loss, e * e / nb);
  } else {
    for (int o = 0; o < H; ++o) { h1[o] = 0.f; h2[o] = 0.f; }
  }
  for (int o = 0; o < H; ++o) { h1s[t][o] = h1[o]; zs[t][o] = h2[o]; }
  dys[t] = dy;
  xs[t][0] = x[0]; xs[t][1] = x[1]; xs[t][2] = x[2];
  __syncthreads();
  {   // output layer: dw3[c], db3
    float a = 0.f;
    for (int q = 0; q < 64; ++q) a += dys[q] * zs[q][t];
    atomicAdd(&g[OW3 + t], a);
    if (t == 0) {
      float b = 0.f;
      for (int q = 0; q < 64; ++q) b += dys[q];
      atomicAdd(&g[OB3], b);
    }
  }
  __syncthreads();
  for (int o = 0; o < H; ++o) zs[t][o] = zs[t][o] > 0.f ? dys[t] * w[OW3 + o] : 0.f;   // dz2
  __syncthreads();
  {   // hidden layer: dW2[o=t][i], db2[t]
    float b = 0.f;
    for (int q = 0; q < 64; ++q) b += zs[q][t];
    atomicAdd(&g[OB2 + t], b);
    for (int i = 0; i < H; ++i) {
      float a = 0.f;
      for (int q = 0; q < 64; ++q) a += zs[q][t] * h1s[q][i];
      atomicAdd(&g[OW2 + t * H + i], a);
    }
  }
  __syncthreads();
  {   // dz1 of this sample (row t), then dW1, db1
    float dz1[H];
    for (int i = 0; i < H; ++i) {
      float a = 0.f;
      for (int o = 0; o < H; ++o) a += w[OW2 + o * H + i] * zs[t][o];
      dz1[i] = h1s[t][i] > 0.f ? a : 0.f;
    }
    __syncthreads();
    for (int i = 0; i < H; ++i) h1s[t][i] = dz1[i];
  }
  __syncthreads();
  {
    float b = 0.f, a0 = 0.f, a1 = 0.f, a2 = 0.f;
    for (int q = 0; q < 64; ++q) {
      float z = h1s[q][t];
      b += z; a0 += z * xs[q][0]; a1 += z * xs[q][1]; a2 += z * xs[q][2];
    }
    atomicAdd(&g[OB1 + t], b);
    atomicAdd(&g[OW1 + t * 3], a0);
    atomicAdd(&g[OW1 + t * 3 + 1], a1);
    atomicAdd(&g[OW1 + t * 3 + 2], a2);
  }
}

// Adam (Kingma & Ba 2015) on the trained entries; clears the gradient
__global__ void k_mlp_adam(float* w, float* g, float* m, float* v, int step, float lr) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= NTRAIN) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float gg = g[k];
  m[k] = b1 * m[k] + (1.f - b1) * gg;
  v[k] = b2 * v[k] + (1.f - b2) * gg * gg;
  float mh = m[k] / (1.f - powf(b1, (float)step)), vh = v[k] / (1.f - powf(b2, (float)step));
  w[k] -= lr * mh / (sqrtf(vh) + eps);
  g[k] = 0.f;
}

// minibatch indices drawn with replacement (SplitMix64 counter stream)
__global__ void k_mlp_sample(long long* idx, int nb, long long n, unsigned long long seed, long long step) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nb) return;
  unsigned long long z = seed + (unsigned long long)(step * (long long)nb + s + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  idx[s] = (long long)(z % (unsigned long long)n);
}

// batch inference: pressure guesses of raw states; mfma = 1 runs the BF16 MFMA
// wavefront path, 0 the FP32 reference path
__global__ void __launch_bounds__(64) k_mlp_infer(const float* __restrict__ w, const bf16x8* __restrict__ w2b,
                                                  const double* __restrict__ st, long long n,
                                                  double* __restrict__ out, int mfma) {
  __shared__ __bf16 lds[64 * H];
  __shared__ float rows[64];
  long long s = (long long)blockIdx.x * 64 + threadIdx.x;
  bool valid = s < n;
  double D = valid ? st[3 * s] : 1., S = valid ? st[3 * s + 1] : 0., tau = valid ? st[3 * s + 2] : 1.;
  float x[3];
  features(w, D, S, tau, x);
  float y;
  if (mfma) y = forward_wave(w, w2b, x, lds, rows);
  else { float h1[H], h2[H]; y = forward_f32(w, x, h1, h2); }
  if (valid) out[s] = pressure_of(w, D, y);
}

// B operands of the hidden layer packed per lane on the device, the layout of
// pack_w2 (one thread per (k-block, n-block, lane), 8 values each), from
// parameters already in device memory
__global__ void __launch_bounds__(512) k_mlp_pack_w2(const float* __restrict__ w, __bf16* __restrict__ out) {
  int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= 8 * 64) return;
  int l = t % 64, blk = t / 64, kb = blk / 2, nb = blk % 2;
  for (int e = 0; e < 8; ++e) {
    int n = nb * 32 + l % 32, k = kb * 16 + 8 * (l / 32) + e;
    out[t * 8 + e] = (__bf16)w[OW2 + n * H + k];
  }
}

__global__ void k_c2p_batch(const double* __restrict__ st, const double* __restrict__ guess,
                            const double* __restrict__ guess2, long long n, double G, int solver, int maxit,
                            double* __restrict__ p_out, int* __restrict__ iters, int* __restrict__ fail,
                            int* __restrict__ used) {
  long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) body_c2p_batch(s, st, guess, guess2, G, solver, maxit, p_out, iters, fail, used);
}

// Update stage with the learned starting pressure.  The network runs per
// wavefront (BF16 MFMA), and only in wavefronts where some cell's previous
// pressure leaves a relative residual above d.mlp_thr; cons2prim then judges
// the two starts by |f|.  Every thread of the grid calls the body (threads
// past the track end with valid = false), as the network needs all 64 lanes.
struct MlpGuess {
  const float* w;
  const bf16x8* w2b;
  __bf16* lds;
  float* rows;
  double thr;
  double G;
  __device__ double operator()(bool want, double D, double S, double tau, double pprev, int geom) const {
    bool hard = false;
    if (want) {
      double pp = geom == SPH ? fmax(1.e-13, pprev) : pprev;
      double f, sc;
      (void)c2p_df(pp, D, S, tau, G, geom, &f, &sc);
      hard = !(fabs(f) <= thr * sc);
    }
    if (__ballot(hard) == 0) return -1.;
    float x[3];
    if (want) features(w, D, S, tau, x);
    else { x[0] = 0.f; x[1] = 0.f; x[2] = 0.f; }
    float y = forward_wave(w, w2b, x, lds, rows);
    return hard ? pressure_of(w, D, y) : -1.;
  }
};

template <bool PK, int ND, int NQ>
__global__ void __launch_bounds__(KB) k_update_mlp(Dev d, int ni, int remap, int stage) {
  __shared__ __bf16 lds[KB / 64][64 * H];
  __shared__ float rows[KB / 64][64];
  int wv = threadIdx.x / 64;
  MlpGuess g{d.mw, (const bf16x8*)d.mw2b, lds[wv], rows[wv], d.mlp_thr, d.P.gamma};
  GH_TILE_LOOP(PK, {
    if (stage > 1 && i < ni && j < d.nj) body_dt(d, j, i);
    body_update_g<ND, NQ, C2P_NEWTON>(d, j, i, stage, i < ni && j < d.nj, g);
  })
}

// x^y as the kernels evaluate it (gh_pow), for numerics checks
__global__ void k_pow_batch(const double* __restrict__ x, const double* __restrict__ y,
                            double* __restrict__ out, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = gh_pow(x[i], y[i]);
}

// tanh / log10 as the kernels evaluate them, for numerics checks
__global__ void k_fn_batch(int fn, const double* __restrict__ x, double* __restrict__ out, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = fn == 0 ? gh_tanh(x[i]) : gh_log10(x[i]);
}

// lane-layout probe of v_mfma_f32_32x32x16_bf16: D = A B for row-major A (32x16), B (16x32)
__global__ void __launch_bounds__(64) k_mfma_probe(const float* A, const float* B, float* Dm) {
  int l = threadIdx.x;
  bf16x8 a, b;
  for (int e = 0; e < 8; ++e) {
    a[e] = (__bf16)A[(l % 32) * 16 + 8 * (l / 32) + e];
    b[e] = (__bf16)B[(8 * (l / 32) + e) * 32 + l % 32];
  }
  floatx16 acc = {};
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  for (int r = 0; r < 16; ++r) Dm[d_row(l, r) * 32 + l % 32] = acc[r];
}
}  // namespace
#endif

// ===========================================================================
// host side
struct KTimer {
  const char* name;
  double ms;
  long long n;
};

struct gammahip_ctx {
  gammahip_config cfg;
  Dev d;
  std::string err;
  int ntr;                 // tracks incl. ghost tracks
  std::vector<void*> allocs;
  // events for kernel timing (0 = update, 1 = faces, 2 = f1faces, 3 = f1grad, 4 = f1sum, 5 = f1cell)
  bool timing = false;
  KTimer kt[16];
  // launch options (gammahip_set_option; the autotuner picks them)
  // defaults: 128-thread workgroups in track groups (two tracks per
  // workgroup), the fastest plain launch measured on the 2D grids (2048^2:
  // 4.99 against 5.07 ms per step for 64 threads in hardware order; 2D jet:
  // 2 % faster); the XCD bands help the uniform grid a little more but cost
  // the uneven ones up to 2x
  int block = 128;         // threads per workgroup of the per-cell kernels (64..256)
  int remap = 2;           // tile order: 0 hardware, 1 XCD bands, 2 track groups
  int use_graph = 0;       // replay steps from a captured hipGraph
  int persist = 0;         // packed launch: at most this many persistent workgroups per CU (0: off),
                           // fewer where the kernel's occupancy allows fewer (occ_blocks)
  int occ = 0;             // bit 0: F1 face kernel built for 4 wavefronts per SIMD; bit 1: update kernel for 5
  int f1cell = 0;          // fused transverse-gradient + own-face pass (k_f1cell) instead of k_f1grad + k_f1faces
  int fuse = 0;            // bit 0: neighbour search inside the transverse-gradient pass (k_nbgrad);
                           // bit 1: MV gradients recomputed in the F1 passes (Dev::gmvre);
                           // bit 2: F1 face pass one thread per owner cell (k_f1own)
  int ncu = 256;           // compute units of the device
#ifndef GH_EMU
  hipStream_t stream = nullptr;
  // event pairs of timed launches, read back when the times are asked for
  // (or the pool is full): no synchronisation between the timed kernels, so
  // each starts behind its predecessor as in an untimed step
  static constexpr int NEV = 128;
  hipEvent_t ev0[NEV], ev1[NEV];
  int ev_id[NEV];
  int nev = 0;
  hipGraphExec_t graph = nullptr;   // two steps (an even number of buffer swaps)
  void* mlp_w = nullptr;            // learned-guess parameters (gammahip_mlp_load)
  void* mlp_w2b = nullptr;
  // workgroups per CU that can be resident, per (kernel, block size, LDS bytes)
  std::map<std::pair<const void*, std::pair<int, size_t>>, int> occ_cache;
#endif
};

namespace {

int fail(gammahip_ctx* c, const std::string& m) {
  if (c) c->err = m;
  return -1;
}

#ifndef GH_EMU
#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) return fail(c, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)
#endif

template <class T>
int dalloc(gammahip_ctx* c, T** p, size_t n) {
  size_t bytes = sizeof(T) * (n ? n : 1);
#ifdef GH_EMU
  *p = (T*)calloc(1, bytes);
  if (!*p) return fail(c, "allocation failed");
#else
  HIPCHK(hipMalloc((void**)p, bytes));
  HIPCHK(hipMemset(*p, 0, bytes));
#endif
  c->allocs.push_back((void*)*p);
  return 0;
}

int h2d(gammahip_ctx* c, void* dst, const void* src, size_t bytes) {
#ifdef GH_EMU
  memcpy(dst, src, bytes);
  return 0;
#else
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return 0;
#endif
}
int d2h(gammahip_ctx* c, void* dst, const void* src, size_t bytes) {
#ifdef GH_EMU
  memcpy(dst, src, bytes);
  return 0;
#else
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return 0;
#endif
}

// ---- launches ----
#ifdef GH_EMU
#define L2(c, body, ntr, ni)                                   \
  for (int j_ = 0; j_ < (ntr); ++j_)                           \
    for (int i_ = 0; i_ < (ni); ++i_) body((c)->d, j_, i_)
#define L2S(c, body, ntr, ni, st)                              \
  for (int j_ = 0; j_ < (ntr); ++j_)                           \
    for (int i_ = 0; i_ < (ni); ++i_) body((c)->d, j_, i_, st)
#define L2P(c, body, kern, targs, ntr, ni) L2(c, body, ntr, ni)
#define L2SP(c, body, kern, targs, ntr, ni, st) L2S(c, body, ntr, ni, st)
#else
// grid and tile order of a plain per-cell launch over ntr tracks x ni cells:
// one workgroup per chunk of the block size, or with the track-group order
// (remap 2) one per 64-cell chunk of block/64 tracks; a launch over fewer
// tracks than the grid that is not a whole number of groups keeps the
// hardware order
struct PlainLaunch { dim3 g; int remap; };
PlainLaunch plain_launch(const gammahip_ctx* c, int ntr, int ni) {
  int bs = c->block, W = bs / 64;
  if (c->remap == 2 && (ntr == c->d.nj || ntr % W == 0))
    return {dim3((ni + 63) / 64, (ntr + W - 1) / W), 2};
  return {dim3((ni + bs - 1) / bs, ntr), c->remap == 2 ? 0 : c->remap};
}
// Grid of a packed (persistent) launch of kernel f: c->persist workgroups per
// CU, capped by how many the kernel's registers and LDS let one CU hold, so
// that a large setting fills every CU without queueing workgroups behind
// resident ones; a multiple of 8 (one band of tiles per XCD).
dim3 persist_grid(gammahip_ctx* c, const void* f, int bs, size_t lds) {
  auto key = std::make_pair(f, std::make_pair(bs, lds));
  auto it = c->occ_cache.find(key);
  int occ;
  if (it != c->occ_cache.end()) {
    occ = it->second;
  } else {
    occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, f, bs, lds) != hipSuccess || occ < 1) occ = 1;
    c->occ_cache[key] = occ;
  }
  int n = c->persist < occ ? c->persist : occ;
  if (n < 1) n = 1;
  return dim3((unsigned)(n * c->ncu));
}
#define L2(c, kern, ntr, ni)                                                         \
  do {                                                                               \
    PlainLaunch pl_ = plain_launch((c), (ntr), (ni));                                \
    hipLaunchKernelGGL(kern, pl_.g, dim3((c)->block), 0, (c)->stream, (c)->d, (int)(ni), pl_.remap); \
  } while (0)
#define L2S(c, kern, ntr, ni, st)                                                    \
  do {                                                                               \
    PlainLaunch pl_ = plain_launch((c), (ntr), (ni));                                \
    hipLaunchKernelGGL(kern, pl_.g, dim3((c)->block), 0, (c)->stream, (c)->d, (int)(ni), pl_.remap, (int)(st)); \
  } while (0)
// kernels with a packed form: kern<PK, targs...> (targs: its other template
// arguments, each after a comma, in parentheses), the packed build on a
// persistent grid of c->persist workgroups per CU, the plain one per tile
#define L2P(c, body, kern, targs, ntr, ni)                                           \
  do {                                                                               \
    int bs_ = (c)->block;                                                            \
    if ((c)->d.packed) {                                                             \
      hipLaunchKernelGGL((kern<true GH_UNPAREN targs>),                               \
                         persist_grid((c), (const void*)&kern<true GH_UNPAREN targs>, bs_, 0), dim3(bs_), 0, \
                         (c)->stream, (c)->d, (int)(ni), (c)->remap);                \
    } else {                                                                         \
      PlainLaunch pl_ = plain_launch((c), (ntr), (ni));                              \
      hipLaunchKernelGGL((kern<false GH_UNPAREN targs>), pl_.g, dim3(bs_), 0,        \
                         (c)->stream, (c)->d, (int)(ni), pl_.remap);                 \
    }                                                                                \
  } while (0)
#define L2SP(c, body, kern, targs, ntr, ni, st)                                      \
  do {                                                                               \
    int bs_ = (c)->block;                                                            \
    if ((c)->d.packed) {                                                             \
      hipLaunchKernelGGL((kern<true GH_UNPAREN targs>),                               \
                         persist_grid((c), (const void*)&kern<true GH_UNPAREN targs>, bs_, 0), dim3(bs_), 0, \
                         (c)->stream, (c)->d, (int)(ni), (c)->remap, (int)(st));     \
    } else {                                                                         \
      PlainLaunch pl_ = plain_launch((c), (ntr), (ni));                              \
      hipLaunchKernelGGL((kern<false GH_UNPAREN targs>), pl_.g, dim3(bs_), 0,        \
                         (c)->stream, (c)->d, (int)(ni), pl_.remap, (int)(st));      \
    }                                                                                \
  } while (0)
#endif

#ifdef GH_EMU
#define KSEL(emu, gpu) emu
#else
#define KSEL(emu, gpu) gpu
#endif

// tile list of the packed launch (after any change of track lengths or block size)
void build_tiles(gammahip_ctx* c) {
#ifndef GH_EMU
  if (c->d.packed) hipLaunchKernelGGL(k_tiles, dim3(1), dim3(1024), 0, c->stream, c->d, c->block);
#else
  (void)c;
#endif
}

// add the recorded launch times to the per-kernel sums
void tcollect(gammahip_ctx* c) {
#ifndef GH_EMU
  if (c->nev == 0) return;
  hipEventSynchronize(c->ev1[c->nev - 1]);
  for (int e = 0; e < c->nev; ++e) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev0[e], c->ev1[e]) != hipSuccess) continue;
    c->kt[c->ev_id[e]].ms += ms;
    c->kt[c->ev_id[e]].n++;
  }
  c->nev = 0;
#else
  (void)c;
#endif
}
void tbeg(gammahip_ctx* c, int id) {
#ifndef GH_EMU
  if (!c->timing) return;
  if (c->nev == gammahip_ctx::NEV) tcollect(c);
  c->ev_id[c->nev] = id;
  hipEventRecord(c->ev0[c->nev], c->stream);
#else
  (void)c; (void)id;
#endif
}
void tend(gammahip_ctx* c, int id) {
#ifndef GH_EMU
  if (!c->timing) return;
  (void)id;
  hipEventRecord(c->ev1[c->nev], c->stream);
  c->nev++;
#else
  (void)c; (void)id;
#endif
}

template <int ND, int NQ>
struct Pipe {
  static void ghosts(gammahip_ctx* c, bool retile = false) {
    Dev& d = c->d;
#ifdef GH_EMU
    if (ND == 2) L2(c, body_ghost_tracks, 2 * d.ngst, d.nax);
    L2(c, (body_ghost_mv<ND, NQ>), d.nj, 2 * d.ngst);
#else
    if (ND == 2) {
      int ncopy = (int)((2LL * d.ngst * d.nax + 127) / 128);
      hipLaunchKernelGGL((k_ghosts2d<NQ>), dim3(ncopy + (d.nj + 1) / 2), dim3(128), 0, c->stream, d, ncopy);
    } else {
      L2(c, (k_ghost_mv<ND, NQ>), d.nj, 2 * d.ngst);
    }
#endif
    if (retile) build_tiles(c);
  }
  // dt = false: the dt candidates of these faces are formed by the update
  // kernel that follows (stages 2 and 3, whose candidates are first read by
  // the next step's collect_dt), which has the cell's faces at hand
  static void prep(gammahip_ctx* c, bool dt = true) {
    Dev& d = c->d;
    tbeg(c, 1);
    if (d.shock) L2P(c, (body_faces<ND, NQ>), k_faces, (, ND, NQ, true), d.nj, d.nax);
    else L2P(c, (body_faces<ND, NQ>), k_faces, (, ND, NQ, false), d.nj, d.nax);
    tend(c, 1);
    if (dt) L2P(c, body_dt, k_dt, (), d.nj, d.nax);
    if (ND == 2) {
      // FUSE bit 0: the neighbour search runs inside the gradient pass below
      // (the slot scan after it needs only the lists)
      bool nbgrad = (c->fuse & 1) && !d.pc && !c->f1cell;
#ifdef GH_EMU
      nbgrad = false;              // the CPU build runs the two bodies as separate sweeps
#endif
      if (nbgrad) {
#ifndef GH_EMU
        tbeg(c, 3);
        int bs = c->block;
        PlainLaunch pl = plain_launch(c, d.nj, d.nax);
        size_t lds = sizeof(double) * bs * f1_kc<NQ>() * NQ;
        bool sph = c->cfg.geom == SPH;
        if (d.packed) {
          if (sph) hipLaunchKernelGGL((k_nbgrad<true, NQ, SPH>), persist_grid(c, (const void*)&k_nbgrad<true, NQ, SPH>, bs, lds),
                                      dim3(bs), lds, c->stream, d, d.nax, c->remap);
          else hipLaunchKernelGGL((k_nbgrad<true, NQ, CART>), persist_grid(c, (const void*)&k_nbgrad<true, NQ, CART>, bs, lds),
                                  dim3(bs), lds, c->stream, d, d.nax, c->remap);
        } else {
          if (sph) hipLaunchKernelGGL((k_nbgrad<false, NQ, SPH>), pl.g, dim3(bs), lds, c->stream, d, d.nax, pl.remap);
          else hipLaunchKernelGGL((k_nbgrad<false, NQ, CART>), pl.g, dim3(bs), lds, c->stream, d, d.nax, pl.remap);
        }
        tend(c, 3);
#endif
      } else {
        L2P(c, body_nb, k_nb, (), d.nj, d.nax);
      }
#ifdef GH_EMU
      for (int j = 0; j < d.nj; ++j) {
        int base = 0, nt = d.ntrack[j];
        for (int i = 0; i < nt; ++i) {
          long long k = CI(d, j, i);
          d.f1off[k] = base;
          int v = d.nbucnt[k] > 0 ? d.nbucnt[k] : 0;
          for (int n = 0; n < v && base + n < d.fcap; ++n) d.f1own[(long long)j * d.fcap + base + n] = i;
          base += v;
        }
        d.f1tot[j] = base;
        if (base > d.fcap) d.cnt[CNT_F1OVF]++;
      }
#else
      hipLaunchKernelGGL(k_f1_scan, dim3(d.nj), dim3(KB), 0, c->stream, d);
#endif
      if (c->f1cell) {
        tbeg(c, 5);
#ifdef GH_EMU
        L2(c, body_f1cell<NQ>, d.nj, d.nax);
#else
        int bs = c->block;
        PlainLaunch pl = plain_launch(c, d.nj, d.nax);
        size_t lds = sizeof(double) * bs * f1_kc<NQ>() * NQ;
        if (d.packed) hipLaunchKernelGGL((k_f1cell<true, NQ>), persist_grid(c, (const void*)&k_f1cell<true, NQ>, bs, lds),
                                         dim3(bs), lds, c->stream, d, d.nax, c->remap);
        else hipLaunchKernelGGL((k_f1cell<false, NQ>), pl.g, dim3(bs), lds, c->stream, d, d.nax, pl.remap);
#endif
        tend(c, 5);
      } else {
        // piecewise constant: the F1 faces take the cells' own states, the
        // transverse gradient is never read
        if (!nbgrad) tbeg(c, 3);
#ifdef GH_EMU
        if (!d.pc) L2(c, body_f1grad<NQ>, d.nj, d.nax);
#else
        if (!d.pc && !nbgrad) {
          int bs = c->block;
          PlainLaunch pl = plain_launch(c, d.nj, d.nax);
          size_t lds = sizeof(double) * bs * f1_kc<NQ>() * NQ;
          bool sph = c->cfg.geom == SPH;
          if (d.packed) {
            if (sph) hipLaunchKernelGGL((k_f1grad<true, NQ, SPH>), persist_grid(c, (const void*)&k_f1grad<true, NQ, SPH>, bs, lds),
                                        dim3(bs), lds, c->stream, d, d.nax, c->remap);
            else hipLaunchKernelGGL((k_f1grad<true, NQ, CART>), persist_grid(c, (const void*)&k_f1grad<true, NQ, CART>, bs, lds),
                                    dim3(bs), lds, c->stream, d, d.nax, c->remap);
          } else {
            if (sph) hipLaunchKernelGGL((k_f1grad<false, NQ, SPH>), pl.g, dim3(bs), lds, c->stream, d, d.nax, pl.remap);
            else hipLaunchKernelGGL((k_f1grad<false, NQ, CART>), pl.g, dim3(bs), lds, c->stream, d, d.nax, pl.remap);
          }
        }
#endif
        if (!nbgrad) tend(c, 3);
        tbeg(c, 2);
        bool f1own = (c->fuse & 4) != 0;
#ifdef GH_EMU
        if (f1own) {
          for (int j = 0; j < d.nj; ++j)
            for (int i = 0; i < d.nax; ++i) {
              if (d.shock) body_f1own<NQ, true>(d, j, i);
              else body_f1own<NQ, false>(d, j, i);
            }
        } else {
          for (int j = 0; j < d.nj; ++j)
            for (int sl = 0; sl < d.fcap; ++sl) body_f1face<NQ>(d, j, sl);
        }
#else
        if (f1own) {
          int bs = c->block;
          PlainLaunch pl = plain_launch(c, d.nj, d.nax);     // one thread per cell slot
          constexpr int WC = NQ > 6 ? 3 : 4;
          bool sph = c->cfg.geom == SPH, cap = (c->occ & 1) != 0;
          auto launch = [&](auto kern) { hipLaunchKernelGGL(kern, pl.g, dim3(bs), 0, c->stream, d, d.nax, pl.remap); };
          if (d.shock) {
            if (sph) { if (cap) launch(k_f1own<NQ, WC, true, SPH>); else launch(k_f1own<NQ, 1, true, SPH>); }
            else { if (cap) launch(k_f1own<NQ, WC, true, CART>); else launch(k_f1own<NQ, 1, true, CART>); }
          } else {
            if (sph) { if (cap) launch(k_f1own<NQ, WC, false, SPH>); else launch(k_f1own<NQ, 1, false, SPH>); }
            else { if (cap) launch(k_f1own<NQ, WC, false, CART>); else launch(k_f1own<NQ, 1, false, CART>); }
          }
        } else {
          int bs = c->block;
          PlainLaunch pl = plain_launch(c, d.nj, d.fcap);  // one thread per F1 slot of the capacity
          // 2D grids have one geometry: the kernel is built for it
          // capped build: 4 wavefronts per SIMD with six fields, 3 with ten
          // (at 4 the ten-field states spill 300 bytes per thread)
          constexpr int WC = NQ > 6 ? 3 : 4;
          bool sph = c->cfg.geom == SPH, cap = (c->occ & 1) != 0;
          auto launch = [&](auto kern) { hipLaunchKernelGGL(kern, pl.g, dim3(bs), 0, c->stream, d, pl.remap); };
          if (d.shock) {
            if (sph) { if (cap) launch(k_f1faces<NQ, WC, true, SPH>); else launch(k_f1faces<NQ, 1, true, SPH>); }
            else { if (cap) launch(k_f1faces<NQ, WC, true, CART>); else launch(k_f1faces<NQ, 1, true, CART>); }
          } else {
            if (sph) { if (cap) launch(k_f1faces<NQ, WC, false, SPH>); else launch(k_f1faces<NQ, 1, false, SPH>); }
            else { if (cap) launch(k_f1faces<NQ, WC, false, CART>); else launch(k_f1faces<NQ, 1, false, CART>); }
          }
        }
#endif
        tend(c, 2);
      }
      if (d.shock) L2P(c, body_f1_shock, k_f1_shock, (), d.nj, d.nax);
      if (!d.f1fused) {
        tbeg(c, 4);
        L2P(c, body_f1sum<NQ>, k_f1sum, (, NQ), d.nj, d.nax);
        tend(c, 4);
      }
      std::swap(d.gf1, d.gf1n);
    }
  }
  // form = true: the stage-1 dt candidates are formed here (prep ran with
  // dt = false)
  static void collect_dt(gammahip_ctx* c, bool form = false) {
    Dev& d = c->d;
#ifdef GH_EMU
    if (form) L2P(c, body_dt, k_dt, (), d.nj, d.nax);
    double gm = 1.e15;
    for (int j = 0; j < d.nj; ++j) {
      double m = 1.e15;
      for (int i = 0; i < d.ntrack[j]; ++i) {
        long long k = CI(d, j, i);
        m = fmin(m, d.dtloc[k]);
        d.dtloc[k] = 1.e15;
      }
      d.tmin[j] = m;
      gm = fmin(gm, m);
    }
    for (int j = 0; j < d.nj; ++j) d.tdt[j] = d.P.cfl * (d.nd == 2 ? gm : d.tmin[j]);
#else
    dim3 g((unsigned)(d.nj * d.dtch));
    if (form) hipLaunchKernelGGL(k_dt_track_min<true>, g, dim3(KB), 0, c->stream, d);
    else hipLaunchKernelGGL(k_dt_track_min<false>, g, dim3(KB), 0, c->stream, d);
    hipLaunchKernelGGL(k_dt_finish, dim3(1), dim3(KB), 0, c->stream, d);
#endif
  }
  static void update(gammahip_ctx* c, int stage) {
    Dev& d = c->d;
    tbeg(c, 0);
#ifdef GH_EMU
    L2S(c, (body_update<ND, NQ>), d.nj, d.nax, stage);
#else
    if (d.mw) L2SP(c, _, k_update_mlp, (, ND, NQ), d.nj, d.nax, stage);
    else {
      // 2D grids have one geometry: the kernel is built for it (1D ensembles
      // mix geometries per track)
      bool cap = (c->occ & 2) != 0;
      auto upd = [&](auto sv) {
        constexpr int SV = decltype(sv)::value;
        if constexpr (ND == 1) {
          if (cap) L2SP(c, _, k_update, (, ND, NQ, SV, 5, -1), d.nj, d.nax, stage);
          else L2SP(c, _, k_update, (, ND, NQ, SV, 1, -1), d.nj, d.nax, stage);
        } else if (c->cfg.geom == SPH) {
          if (cap) L2SP(c, _, k_update, (, ND, NQ, SV, 5, SPH), d.nj, d.nax, stage);
          else L2SP(c, _, k_update, (, ND, NQ, SV, 1, SPH), d.nj, d.nax, stage);
        } else {
          if (cap) L2SP(c, _, k_update, (, ND, NQ, SV, 5, CART), d.nj, d.nax, stage);
          else L2SP(c, _, k_update, (, ND, NQ, SV, 1, CART), d.nj, d.nax, stage);
        }
      };
      if (d.solver == C2P_NEWTON) upd(std::integral_constant<int, C2P_NEWTON>{});
      else upd(std::integral_constant<int, C2P_BRENT>{});
    }
#endif
    tend(c, 0);
    std::swap(d.fx, d.fxn);
  }
  static void regrid(gammahip_ctx* c) {
    Dev& d = c->d;
    if (!d.rg_mode) return;
    L2(c, KSEL(body_rg_init, k_rg_init), d.nj, 1);
    if (d.rg_mode & (1 << RG_BM1D)) L2P(c, body_rg_shockpos, k_rg_shockpos, (), d.nj, d.nax);
    L2P(c, (body_rg_screen<ND, NQ>), k_rg_screen, (, ND, NQ), d.nj, d.nax);
#ifdef GH_EMU
    for (int j = 0; j < d.nj; ++j) body_rg_apply<ND, NQ>(d, j, 0);
#else
    hipLaunchKernelGGL((k_rg_apply<ND, NQ>), dim3(d.nj), dim3(RGB), 0, c->stream, d);
#endif
  }
  static void apply(gammahip_ctx* c, int what) {
    Dev& d = c->d;
    switch (what) {
      case GAMMAHIP_RESET_SHOCK: if (d.shock) L2(c, KSEL(body_shock_reset, k_shock_reset), d.nj, d.nax); break;
      case GAMMAHIP_DETECT_SHOCK: if (d.shock) L2(c, KSEL(body_shock_detect, k_shock_detect), d.nj, d.nax); break;
      case GAMMAHIP_INJECT: if (NQ == 10 && d.sync) L2P(c, (body_inject<ND, NQ>), k_inject, (, ND, NQ), d.nj, d.nax); break;
      case GAMMAHIP_INIT_GAMMAS: if (NQ == 10 && d.sync) L2(c, KSEL((body_init_gammas<ND, NQ>), (k_init_gammas<ND, NQ>)), d.nj, d.nax); break;
      case GAMMAHIP_ADVANCE_TIME: L2(c, KSEL(body_advance_time, k_advance_time), d.nj, 1); break;
      default: break;
    }
  }
  // Grid::evolve (rk3.cpp:12-195)
  static void evolve(gammahip_ctx* c) {
    update(c, 1);
    if (c->d.euler) return;              // Grid::evolve of euler.cpp:12-19
    ghosts(c); prep(c, false); update(c, 2);
    ghosts(c); prep(c, false); update(c, 3);
  }
  // one iteration of Simu::run (simu.cpp:47-77)
  static void step(gammahip_ctx* c) {
    regrid(c);
    apply(c, GAMMAHIP_RESET_SHOCK);
    ghosts(c, true);          // track lengths settle here: rebuild the tile list
    prep(c, false);           // the stage-1 dt candidates: formed by collect_dt
    apply(c, GAMMAHIP_DETECT_SHOCK);
    collect_dt(c, true);
    evolve(c);
    apply(c, GAMMAHIP_ADVANCE_TIME);
    apply(c, GAMMAHIP_INJECT);
  }
};

template <class F>
int dispatch(gammahip_ctx* c, F&& f) {
  int nd = c->cfg.nd, nq = c->cfg.nq;
  if (nd == 1 && nq == 6) { f(Pipe<1, 6>()); return 0; }
  if (nd == 1 && nq == 10) { f(Pipe<1, 10>()); return 0; }
  if (nd == 2 && nq == 6) { f(Pipe<2, 6>()); return 0; }
  if (nd == 2 && nq == 10) { f(Pipe<2, 10>()); return 0; }
  return fail(c, "unsupported (nd, nq)");
}

int post(gammahip_ctx* c) {
#ifndef GH_EMU
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, std::string("kernel launch: ") + hipGetErrorString(e));
#endif
  (void)c;
  return 0;
}

}  // namespace

// ===========================================================================
// C-ABI
static void drop_graph(gammahip_ctx* c) {
#ifndef GH_EMU
  if (c->graph) { (void)hipGraphExecDestroy(c->graph); c->graph = nullptr; }
#else
  (void)c;
#endif
}

#ifndef GH_EMU
namespace {
struct DevBuf {
  std::vector<void*> p;
  ~DevBuf() { for (void* q : p) (void)hipFree(q); }
  template <class T> T* get(size_t n) {
    void* q = nullptr;
    if (hipMalloc(&q, sizeof(T) * (n ? n : 1)) != hipSuccess) return nullptr;
    p.push_back(q);
    return (T*)q;
  }
};
int plain_fail(const char* m) { fprintf(stderr, "gammahip: %s\n", m); return -1; }
#define CK(x) do { if ((x) != hipSuccess) return plain_fail(#x); } while (0)
}  // namespace

// Device memory of gammahip_mlp_infer_noalloc, in place of the buffers
// gammahip_mlp_infer allocates per call: the hidden layer's packed BF16 B
// operands (8 x 64 lanes x 8 values = 8 KiB).  Each device holds its own copy
// until the process ends.  In host code `scratch` names a host-side stand-in:
// kernels get the device address from noalloc_scratch() instead.
constexpr size_t NOALLOC_SCRATCH_BYTES = 16384;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mu;
void* noalloc_ptr[NOALLOC_MAX_DEVICES];

// device address of the current device's copy of `scratch`, looked up once per
// device (with that device current) and kept
int noalloc_scratch(void** out) {
  int dev = -1;
  CK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return plain_fail("noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(noalloc_mu);
  if (!noalloc_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return plain_fail("noalloc: hipGetSymbolAddress(scratch) failed");
    noalloc_ptr[dev] = p;
  }
  *out = noalloc_ptr[dev];
  return 0;
}
}  // namespace
#endif

extern "C" {

int gammahip_create(const gammahip_config* cfg, gammahip_ctx** out) {
  if (!cfg || !out) return -1;
  gammahip_ctx* c = new gammahip_ctx();
  *out = c;
  c->cfg = *cfg;
  for (int k = 0; k < 16; ++k) c->kt[k] = KTimer{"", 0., 0};
  if (cfg->nd != 1 && cfg->nd != 2) return fail(c, "nd must be 1 or 2");
  if (cfg->nq != 6 && cfg->nq != 10) return fail(c, "nq must be 6 or 10");
  if (cfg->sync && cfg->nq != 10) return fail(c, "synchrotron tracers need nq == 10");
  if (cfg->ngst != 2) return fail(c, "ngst must be 2");
  if (cfg->ncell_mv < 8 || cfg->nmax < cfg->ncell_mv) return fail(c, "bad ncell_mv / nmax");
  if (cfg->ntracks_active < 1) return fail(c, "bad ntracks_active");
  if (cfg->nd == 2 && cfg->geom < 0) return fail(c, "2D needs a single geometry");
  if (cfg->rg_mode < RG_NONE || cfg->rg_mode > RG_USER) return fail(c, "unknown rg_mode");
  if (cfg->timestep != 0 && cfg->timestep != 1) return fail(c, "timestep must be 0 (RK3) or 1 (Euler)");
  if (cfg->recon != 0 && cfg->recon != 1) return fail(c, "recon must be 0 (piecewise linear) or 1 (piecewise constant)");
  if (cfg->eos != 0 && cfg->eos != 1) return fail(c, "eos must be 0 (ideal) or 1 (Synge)");
  if (cfg->circ_regrid != 0 && cfg->circ_regrid != 1) return fail(c, "circ_regrid must be 0 or 1");
  if (cfg->eos == 1 && cfg->c2p_solver != 0)
    return fail(c, "the Synge equation of state runs with the Brent cons2prim (the Newton residual assumes a constant index)");
  if (cfg->c2p_guess != 0) return fail(c, "c2p_guess must be 0 at creation; gammahip_mlp_load switches the learned guess on");
#ifndef GH_EMU
  {
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev < 1) return fail(c, "no HIP device");
    if (hipStreamCreate(&c->stream) != hipSuccess) return fail(c, "stream creation failed");
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
      c->ncu = (ncu + 7) / 8 * 8;     // persistent grids are multiples of 8 (one band per XCD)
    for (int k = 0; k < gammahip_ctx::NEV; ++k) {
      hipEventCreate(&c->ev0[k]);
      hipEventCreate(&c->ev1[k]);
    }
  }
#endif
  Dev& d = c->d;
  memset(&d, 0, sizeof(d));
  d.nd = cfg->nd; d.nq = cfg->nq; d.nqs = cfg->nd == 1 ? cfg->nq - 1 : cfg->nq;
  d.ngst = cfg->ngst;
  d.nj = cfg->nd == 2 ? cfg->ntracks_active + 2 * cfg->ngst : cfg->ntracks_active;
  d.nax = cfg->nmax + 2 * cfg->ngst;
  d.jL = cfg->ngst - 1;
  d.jR = d.nj - cfg->ngst;
  d.fcap = 4 * d.nax;
  d.shock = cfg->shock; d.sync = cfg->sync;
  d.solver = cfg->c2p_solver; d.guess = cfg->c2p_guess;
  d.euler = cfg->timestep == 1; d.pc = cfg->recon == 1; d.vpspec = cfg->vpspec != 0;
  d.circ = cfg->circ_regrid != 0;
  d.mw = nullptr; d.mw2b = nullptr; d.mlp_thr = 1e-3;
  d.maxit = cfg->c2p_maxiter > 0 ? cfg->c2p_maxiter : 1000;
  d.rg_mode = 0;            // set from the tracks' rules by gammahip_load
  for (int k = 0; k < 6; ++k) d.rg[k] = cfg->rg[k];
  d.P = Phys{cfg->gamma, cfg->cfl, cfg->vi, cfg->shock_thr, cfg->gamma_max_init,
             cfg->Nalpha, cfg->Nmp, cfg->Nme, cfg->eos == 1};
  d.cs = (long long)d.nj * d.nax;
  d.fs1 = (long long)d.nj * d.fcap;
  c->ntr = d.nj;
  size_t nj = d.nj, cs = d.cs, nqs = d.nqs;
  int e = 0;
  e |= dalloc(c, &d.tiles, (size_t)2 * nj * ((d.nax + 63) / 64)); e |= dalloc(c, &d.ntiles, 1);
  e |= dalloc(c, &d.ttrig, (size_t)7 * nj);
  e |= dalloc(c, &d.upar, (size_t)GH_USER_NPAR);
  e |= dalloc(c, &d.ntrack, nj); e |= dalloc(c, &d.trg, nj); e |= dalloc(c, &d.tgeom, nj); e |= dalloc(c, &d.tkin, nj);
  e |= dalloc(c, &d.tbnd, nj); e |= dalloc(c, &d.tfirst, nj); e |= dalloc(c, &d.tiS, nj);
  e |= dalloc(c, &d.tlast, nj);
  e |= dalloc(c, &d.rgflag, (size_t)cs);
  e |= dalloc(c, &d.tvin, nj); e |= dalloc(c, &d.tvout, nj); e |= dalloc(c, &d.tbm, 4 * nj);
  e |= dalloc(c, &d.ty, nj); e |= dalloc(c, &d.tdy, nj); e |= dalloc(c, &d.tdt, nj);
  e |= dalloc(c, &d.tt, nj); e |= dalloc(c, &d.tmin, nj);
  d.dtch = (d.nax + DT_CH - 1) / DT_CH;
  e |= dalloc(c, &d.tpart, nj * (size_t)d.dtch);
  e |= dalloc(c, &d.prim, nqs * cs); e |= dalloc(c, &d.cx, cs); e |= dalloc(c, &d.cdx, cs);
  e |= dalloc(c, &d.dtloc, cs); e |= dalloc(c, &d.U0, nqs * cs); e |= dalloc(c, &d.dV0, cs);
  e |= dalloc(c, &d.fx, cs); e |= dalloc(c, &d.fxn, cs); e |= dalloc(c, &d.fx0, cs);
  e |= dalloc(c, &d.fv, cs); e |= dalloc(c, &d.flL, cs); e |= dalloc(c, &d.flR, cs);
  e |= dalloc(c, &d.fflux, nqs * cs);
  e |= dalloc(c, &d.cnt, (size_t)CNT_N * CNT_PARTS); e |= dalloc(c, &d.nsplit, nj); e |= dalloc(c, &d.nmerge, nj);
  if (d.nd == 2) {
    e |= dalloc(c, &d.gmv, nqs * cs); e |= dalloc(c, &d.gf1, nqs * cs); e |= dalloc(c, &d.gf1n, nqs * cs);
    e |= dalloc(c, &d.nbdj, cs); e |= dalloc(c, &d.nbdlo, cs); e |= dalloc(c, &d.nbdcnt, cs);
    e |= dalloc(c, &d.nbuj, cs); e |= dalloc(c, &d.nbulo, cs); e |= dalloc(c, &d.nbucnt, cs);
    e |= dalloc(c, &d.f1off, cs); e |= dalloc(c, &d.f1tot, nj);
    e |= dalloc(c, &d.f1own, (size_t)nj * d.fcap);
    e |= dalloc(c, &d.ff1, nqs * (size_t)d.fs1);
    if (d.shock) e |= dalloc(c, &d.ff1sd, 4 * (size_t)d.fs1);
    e |= dalloc(c, &d.f1net, nqs * cs);
  }
  if (d.shock) { e |= dalloc(c, &d.sd, cs); e |= dalloc(c, &d.pspec, cs); e |= dalloc(c, &d.shk, cs); }
  if (e) return -1;
  return 0;
}

int gammahip_destroy(gammahip_ctx* c) {
  if (!c) return 0;
#ifdef GH_EMU
  for (void* p : c->allocs) free(p);
#else
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->mlp_w) (void)hipFree(c->mlp_w);
  if (c->mlp_w2b) (void)hipFree(c->mlp_w2b);
  if (c->stream) {
    for (int k = 0; k < gammahip_ctx::NEV; ++k) { hipEventDestroy(c->ev0[k]); hipEventDestroy(c->ev1[k]); }
    hipStreamDestroy(c->stream);
  }
#endif
  delete c;
  return 0;
}

const char* gammahip_last_error(gammahip_ctx* c) { return c ? c->err.c_str() : "null context"; }

int gammahip_load(gammahip_ctx* c, const double* x, const double* dx, const double* y,
                  const double* dy, const double* prim, const gammahip_track_hooks* hooks,
                  double t0) {
  if (!c) return -1;
  Dev& d = c->d;
  const gammahip_config& cf = c->cfg;
  int nj = d.nj, nax = d.nax, ng = d.ngst, nm = cf.ncell_mv, na = cf.ntracks_active;
  size_t cs = d.cs;
  // host images of the initial state (Cinit + mpi_distribute, 1d.cpp:956-969, 2d.cpp:1175-1195)
  std::vector<int> ntrack(nj, nm + 2 * ng), tgeom(nj), tkin(nj), tbnd(nj), trg(nj, RG_NONE);
  std::vector<double> tvin(nj), tvout(nj), tbm(4 * nj), ty(nj), tdy(nj), tt(nj, t0);
  std::vector<double> hp((size_t)d.nqs * cs, 0.), hx(cs, 0.), hdx(cs, 0.), hfx(cs, 0.), hdt(cs, 1.e15);
  auto trk = [&](int a) { return cf.nd == 2 ? a + ng : a; };
  for (int a = 0; a < na; ++a) {
    int j = trk(a);
    const gammahip_track_hooks& h = hooks[a];
    int nc = h.ncell > 0 ? h.ncell : nm;
    if (nc < 2 || nc > nm) return fail(c, "track " + std::to_string(a) + ": ncell outside 2..ncell_mv");
    if (h.rg_mode < -1 || h.rg_mode > RG_USER) return fail(c, "track " + std::to_string(a) + ": unknown rg_mode");
    if (h.kin_mode < KIN_NONE || h.kin_mode > KIN_USER) return fail(c, "track " + std::to_string(a) + ": unknown kin_mode");
    if (h.bnd_mode < BND_NONE || h.bnd_mode > BND_USER) return fail(c, "track " + std::to_string(a) + ": unknown bnd_mode");
    ntrack[j] = nc + 2 * ng;
    trg[j] = h.rg_mode >= 0 ? h.rg_mode : cf.rg_mode;
    d.rg_mode |= trg[j] != RG_NONE ? (1 << trg[j]) : 0;
    tgeom[j] = cf.geom >= 0 ? cf.geom : h.geom;
    tkin[j] = h.kin_mode; tvin[j] = h.v_in; tvout[j] = h.v_out; tbnd[j] = h.bnd_mode;
    for (int q = 0; q < 4; ++q) tbm[4 * j + q] = h.bm[q];
    ty[j] = cf.nd == 2 ? y[a] : 0.;
    tdy[j] = cf.nd == 2 ? dy[a] : 0.;
    for (int ii = 0; ii < nc; ++ii) {
      size_t k = (size_t)j * nax + ii + ng, s = (size_t)a * nm + ii;
      hx[k] = x[s]; hdx[k] = dx[s];
      double pv[10];
      for (int q = 0; q < cf.nq; ++q) pv[q] = prim[((size_t)q * na + a) * nm + ii];
      // Grid::v2u (1d.cpp:934-951)
      double v = 0;
      for (int dd = 0; dd < cf.nd; ++dd) v += pv[UU1 + dd] * pv[UU1 + dd];
      v = sqrt(v);
      double lf = 1. / sqrt(1. - v * v);
      for (int dd = 0; dd < cf.nd; ++dd) pv[UU1 + dd] *= lf;
      for (int q = 0; q < cf.nq; ++q) {
        if (cf.nd == 1 && q == 3) continue;
        int s2 = (cf.nd == 1 && q > 3) ? q - 1 : q;
        hp[(size_t)s2 * cs + k] = pv[q];
      }
    }
  }
  if (cf.nd == 2) {
    // ghost tracks: positions shifted from the edge tracks (2d.cpp:259, 281)
    double dyL = tdy[ng], dyR = tdy[nj - ng - 1];
    for (int j = 0; j <= d.jL; ++j) {
      int js = d.jL + 1;
      ty[j] = ty[js] - (d.jL - j + 1) * dyL; tdy[j] = tdy[js];
      tgeom[j] = tgeom[js]; tkin[j] = tkin[js]; tvin[j] = tvin[js]; tvout[j] = tvout[js]; tbnd[j] = tbnd[js];
      ntrack[j] = ntrack[js]; trg[j] = trg[js];
      for (int q = 0; q < 4; ++q) tbm[4 * j + q] = tbm[4 * js + q];
    }
    for (int j = d.jR; j < nj; ++j) {
      int js = d.jR - 1;
      ty[j] = ty[js] + (j - d.jR + 1) * dyR; tdy[j] = tdy[js];
      tgeom[j] = tgeom[js]; tkin[j] = tkin[js]; tvin[j] = tvin[js]; tvout[j] = tvout[js]; tbnd[j] = tbnd[js];
      ntrack[j] = ntrack[js]; trg[j] = trg[js];
      for (int q = 0; q < 4; ++q) tbm[4 * j + q] = tbm[4 * js + q];
    }
  }
  // Grid::interfaceGeomFromCellPos (1d.cpp:844-856, 2d.cpp:1039-1058), active tracks
  for (int a = 0; a < na; ++a) {
    int j = trk(a), iL = ng - 1, nt = ntrack[j];
    for (int i = 0; i < nt - 1; ++i) {
      size_t k = (size_t)j * nax + i;
      if (i != iL) hfx[k] = hx[k] + hdx[k] / 2.;
      else hfx[k] = hx[k + 1] - hdx[k + 1] / 2.;
    }
  }
  int e = 0;
  e |= h2d(c, d.ntrack, ntrack.data(), sizeof(int) * nj);
  e |= h2d(c, d.trg, trg.data(), sizeof(int) * nj);
  e |= h2d(c, d.tgeom, tgeom.data(), sizeof(int) * nj);
  e |= h2d(c, d.tkin, tkin.data(), sizeof(int) * nj);
  e |= h2d(c, d.tbnd, tbnd.data(), sizeof(int) * nj);
  e |= h2d(c, d.tvin, tvin.data(), sizeof(double) * nj);
  e |= h2d(c, d.tvout, tvout.data(), sizeof(double) * nj);
  e |= h2d(c, d.tbm, tbm.data(), sizeof(double) * 4 * nj);
  e |= h2d(c, d.ty, ty.data(), sizeof(double) * nj);
  {
    std::vector<double> tt7((size_t)7 * nj, 0.);
    for (int j = 0; j < nj; ++j) {
      Trig t = trig_of(ty[j], tdy[j]);
      double* a = &tt7[(size_t)7 * j];
      a[0] = t.s_hdy; a[1] = t.s_y; a[2] = t.c_y; a[3] = t.s_1; a[4] = t.s_2; a[5] = t.c_1; a[6] = t.c_2;
    }
    e |= h2d(c, d.ttrig, tt7.data(), sizeof(double) * 7 * nj);
  }
  e |= h2d(c, d.tdy, tdy.data(), sizeof(double) * nj);
  e |= h2d(c, d.tt, tt.data(), sizeof(double) * nj);
  e |= h2d(c, d.prim, hp.data(), sizeof(double) * hp.size());
  e |= h2d(c, d.cx, hx.data(), sizeof(double) * cs);
  e |= h2d(c, d.cdx, hdx.data(), sizeof(double) * cs);
  e |= h2d(c, d.fx, hfx.data(), sizeof(double) * cs);
  e |= h2d(c, d.dtloc, hdt.data(), sizeof(double) * cs);
  if (e) return -1;
  if (d.shock) {
    std::vector<double> m1(cs, -1.);
    e |= h2d(c, d.sd, m1.data(), sizeof(double) * cs);
  }
  if (e) return -1;
  // radiation_initGammas (simu.cpp:34-36), then the first updateGhosts (simu.cpp:43)
  int r = dispatch(c, [&](auto P) {
    P.apply(c, GAMMAHIP_INIT_GAMMAS);
    P.ghosts(c, true);
  });
  if (r) return r;
  return post(c);
}

int gammahip_mlp_load(gammahip_ctx* c, const float* params) {
  if (!c) return -1;
  Dev& d = c->d;
#ifdef GH_EMU
  if (params) return fail(c, "the learned guess needs the HIP build");
  return 0;
#else
  drop_graph(c);
  if (!params) {
    d.mw = nullptr; d.mw2b = nullptr; d.guess = 0;
    return 0;
  }
  if (d.solver != gh::C2P_NEWTON) return fail(c, "the learned guess seeds the Newton solver (c2p_solver = 1)");
  if (!c->mlp_w) {
    HIPCHK(hipMalloc(&c->mlp_w, sizeof(float) * gh::mlp::NPARAM));
    HIPCHK(hipMalloc(&c->mlp_w2b, sizeof(__bf16) * 8 * 64 * 8));
  }
  std::vector<__bf16> hb(8 * 64 * 8);
  gh::mlp::pack_w2(params, hb.data());
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(c->mlp_w, params, sizeof(float) * gh::mlp::NPARAM, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->mlp_w2b, hb.data(), sizeof(__bf16) * hb.size(), hipMemcpyHostToDevice));
  d.mw = (const float*)c->mlp_w;
  d.mw2b = c->mlp_w2b;
  d.guess = 1;
  return 0;
#endif
}

int gammahip_set_user_params(gammahip_ctx* c, const double* par, int n) {
  if (!c) return -1;
  if (n < 0 || n > GH_USER_NPAR || (n > 0 && !par))
    return fail(c, "user parameters: n must be in 0.." + std::to_string((int)GH_USER_NPAR));
  double v[GH_USER_NPAR] = {0.};
  for (int q = 0; q < n; ++q) v[q] = par[q];
#ifndef GH_EMU
  HIPCHK(hipStreamSynchronize(c->stream));
#endif
  return h2d(c, c->d.upar, v, sizeof(v));
}

int gammahip_set_option(gammahip_ctx* c, int key, long long value) {
  if (!c) return -1;
  drop_graph(c);
  switch (key) {
    case GAMMAHIP_OPT_BLOCK:
      if (value != 64 && value != 128 && value != 256) return fail(c, "block must be 64, 128 or 256");
      c->block = (int)value;
      build_tiles(c);
      return post(c);
    case GAMMAHIP_OPT_F1FUSE: c->d.f1fused = value ? 1 : 0; return 0;
    case GAMMAHIP_OPT_F1CELL: c->f1cell = value ? 1 : 0; return 0;
    case GAMMAHIP_OPT_FUSE:
      if (value < 0 || value > 7) return fail(c, "fuse option must be in 0..7");
      c->fuse = (int)value;
      c->d.gmvre = (value & 2) ? 1 : 0;
      return 0;
    case GAMMAHIP_OPT_OCC:
      if (value < 0 || value > 3) return fail(c, "occupancy option must be in 0..3");
      c->occ = (int)value; return 0;
    case GAMMAHIP_OPT_PERSIST:
      if (value < 0 || value > 32) return fail(c, "persistent workgroups per CU must be in 0..32");
      c->persist = (int)value;
      c->d.packed = value > 0;
      build_tiles(c);
      return post(c);
    case GAMMAHIP_OPT_SOLVER:
      if (value != gh::C2P_BRENT && value != gh::C2P_NEWTON) return fail(c, "solver must be 0 (Brent) or 1 (Newton)");
      if (value == gh::C2P_BRENT && c->d.mw) return fail(c, "the learned guess is loaded: it seeds the Newton solver");
      if (value == gh::C2P_NEWTON && c->d.P.synge) return fail(c, "the Synge equation of state runs with the Brent cons2prim");
      c->d.solver = (int)value; return 0;
    case GAMMAHIP_OPT_REMAP:
      if (value < 0 || value > 2) return fail(c, "remap must be 0 (hardware order), 1 (XCD bands) or 2 (track groups)");
      c->remap = value;
      return 0;
    case GAMMAHIP_OPT_GRAPH: c->use_graph = value ? 1 : 0; return 0;
    case GAMMAHIP_OPT_USER_CELL:
      if (value < 0 || value > 3) return fail(c, "user cell hooks option must be in 0..3");
      c->d.ucell = (int)value; return 0;
    case GAMMAHIP_OPT_MLP_THR:
      if (value < 0 || value > 16) return fail(c, "network threshold exponent must be in 0..16");
      c->d.mlp_thr = pow(10., -(double)value); return 0;
    default: return fail(c, "unknown option");
  }
}

int gammahip_step(gammahip_ctx* c, int nsteps) {
  if (!c) return -1;
  int s = 0;
#ifndef GH_EMU
  if (c->use_graph && !c->timing && nsteps >= 2) {
    if (!c->graph) {
      // two steps = six stage updates: the face-position and transverse-gradient
      // double buffers come back to their starting roles, so the pointers baked
      // into the graph stay valid for every replay
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeGlobal));
      int r = dispatch(c, [&](auto P) { P.step(c); P.step(c); });
      hipError_t e = hipStreamEndCapture(c->stream, &g);
      if (r || e != hipSuccess) return fail(c, "graph capture failed");
      e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) { c->graph = nullptr; return fail(c, "graph instantiation failed"); }
    }
    for (; s + 1 < nsteps; s += 2) HIPCHK(hipGraphLaunch(c->graph, c->stream));
  }
#endif
  for (; s < nsteps; ++s) {
    drop_graph(c);       // a single step flips the buffer roles the graph baked in
    int r = dispatch(c, [&](auto P) { P.step(c); });
    if (r) return r;
    if (post(c)) return -1;
  }
  return 0;
}

int gammahip_regrid(gammahip_ctx* c) { drop_graph(c); return dispatch(c, [&](auto P) { P.regrid(c); }) || post(c); }
int gammahip_update_ghosts(gammahip_ctx* c) { drop_graph(c); return dispatch(c, [&](auto P) { P.ghosts(c, true); }) || post(c); }
int gammahip_prep_for_update(gammahip_ctx* c) { drop_graph(c); return dispatch(c, [&](auto P) { P.prep(c); }) || post(c); }
int gammahip_evolve(gammahip_ctx* c) { drop_graph(c); return dispatch(c, [&](auto P) { P.evolve(c); }) || post(c); }
int gammahip_apply(gammahip_ctx* c, int what) { drop_graph(c); return dispatch(c, [&](auto P) { P.apply(c, what); }) || post(c); }

int gammahip_collect_dt(gammahip_ctx* c, double* dt_out) {
  drop_graph(c);
  if (dispatch(c, [&](auto P) { P.collect_dt(c); }) || post(c)) return -1;
  if (dt_out) return d2h(c, dt_out, c->d.tdt, sizeof(double));
  return 0;
}

int gammahip_collect_dt_min(gammahip_ctx* c, double* min_out) {
  if (!c || !min_out) return -1;
  if (c->cfg.nd == 1 && c->d.nj != 1)
    return fail(c, "collect_dt_min: a 1D context holding several runs has one minimum per run "
                   "(use gammahip_collect_dt and gammahip_time)");
  drop_graph(c);
  if (dispatch(c, [&](auto P) { P.collect_dt(c); }) || post(c)) return -1;
  std::vector<double> tm(c->d.nj);
  if (d2h(c, tm.data(), c->d.tmin, sizeof(double) * c->d.nj)) return -1;
  double m = 1.e15;
  for (double v : tm) m = fmin(m, v);   // the order of a minimum does not change it
  *min_out = m;
  return 0;
}

int gammahip_evolve_dt(gammahip_ctx* c, double dt) {
  if (!c) return -1;
  if (c->cfg.nd == 1 && c->d.nj != 1)
    return fail(c, "evolve_dt: a 1D context holding several runs advances each by its own dt "
                   "(use gammahip_evolve)");
  if (!(dt > 0.) || !(dt < 1.e300)) return fail(c, "evolve_dt: dt must be positive and finite");
  drop_graph(c);
#ifndef GH_EMU
  HIPCHK(hipStreamSynchronize(c->stream));
#endif
  std::vector<double> v(c->d.nj, dt);
  if (h2d(c, c->d.tdt, v.data(), sizeof(double) * c->d.nj)) return -1;
  return gammahip_evolve(c);
}

int gammahip_sync(gammahip_ctx* c) {
#ifndef GH_EMU
  HIPCHK(hipStreamSynchronize(c->stream));
#endif
  (void)c;
  return 0;
}

int gammahip_dims(gammahip_ctx* c, int* ntracks, int* nax) {
  *ntracks = c->d.nj;
  *nax = c->d.nax;
  return 0;
}

// ---- exact state (binary resume) ----
namespace {
constexpr int ST_MAGIC = 0x54534847;   // "GHST"
constexpr int ST_VERSION = 5;
constexpr int ST_HDR = 16;             // ints
// bytes: the array's size; pad: its size in the buffer (a multiple of 8)
struct StatePart { void* p; size_t bytes, pad; };
std::vector<StatePart> state_parts(const Dev& d) {
  size_t cs = (size_t)d.cs, nj = (size_t)d.nj, nqs = (size_t)d.nqs, fs1 = (size_t)d.fs1;
  std::vector<StatePart> v;
  auto add = [&](void* p, size_t bytes) { v.push_back({p, bytes, (bytes + 7) & ~(size_t)7}); };
  add(d.tt, sizeof(double) * nj); add(d.tdt, sizeof(double) * nj); add(d.ntrack, sizeof(int) * nj);
  add(d.prim, sizeof(double) * nqs * cs); add(d.cx, sizeof(double) * cs); add(d.cdx, sizeof(double) * cs);
  add(d.fx, sizeof(double) * cs); add(d.fv, sizeof(double) * cs);
  add(d.dtloc, sizeof(double) * cs);   // the next step's CFL minimum still sees the last stage's
  if (d.shock) {
    add(d.sd, sizeof(double) * cs); add(d.pspec, sizeof(double) * cs); add(d.shk, sizeof(int) * cs);
  }
  if (d.nd == 2) {
    // 2D keeps gradients, F1 neighbour lists and F1 face values in memory,
    // and reads some of them a stage later: the face pass takes the upper
    // track's transverse gradient of the previous stage (2d.cpp:901 loops
    // tracks upward), ghost tracks keep the lists of earlier stages
    add(d.gmv, sizeof(double) * nqs * cs); add(d.gf1, sizeof(double) * nqs * cs);
    for (int* q : {d.nbdj, d.nbdlo, d.nbdcnt, d.nbuj, d.nbulo, d.nbucnt, d.f1off})
      add(q, sizeof(int) * cs);
    add(d.f1tot, sizeof(int) * nj); add(d.f1own, sizeof(int) * nj * d.fcap);
    add(d.ff1, sizeof(double) * nqs * fs1);
    if (d.shock) add(d.ff1sd, sizeof(double) * 4 * fs1);
  }
  return v;
}
void state_header(const Dev& d, int* h) {
  for (int k = 0; k < ST_HDR; ++k) h[k] = 0;
  h[0] = ST_MAGIC; h[1] = ST_VERSION; h[2] = d.nd; h[3] = d.nq; h[4] = d.nqs;
  h[5] = d.nj; h[6] = d.nax; h[7] = d.shock;
}
}  // namespace

int gammahip_state_size(gammahip_ctx* c, unsigned long long* nbytes) {
  if (!c || !nbytes) return -1;
  size_t n = sizeof(int) * ST_HDR;
  for (const StatePart& s : state_parts(c->d)) n += s.pad;
  *nbytes = n;
  return 0;
}

int gammahip_state_save(gammahip_ctx* c, void* buf, unsigned long long nbytes) {
  if (!c || !buf) return -1;
  unsigned long long need = 0;
  gammahip_state_size(c, &need);
  if (nbytes < need) return fail(c, "state buffer smaller than gammahip_state_size");
  char* o = (char*)buf;
  state_header(c->d, (int*)o);
  o += sizeof(int) * ST_HDR;
  for (const StatePart& s : state_parts(c->d)) {
    if (s.bytes < s.pad) memset(o + s.bytes, 0, s.pad - s.bytes);
    if (d2h(c, o, s.p, s.bytes)) return -1;
    o += s.pad;
  }
  return 0;
}

int gammahip_state_restore(gammahip_ctx* c, const void* buf, unsigned long long nbytes) {
  if (!c || !buf) return -1;
  Dev& d = c->d;
  unsigned long long need = 0;
  gammahip_state_size(c, &need);
  if (nbytes != need) return fail(c, "state buffer size differs from gammahip_state_size");
  int want[ST_HDR];
  state_header(d, want);
  const int* h = (const int*)buf;
  if (h[0] != ST_MAGIC || h[1] != ST_VERSION) return fail(c, "not a gammahip state buffer (magic/version)");
  for (int k = 2; k < 8; ++k)
    if (h[k] != want[k]) return fail(c, "state buffer dimensions differ from the context's (nd, nq, tracks, nax, shock)");
  const char* o = (const char*)buf + sizeof(int) * ST_HDR;
  // track lengths must fit the context's slots
  const int* nt = (const int*)(o + 2 * sizeof(double) * d.nj);
  for (int j = 0; j < d.nj; ++j)
    if (nt[j] < 2 * d.ngst + 2 || nt[j] > d.nax) return fail(c, "state buffer: track length out of range");
#ifndef GH_EMU
  HIPCHK(hipStreamSynchronize(c->stream));
#endif
  drop_graph(c);
  for (const StatePart& s : state_parts(d)) {
    if (h2d(c, s.p, o, s.bytes)) return -1;
    o += s.pad;
  }
  // the restored track lengths replace the loaded ones: the packed launch's
  // tile list must cover them before the next step's regrid screens the cells
  build_tiles(c);
  return post(c);
}

int gammahip_track_info(gammahip_ctx* c, int* ntrack, int* nact, int* iL, int* iR) {
  Dev& d = c->d;
  std::vector<int> nt(d.nj);
  if (d2h(c, nt.data(), d.ntrack, sizeof(int) * d.nj)) return -1;
  for (int j = 0; j < d.nj; ++j) {
    if (ntrack) ntrack[j] = nt[j];
    if (nact) nact[j] = nt[j] - 2 * d.ngst;
    if (iL) iL[j] = d.ngst - 1;
    if (iR) iR[j] = nt[j] - d.ngst;
  }
  return 0;
}

int gammahip_time(gammahip_ctx* c, double* t, double* dt) {
  Dev& d = c->d;
  if (t && d2h(c, t, d.tt, sizeof(double) * d.nj)) return -1;
  if (dt && d2h(c, dt, d.tdt, sizeof(double) * d.nj)) return -1;
  return 0;
}

int gammahip_download(gammahip_ctx* c, int field, double* out) {
  Dev& d = c->d;
  size_t cs = d.cs;
  int nq = c->cfg.nq, nd = d.nd;
  auto sidx = [&](int q) { return (nd == 1 && q > 3) ? q - 1 : q; };
  auto plane = [&](const double* src) { return d2h(c, out, src, sizeof(double) * cs); };
  if (field >= 0 && field < nq) {
    if (nd == 1 && field == 3) { memset(out, 0, sizeof(double) * cs); return 0; }
    return plane(d.prim + (size_t)sidx(field) * cs);
  }
  if (field >= 100 && field < 100 + nq) {
    int q = field - 100;
    std::vector<double> p((size_t)d.nqs * cs), x(cs), dx(cs), ty(d.nj), tdy(d.nj);
    std::vector<int> tg(d.nj);
    if (d2h(c, p.data(), d.prim, sizeof(double) * p.size()) || d2h(c, x.data(), d.cx, sizeof(double) * cs) ||
        d2h(c, dx.data(), d.cdx, sizeof(double) * cs) || d2h(c, ty.data(), d.ty, sizeof(double) * d.nj) ||
        d2h(c, tdy.data(), d.tdy, sizeof(double) * d.nj) || d2h(c, tg.data(), d.tgeom, sizeof(int) * d.nj))
      return -1;
    for (int j = 0; j < d.nj; ++j)
      for (int i = 0; i < d.nax; ++i) {
        size_t k = (size_t)j * d.nax + i;
        double pv[10], cv[10];
        for (int qq = 0; qq < nq; ++qq) pv[qq] = (nd == 1 && qq == 3) ? 0. : p[(size_t)sidx(qq) * cs + k];
        if (nd == 1) {
          Geo g = cell_geo<1>(x[k], dx[k], ty[j], tdy[j], tg[j], Trig{});
          if (nq == 6) prim2cons<1, 6>(pv, cv, g.cen0, tg[j], d.P.gamma, d.P.synge);
          else prim2cons<1, 10>(pv, cv, g.cen0, tg[j], d.P.gamma, d.P.synge);
        } else {
          if (nq == 6) prim2cons<2, 6>(pv, cv, x[k], tg[j], d.P.gamma, d.P.synge);
          else prim2cons<2, 10>(pv, cv, x[k], tg[j], d.P.gamma, d.P.synge);
        }
        out[k] = cv[q];
      }
    return 0;
  }
  switch (field) {
    case 200: return plane(d.cx);
    case 201: return plane(d.cdx);
    case 206: return plane(d.dtloc);
    case 207: if (d.shock) return plane(d.sd); break;
    case 209: if (d.shock) return plane(d.pspec); break;
    case 1000: return plane(d.fx);
    case 1001: return plane(d.fv);
    case 1002: return plane(d.flL);
    case 1003: return plane(d.flR);
    default: break;
  }
  if (field == 210 && nd == 2) {       // F1 faces above each cell (its upper-neighbour count)
    std::vector<int> s(cs);
    if (d2h(c, s.data(), d.nbucnt, sizeof(int) * cs)) return -1;
    for (size_t k = 0; k < cs; ++k) out[k] = s[k] > 0 ? s[k] : 0;
    return 0;
  }
  if (field == 208 && d.shock) {
    std::vector<int> s(cs);
    if (d2h(c, s.data(), d.shk, sizeof(int) * cs)) return -1;
    for (size_t k = 0; k < cs; ++k) out[k] = s[k];
    return 0;
  }
  if (field == 202 || field == 203) {
    std::vector<double> t(d.nj);
    if (d2h(c, t.data(), field == 202 ? d.ty : d.tdy, sizeof(double) * d.nj)) return -1;
    for (int j = 0; j < d.nj; ++j)
      for (int i = 0; i < d.nax; ++i) out[(size_t)j * d.nax + i] = t[j];
    return 0;
  }
  if (field == 204 || field == 205) {
    std::vector<double> x(cs), dx(cs), ty(d.nj), tdy(d.nj);
    std::vector<int> tg(d.nj);
    if (d2h(c, x.data(), d.cx, sizeof(double) * cs) || d2h(c, dx.data(), d.cdx, sizeof(double) * cs) ||
        d2h(c, ty.data(), d.ty, sizeof(double) * d.nj) || d2h(c, tdy.data(), d.tdy, sizeof(double) * d.nj) ||
        d2h(c, tg.data(), d.tgeom, sizeof(int) * d.nj))
      return -1;
    for (int j = 0; j < d.nj; ++j)
      for (int i = 0; i < d.nax; ++i) {
        size_t k = (size_t)j * d.nax + i;
        Geo g = nd == 1 ? cell_geo<1>(x[k], dx[k], ty[j], tdy[j], tg[j], Trig{})
                        : cell_geo<2>(x[k], dx[k], ty[j], tdy[j], tg[j], trig_of(ty[j], tdy[j]));
        out[k] = field == 204 ? g.dV : g.cen0;
      }
    return 0;
  }
  if ((field >= 300 && field < 300 + nq) || (field >= 400 && field < 400 + nq)) {
    if (nd != 2) { memset(out, 0, sizeof(double) * cs); return 0; }
    int q = field % 100;
    return plane((field < 400 ? d.gmv : d.gf1) + (size_t)q * cs);
  }
  if (field >= 1010 && field < 1010 + nq) {
    int q = field - 1010;
    if (nd == 1 && q == 3) { memset(out, 0, sizeof(double) * cs); return 0; }
    return plane(d.fflux + (size_t)sidx(q) * cs);
  }
  return fail(c, "unknown field");
}

int gammahip_stats_get(gammahip_ctx* c, gammahip_stats* st) {
  std::vector<unsigned long long> parts((size_t)CNT_N * CNT_PARTS);
#ifdef GH_EMU
  if (d2h(c, parts.data(), c->d.cnt, sizeof(unsigned long long) * CNT_N)) return -1;
  const int np = 1;
#else
  if (d2h(c, parts.data(), c->d.cnt, sizeof(unsigned long long) * parts.size())) return -1;
  const int np = CNT_PARTS;
#endif
  unsigned long long h[CNT_N] = {};
  for (int pp = 0; pp < np; ++pp)
    for (int q = 0; q < CNT_N; ++q) h[q] += parts[(size_t)pp * CNT_N + q];
  memset(st, 0, sizeof(*st));
  st->c2p_calls = h[CNT_C2P_CALLS]; st->c2p_iters = h[CNT_C2P_ITERS];
  st->c2p_nobracket = h[CNT_NOBRACKET]; st->c2p_maxiter = h[CNT_MAXITER]; st->c2p_phicap = h[CNT_PHICAP];
  st->negative_dt = h[CNT_NEGDT]; st->nan_flux = h[CNT_NANFLUX]; st->f1_overflow = h[CNT_F1OVF];
  std::vector<long long> s(c->d.nj), m(c->d.nj);
  if (d2h(c, s.data(), c->d.nsplit, sizeof(long long) * c->d.nj)) return -1;
  if (d2h(c, m.data(), c->d.nmerge, sizeof(long long) * c->d.nj)) return -1;
  for (int j = 0; j < c->d.nj; ++j) { st->splits += s[j]; st->merges += m[j]; }
  st->mlp_used = h[CNT_MLP];
  st->regrid_clash = h[CNT_RGCLASH];
  return 0;
}

int gammahip_regrid_counts(gammahip_ctx* c, long long* splits, long long* merges) {
  if (d2h(c, splits, c->d.nsplit, sizeof(long long) * c->d.nj)) return -1;
  return d2h(c, merges, c->d.nmerge, sizeof(long long) * c->d.nj);
}


// ---------------------------------------------------------------------------
// learned cons2prim guess (no counterpart in the reference)

int gammahip_mlp_nparams(void) { return gh::mlp::NPARAM; }

int gammahip_mlp_train(const double* states, long long n, float* params, int steps, int batch,
                       float lr, unsigned long long seed, float* loss_hist) {
#ifdef GH_EMU
  (void)states; (void)n; (void)params; (void)steps; (void)batch; (void)lr; (void)seed; (void)loss_hist;
  return -1;
#else
  using namespace gh::mlp;
  if (n < 1 || batch < 64 || steps < 1) return plain_fail("bad training arguments");
  DevBuf B;
  double* st = B.get<double>(4 * n);
  float *w = B.get<float>(NPARAM), *g = B.get<float>(NPARAM), *m = B.get<float>(NPARAM), *v = B.get<float>(NPARAM);
  float *X = B.get<float>(3 * n), *T = B.get<float>(n), *loss = B.get<float>(1);
  long long* idx = B.get<long long>(batch);
  if (!st || !w || !g || !m || !v || !X || !T || !loss || !idx) return plain_fail("allocation failed");
  CK(hipMemcpy(st, states, sizeof(double) * 4 * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(w, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
  CK(hipMemset(g, 0, sizeof(float) * NPARAM)); CK(hipMemset(m, 0, sizeof(float) * NPARAM));
  CK(hipMemset(v, 0, sizeof(float) * NPARAM));
  hipLaunchKernelGGL(k_mlp_prepare, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, w, st, n, X, T);
  for (int k = 0; k < steps; ++k) {
    CK(hipMemset(loss, 0, sizeof(float)));
    hipLaunchKernelGGL(k_mlp_sample, dim3((batch + 255) / 256), dim3(256), 0, 0, idx, batch, n, seed, (long long)k);
    hipLaunchKernelGGL(k_mlp_grad, dim3((batch + 63) / 64), dim3(64), 0, 0, w, X, T, idx, batch, g, loss);
    hipLaunchKernelGGL(k_mlp_adam, dim3((NTRAIN + 255) / 256), dim3(256), 0, 0, w, g, m, v, k + 1, lr);
    if (loss_hist) CK(hipMemcpy(&loss_hist[k], loss, sizeof(float), hipMemcpyDeviceToHost));
  }
  CK(hipGetLastError());
  CK(hipMemcpy(params, w, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  return 0;
#endif
}

int gammahip_mlp_grad(const float* params, const double* states, int n, float* grad, float* loss_out) {
#ifdef GH_EMU
  (void)params; (void)states; (void)n; (void)grad; (void)loss_out;
  return -1;
#else
  using namespace gh::mlp;
  DevBuf B;
  double* st = B.get<double>(4 * (size_t)n);
  float *w = B.get<float>(NPARAM), *g = B.get<float>(NPARAM), *X = B.get<float>(3 * (size_t)n);
  float *T = B.get<float>(n), *loss = B.get<float>(1);
  long long* idx = B.get<long long>(n);
  if (!st || !w || !g || !X || !T || !loss || !idx) return plain_fail("allocation failed");
  std::vector<long long> hi(n);
  for (int k = 0; k < n; ++k) hi[k] = k;
  CK(hipMemcpy(st, states, sizeof(double) * 4 * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(w, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
  CK(hipMemcpy(idx, hi.data(), sizeof(long long) * n, hipMemcpyHostToDevice));
  CK(hipMemset(g, 0, sizeof(float) * NPARAM)); CK(hipMemset(loss, 0, sizeof(float)));
  hipLaunchKernelGGL(k_mlp_prepare, dim3((n + 255) / 256), dim3(256), 0, 0, w, st, (long long)n, X, T);
  hipLaunchKernelGGL(k_mlp_grad, dim3((n + 63) / 64), dim3(64), 0, 0, w, X, T, idx, n, g, loss);
  CK(hipGetLastError());
  CK(hipMemcpy(grad, g, sizeof(float) * NTRAIN, hipMemcpyDeviceToHost));
  CK(hipMemcpy(loss_out, loss, sizeof(float), hipMemcpyDeviceToHost));
  return 0;
#endif
}

int gammahip_mlp_infer(const float* params, const double* states, long long n, double* p_out, int mfma) {
#ifdef GH_EMU
  (void)params; (void)states; (void)n; (void)p_out; (void)mfma;
  return -1;
#else
  using namespace gh::mlp;
  DevBuf B;
  double *st = B.get<double>(3 * n), *out = B.get<double>(n);
  float* w = B.get<float>(NPARAM);
  bf16x8* w2b = B.get<bf16x8>(8 * 64);
  if (!st || !out || !w || !w2b) return plain_fail("allocation failed");
  std::vector<__bf16> hb(8 * 64 * 8);
  pack_w2(params, hb.data());
  CK(hipMemcpy(st, states, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(w, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
  CK(hipMemcpy(w2b, hb.data(), sizeof(__bf16) * hb.size(), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, 0, w, w2b, st, n, out, mfma);
  CK(hipGetLastError());
  CK(hipMemcpy(p_out, out, sizeof(double) * n, hipMemcpyDeviceToHost));
  return 0;
#endif
}

int gammahip_mlp_infer_noalloc(const float* params, const double* states, long long n, double* p_out, int mfma,
                               void* stream) {
#ifdef GH_EMU
  (void)params; (void)states; (void)n; (void)p_out; (void)mfma; (void)stream;
  return -1;
#else
  using namespace gh::mlp;
  constexpr size_t need = sizeof(bf16x8) * 8 * 64;
  static_assert(need <= NOALLOC_SCRATCH_BYTES, "packed operands exceed the scratch array");
  if (!params || !states || !p_out) return plain_fail("mlp_infer_noalloc: null device pointer");
  if (mfma != 0 && mfma != 1) return plain_fail("mlp_infer_noalloc: mfma must be 0 or 1");
  if (n < 1 || (n + 63) / 64 > (long long)INT_MAX)
    return plain_fail("mlp_infer_noalloc: n must be in [1, 64 * INT_MAX]");
  void* sp = nullptr;
  if (noalloc_scratch(&sp)) return -1;
  hipStream_t s = (hipStream_t)stream;
  bf16x8* w2b = (bf16x8*)sp;
  hipLaunchKernelGGL(k_mlp_pack_w2, dim3(1), dim3(512), 0, s, params, (__bf16*)w2b);
  hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, params, w2b, states, n, p_out,
                     mfma);
  CK(hipGetLastError());
  return 0;
#endif
}

int gammahip_c2p_batch(const double* states, const double* guess, const double* guess2, long long n,
                       double gamma, int solver, int maxit, double* p_out, int* iters, int* fail,
                       int* used) {
  if (n < 1 || !(gamma > 1.) || (solver != gh::C2P_BRENT && solver != gh::C2P_NEWTON) || maxit < 1) {
    fprintf(stderr, "gammahip: bad c2p_batch arguments\n");
    return -1;
  }
#ifdef GH_EMU
  for (long long s = 0; s < n; ++s)
    body_c2p_batch(s, states, guess, guess2, gamma, solver, maxit, p_out, iters, fail, used);
  return 0;
#else
  DevBuf B;
  double *st = B.get<double>(3 * n), *g = B.get<double>(n), *g2 = guess2 ? B.get<double>(n) : nullptr;
  double* p = B.get<double>(n);
  int *it = B.get<int>(n), *fl = B.get<int>(n), *us = B.get<int>(n);
  if (!st || !g || (guess2 && !g2) || !p || !it || !fl || !us) return plain_fail("allocation failed");
  CK(hipMemcpy(st, states, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(g, guess, sizeof(double) * n, hipMemcpyHostToDevice));
  if (guess2) CK(hipMemcpy(g2, guess2, sizeof(double) * n, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_c2p_batch, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, st, g, g2, n, gamma,
                     solver, maxit, p, it, fl, us);
  CK(hipGetLastError());
  CK(hipMemcpy(p_out, p, sizeof(double) * n, hipMemcpyDeviceToHost));
  CK(hipMemcpy(iters, it, sizeof(int) * n, hipMemcpyDeviceToHost));
  CK(hipMemcpy(fail, fl, sizeof(int) * n, hipMemcpyDeviceToHost));
  if (used) CK(hipMemcpy(used, us, sizeof(int) * n, hipMemcpyDeviceToHost));
  return 0;
#endif
}

int gammahip_pow_batch(const double* x, const double* y, double* out, long long n) {
  if (n < 1) return -1;
#ifdef GH_EMU
  for (long long i = 0; i < n; ++i) out[i] = gh_pow(x[i], y[i]);
  return 0;
#else
  DevBuf B;
  double *dx = B.get<double>(n), *dy = B.get<double>(n), *dz = B.get<double>(n);
  if (!dx || !dy || !dz) return plain_fail("allocation failed");
  CK(hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(dy, y, sizeof(double) * n, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_pow_batch, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, dx, dy, dz, n);
  CK(hipGetLastError());
  CK(hipMemcpy(out, dz, sizeof(double) * n, hipMemcpyDeviceToHost));
  return 0;
#endif
}

int gammahip_fn_batch(int fn, const double* x, double* out, long long n) {
  if (n < 1 || (fn != 0 && fn != 1)) return -1;
#ifdef GH_EMU
  for (long long i = 0; i < n; ++i) out[i] = fn == 0 ? gh_tanh(x[i]) : gh_log10(x[i]);
  return 0;
#else
  DevBuf B;
  double *dx = B.get<double>(n), *dz = B.get<double>(n);
  if (!dx || !dz) return plain_fail("allocation failed");
  CK(hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_fn_batch, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, fn, dx, dz, n);
  CK(hipGetLastError());
  CK(hipMemcpy(out, dz, sizeof(double) * n, hipMemcpyDeviceToHost));
  return 0;
#endif
}

int gammahip_mfma_probe(const float* A, const float* Bm, float* Dm) {
#ifdef GH_EMU
  (void)A; (void)Bm; (void)Dm;
  return -1;
#else
  DevBuf B;
  float *a = B.get<float>(32 * 16), *b = B.get<float>(16 * 32), *d = B.get<float>(32 * 32);
  if (!a || !b || !d) return plain_fail("allocation failed");
  CK(hipMemcpy(a, A, sizeof(float) * 512, hipMemcpyHostToDevice));
  CK(hipMemcpy(b, Bm, sizeof(float) * 512, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, 0, a, b, d);
  CK(hipGetLastError());
  CK(hipMemcpy(Dm, d, sizeof(float) * 1024, hipMemcpyDeviceToHost));
  return 0;
#endif
}

#ifdef GH_EMU
// CPU build only: capture the update's cons2prim solves into buf ([cap][7]
// doubles, see C2PCapture), one in every `stride`; buf = null stops it.
// Returns the rows written so far.
long long gammahip_emu_capture(double* buf, long long cap, long long stride) {
  long long n = g_c2p_capture.n;
  if (buf) g_c2p_capture = {buf, cap, 0, stride > 0 ? stride : 1, 0};
  else g_c2p_capture.buf = nullptr;
  return n;
}
#endif

int gammahip_kernel_times(gammahip_ctx* c, double* ms, long long* n) {
  c->timing = true;     // the first call switches event timing on
  tcollect(c);
  for (int k = 0; k < 16; ++k) {
    ms[k] = c->kt[k].n ? c->kt[k].ms / c->kt[k].n : 0.;
    n[k] = c->kt[k].n;
    c->kt[k].ms = 0.;
    c->kt[k].n = 0;
  }
  return 0;
}

}  // extern "C"
