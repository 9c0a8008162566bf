// shock_tube.cpp -- a C++ caller of the hot path: the relativistic shock tube
// (Marti & Mueller problem 1: rho 10 | 1, p 13.33 | 1e-6, gamma 5/3, static
// boundary faces) set up as the reference's setup hooks would
// (initialGeometry / initialValues / userKinematics), advanced by
// gammahip::Simu, i.e. by the reference's Grid method sequence.
//
// usage: shock_tube [ncells=200] [nsteps=100] [mode=methods|step|split] [dir]
// prints the final time and, per active cell, x rho p u as C99 hex floats;
// mode split runs nsteps/2 steps, saves the exact state, restores it into a
// second grid and runs the rest there; with dir, Grid::printCols writes
// dir/phys%010d.out of the final state.
#include "gammahip_grid.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>

int main(int argc, char** argv) {
  int n = argc > 1 ? atoi(argv[1]) : 200;
  long nsteps = argc > 2 ? atol(argv[2]) : 100;
  bool by_step = argc > 3 && strcmp(argv[3], "step") == 0;
  bool split = argc > 3 && strcmp(argv[3], "split") == 0;
  const char* dir = argc > 4 ? argv[4] : nullptr;
  if (n < 8 || nsteps < 0) { fprintf(stderr, "bad arguments\n"); return 2; }
  gammahip_config c{};
  c.nd = 1; c.nq = 6; c.geom = 0; c.shock = 0; c.sync = 0; c.ngst = 2;
  c.nmax = n + 8; c.ncell_mv = n; c.ntracks_active = 1;
  c.gamma = 5. / 3.; c.cfl = 0.2; c.vi = 1.; c.shock_thr = 0.01; c.gamma_max_init = 1.e8;
  c.Nalpha = 0.; c.Nmp = 1.; c.Nme = 1.;
  c.rg_mode = 0; c.c2p_solver = 0; c.c2p_guess = 0; c.c2p_maxiter = 1000;
  try {
    gammahip::Grid g(c);
    // initialGeometry / initialValues
    double dxc = 1. / n;
    std::vector<double> x(n), dx(n, dxc), y(1, 0.), dy(1, 1.), prim((size_t)6 * n, 0.);
    for (int i = 0; i < n; ++i) {
      x[i] = (i + 0.5) * dxc;
      bool left = x[i] <= 0.5;
      prim[0 * (size_t)n + i] = left ? 10. : 1.;       // rho
      prim[1 * (size_t)n + i] = left ? 13.33 : 1.e-6;  // p
      prim[4 * (size_t)n + i] = left ? 1. : 2.;        // tracer
    }
    // userKinematics: boundary faces at rest (2DST.cpp:261-278)
    gammahip_track_hooks h{};
    h.kin_mode = 1; h.v_in = 0.; h.v_out = 0.; h.rg_mode = -1;
    g.prepForRun(x, dx, y, dy, prim, {h}, 0.);
    gammahip::Simu sim{g, false, false, c.cfl};
    gammahip::Grid g2(c);
    gammahip::Grid* out = &g;
    if (by_step) g.step((int)nsteps);
    else if (split) {
      g.step((int)(nsteps / 2));
      g2.prepForRun(x, dx, y, dy, prim, {h}, 0.);
      g2.restoreState(g.saveState());
      g2.step((int)(nsteps - nsteps / 2));
      out = &g2;
    } else sim.run(nsteps);
    std::vector<int> nt, na, iL, iR;
    out->tracks(nt, na, iL, iR);
    auto X = out->download(200), R = out->download(0), P = out->download(1), U = out->download(2);
    double t[1], dt[1];
    if (gammahip_time(out->handle(), t, dt) != 0) throw gammahip::Error("gammahip_time");
    printf("t %a\n", t[0]);
    for (int i = iL[0] + 1; i < iR[0]; ++i) printf("%a %a %a %a\n", X[i], R[i], P[i], U[i]);
    if (dir) out->printCols((int)nsteps, t[0], dir);
  } catch (const std::exception& e) {
    fprintf(stderr, "shock_tube: %s\n", e.what());
    return 1;
  }
  return 0;
}
