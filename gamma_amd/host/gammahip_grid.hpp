// gammahip_grid.hpp -- C++ host side of the hot path over the C-ABI.
//
// The reference drives its hot path through Grid methods (src/grid.h:61-112)
// called from Simu::run (src/simu.cpp:41-78).  This header gives a C++ caller
// the same methods on the MI355X backend: gammahip::Grid wraps one device
// context (include/gammahip.h) and keeps the reference's method names and
// call order; gammahip::Simu runs the loop body of Simu::run.  Errors become
// exceptions carrying the library's message.  Header-only; link with
// libgammahip.so.
#pragma once
#include "../../include/gammahip.h"

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace gammahip {

struct Error : std::runtime_error {
  explicit Error(const std::string& m) : std::runtime_error(m) {}
};

class Grid {
 public:
  // Grid::Grid + Grid::initialise (1d.cpp:53-81, 2d.cpp:59-96)
  explicit Grid(const gammahip_config& cfg) : cfg_(cfg), cfl_(cfg.cfl) {
    if (gammahip_create(&cfg, &ctx_) != 0) {
      std::string m = gammahip_last_error(ctx_);
      gammahip_destroy(ctx_);
      throw Error("gammahip_create: " + m);
    }
  }
  ~Grid() { gammahip_destroy(ctx_); }
  Grid(const Grid&) = delete;
  Grid& operator=(const Grid&) = delete;

  // initialGeometry/initialValues output + Grid::prepForRun (grid.cpp:20-26):
  // x, dx [tracks][ncell_mv], y, dy [tracks], prim [nq][tracks][ncell_mv]
  // (3-velocities), one hooks entry per track
  void prepForRun(const std::vector<double>& x, const std::vector<double>& dx, const std::vector<double>& y,
                  const std::vector<double>& dy, const std::vector<double>& prim,
                  const std::vector<gammahip_track_hooks>& hooks, double t0) {
    check(gammahip_load(ctx_, x.data(), dx.data(), y.data(), dy.data(), prim.data(), hooks.data(), t0));
  }

  // parameters of the user setup's hooks (gamma_amd/csrc/user_hooks.h), set
  // before prepForRun: its first ghost update calls them
  void setUserParams(const std::vector<double>& par) {
    check(gammahip_set_user_params(ctx_, par.data(), (int)par.size()));
  }

  void regrid() { check(gammahip_regrid(ctx_)); }                    // 1d.cpp:484, 2d.cpp:609
  void updateGhosts() { check(gammahip_update_ghosts(ctx_)); }       // 1d.cpp:217, 2d.cpp:241
  void prepForUpdate() { check(gammahip_prep_for_update(ctx_)); }    // grid.cpp:28-37
  // Grid::collect_dt (1d.cpp:779-795): the minimum candidate; the device keeps
  // CFL times it for evolve()
  double collect_dt() {
    double dt = 0.;
    check(gammahip_collect_dt(ctx_, &dt));
    return dt / cfl_;
  }
  void evolve() { check(gammahip_evolve(ctx_)); }                    // rk3.cpp:98-195
  // Grid::apply(&Cell::resetShock / detectShock / radiation_injectParticles)
  void apply(int what) { check(gammahip_apply(ctx_, what)); }
  // whole loop-body iterations without host round trips
  void step(int n) { check(gammahip_step(ctx_, n)); }

  // state access (what Grid::printCols reads)
  void dims(int& ntracks, int& nax) const {
    if (gammahip_dims(ctx_, &ntracks, &nax) != 0) throw Error(gammahip_last_error(ctx_));
  }
  std::vector<double> download(int field) const {
    int nj = 0, nax = 0;
    dims(nj, nax);
    std::vector<double> out((size_t)nj * nax);
    if (gammahip_download(ctx_, field, out.data()) != 0) throw Error(gammahip_last_error(ctx_));
    return out;
  }
  void tracks(std::vector<int>& ntrack, std::vector<int>& nact, std::vector<int>& iL, std::vector<int>& iR) const {
    int nj = 0, nax = 0;
    dims(nj, nax);
    ntrack.resize(nj); nact.resize(nj); iL.resize(nj); iR.resize(nj);
    if (gammahip_track_info(ctx_, ntrack.data(), nact.data(), iL.data(), iR.data()) != 0)
      throw Error(gammahip_last_error(ctx_));
  }
  gammahip_stats stats() const {
    gammahip_stats s{};
    if (gammahip_stats_get(ctx_, &s) != 0) throw Error(gammahip_last_error(ctx_));
    return s;
  }
  gammahip_ctx* handle() { return ctx_; }

  // Grid::printCols (text1d.cpp:135-243, text2d.cpp:146-283): writes
  // dir/phys%010d.out with the reference's header, rows and conversions.  The
  // second tracer column shows isShocked (1D) or Sd (2D) under shock
  // detection, as the reference prints it (the state is not modified); with
  // synchrotron tracers gmin/gmax pass through radiation_apply_trac2gammae.
  // A 1D context holding several runs prints the run on `track`.
  void printCols(int it, double t, const std::string& dir, int track = 0) const {
    int nj = 0, nax = 0;
    dims(nj, nax);
    const int nd = cfg_.nd, nq = cfg_.nq, ng = cfg_.ngst;
    const bool sync = cfg_.sync != 0, shock = cfg_.shock != 0;
    std::vector<int> nt, na, iL, iR;
    tracks(nt, na, iL, iR);
    std::vector<std::vector<double>> P(nq);
    for (int q = 0; q < nq; ++q) P[q] = download(q);
    auto D = download(100), TAU = download(101), S1 = download(102);
    std::vector<double> S2 = nd == 2 ? download(103) : std::vector<double>();
    auto X = download(200), DX = download(201);
    std::vector<double> Y, DY;
    if (nd == 2) { Y = download(202); DY = download(203); }
    std::vector<double> second = P[5];
    if (shock) second = download(nd == 1 ? 208 : 207);
    char name[32];
    snprintf(name, sizeof name, "phys%010d.out", it);
    std::string path = dir + "/" + name;
    FILE* f = fopen(path.c_str(), "w");
    if (!f) throw Error("cannot open " + path);
    if (nd == 1) fprintf(f, sync ? "t nact i x dx dlx rho vx p D sx tau trac Sd gmin gmax\n"
                                 : "t nact i x dx dlx rho vx p D sx tau trac Sd\n");
    else fprintf(f, sync ? "t nact j i x y dx dy dlx dly rho vx vy p D sx sy tau trac Sd gmin gmax gmid psyn\n"
                         : "t nact j i x y dx dy dlx dly rho vx vy p D sx sy tau trac Sd\n");
    int j0 = nd == 1 ? track : ng, j1 = nd == 1 ? track + 1 : nj - ng;
    for (int j = j0; j < j1; ++j) {
      int nact = nt[j] - 2 * ng;
      for (int i = ng; i < nt[j] - ng; ++i) {
        size_t k = (size_t)j * nax + i;
        double u1 = P[2][k], u2 = nd == 2 ? P[3][k] : 0.;
        double u = std::sqrt(nd == 2 ? u1 * u1 + u2 * u2 : u1 * u1);     // FluidState::lfac
        double lf = std::sqrt(1 + u * u);
        double rho = P[0][k];
        double gmin = 0., gmax = 0.;
        if (sync) {   // Cell::radiation_apply_trac2gammae
          double a = P[6][k] * (lf * rho), b = P[7][k] * (lf * rho);
          gmin = a != 0. ? lf * std::pow(rho, 4. / 3.) / a : 0.;
          gmax = b != 0. ? lf * std::pow(rho, 4. / 3.) / b : 0.;
        }
        if (nd == 1) {
          fprintf(f, "%1.15le %d %d %1.15le %1.15le %1.15le %1.15le %1.15le %1.15le %1.15le %1.15le %1.15le %1.15le %1.15le",
                  t, nact, i - ng, X[k], DX[k], DX[k], rho, u1 / lf, P[1][k], D[k], S1[k], TAU[k], P[4][k], second[k]);
          if (sync) fprintf(f, " %1.15le %1.15le", gmin, gmax);
          fprintf(f, "\n");
        } else {
          double dly = cfg_.geom == 1 ? X[k] * DY[k] : DY[k];   // Cell::computedl
          if (sync)
            fprintf(f, "%1.15le %d %d %d %1.15le %1.15le %1.15le %1.15le %1.15le %le %le %le %le %le %le %le %le %le %le %le %le %le %le %le\n",
                    t, nact, j - ng, i - ng, X[k], Y[k], DX[k], DY[k], DX[k], dly, rho, u1 / lf, u2 / lf, P[1][k],
                    D[k], S1[k], S2[k], TAU[k], P[4][k], second[k], gmin, gmax, P[8][k], P[9][k]);
          else
            fprintf(f, "%1.15le %d %d %d %1.15le %1.15le %1.15le %1.15le %1.15le %1.15le %le %le %le %le %le %le %le %le %le %le\n",
                    t, nact, j - ng, i - ng, X[k], Y[k], DX[k], DY[k], DX[k], dly, rho, u1 / lf, u2 / lf, P[1][k],
                    D[k], S1[k], S2[k], TAU[k], P[4][k], second[k]);
        }
      }
    }
    fclose(f);
  }

  // exact device state between steps (gammahip_state_save / restore)
  std::vector<unsigned char> saveState() const {
    unsigned long long n = 0;
    check(gammahip_state_size(ctx_, &n));
    std::vector<unsigned char> b(n);
    check(gammahip_state_save(ctx_, b.data(), n));
    return b;
  }
  void restoreState(const std::vector<unsigned char>& b) {
    check(gammahip_state_restore(ctx_, b.data(), b.size()));
  }

 private:
  void check(int r) const {
    if (r != 0) throw Error(gammahip_last_error(ctx_));
  }
  gammahip_ctx* ctx_ = nullptr;
  gammahip_config cfg_;
  double cfl_;
};

// Simu::run (simu.cpp:41-78) without its IO hooks (dataDump, runInfo,
// evalEnd): nsteps iterations of the loop body (the first updateGhosts of
// simu.cpp:43 is part of Grid::prepForRun here).
struct Simu {
  Grid& grid;
  bool shock, sync;
  double cfl;
  long it = 0;
  double t = 0.;
  void run(long nsteps) {
    for (long n = 0; n < nsteps; ++n) {
      grid.regrid();
      if (shock) grid.apply(GAMMAHIP_RESET_SHOCK);
      grid.updateGhosts();
      grid.prepForUpdate();
      if (shock) grid.apply(GAMMAHIP_DETECT_SHOCK);
      double dt = cfl * grid.collect_dt();
      grid.evolve();
      grid.apply(GAMMAHIP_ADVANCE_TIME);     // the device keeps t for its hooks
      t += dt;
      it++;
      if (sync) grid.apply(GAMMAHIP_INJECT);
    }
  }
};

}  // namespace gammahip
