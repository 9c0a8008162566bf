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
  explicit Grid(const gammahip_config& cfg) : cfl_(cfg.cfl) {
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

 private:
  void check(int r) const {
    if (r != 0) throw Error(gammahip_last_error(ctx_));
  }
  gammahip_ctx* ctx_ = nullptr;
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
