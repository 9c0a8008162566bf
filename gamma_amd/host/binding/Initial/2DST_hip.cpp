// The device hooks of the 2DST setup (src/Initial/Tests/2DST.cpp), to be
// appended to that file when it runs on the MI355X backend: its
// userKinematics holds the ngst outer faces of every track at v = 0
// (2DST.cpp:107-125), userBoundaries does nothing (:142-147) and
// checkCellForRegrid splits at dx/dy > 3 and merges below 0.2 (:150-168).
#include "../../environment.h"
#include "../../grid.h"
#include "gammahip_binding.h"

void gammahip_setup_hooks(const Grid& g, gammahip_config* c, gammahip_track_hooks* h) {
  c->geom = 0;                          // rel_cart + cartesian
  c->rg_mode = 1;                       // aspect ratio dx/dy
  c->rg[0] = 3.;
  c->rg[1] = 0.2;
  for (int a = 0; a < g.ncell[F1]; ++a) {
    h[a].kin_mode = 1;
    h[a].v_in = 0.;
    h[a].v_out = 0.;
    h[a].bnd_mode = 0;
    h[a].rg_mode = -1;                  // the configuration's rule
    h[a].ncell = 0;
  }
}
