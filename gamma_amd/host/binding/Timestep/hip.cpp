// src/Timestep/hip.cpp -- the Timestep module of GAMMA on the MI355X backend
// (Makefile: TIMESTEP = hip, with DIMENSIONS = 1d_hip or 2d_hip).  Grid::evolve
// of src/Timestep/rk3.cpp:13, 99 (and euler.cpp:12-19 when the setup's hooks
// select timestep = 1) advances the device grid by the caller's dt: a driver
// that clamps dt is followed.
#include "../environment.h"
#include "../grid.h"
#include "gammahip_binding.h"

void Grid::evolve(int it, double t, double dt) {
  (void)it;
  (void)t;
  gammahip_binding_check(gammahip_evolve_dt(gammahip_binding_ctx(), dt));
}
