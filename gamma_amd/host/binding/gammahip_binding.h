// gammahip_binding.h -- shared pieces of the reference-side binding of the
// MI355X backend: the Dimensions modules Dimensions/1d_hip.cpp and
// Dimensions/2d_hip.cpp (Makefile: DIMENSIONS = 1d_hip / 2d_hip) and the
// Timestep module Timestep/hip.cpp (TIMESTEP = hip), which define the Grid
// methods of src/grid.h:61-112 by calls into include/gammahip.h.
//
// Include after the reference's own headers (environment.h, grid.h).
#ifndef GAMMAHIP_BINDING_H_
#define GAMMAHIP_BINDING_H_

#include "gammahip.h"

class Grid;

// The context of the running Grid (one Grid per process, one GPU).
gammahip_ctx* gammahip_binding_ctx();

// Exits like the reference's hard errors when a backend call fails, with the
// library's message.
void gammahip_binding_check(int status);

// The setup's hooks as the device evaluates them (userKinematics,
// userBoundaries and checkCellForRegrid select a kind and its parameters).
// Hooks outside the built-in kinds are written as the user setup's functions
// (gamma_amd/csrc/user_hooks.h, or the file GAMMAHIP_USER_HOOKS names when
// the library is built) and selected with the user kinds
// (GAMMAHIP_KIN_USER, GAMMAHIP_BND_USER, GAMMAHIP_RG_USER); the per-cell
// hooks cons2prim_user and userSourceTerms live in the same file.  A setup
// run on the HIP backend defines this next to its own hooks, e.g. for
// src/Initial/Tests/2DST.cpp (2DST.cpp:107-168):
//
//   void gammahip_setup_hooks(const Grid& g, gammahip_config* c, gammahip_track_hooks* h) {
//     c->rg_mode = 1; c->rg[0] = 3.; c->rg[1] = 0.2;        // checkCellForRegrid
//     for (int a = 0; a < g.ncell[F1]; ++a) {                // userKinematics: v = 0
//       h[a].kin_mode = 1; h[a].v_in = 0.; h[a].v_out = 0.; h[a].rg_mode = -1;
//     }
//   }
//
// cfg arrives with every field the reference's compile-time environment
// fixes (environment.h, constants.h) already filled; hooks holds one zeroed
// entry per active track (ncell[F1] in 2D, one in 1D).
void gammahip_setup_hooks(const Grid& g, gammahip_config* cfg, gammahip_track_hooks* hooks);

// Called once the context exists and before the cells are loaded (whose first
// ghost update already calls the user hooks): a setup with user hooks sets
// their parameters and switches its per-cell hooks on here, e.g.
//
//   void gammahip_setup_user(const Grid& g, gammahip_ctx* ctx) {
//     double par[2] = {0.5, 1e-3};
//     gammahip_set_user_params(ctx, par, 2);
//     gammahip_set_option(ctx, GAMMAHIP_OPT_USER_CELL, 3);   // userSourceTerms, cons2prim_user
//   }
//
// The binding's own definition is weak and does nothing.
void gammahip_setup_user(const Grid& g, gammahip_ctx* ctx);

// Copies the device state into the host cells (Ctot, Itot, the track
// lengths) so that host code reading them -- Grid::printCols of the IO
// module -- sees the current step.  Call it in Simu::dataDump before
// grid.printCols(it, t).  Synchronous.
void gammahip_pull(Grid* g);

#endif
