// src/Dimensions/2d_hip.cpp -- the 2D Dimensions module of GAMMA on the
// MI355X backend (Makefile: DIMENSIONS = 2d_hip, TIMESTEP = hip).  It defines
// the Grid methods of src/Dimensions/2d.cpp (and mpi_distribute) through the
// C-ABI of include/gammahip.h; Simu::run, the setups and the IO module stay
// as they are.  Build: add -I<gamma_amd>/include -I<gamma_amd>/gamma_amd/host/binding
// and link -L<gamma_amd>/gamma_amd -lgammahip.
#include "../environment.h"
#include "../grid.h"
#include "../cell.h"
#include "../constants.h"
#include "../array_tools.h"
#include "../mpisetup.h"
#include "gammahip_binding.h"

#if NUM_D != 2
#error "2d_hip.cpp is the NUM_D == 2 module"
#endif

#include "grid_hip.inc"
