// Test infrastructure: the reference-side binding of the device path.
//
// Compiled against the reference's own headers (include/grid.h,
// include/sys_mem.h, include/params.h) and linked ahead of the reference's
// solve.C object into _ref/ref_driver_gpu, this translation unit replaces the
// reference's solve() (include/solve.h:21) and Compute_Body_Forces()
// (include/grid_io.h:36) with the device path through the
// reference-shaped header include/fs_reference_shim.hpp: the reference's own
// driver loop then runs every pseudo-time step on the GPU.  It is the stub a
// maintainer adds to the reference tree to switch it over.
#include "Defined_Variables.h"
#include "grid.h"
#include "params.h"
#include "sys_mem.h"
#include "solve.h"
#include "grid_io.h"

#include "fs_reference_shim.hpp"

double solve(GRID* grid, SYS_MEM* smem, PARAMS p, int ts) { return fsref::solve(grid, smem, p, ts); }
void Compute_Body_Forces(GRID* grid, PARAMS p, double* Forces) { fsref::Compute_Body_Forces(grid, p, Forces); }
