// Compiled grid-step plans with 2 radial levels per workgroup (the level
// block the tuner may choose instead of the plan's own; see gridstep.hip).
#include "gridstep_kernel.h"

bool grid_launch_rb2(const GridArgs& a, const GPlan& P, dim3 gd, dim3 bd, size_t lds,
                     hipStream_t st) {
  return a.fprec ? launch_compiled<2, float>(a, P, gd, bd, lds, st)
                 : launch_compiled<2, double>(a, P, gd, bd, lds, st);
}
