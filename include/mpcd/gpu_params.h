/* gpu_params.h -- the POD image of mpcd::Params that the C ABI of
 * libmpcd_gpu.so takes (include/mpcd_gpu.h).  No counterpart in the
 * reference: C++ callers of the resident simulation (mpcd_gpu_sim_*) use it
 * to turn a Params read by read_params_kv into an mpcd_params. */
#pragma once

#include "mpcd_gpu.h"
#include "types.h"

namespace mpcd {

/* boundary_* strings become MPCD_BC_* (an unknown mode is an error only when
 * a particle reaches that wall, as in apply_bc_general); highMode/lowMode
 * "1+coef" become the 0/1 flags */
mpcd_params to_gpu_params(const Params& params);

}  // namespace mpcd
