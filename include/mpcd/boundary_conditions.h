/* boundary_conditions.h -- wall boundary conditions on the GPU (drop-in for
 * the reference's include/boundary_conditions.h:8-13). */
#pragma once

#include <string>

#include "types.h"

namespace mpcd {

/* axis_name "x" or "y": both walls of the axis periodic */
bool is_periodic_pair(const Params& params, const std::string& axis_name);

/* wall sweeps and final wrap/clamp of x, v (interleaved, in place); the
 * thermalising draws are keyed by rng_seed, particle id, sweep and wall */
void apply_bc_general(std::vector<double>& x, std::vector<double>& v, const Params& params, WallInfo& info,
                      std::uint64_t rng_seed);

}  // namespace mpcd
