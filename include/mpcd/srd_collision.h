/* srd_collision.h -- random-shift binning and SRD collision on the GPU
 * (drop-in for the reference's include/srd_collision.h:9-15). */
#pragma once

#include <cstdint>
#include <vector>

#include "types.h"

namespace mpcd {

std::vector<int> srd_cell_id_with_random_shift(const std::vector<double>& x, const Params& params,
                                               std::uint64_t rng_seed);
void srd_collision_step(std::vector<double>& v, const std::vector<int>& cid, const Params& params,
                        std::uint64_t rng_seed);

/* the reference's OpenMP variants (their thread count has no meaning here) */
std::vector<int> srd_cell_id_with_random_shift_openmp(const std::vector<double>& x, const Params& params,
                                                      std::uint64_t rng_seed, int nThreads);
void srd_collision_step_openmp(std::vector<double>& v, const std::vector<int>& cid, const Params& params,
                               std::uint64_t rng_seed, int nThreads);

}  // namespace mpcd
