/* zone_pass.h -- zone layouts, the zone kernel and the in-place
 * (Gauss-Seidel) zone pass on the GPU (drop-in for the reference's
 * include/zone_pass.h:10-77). */
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "types.h"

namespace mpcd {

/* One zone: 1-based id, its cells (iy + Ny*ix) and a per-cell membership
 * mask over the domain. */
struct ZoneDescriptor {
    int id = 0;
    std::vector<int> cellIds;
    std::vector<char> maskVec;
};

struct ZoneLayout {
    std::string mode; /* "base", "shifted" (or the single-zone layout) */
    std::vector<ZoneDescriptor> zones;
    std::vector<std::int32_t> ownerMap; /* per cell: id of the last zone covering it */
};

struct ZoneKernelStats {
    std::array<double, 6> stats{}; /* nHigh, nLow, moved, moved, rms du, max du */
    std::array<double, 10> diag{};
    int movedDense = 0;
    int movedSparse = 0;
};

struct ZonePassMetrics {
    double occStdBefore = 0.0;
    double occStdAfter = 0.0;
    double outBandBefore = 0.0;
    double outBandAfter = 0.0;
    int nDenseBefore = 0;
    int nDenseAfter = 0;
    int nSparseBefore = 0;
    int nSparseAfter = 0;
    int nEmptyActiveBefore = 0;
    int nEmptyActiveAfter = 0;
    int nParticlesMovedDense = 0;
    int nParticlesMovedSparse = 0;
    int nZonesExecuted = 0;
    double corrRmsDU = 0.0;
    double corrMaxDU = 0.0;
    double momentumDeltaX = 0.0;
    double momentumDeltaY = 0.0;
    std::string layoutMode;
};

struct ZonePassResult {
    State stateOut;
    CellFields beforeFields;
    CellFields afterFields;
    ZonePassMetrics metrics;
    std::vector<std::int32_t> zoneOwnerMap;
};

ZoneLayout build_zone_layout(const Params& params, const std::string& mode);

/* one zone in place; the zone must be a rectangle of cells */
ZoneKernelStats run_zone_kernel_inplace(std::vector<double>& x, std::vector<double>& v, const Params& params,
                                        const ZoneDescriptor& zone, std::uint64_t rngSeed);

std::vector<int> build_particle_ids_in_zone(const std::vector<double>& x, const Params& params,
                                            const ZoneDescriptor& zone);

/* zones applied one after another on the evolving state */
ZonePassResult run_zone_pass(const State& stateIn, const Params& params, const std::string& layoutMode,
                             std::uint64_t rngSeedBase);

void write_i32_grid_bin(const std::string& filepath, int Nx, int Ny, const std::vector<std::int32_t>& data);

}  // namespace mpcd
