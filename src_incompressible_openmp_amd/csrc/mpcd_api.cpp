/*
 * mpcd_api.cpp -- the mpcd:: C++ API (include/mpcd/) over the C ABI
 * (include/mpcd_gpu.h): std::vector in, std::vector out, errors rethrown as
 * std::runtime_error with the library's message, as the reference's
 * functions throw.  Each function names the reference function it stands
 * in for.
 */
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "boundary_conditions.h"
#include "common_grid.h"
#include "gpu_params.h"
#include "liquid_closure.h"
#include "mpcd_gpu.h"
#include "srd_collision.h"
#include "zone_pass.h"
#include "zone_pass_openmp.h"

namespace mpcd {
namespace {

void ok(int rc) {
    if (rc != MPCD_OK) {
        const char* m = mpcd_gpu_last_error();
        throw std::runtime_error(m && *m ? m : "libmpcd_gpu call failed");
    }
}

int bc_mode(const std::string& s) {
    if (s == "periodic") return MPCD_BC_PERIODIC;
    if (s == "thermalize") return MPCD_BC_THERMALIZE;
    if (s == "bounceback") return MPCD_BC_BOUNCEBACK;
    if (s == "specular") return MPCD_BC_SPECULAR;
    return MPCD_BC_INVALID; /* an error only if a particle reaches that wall */
}

}  // namespace

mpcd_params to_gpu_params(const Params& q) {
    mpcd_params p{};
    p.Nx = q.Nx;
    p.Ny = q.Ny;
    p.Nc = q.Nc;
    p.n = q.n;
    p.Lx = q.Lx;
    p.Ly = q.Ly;
    p.dt = q.dt;
    p.a0 = q.a0;
    p.alpha = q.alpha;
    p.kBT = q.kBT;
    p.g = q.g;
    p.bodyForceX = q.bodyForceX;
    p.useThermostat = q.useThermostat;
    p.keepMeanFlow = q.keepMeanFlow;
    p.bc_left = bc_mode(q.boundary_left);
    p.bc_right = bc_mode(q.boundary_right);
    p.bc_bottom = bc_mode(q.boundary_bottom);
    p.bc_top = bc_mode(q.boundary_top);
    p.Utop = q.Utop;
    p.Ubottom = q.Ubottom;
    p.wallSigma = q.wallSigma;
    p.noSurfaceCase = q.noSurfaceCase;
    p.redistributionEnableSurfaceTopology = q.redistributionEnableSurfaceTopology;
    p.redistributionWallWettingEnabled = q.redistributionWallWettingEnabled;
    p.useInterfaceVelocityReorientation = q.useInterfaceVelocityReorientation;
    p.useIncompressibleRedistribution = q.useIncompressibleRedistribution;
    p.redistribAfterCollision = q.redistribAfterCollision;
    p.useZoneRedistribution = q.useZoneRedistribution;
    p.zoneTileNx = q.zoneTileNx;
    p.zoneTileNy = q.zoneTileNy;
    p.zoneUseShiftedSecondPass = q.zoneUseShiftedSecondPass;
    p.zonePassthroughMode = q.zonePassthroughMode;
    p.zoneSingleFullDomainMode = q.zoneSingleFullDomainMode;
    p.enableMomentumCorrectionPostRedistribution = q.enableMomentumCorrectionPostRedistribution;
    p.gamma = q.gamma;
    p.coef = q.coef;
    p.high_mode_one_plus_coef = q.highMode == "1+coef";
    p.low_mode_one_plus_coef = q.lowMode == "1+coef";
    p.maxRedistribPasses = q.maxRedistribPasses;
    p.useLocalFluidFractionThresholds = q.useLocalFluidFractionThresholds;
    p.fluidFracGain = q.fluidFracGain;
    p.lowThrFloor = q.lowThrFloor;
    p.highThrFloor = q.highThrFloor;
    p.lowThrBulkOverride = q.lowThrBulkOverride;
    p.useLiquidClosure = q.useLiquidClosure;
    p.useOptimalBetaRepair = q.useOptimalBetaRepair;
    p.betaRepair = q.betaRepair;
    p.betaEOS = q.betaEOS;
    p.Kvirial = q.Kvirial;
    p.smoothPdriveAtInterzone = q.smoothPdriveAtInterzone;
    p.smoothPdriveInterzoneWidthCells = q.smoothPdriveInterzoneWidthCells;
    p.smoothPdriveInterzoneBlend = q.smoothPdriveInterzoneBlend;
    p.smoothPdriveIncludeShiftedLayouts = q.smoothPdriveIncludeShiftedLayouts;
    return p;
}

namespace {

mpcd_params to_c(const Params& q) { return to_gpu_params(q); }

/* "single_full" (or the params flag), "shifted", anything else the base
 * layout (zone_pass.cpp:177-196) */
int layout_mode(const Params& q, const std::string& mode) {
    if (mode == "single_full" || q.zoneSingleFullDomainMode) return MPCD_LAYOUT_SINGLE_FULL;
    if (mode == "shifted") return MPCD_LAYOUT_SHIFTED;
    return MPCD_LAYOUT_BASE;
}

void need_size(const std::vector<double>& a, const Params& q, const char* what) {
    if (a.size() != 2 * (size_t)q.n) throw std::runtime_error(std::string(what) + ": unexpected state size");
}

double rho_target(const Params& q) { return q.gamma / ((q.Lx / (double)q.Nx) * (q.Ly / (double)q.Ny)); }

CellFields fields_of(const std::vector<double>& x, const std::vector<double>& v, const Params& q, double rhoT) {
    const mpcd_params p = to_c(q);
    CellFields G{};
    G.Nx = q.Nx;
    G.Ny = q.Ny;
    const size_t Nc = (size_t)q.Nc;
    G.N.assign(Nc, 0);
    for (std::vector<double>* f : {&G.Ux, &G.Uy, &G.Px, &G.Py, &G.T, &G.rho, &G.Pkin, &G.Pvir, &G.P, &G.rhoTarget})
        f->assign(Nc, 0.0);
    mpcd_cell_fields c{G.N.data(), G.Ux.data(), G.Uy.data(), G.Px.data(), G.Py.data(), G.T.data(),
                       G.rho.data(), G.Pkin.data(), G.Pvir.data(), G.P.data(), G.rhoTarget.data()};
    ok(mpcd_gpu_compute_cell_fields(&p, x.data(), v.data(), rhoT, &c));
    return G;
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

OpenMPPassResult snapshot_pass(const State& in, const Params& q, const std::string& mode, std::uint64_t seed,
                               int threads) {
    if (!q.useIncompressibleRedistribution || !q.useZoneRedistribution)
        throw std::runtime_error(
            "run_zone_pass requires useIncompressibleRedistribution=true and useZoneRedistribution=true");
    need_size(in.x, q, "run_zone_pass_openmp");
    const auto t0 = std::chrono::steady_clock::now();
    OpenMPPassResult out{};
    out.stateOut = in;
    out.zoneOwnerMap.assign((size_t)q.Nc, 0);
    const mpcd_params p = to_c(q);
    mpcd_zone_pass_metrics m{};
    ok(mpcd_gpu_zone_pass(&p, out.stateOut.x.data(), out.stateOut.v.data(), layout_mode(q, mode), seed, &m,
                          out.zoneOwnerMap.data()));
    out.metrics.nZones = m.nZones;
    out.metrics.nThreadsRequested = threads;
    out.metrics.nThreadsUsed = 1;
    out.metrics.nParticlesMovedDense = m.nParticlesMovedDense;
    out.metrics.nParticlesMovedSparse = m.nParticlesMovedSparse;
    out.metrics.corrRmsDU = m.corrRmsDU;
    out.metrics.corrMaxDU = m.corrMaxDU;
    out.metrics.timeKernel = seconds_since(t0);
    return out;
}

}  // namespace

/* ---- boundary_conditions.h ---- */

bool is_periodic_pair(const Params& params, const std::string& axis_name) {
    if (axis_name == "x") return params.boundary_left == "periodic" && params.boundary_right == "periodic";
    if (axis_name == "y") return params.boundary_bottom == "periodic" && params.boundary_top == "periodic";
    throw std::runtime_error("Unknown axis in is_periodic_pair: " + axis_name);
}

void apply_bc_general(std::vector<double>& x, std::vector<double>& v, const Params& params, WallInfo& info,
                      std::uint64_t rng_seed) {
    need_size(x, params, "apply_bc_general");
    need_size(v, params, "apply_bc_general");
    const mpcd_params p = to_c(params);
    mpcd_wall_info w{};
    ok(mpcd_gpu_apply_bc_general(&p, x.data(), v.data(), rng_seed, &w));
    info.nBot += w.nBot;
    info.nTop += w.nTop;
    info.nLeft += w.nLeft;
    info.nRight += w.nRight;
    info.dEwall += w.dEwall;
    info.dPxBot += w.dPxBot;
    info.dPxTop += w.dPxTop;
    info.dPyBot += w.dPyBot;
    info.dPyTop += w.dPyTop;
    info.dPxLeft += w.dPxLeft;
    info.dPxRight += w.dPxRight;
    info.dPyLeft += w.dPyLeft;
    info.dPyRight += w.dPyRight;
}

/* ---- srd_collision.h ---- */

std::vector<int> srd_cell_id_with_random_shift(const std::vector<double>& x, const Params& params,
                                               std::uint64_t rng_seed) {
    need_size(x, params, "srd_cell_id_with_random_shift");
    const mpcd_params p = to_c(params);
    std::vector<int> cid((size_t)params.n);
    ok(mpcd_gpu_srd_cell_id_with_random_shift(&p, x.data(), rng_seed, cid.data()));
    return cid;
}

void srd_collision_step(std::vector<double>& v, const std::vector<int>& cid, const Params& params,
                        std::uint64_t rng_seed) {
    if (v.size() != 2 * (size_t)params.n || cid.size() != (size_t)params.n)
        throw std::runtime_error("srd_collision_step: size mismatch");
    const mpcd_params p = to_c(params);
    ok(mpcd_gpu_srd_collision_step(&p, v.data(), cid.data(), rng_seed));
}

std::vector<int> srd_cell_id_with_random_shift_openmp(const std::vector<double>& x, const Params& params,
                                                      std::uint64_t rng_seed, int) {
    return srd_cell_id_with_random_shift(x, params, rng_seed);
}

void srd_collision_step_openmp(std::vector<double>& v, const std::vector<int>& cid, const Params& params,
                               std::uint64_t rng_seed, int) {
    srd_collision_step(v, cid, params, rng_seed);
}

/* ---- common_grid.h ---- */

int cell_id_from_position(double x, double y, const Params& params) {
    const double dx = params.Lx / (double)params.Nx, dy = params.Ly / (double)params.Ny;
    int ix = (int)std::floor(x / dx), iy = (int)std::floor(y / dy);
    ix = ix < 0 ? 0 : (ix >= params.Nx ? params.Nx - 1 : ix);
    iy = iy < 0 ? 0 : (iy >= params.Ny ? params.Ny - 1 : iy);
    return iy + params.Ny * ix;
}

std::vector<int> build_cell_ids(const std::vector<double>& x, const Params& params) {
    if (x.size() != 2 * (size_t)params.n) throw std::runtime_error("build_cell_ids: unexpected x size");
    std::vector<int> cid((size_t)params.n);
    for (size_t i = 0; i < cid.size(); ++i) cid[i] = cell_id_from_position(x[2 * i], x[2 * i + 1], params);
    return cid;
}

CellFields compute_cell_fields(const std::vector<double>& x, const std::vector<double>& v, const Params& params,
                               double rhoTargetScalar) {
    if (x.size() != 2 * (size_t)params.n || v.size() != 2 * (size_t)params.n)
        throw std::runtime_error("compute_cell_fields: unexpected state size");
    return fields_of(x, v, params, rhoTargetScalar);
}

/* ---- zone_pass.h ---- */

ZoneLayout build_zone_layout(const Params& params, const std::string& mode) {
    const mpcd_params p = to_c(params);
    const int lm = layout_mode(params, mode);
    ZoneLayout L{};
    L.mode = mode;
    L.ownerMap.assign((size_t)params.Nc, 0);
    int32_t nz = 0;
    ok(mpcd_gpu_build_zone_layout(&p, lm, L.ownerMap.data(), &nz, nullptr));
    std::vector<int32_t> r4(4 * (size_t)nz);
    ok(mpcd_gpu_build_zone_layout(&p, lm, L.ownerMap.data(), &nz, r4.data()));
    L.zones.resize((size_t)nz);
    for (int k = 0; k < nz; ++k) {
        ZoneDescriptor& z = L.zones[(size_t)k];
        z.id = k + 1;
        z.maskVec.assign((size_t)params.Nc, 0);
        for (int ix = r4[4 * (size_t)k]; ix < r4[4 * (size_t)k + 1]; ++ix)
            for (int iy = r4[4 * (size_t)k + 2]; iy < r4[4 * (size_t)k + 3]; ++iy) {
                const int c = iy + params.Ny * ix;
                z.cellIds.push_back(c);
                z.maskVec[(size_t)c] = 1;
            }
    }
    return L;
}

ZoneKernelStats run_zone_kernel_inplace(std::vector<double>& x, std::vector<double>& v, const Params& params,
                                        const ZoneDescriptor& zone, std::uint64_t rngSeed) {
    need_size(x, params, "run_zone_kernel_inplace");
    need_size(v, params, "run_zone_kernel_inplace");
    /* the zone is its maskVec, as the original's kernel reads it
     * (zone_pass.cpp:676); a descriptor without a mask of Nc entries gives
     * its cellIds instead */
    std::vector<std::uint8_t> mask((size_t)params.Nc, 0);
    if ((int)zone.maskVec.size() == params.Nc) {
        for (int c = 0; c < params.Nc; ++c) mask[(size_t)c] = zone.maskVec[(size_t)c] ? 1 : 0;
    } else {
        for (int c : zone.cellIds) {
            if (c < 0 || c >= params.Nc) throw std::runtime_error("run_zone_kernel_inplace: cell id out of range");
            mask[(size_t)c] = 1;
        }
    }
    const mpcd_params p = to_c(params);
    ZoneKernelStats out{};
    int32_t moved[2] = {0, 0};
    ok(mpcd_gpu_run_zone_kernel_mask(&p, x.data(), v.data(), mask.data(), rngSeed, out.stats.data(), moved,
                                     out.diag.data()));
    out.movedDense = moved[0];
    out.movedSparse = moved[1];
    return out;
}

std::vector<int> build_particle_ids_in_zone(const std::vector<double>& x, const Params& params,
                                            const ZoneDescriptor& zone) {
    const auto cid = build_cell_ids(x, params);
    std::vector<int> ids;
    for (int i = 0; i < params.n; ++i)
        if (zone.maskVec[(size_t)cid[(size_t)i]]) ids.push_back(i);
    return ids;
}

ZonePassResult run_zone_pass(const State& stateIn, const Params& params, const std::string& layoutMode,
                             std::uint64_t rngSeedBase) {
    if (!params.useIncompressibleRedistribution || !params.useZoneRedistribution)
        throw std::runtime_error(
            "run_zone_pass requires useIncompressibleRedistribution=true and useZoneRedistribution=true");
    need_size(stateIn.x, params, "run_zone_pass");
    ZonePassResult out{};
    out.stateOut = stateIn;
    const double rhoT = rho_target(params);
    out.beforeFields = fields_of(stateIn.x, stateIn.v, params, rhoT);
    out.zoneOwnerMap.assign((size_t)params.Nc, 0);
    const mpcd_params p = to_c(params);
    mpcd_zone_pass_metrics m{};
    ok(mpcd_gpu_zone_pass_gauss_seidel(&p, out.stateOut.x.data(), out.stateOut.v.data(),
                                       layout_mode(params, layoutMode), rngSeedBase, &m, out.zoneOwnerMap.data()));
    out.afterFields = fields_of(out.stateOut.x, out.stateOut.v, params, rhoT);
    mpcd_occupancy_stats sb{}, sa{};
    ok(mpcd_gpu_occupancy_stats(&p, out.beforeFields.N.data(), &sb));
    ok(mpcd_gpu_occupancy_stats(&p, out.afterFields.N.data(), &sa));
    out.metrics.occStdBefore = sb.occStd;
    out.metrics.occStdAfter = sa.occStd;
    out.metrics.outBandBefore = sb.outBand;
    out.metrics.outBandAfter = sa.outBand;
    out.metrics.nDenseBefore = sb.nDense;
    out.metrics.nDenseAfter = sa.nDense;
    out.metrics.nSparseBefore = sb.nSparse;
    out.metrics.nSparseAfter = sa.nSparse;
    out.metrics.nEmptyActiveBefore = sb.nEmptyActive;
    out.metrics.nEmptyActiveAfter = sa.nEmptyActive;
    out.metrics.layoutMode = layoutMode;
    out.metrics.nZonesExecuted = m.nZones;
    out.metrics.nParticlesMovedDense = m.nParticlesMovedDense;
    out.metrics.nParticlesMovedSparse = m.nParticlesMovedSparse;
    out.metrics.corrRmsDU = m.corrRmsDU;
    out.metrics.corrMaxDU = m.corrMaxDU;
    double bx = 0.0, by = 0.0, ax = 0.0, ay = 0.0;
    for (int i = 0; i < params.n; ++i) {
        bx += stateIn.v[2 * (size_t)i];
        by += stateIn.v[2 * (size_t)i + 1];
        ax += out.stateOut.v[2 * (size_t)i];
        ay += out.stateOut.v[2 * (size_t)i + 1];
    }
    out.metrics.momentumDeltaX = ax - bx;
    out.metrics.momentumDeltaY = ay - by;
    return out;
}

/* ---- zone_pass_openmp.h ---- */

OpenMPPassResult run_zone_pass_openmp(const State& stateIn, const Params& params, const std::string& layoutMode,
                                      std::uint64_t rngSeedBase, int nThreadsRequested) {
    return snapshot_pass(stateIn, params, layoutMode, rngSeedBase, nThreadsRequested);
}

OpenMPPassResult run_zone_pass_openmp_base(const State& stateIn, const Params& params, std::uint64_t rngSeedBase,
                                           int nThreadsRequested) {
    return snapshot_pass(stateIn, params, "base", rngSeedBase, nThreadsRequested);
}

OpenMPPassResult run_zone_pass_sync_serial(const State& stateIn, const Params& params,
                                           const std::string& layoutMode, std::uint64_t rngSeedBase) {
    return snapshot_pass(stateIn, params, layoutMode, rngSeedBase, 1);
}

OpenMPPassResult run_zone_pass_sync_serial_base(const State& stateIn, const Params& params,
                                                std::uint64_t rngSeedBase) {
    return snapshot_pass(stateIn, params, "base", rngSeedBase, 1);
}

/* ---- liquid_closure.h ---- */

LiquidClosureResult run_liquid_closure(const State& stateRef, const State& stateRed, const Params& params) {
    if (!params.useLiquidClosure) throw std::runtime_error("run_liquid_closure requires useLiquidClosure=true");
    if (stateRef.x.size() != stateRed.x.size() || stateRef.v.size() != stateRed.v.size())
        throw std::runtime_error("run_liquid_closure requires ref/red states with matching sizes");
    need_size(stateRed.x, params, "run_liquid_closure");
    const auto t0 = std::chrono::steady_clock::now();
    const mpcd_params p = to_c(params);
    const size_t Nc = (size_t)params.Nc;
    LiquidClosureResult out{};
    out.stateOut = stateRed;
    for (std::vector<double>* g : {&out.pdriveRaw, &out.pdrive, &out.duRepairX, &out.duRepairY, &out.duEOSX,
                                   &out.duEOSY})
        g->assign(Nc, 0.0);
    out.interzoneMask.assign(Nc, 0);
    out.stateRepaired = stateRed;
    mpcd_closure_grids g{out.pdriveRaw.data(), out.pdrive.data(), out.duRepairX.data(), out.duRepairY.data(),
                         out.duEOSX.data(),   out.duEOSY.data(),  out.interzoneMask.data(), out.stateRepaired.v.data()};
    mpcd_closure_metrics m{};
    ok(mpcd_gpu_liquid_closure(&p, stateRef.x.data(), stateRef.v.data(), stateRed.x.data(), stateRed.v.data(),
                               out.stateOut.x.data(), out.stateOut.v.data(), &m, &g));
    out.metrics.betaRepairOpt = m.betaRepairOpt;
    out.metrics.betaRepairApplied = m.betaRepairApplied;
    out.metrics.betaRepairNum = m.betaRepairNum;
    out.metrics.betaRepairDen = m.betaRepairDen;
    out.metrics.betaEOSApplied = m.betaEOSApplied;
    out.metrics.meanPdriveRaw = m.meanPdriveRaw;
    out.metrics.meanPdrive = m.meanPdrive;
    out.metrics.rmsRepairDU = m.rmsRepairDU;
    out.metrics.meanAbsRepairDU = m.meanAbsRepairDU;
    out.metrics.maxRepairDU = m.maxRepairDU;
    out.metrics.rmsEOSDU = m.rmsEOSDU;
    out.metrics.maxEOSDU = m.maxEOSDU;
    out.metrics.momentumDeltaX = m.momentumDeltaX;
    out.metrics.momentumDeltaY = m.momentumDeltaY;
    out.metrics.nMaskedInterzoneCells = m.nMaskedInterzoneCells;
    const double rhoT = rho_target(params);
    out.refFields = fields_of(stateRef.x, stateRef.v, params, rhoT);
    out.redFields = fields_of(stateRed.x, stateRed.v, params, rhoT);
    out.repairedFields = fields_of(out.stateRepaired.x, out.stateRepaired.v, params, rhoT);
    out.outFields = fields_of(out.stateOut.x, out.stateOut.v, params, rhoT);
    out.metrics.timeRepairApply = seconds_since(t0);
    return out;
}

}  // namespace mpcd
