/*
 * types.h -- data types of the mpcd:: C++ API (drop-in for the reference's
 * include/types.h:10-123).  Field names, types and order follow the
 * reference so its drivers compile unchanged against this library.
 */
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace mpcd {

/* Run parameters (params.kv keys of the same names). */
struct Params {
    int Nx;
    int Ny;
    int Nc;
    int n;
    double Lx;
    double Ly;
    double dt;
    double a0;

    double alpha;
    double kBT;
    double g;
    double bodyForceX;
    bool useThermostat;
    bool keepMeanFlow;

    std::string boundary_left; /* "periodic", "thermalize", "bounceback", "specular" */
    std::string boundary_right;
    std::string boundary_bottom;
    std::string boundary_top;
    double Utop;
    double Ubottom;
    double wallSigma;

    bool noSurfaceCase;
    bool redistributionEnableSurfaceTopology;
    bool redistributionWallWettingEnabled;
    bool useInterfaceVelocityReorientation;
    bool useIncompressibleRedistribution;
    bool redistribAfterCollision;
    bool useZoneRedistribution;
    int zoneTileNx;
    int zoneTileNy;
    bool zoneUseShiftedSecondPass;
    bool zonePassthroughMode;
    bool zoneSingleFullDomainMode;
    bool enableMomentumCorrectionPostRedistribution;
    double gamma;
    double coef;
    std::string highMode; /* "1+coef" or additive */
    std::string lowMode;
    int maxRedistribPasses;
    bool useLocalFluidFractionThresholds;
    double fluidFracGain;
    double lowThrFloor;
    double highThrFloor;
    double lowThrBulkOverride; /* NaN: derived from gamma and coef */

    bool useLiquidClosure;
    bool useOptimalBetaRepair;
    double betaRepair;
    double betaEOS;
    double Kvirial;
    bool smoothPdriveAtInterzone;
    int smoothPdriveInterzoneWidthCells;
    double smoothPdriveInterzoneBlend;
    bool smoothPdriveIncludeShiftedLayouts;
};

/* Particle state: interleaved x and v ([x1,y1,x2,y2,...]), optional type
 * and reference positions carried through untouched. */
struct State {
    std::vector<double> x;
    std::vector<double> v;
    std::vector<std::uint8_t> type;
    std::vector<double> r0;
};

/* Per-cell fields, cell index iy + Ny*ix. */
struct CellFields {
    int Nx;
    int Ny;
    std::vector<std::int32_t> N;
    std::vector<double> Ux;
    std::vector<double> Uy;
    std::vector<double> Px;
    std::vector<double> Py;
    std::vector<double> T;
    std::vector<double> rho;
    std::vector<double> Pkin;
    std::vector<double> Pvir;
    std::vector<double> P;
    std::vector<double> rhoTarget;
};

/* Wall hits and the energy / momentum they exchanged. */
struct WallInfo {
    int nBot = 0;
    int nTop = 0;
    int nLeft = 0;
    int nRight = 0;
    double dEwall = 0.0;
    double dPxBot = 0.0;
    double dPxTop = 0.0;
    double dPyBot = 0.0;
    double dPyTop = 0.0;
    double dPxLeft = 0.0;
    double dPxRight = 0.0;
    double dPyLeft = 0.0;
    double dPyRight = 0.0;
};

/* What write_runout_kv records for a tranche step. */
struct RunOutStep {
    std::string inputTag;
    std::string outputTag;
    double pBotInst = 0.0;
    double pTopInst = 0.0;
    double pMeanInst = 0.0;
    std::array<double, 6> redistribStats{};
    std::array<double, 10> redistDiag{};
    std::string layoutMode;
    int nZonesExecuted = 0;
};

}  // namespace mpcd
