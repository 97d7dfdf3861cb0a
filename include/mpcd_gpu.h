/*
 * mpcd_gpu.h -- C ABI of the gfx950 SRD/MPCD step library (libmpcd_gpu.so).
 *
 * Plain pointers and sizes only.  Every entry point returns 0 on success or a
 * negative MPCD_E* code; mpcd_gpu_last_error() then returns the message the
 * reference would have thrown as std::runtime_error (the C++ layer in
 * mpcd/*.h rethrows it).  Particle arrays at this boundary use the
 * reference's interleaved layout [x1,y1,x2,y2,...] in particle-id order
 * (types.h:74-79); inside the library the state is FP64 structure-of-arrays.
 *
 * Counter-based randomness: every rng_seed argument is the seed the
 * reference passes to the same function; draws are keyed as mpcd_rng.h
 * documents.
 */
#pragma once

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MPCD_OK = 0,
    MPCD_E_INVALID = -1,  /* bad argument / size mismatch / disallowed mode */
    MPCD_E_HIP = -2,      /* HIP runtime failure */
    MPCD_E_BC_MODE = -3,  /* a particle hit a wall whose mode is unsupported */
    MPCD_E_CAPACITY = -4, /* a zone exceeds the library's working-set limits */
};

/* Wall modes (boundary_conditions.cpp:47-172). */
enum {
    MPCD_BC_INVALID = -1,
    MPCD_BC_PERIODIC = 0,
    MPCD_BC_THERMALIZE = 1,
    MPCD_BC_BOUNCEBACK = 2,
    MPCD_BC_SPECULAR = 3,
};

/* Zone layout modes (zone_pass.cpp:173-221). */
enum {
    MPCD_LAYOUT_BASE = 0,
    MPCD_LAYOUT_SHIFTED = 1,
    MPCD_LAYOUT_SINGLE_FULL = 2,
};

/* POD image of mpcd::Params (types.h:10-72).  Strings become enums:
 * bc_* from boundary_* ("periodic", "thermalize", "bounceback",
 * "specular"; anything else MPCD_BC_INVALID), high/low_mode_one_plus_coef
 * is 1 when highMode/lowMode == "1+coef".  Booleans are 0/1. */
typedef struct mpcd_params {
    int32_t Nx, Ny, Nc, n;
    double Lx, Ly, dt, a0;
    double alpha, kBT, g, bodyForceX;
    int32_t useThermostat, keepMeanFlow;
    int32_t bc_left, bc_right, bc_bottom, bc_top;
    double Utop, Ubottom, wallSigma;
    int32_t noSurfaceCase, redistributionEnableSurfaceTopology, redistributionWallWettingEnabled,
        useInterfaceVelocityReorientation, useIncompressibleRedistribution, redistribAfterCollision,
        useZoneRedistribution;
    int32_t zoneTileNx, zoneTileNy;
    int32_t zoneUseShiftedSecondPass, zonePassthroughMode, zoneSingleFullDomainMode,
        enableMomentumCorrectionPostRedistribution;
    double gamma, coef;
    int32_t high_mode_one_plus_coef, low_mode_one_plus_coef;
    int32_t maxRedistribPasses, useLocalFluidFractionThresholds;
    double fluidFracGain, lowThrFloor, highThrFloor, lowThrBulkOverride;
    int32_t useLiquidClosure, useOptimalBetaRepair;
    double betaRepair, betaEOS, Kvirial;
    int32_t smoothPdriveAtInterzone, smoothPdriveInterzoneWidthCells;
    double smoothPdriveInterzoneBlend;
    int32_t smoothPdriveIncludeShiftedLayouts;
    int32_t _pad;
} mpcd_params;

/* WallInfo (types.h:97-111). */
typedef struct mpcd_wall_info {
    int32_t nBot, nTop, nLeft, nRight;
    double dEwall, dPxBot, dPxTop, dPyBot, dPyTop, dPxLeft, dPxRight, dPyLeft, dPyRight;
} mpcd_wall_info;

/* Per-pass metrics of the snapshot zone pass (zone_pass_openmp.h:12-27,
 * the fields the GPU produces). */
typedef struct mpcd_zone_pass_metrics {
    int32_t nZones;
    int32_t nParticlesMovedDense;
    int32_t nParticlesMovedSparse;
    int32_t _pad;
    double corrRmsDU;
    double corrMaxDU;
} mpcd_zone_pass_metrics;

/* LiquidClosureMetrics (liquid_closure.h:11-27). */
typedef struct mpcd_closure_metrics {
    double betaRepairOpt, betaRepairApplied, betaRepairNum, betaRepairDen, betaEOSApplied;
    double meanPdriveRaw, meanPdrive;
    double rmsRepairDU, meanAbsRepairDU, maxRepairDU, rmsEOSDU, maxEOSDU;
    double momentumDeltaX, momentumDeltaY;
    int32_t nMaskedInterzoneCells;
    int32_t _pad;
} mpcd_closure_metrics;

/* CellFields (types.h:81-95): caller-owned arrays of Nc entries each; any
 * pointer may be NULL to skip that field. */
typedef struct mpcd_cell_fields {
    int32_t* N;
    double *Ux, *Uy, *Px, *Py, *T, *rho, *Pkin, *Pvir, *P, *rhoTarget;
} mpcd_cell_fields;

#ifdef __cplusplus
}
#endif
