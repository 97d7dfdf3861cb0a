/*
 * mpcd_gpu.h -- C ABI of the gfx950 SRD/MPCD step library (libmpcd_gpu.so).
 *
 * Plain pointers and sizes only.  Every entry point returns 0 on success or a
 * negative MPCD_E* code; mpcd_gpu_last_error() then returns the message the
 * reference would have thrown as std::runtime_error (the C++ layer in
 * include/mpcd/ rethrows it).  Particle arrays at this boundary use the
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

/* The closure grids of LiquidClosureResult (liquid_closure.h:36-42) and,
 * in vRepaired (2n, interleaved, input particle order), the velocities of
 * its stateRepaired (liquid_closure.h:31: the red state kicked by
 * beta*duRepair, liquid_closure.cpp:379-387); any pointer may be NULL. */
typedef struct mpcd_closure_grids {
    double *pdriveRaw, *pdrive, *duRepairX, *duRepairY, *duEOSX, *duEOSY;
    int32_t* interzoneMask;
    double* vRepaired;
} mpcd_closure_grids;

/* Occupancy statistics of a pass (ZonePassMetrics' before/after fields,
 * zone_pass.cpp:975-1044). */
typedef struct mpcd_occupancy_stats {
    double occStd, outBand;
    int32_t nDense, nSparse, nEmptyActive, _pad;
} mpcd_occupancy_stats;

/* Diagnostics of one resident time step, reduced on the device: the
 * occupancy statistics after each zone pass (the driver's baseOccStd,
 * shiftedOccStd, ... columns and ZonePassMetrics' after-fields), the
 * particle means of the step's output state (mean_kinetic, mean_component,
 * main_benchmark_poiseuille_openmp_profiled_shifted_profiled.cpp:105-116)
 * and the means of its cell fields (the closure's outFields; Qx =
 * mean(Ux) Ly, :124-126). */
typedef struct mpcd_step_diag {
    mpcd_occupancy_stats base;
    mpcd_occupancy_stats shifted;
    double meanKinetic, meanVelocityX, meanVelocityY;
    double meanUx, meanUy, Qx;
} mpcd_step_diag;

/* Metrics of one resident time step. */
typedef struct mpcd_step_metrics {
    mpcd_wall_info wall;
    mpcd_zone_pass_metrics base;
    mpcd_zone_pass_metrics shifted;
    mpcd_closure_metrics closure;
} mpcd_step_metrics;

/* Kernel classes timed by mpcd_gpu_sim_profile (events recorded on the
 * simulation's stream around each launch group). */
enum {
    MPCD_PROF_STREAM_BC = 0,   /* force + stream + wall BC + shifted cell */
    MPCD_PROF_BIN = 1,         /* binnings (2 counting sorts, 2 updates per step) */
    MPCD_PROF_COLLIDE = 2,     /* SRD rotation + thermostat */
    MPCD_PROF_CELL_FIELDS = 3, /* ref cell sums for the closure */
    MPCD_PROF_ZONE_PASS = 4,   /* a whole zone pass (2 per step) */
    MPCD_PROF_ZONE_KERNEL = 5, /* the zone kernel launches (passes, and the correction when fused; 2 per step) */
    MPCD_PROF_CLOSURE = 6,     /* closure: fields, beta, stencil, kick */
    MPCD_PROF_KICK = 7,        /* the closure's per-particle kick alone */
    MPCD_PROF_ZONE_FINISH = 8, /* the separate momentum correction kernel (zones the fused kernel does not take) */
    MPCD_PROF_KICK_STREAM = 9, /* the kick fused with the next step's stream (steps of one call) */
    MPCD_PROF_CLASSES = 10
};

/* Opaque device-resident simulation (state, binnings, layouts). */
typedef struct mpcd_gpu_sim mpcd_gpu_sim;

/* ---- errors / device ---- */
const char* mpcd_gpu_last_error(void);
int mpcd_gpu_device_count(int* count);

/* ---- stateless stage calls (host buffers in, host buffers out) ---- */

/* replaces mpcd::apply_bc_general (boundary_conditions.h:9-13) */
int mpcd_gpu_apply_bc_general(const mpcd_params* p, double* x, double* v, uint64_t rng_seed, mpcd_wall_info* info);

/* body force + streaming + apply_bc_general in one pass: the driver loop
 * main_benchmark_poiseuille_openmp_profiled_shifted_profiled.cpp:242-249
 * (do_stream = 0 gives apply_bc_general alone) */
int mpcd_gpu_force_stream_bc(const mpcd_params* p, double* x, double* v, uint64_t rng_seed, int do_stream,
                             mpcd_wall_info* info);

/* replaces mpcd::srd_cell_id_with_random_shift (srd_collision.h:9-11) */
int mpcd_gpu_srd_cell_id_with_random_shift(const mpcd_params* p, const double* x, uint64_t rng_seed, int32_t* cid);

/* replaces mpcd::srd_collision_step (srd_collision.h:12-15) */
int mpcd_gpu_srd_collision_step(const mpcd_params* p, double* v, const int32_t* cid, uint64_t rng_seed);

/* replaces mpcd::compute_cell_fields (common_grid.h:10-13) */
int mpcd_gpu_compute_cell_fields(const mpcd_params* p, const double* x, const double* v, double rhoTarget,
                                 mpcd_cell_fields* out);

/* No counterpart in the reference: compute_cell_fields (as
 * mpcd_gpu_compute_cell_fields above, same results bit for bit) for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  x, v (interleaved, 2n
 * doubles) and every non-NULL array of `out` are DEVICE pointers the caller
 * owns; the work is enqueued on `stream` (a hipStream_t, NULL for the null
 * stream) and the call returns without waiting.  It allocates nothing: its
 * working memory is a static device array of the library (about 39 MB per
 * GPU, held until the process ends) that bounds the problem to
 * MPCD_NOALLOC_MAX_PARTICLES particles and MPCD_NOALLOC_MAX_CELLS cells
 * (larger sizes return MPCD_E_INVALID before anything is enqueued).  Every
 * call on one GPU shares that array: calls that overlap on one GPU (on
 * different streams, or from several threads) race. */
#define MPCD_NOALLOC_MAX_PARTICLES (1 << 20)
#define MPCD_NOALLOC_MAX_CELLS (1 << 18)
int mpcd_gpu_compute_cell_fields_noalloc(const mpcd_params* p, const double* x, const double* v, double rhoTarget,
                                         const mpcd_cell_fields* out, void* stream);

/* replaces mpcd::build_zone_layout (zone_pass.h:60); rects4 (4 per zone:
 * xs, xe, ys, ye) may be NULL */
int mpcd_gpu_build_zone_layout(const mpcd_params* p, int mode, int32_t* ownerMap, int32_t* nZones, int32_t* rects4);

/* replaces mpcd::run_zone_kernel_inplace (zone_pass.h:62-66) for a
 * rectangular zone [xs,xe) x [ys,ye); stats6 = ZoneKernelStats::stats,
 * moved2 = movedDense, movedSparse, diag10 = ZoneKernelStats::diag
 * (zone_pass.cpp:924-970, reduced on the GPU: occupancy spread / gamma and
 * out-of-band fraction of the final counts over the whole domain, |change
 * of total momentum|, RMS of the cells' nonzero x and y momentum defects
 * before the correction, change of kinetic energy, NaN, 0, RMS of the zone
 * cells' final x and y velocities over the occupied cells).  Any of the
 * three may be NULL; a NULL diag10 skips the diagnostics' work. */
int mpcd_gpu_run_zone_kernel(const mpcd_params* p, double* x, double* v, int xs, int xe, int ys, int ye,
                             uint64_t rng_seed, double* stats6, int32_t* moved2, double* diag10);
/* replaces mpcd::run_zone_kernel_inplace (zone_pass.h:62-66) for a zone of
 * any shape, given as ZoneDescriptor::maskVec (zone_pass.h:12-16): zone_mask
 * holds Nc bytes, nonzero for the zone's cells (at least one).  The kernel
 * runs over the mask's bounding rectangle; cells of it outside the mask keep
 * their particles and are never in the dense or low lists, a target, a
 * neighbour or a donor (zone_pass.cpp:676-678, :611, :625, :263-278, :819).
 * Outputs as mpcd_gpu_run_zone_kernel. */
int mpcd_gpu_run_zone_kernel_mask(const mpcd_params* p, double* x, double* v, const uint8_t* zone_mask,
                                  uint64_t rng_seed, double* stats6, int32_t* moved2, double* diag10);

/* replaces mpcd::run_zone_pass_openmp / run_zone_pass_sync_serial
 * (zone_pass_openmp.h:35-53): snapshot semantics, zones on the GPU */
int mpcd_gpu_zone_pass(const mpcd_params* p, double* x, double* v, int layoutMode, uint64_t rngSeedBase,
                       mpcd_zone_pass_metrics* m, int32_t* ownerMap);

/* replaces mpcd::run_zone_pass (zone_pass.h:71-74): the zones of the layout
 * applied one after another to the evolving state (Gauss-Seidel), each
 * seeded rngSeedBase + 104729 (zone index + 1) */
int mpcd_gpu_zone_pass_gauss_seidel(const mpcd_params* p, double* x, double* v, int layoutMode,
                                    uint64_t rngSeedBase, mpcd_zone_pass_metrics* m, int32_t* ownerMap);

/* occupancy statistics from Nc cell counts (compute_occ_std_from_counts,
 * compute_out_band_from_counts, count_dense/sparse_from_counts,
 * count_empty_active: zone_pass.cpp:975-1044), reduced on the GPU */
int mpcd_gpu_occupancy_stats(const mpcd_params* p, const int32_t* counts, mpcd_occupancy_stats* out);

/* replaces mpcd::run_liquid_closure (liquid_closure.h:46-48) */
int mpcd_gpu_liquid_closure(const mpcd_params* p, const double* xRef, const double* vRef, const double* xRed,
                            const double* vRed, double* xOut, double* vOut, mpcd_closure_metrics* m,
                            mpcd_closure_grids* grids);

/* ---- resident stepping: the driver loop body
 * (main_benchmark_poiseuille_openmp_profiled_shifted_profiled.cpp:240-279)
 * with the state kept in HBM between steps ---- */
int mpcd_gpu_sim_create(const mpcd_params* p, mpcd_gpu_sim** out);
int mpcd_gpu_sim_destroy(mpcd_gpu_sim* s);
/* x or v may be NULL: that array keeps each particle's current value */
int mpcd_gpu_sim_set_state(mpcd_gpu_sim* s, const double* x, const double* v);
int mpcd_gpu_sim_get_state(mpcd_gpu_sim* s, double* x, double* v);
/* steps first_step .. first_step+n_steps-1 with the driver's per-step seeds;
 * `last` (nullable) receives the last step's metrics */
int mpcd_gpu_sim_step(mpcd_gpu_sim* s, int first_step, int n_steps, mpcd_step_metrics* last);
/* same, enqueued without a host wait; mpcd_gpu_sim_check then waits,
 * reports an error (wall mode, capacity, binning update) raised by any step
 * enqueued since the last check, and clears it */
int mpcd_gpu_sim_step_async(mpcd_gpu_sim* s, int first_step, int n_steps);
int mpcd_gpu_sim_check(mpcd_gpu_sim* s);
/* one step that also returns the ref / after-base / red states */
int mpcd_gpu_sim_step_stages(mpcd_gpu_sim* s, int step, double* x_ref, double* v_ref, double* x_base, double* v_base,
                             double* x_red, double* v_red, mpcd_step_metrics* m);
/* one step (metrics as mpcd_gpu_sim_step_stages, m nullable) that also
 * reduces its diagnostics on the device: the metrics row of the benchmark
 * driver (main_benchmark_poiseuille_openmp_profiled_shifted_profiled.cpp:285-337) */
int mpcd_gpu_sim_step_diag(mpcd_gpu_sim* s, int step, mpcd_step_metrics* m, mpcd_step_diag* d);
/* compute_cell_fields (common_grid.h:10-13) of the resident state: every
 * field of `out` that is not NULL, rhoTarget filled with the given value */
int mpcd_gpu_sim_cell_fields(mpcd_gpu_sim* s, double rhoTarget, mpcd_cell_fields* out);
int mpcd_gpu_sim_synchronize(mpcd_gpu_sim* s);
int mpcd_gpu_device_synchronize(void);
/* timing with events on the simulation's stream: enable is a mask of
 * MPCD_PROFILE_KERNELS (per-class kernel times, mpcd_gpu_sim_profile_read)
 * and MPCD_PROFILE_PHASES (per-phase step times, mpcd_gpu_sim_phase_read);
 * 0 disables; every call resets the totals */
enum { MPCD_PROFILE_KERNELS = 1, MPCD_PROFILE_PHASES = 2 };
int mpcd_gpu_sim_profile(mpcd_gpu_sim* s, int enable);
/* total milliseconds and launch-group counts per MPCD_PROF_* class since the
 * last reset (arrays of MPCD_PROF_CLASSES entries) */
int mpcd_gpu_sim_profile_read(mpcd_gpu_sim* s, double* total_ms, int32_t* launches);
/* milliseconds, summed over the steps profiled since the last reset, of the
 * step's phases as the benchmark driver times them
 * (main_benchmark_poiseuille_openmp_profiled_shifted_profiled.cpp:240-279):
 * ms4 = reference build (force, stream, walls, collision, reference
 * fields), base pass, shifted pass, closure (with the next step's stream
 * when the kick performs it); each pass includes its binning update */
int mpcd_gpu_sim_phase_read(mpcd_gpu_sim* s, double* ms4);
/* the hipStream_t the simulation launches on */
int mpcd_gpu_sim_stream(mpcd_gpu_sim* s, void** stream);

/* ---- closure surrogate: a 4-layer 3x3 convolutional network
 * (4 -> 32 -> 32 -> 32 -> 2 channels, ReLU after the first three) over the
 * cell grid, bfloat16 weights and activations on the matrix cores with
 * float32 sums.  No counterpart in the reference: it stands in for the
 * run_liquid_closure call of the benchmark driver
 * (main_benchmark_poiseuille_openmp_profiled_shifted_profiled.cpp:277) on
 * the steps the plain closure skips.  Grids: B samples of Nx*Ny cells,
 * cell index iy + Ny*ix; features [sample][cell][4], outputs and targets
 * [sample][cell][2]; parameters (float64) per layer W[co][tap][ci] with
 * tap = 3*(dx+1) + (dy+1), then b[co]. ---- */
typedef struct mpcd_gpu_net mpcd_gpu_net;
int mpcd_gpu_net_num_params(void);
int mpcd_gpu_net_create(int Nx, int Ny, int perx, int pery, int maxBatch, mpcd_gpu_net** out);
int mpcd_gpu_net_destroy(mpcd_gpu_net* net);
/* sets the parameters and resets the optimiser state */
int mpcd_gpu_net_set_params(mpcd_gpu_net* net, const double* theta);
int mpcd_gpu_net_get_params(mpcd_gpu_net* net, double* theta);
int mpcd_gpu_net_forward(mpcd_gpu_net* net, const float* feat, int B, float* out);
/* mean squared error over B samples and its gradient (either may be NULL) */
int mpcd_gpu_net_loss_grad(mpcd_gpu_net* net, const float* feat, const float* target, int B, double* loss,
                           double* grad);
/* one Adam step (lr; betas 0.9, 0.999; eps 1e-8) with the given gradient, or
 * with the last computed one when grad is NULL */
int mpcd_gpu_net_adam_step(mpcd_gpu_net* net, const double* grad, double lr);
/* iters Adam steps on mini-batches of `batch` samples drawn (counter-based,
 * seeded) from nSamples samples; losses[iters] (nullable) */
int mpcd_gpu_net_train(mpcd_gpu_net* net, const float* feat, const float* target, int nSamples, int batch,
                       int iters, double lr, uint64_t seed, double* losses);
/* the resident loop with the surrogate: steps s with s % K != 0 take the
 * surrogate's correction (outScale * network output) instead of the plain
 * closure, unless the residual check fails: a prediction beyond duCap (or
 * not finite), a red state whose occupancy spread exceeds occTol times the
 * last plain step's, or a corrected velocity whose divergence residual
 * exceeds divTol times the plain closure's (mpcd_gpu_sim_set_surrogate_check;
 * defaults 1.1 and 2.0).  The check runs on the device and a failed step
 * takes the plain closure without a host round trip.  K <= 0 or net NULL:
 * plain closure every step. */
int mpcd_gpu_sim_set_surrogate(mpcd_gpu_sim* s, mpcd_gpu_net* net, int K, double outScale, double duCap);
int mpcd_gpu_sim_set_surrogate_check(mpcd_gpu_sim* s, double occTol, double divTol);
/* row profiles of the state after each step, summed on the device over the
 * steps since enabling (for the Poiseuille profile): enable 1 clears and
 * starts, 0 stops, -1 leaves the state as it is; rows3 (nullable, 3 x Ny)
 * receives per cell row the sums of vx, of vx^2 + vy^2 and of the particle
 * count, steps (nullable) the number of steps summed.  Steps of one call
 * stream on their own while enabled. */
int mpcd_gpu_sim_row_profile(mpcd_gpu_sim* s, int enable, double* rows3, int32_t* steps);
/* steps taken with the surrogate, with the plain closure, and fallbacks */
int mpcd_gpu_sim_surrogate_counts(mpcd_gpu_sim* s, int32_t* counts3);
/* training samples: with enable, every plain closure also stores its
 * features [cell][4] (those of the step's reference state: the cells after
 * the collision, before the zone passes -- the state the closure's EOS drive
 * is computed from) and its applied correction / outScale [cell][2];
 * feat/target (nullable) receive the last stored sample */
int mpcd_gpu_sim_closure_sample(mpcd_gpu_sim* s, int enable, double outScale, float* feat, float* target);
/* Training data of the closure surrogate in the reference's closure formats
 * (the CellFields and grids of LiquidClosureResult, liquid_closure.h:29-44,
 * as main_step04_liquid_closure.cpp:61-71 writes them): with enable = 1
 * every plain closure of the following steps keeps a snapshot of its input
 * and output on the device -- the reference state's and the red state's
 * CellFields (compute_cell_fields, common_grid.h:10-13; rhoTarget filled
 * with the simulation's target density), the grids pdriveRaw, pdrive,
 * duRepairX/Y, duEOSX/Y and betaRepairApplied (interzoneMask, a constant of
 * the layouts, is copied from the simulation's own).  The call
 * copies the last snapshot kept into every non-NULL array of ref, red and
 * grids and into *beta, and sets *step to its step (-1 when none was kept
 * yet; then nothing else is written).  enable = 0 stops keeping them. */
int mpcd_gpu_sim_closure_snapshot(mpcd_gpu_sim* s, int enable, const mpcd_cell_fields* ref,
                                  const mpcd_cell_fields* red, const mpcd_closure_grids* grids, double* beta,
                                  int32_t* step);
/* The surrogate's training pair of one closure snapshot (as the files of
 * mpcd_gpu_sim_closure_snapshot / main_step04_liquid_closure hold it),
 * computed on the GPU by the kernels of mpcd_gpu_sim_closure_sample:
 * feat[Nc][4] from the reference state's N, Ux, Uy, T, and target[Nc][2] =
 * (beta duRepair + duEOS) / outScale.  feat or target may be NULL. */
int mpcd_gpu_closure_training_pair(const mpcd_params* p, const int32_t* refN, const double* refUx,
                                   const double* refUy, const double* refT, const double* duRepairX,
                                   const double* duRepairY, const double* duEOSX, const double* duEOSY, double beta,
                                   double outScale, float* feat, float* target);
/* ---- batched ensembles ----
 * No counterpart in the reference, which runs one channel per process
 * (main_benchmark_poiseuille_openmp_profiled_shifted_profiled.cpp:186-237)
 * and its ensembles as separate runs.  An ensemble advances many
 * independent channels -- replicas with their own parameters (grid, particle
 * count, forcing, walls, thresholds) and seed keys -- together: every stage
 * of the step is one launch over all replicas, and the step is captured
 * into a HIP graph and replayed.  Replica r follows, bit for bit, a single
 * simulation of its parameters given mpcd_gpu_sim_set_seed_key(key[r]).
 * Replicas run the full plain step (zones and the liquid closure on;
 * keepMeanFlow and the surface modes are refused). */
typedef struct mpcd_gpu_ensemble mpcd_gpu_ensemble;

/* the seeds of every following step of s become the benchmark driver's plus
 * key * 0x9E3779B97F4A7C15 (key 0: the driver's own seeds, the default) */
int mpcd_gpu_sim_set_seed_key(mpcd_gpu_sim* s, uint64_t key);

/* params[nReplicas]; seedKeys[nReplicas] or NULL (every key 0) */
int mpcd_gpu_ensemble_create(const mpcd_params* params, const uint64_t* seedKeys, int nReplicas,
                             mpcd_gpu_ensemble** out);
int mpcd_gpu_ensemble_destroy(mpcd_gpu_ensemble* e);
/* total particles and cells, and the zones of the base and shifted layouts */
int mpcd_gpu_ensemble_info(mpcd_gpu_ensemble* e, int64_t* particles, int32_t* cells, int32_t* zones2);
/* replica r's state, interleaved host arrays in particle-id order (as
 * mpcd_gpu_sim_set_state / _get_state) */
int mpcd_gpu_ensemble_set_state(mpcd_gpu_ensemble* e, int replica, const double* x, const double* v);
int mpcd_gpu_ensemble_get_state(mpcd_gpu_ensemble* e, int replica, double* x, double* v);
/* enqueues steps firstStep .. firstStep+nSteps-1 of every replica on the
 * ensemble's stream and returns; useGraph: replay one captured graph per
 * step (captured at the first use) instead of launching each kernel */
int mpcd_gpu_ensemble_step(mpcd_gpu_ensemble* e, int firstStep, int nSteps, int useGraph);
/* waits for the enqueued steps; reports a wall-mode or capacity error */
int mpcd_gpu_ensemble_synchronize(mpcd_gpu_ensemble* e);
/* flow-rate statistics: with enable, every following step adds each
 * replica's Qx = Ly * mean cell Ux of the state after the step (the
 * driver's Qx, :105-155) to qSum and its square to qSumSq (cleared when
 * enabled); the call first copies the sums so far (any pointer may be
 * NULL; qSum/qSumSq have nReplicas entries) and the step count */
int mpcd_gpu_ensemble_flow(mpcd_gpu_ensemble* e, int enable, double* qSum, double* qSumSq, int32_t* nSteps);

/* one 16x16x32 bfloat16 matrix-core product of 64 lanes x 8 operands each
 * (a, b: 512 raw bfloat16 values, d: 256 floats), to check the lane layout */
int mpcd_gpu_net_mfma_probe(const uint16_t* a, const uint16_t* b, float* d);

#ifdef __cplusplus
}
#endif
