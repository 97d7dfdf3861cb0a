/*
 * host_diag.cpp -- occupancy statistics of a zone pass from cell counts
 * (compute_occ_std_from_counts, compute_out_band_from_counts,
 * count_dense_from_counts, count_sparse_from_counts, count_empty_active:
 * zone_pass.cpp:975-1044), on the host: they read Nc counts once per pass.
 */
#include <hip/hip_runtime.h> /* the library's sources are compiled as HIP; the shared headers use its intrinsics */

#include <cmath>
#include <cstdint>
#include <vector>

#include "mpcd_gpu.h"
#include "mpcd_host.h"

using namespace mpcdk;

extern "C" int mpcd_gpu_occupancy_stats(const mpcd_params* p, const int32_t* counts, mpcd_occupancy_stats* out) {
    if (!p || !counts || !out || p->Nc != p->Nx * p->Ny || p->Nc <= 0) return MPCD_E_INVALID;
    const ZoneParams zp = make_zone_params(*p);
    const int Nx = p->Nx, Ny = p->Ny, Nc = p->Nc;
    /* every cell is active (zone_pass.cpp:128-130) */
    double mean = 0.0;
    for (int c = 0; c < Nc; ++c) mean += (double)counts[c];
    mean /= (double)Nc;
    double var = 0.0;
    for (int c = 0; c < Nc; ++c) {
        const double d = (double)counts[c] - mean;
        var += d * d;
    }
    out->occStd = std::sqrt(var / (double)Nc) / dmax(p->gamma, 1e-30);
    int nOut = 0, nDense = 0, nSparse = 0, nEmpty = 0;
    for (int ix = 0; ix < Nx; ++ix)
        for (int iy = 0; iy < Ny; ++iy) {
            const int c = iy + Ny * ix;
            const int m = counts[c];
            /* empty-cell thresholds from the 3x3 binomial smoothing (no wrap) */
            double lowLoc = zp.lowLoc1, highLoc = zp.highLoc1;
            if (zp.useLocal) {
                double acc = 0.0, ws = 0.0;
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int jy = iy + dy, jx = ix + dx;
                        if (jy < 0 || jy >= Ny || jx < 0 || jx >= Nx) continue;
                        const double w = (double)((dx == 0 ? 2 : 1) * (dy == 0 ? 2 : 1)) / 16.0;
                        acc += w * (double)counts[jy + Ny * jx];
                        ws += w;
                    }
                const double t = zp.fluidFracGain * ((acc / dmax(ws, 1e-30)) / dmax(zp.gamma, 1e-12));
                const double gl = zp.gamma * ((t < 0.0) ? 0.0 : ((1.0 < t) ? 1.0 : t));
                lowLoc = dmax(zp.lowFrac * gl, zp.lowThrFloor);
                highLoc = dmax(zp.highFrac * gl, zp.highThrFloor);
            }
            const bool bulk = m > 0;
            const double lo = bulk ? zp.lowThrBulk : lowLoc, hi = bulk ? zp.highThr : highLoc;
            if ((double)m > hi || (m > 0 && (double)m < lo)) ++nOut;
            if ((double)m > hi) ++nDense;
            if (m > 0 && (double)m < lo) ++nSparse;
            if (m == 0) ++nEmpty;
        }
    out->outBand = (double)nOut / (double)Nc;
    out->nDense = nDense;
    out->nSparse = nSparse;
    out->nEmptyActive = nEmpty;
    out->_pad = 0;
    return MPCD_OK;
}
