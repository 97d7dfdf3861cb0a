/*
 * mpcd_host.h -- host-side set-up shared by the engine: geometry, kernel
 * constants, zone layouts and the interzone mask, derived from mpcd_params.
 */
#pragma once

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "mpcd_core.h"
#include "mpcd_zone.h"

namespace mpcdk {

struct ApiError : std::runtime_error {
    int code;
    ApiError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline std::vector<int> start_indices(int N, int tile, int offset) {
    std::vector<int> s{0};
    for (int k = offset; k < N; k += tile)
        if (k > 0 && k < N) s.push_back(k);
    std::sort(s.begin(), s.end());
    s.erase(std::unique(s.begin(), s.end()), s.end());
    return s;
}

/* zone_pass.cpp:173-221: zones as rectangles in id order, plus owner map */
inline std::vector<ZoneRect> layout_rects(const mpcd_params& p, int mode, std::vector<int32_t>* owner) {
    std::vector<ZoneRect> z;
    if (owner) owner->assign((size_t)p.Nc, 0);
    if (mode == MPCD_LAYOUT_SINGLE_FULL || p.zoneSingleFullDomainMode) {
        z.push_back(ZoneRect{0, p.Nx, 0, p.Ny});
        if (owner) std::fill(owner->begin(), owner->end(), 1);
        return z;
    }
    if (p.zoneTileNx <= 0 || p.zoneTileNy <= 0) throw ApiError(MPCD_E_INVALID, "zone tiles must be positive");
    const int offx = (mode == MPCD_LAYOUT_SHIFTED) ? p.zoneTileNx / 2 : 0;
    const int offy = (mode == MPCD_LAYOUT_SHIFTED) ? p.zoneTileNy / 2 : 0;
    const auto xs = start_indices(p.Nx, p.zoneTileNx, offx);
    const auto ys = start_indices(p.Ny, p.zoneTileNy, offy);
    int id = 0;
    for (int y0 : ys)
        for (int x0 : xs) {
            const ZoneRect r{x0, std::min(x0 + p.zoneTileNx, p.Nx), y0, std::min(y0 + p.zoneTileNy, p.Ny)};
            z.push_back(r);
            ++id;
            if (owner)
                for (int ix = r.xs; ix < r.xe; ++ix)
                    for (int iy = r.ys; iy < r.ye; ++iy) (*owner)[(size_t)(iy + p.Ny * ix)] = id;
        }
    return z;
}

/* interzone mask (liquid_closure.cpp:143-213) */
inline std::vector<int32_t> interzone_mask_host(const mpcd_params& p) {
    std::vector<int32_t> m((size_t)p.Nc, 0);
    if (!p.smoothPdriveAtInterzone || !p.useZoneRedistribution || p.zonePassthroughMode || p.zoneSingleFullDomainMode)
        return m;
    const bool px = p.bc_left == MPCD_BC_PERIODIC && p.bc_right == MPCD_BC_PERIODIC;
    const bool py = p.bc_bottom == MPCD_BC_PERIODIC && p.bc_top == MPCD_BC_PERIODIC;
    auto boundary = [&](const std::vector<int32_t>& own) {
        std::vector<int32_t> b((size_t)p.Nc, 0);
        for (int ix = 0; ix < p.Nx; ++ix)
            for (int iy = 0; iy < p.Ny; ++iy) {
                const int c = iy + p.Ny * ix;
                const int right = (ix + 1 < p.Nx) ? iy + p.Ny * (ix + 1) : (px ? iy : -1);
                const int up = (iy + 1 < p.Ny) ? (iy + 1) + p.Ny * ix : (py ? p.Ny * ix : -1);
                if ((right >= 0 && own[(size_t)right] != own[(size_t)c]) || (up >= 0 && own[(size_t)up] != own[(size_t)c]))
                    b[(size_t)c] = 1;
            }
        return b;
    };
    auto dilate = [&](const std::vector<int32_t>& in, int w) {
        if (w <= 0) return in;
        std::vector<int32_t> out((size_t)p.Nc, 0);
        for (int ix = 0; ix < p.Nx; ++ix)
            for (int iy = 0; iy < p.Ny; ++iy) {
                bool on = false;
                for (int dx = -w; dx <= w && !on; ++dx)
                    for (int dy = -w; dy <= w; ++dy) {
                        int jx = ix + dx, jy = iy + dy;
                        if (px) jx = wrapi(jx, p.Nx);
                        else if (jx < 0 || jx >= p.Nx) continue;
                        if (py) jy = wrapi(jy, p.Ny);
                        else if (jy < 0 || jy >= p.Ny) continue;
                        if (in[(size_t)(jy + p.Ny * jx)] != 0) {
                            on = true;
                            break;
                        }
                    }
                if (on) out[(size_t)(iy + p.Ny * ix)] = 1;
            }
        return out;
    };
    std::vector<int32_t> own;
    layout_rects(p, MPCD_LAYOUT_BASE, &own);
    m = dilate(boundary(own), p.smoothPdriveInterzoneWidthCells);
    if (p.smoothPdriveIncludeShiftedLayouts) {
        layout_rects(p, MPCD_LAYOUT_SHIFTED, &own);
        const auto s = dilate(boundary(own), p.smoothPdriveInterzoneWidthCells);
        for (size_t c = 0; c < (size_t)p.Nc; ++c) m[c] = (m[c] || s[c]) ? 1 : 0;
    }
    return m;
}

inline Geom make_geom(const mpcd_params& p) {
    Geom g;
    g.Nx = p.Nx;
    g.Ny = p.Ny;
    g.Nc = p.Nc;
    g.n = p.n;
    g.perx = p.bc_left == MPCD_BC_PERIODIC && p.bc_right == MPCD_BC_PERIODIC;
    g.pery = p.bc_bottom == MPCD_BC_PERIODIC && p.bc_top == MPCD_BC_PERIODIC;
    g.Lx = p.Lx;
    g.Ly = p.Ly;
    g.a0 = p.a0;
    g.dxc = p.Lx / (double)p.Nx;
    g.dyc = p.Ly / (double)p.Ny;
    return g;
}

inline void validate(const mpcd_params& p) {
    if (p.Nx <= 0 || p.Ny <= 0 || p.n < 0) throw ApiError(MPCD_E_INVALID, "invalid grid or particle count");
    if ((long long)p.Nx * (long long)p.Ny != (long long)p.Nc)
        throw ApiError(MPCD_E_INVALID, "Nc must equal Nx*Ny");
    if ((long long)p.Nc >= (1LL << 30)) throw ApiError(MPCD_E_INVALID, "too many cells");
}


inline ZoneParams make_zone_params(const mpcd_params& p) {
    ZoneParams zp{};
    zp.highThr = p.high_mode_one_plus_coef ? p.gamma * (1.0 + p.coef) : p.gamma + p.gamma * p.coef;
    zp.lowThrBulk = p.low_mode_one_plus_coef ? p.gamma / (1.0 + p.coef) : p.gamma - p.gamma * p.coef;
    if (std::isfinite(p.lowThrBulkOverride)) zp.lowThrBulk = p.lowThrBulkOverride;
    zp.lowFrac = 1.0 / (1.0 + p.coef);
    zp.highFrac = 1.0 + p.coef;
    zp.gamma = p.gamma;
    zp.fluidFracGain = p.fluidFracGain;
    zp.lowThrFloor = p.lowThrFloor;
    zp.highThrFloor = p.highThrFloor;
    zp.useLocal = p.useLocalFluidFractionThresholds;
    zp.maxPasses = p.maxRedistribPasses;
    /* a particle's move record holds the dense pass that last moved it in a byte */
    if (p.maxRedistribPasses > 254) throw ApiError(MPCD_E_INVALID, "maxRedistribPasses above 254 is not supported");
    zp.corrEnabled = p.enableMomentumCorrectionPostRedistribution;
    zp.dedupe = ((p.bc_left == MPCD_BC_PERIODIC && p.bc_right == MPCD_BC_PERIODIC && p.Nx < 5) ||
                 (p.bc_bottom == MPCD_BC_PERIODIC && p.bc_top == MPCD_BC_PERIODIC && p.Ny < 5))
                    ? 1
                    : 0;
    { /* empty-cell thresholds at fluid fraction 1 (zone_pass.cpp:334-342) */
        const double gl = p.gamma * 1.0;
        const double lo = zp.lowFrac * gl, hi = zp.highFrac * gl;
        zp.lowLoc1 = (lo < p.lowThrFloor) ? p.lowThrFloor : lo;
        zp.highLoc1 = (hi < p.highThrFloor) ? p.highThrFloor : hi;
    }
    /* angular bins of rank_dense_targets_unified (zone_pass.cpp:452-454) */
    for (int dy = -2; dy <= 2; ++dy)
        for (int dx = -2; dx <= 2; ++dx) {
            const double ang = std::atan2((double)dy, (double)dx);
            int ib = (int)std::floor((ang + M_PI) / (2.0 * M_PI / 8.0));
            ib = ((ib % 8) + 8) % 8;
            const int idx = (dy + 2) * 5 + (dx + 2);
            if (idx < 21) zp.binLo |= (unsigned long long)ib << (3 * idx);
            else zp.binHi |= (unsigned long long)ib << (3 * (idx - 21));
        }
    return zp;
}

inline CollideParams make_collide_params(const mpcd_params& p) {
    CollideParams cp{};
    cp.kBT = p.kBT;
    cp.thermostat = p.useThermostat;
    const double th = p.alpha * 1.0, thm = p.alpha * -1.0;
    cp.rot[0] = std::cos(th);
    cp.rot[1] = std::sin(th);
    cp.rot[2] = std::cos(thm);
    cp.rot[3] = std::sin(thm);
    return cp;
}

inline BCParams make_bc_params(const mpcd_params& p) {
    BCParams bc{};
    bc.dt = p.dt;
    bc.Fx = p.bodyForceX;
    bc.Fy = p.g;
    bc.Utop = p.Utop;
    bc.Ubottom = p.Ubottom;
    bc.sigma = p.wallSigma;
    bc.left = p.bc_left;
    bc.right = p.bc_right;
    bc.bottom = p.bc_bottom;
    bc.top = p.bc_top;
    return bc;
}

}  // namespace mpcdk
