/*
 * oracle.cpp -- CPU restatement of the reference SRD/MPCD step (TEST
 * INFRASTRUCTURE ONLY: the checker for the GPU library, never part of it).
 * Each function cites the reference file:line it restates.
 */
#include "oracle.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <unordered_set>

namespace orc {

/* ================= wall BC: boundary_conditions.cpp:19-203 ================= */

template <class R>
void apply_bc(const P& p, std::vector<double>& x, std::vector<double>& v, mpcd_wall_info& info, R& rng) {
    const double Lx = p.Lx, Ly = p.Ly, s = p.wallSigma;
    for (int k = 0; k < 4; ++k) {
        for (int i = 0; i < p.n; ++i) {
            double& xi = x[2 * (size_t)i];
            double& yi = x[2 * (size_t)i + 1];
            double& vx = v[2 * (size_t)i];
            double& vy = v[2 * (size_t)i + 1];
            if (xi < 0.0) { /* :44-69 */
                const double ox = vx, oy = vy;
                if (p.bc_left == MPCD_BC_PERIODIC) {
                    xi += Lx;
                } else if (p.bc_left >= MPCD_BC_THERMALIZE && p.bc_left <= MPCD_BC_SPECULAR) {
                    xi = -xi;
                    if (p.bc_left == MPCD_BC_THERMALIZE) {
                        vx = std::abs(s * rng.first(i, k, 0));
                        vy = s * rng.second(i, k, 0);
                    } else if (p.bc_left == MPCD_BC_BOUNCEBACK) {
                        vx = std::abs(vx);
                        vy = -vy;
                    } else {
                        vx = std::abs(vx);
                    }
                    info.nLeft += 1;
                    info.dEwall += 0.5 * ((vx * vx + vy * vy) - (ox * ox + oy * oy));
                    info.dPxLeft += std::abs(vx - ox);
                    info.dPyLeft += vy - oy;
                } else {
                    throw std::runtime_error("Unsupported boundary_left mode");
                }
            }
            if (xi > Lx) { /* :71-96 */
                const double ox = vx, oy = vy;
                if (p.bc_right == MPCD_BC_PERIODIC) {
                    xi -= Lx;
                } else if (p.bc_right >= MPCD_BC_THERMALIZE && p.bc_right <= MPCD_BC_SPECULAR) {
                    xi = 2.0 * Lx - xi;
                    if (p.bc_right == MPCD_BC_THERMALIZE) {
                        vx = -std::abs(s * rng.first(i, k, 1));
                        vy = s * rng.second(i, k, 1);
                    } else if (p.bc_right == MPCD_BC_BOUNCEBACK) {
                        vx = -std::abs(vx);
                        vy = -vy;
                    } else {
                        vx = -std::abs(vx);
                    }
                    info.nRight += 1;
                    info.dEwall += 0.5 * ((vx * vx + vy * vy) - (ox * ox + oy * oy));
                    info.dPxRight += std::abs(vx - ox);
                    info.dPyRight += vy - oy;
                } else {
                    throw std::runtime_error("Unsupported boundary_right mode");
                }
            }
            if (yi < 0.0) { /* :98-135 */
                const double ox = vx, oy = vy;
                if (p.bc_bottom == MPCD_BC_PERIODIC) {
                    yi += Ly;
                } else if (p.bc_bottom >= MPCD_BC_THERMALIZE && p.bc_bottom <= MPCD_BC_SPECULAR) {
                    yi = -yi;
                    if (p.bc_bottom == MPCD_BC_THERMALIZE) {
                        vx = p.Ubottom + s * rng.first(i, k, 2);
                        vy = std::abs(s * rng.second(i, k, 2));
                    } else if (p.bc_bottom == MPCD_BC_BOUNCEBACK) {
                        vx = 2.0 * p.Ubottom - vx;
                        vy = std::abs(vy);
                    } else {
                        vy = std::abs(vy);
                    }
                    info.nBot += 1;
                    info.dEwall += 0.5 * ((vx * vx + vy * vy) - (ox * ox + oy * oy));
                    info.dPxBot += vx - ox;
                    info.dPyBot += std::abs(vy - oy);
                } else {
                    throw std::runtime_error("Unsupported boundary_bottom mode");
                }
            }
            if (yi > Ly) { /* :137-174 */
                const double ox = vx, oy = vy;
                if (p.bc_top == MPCD_BC_PERIODIC) {
                    yi -= Ly;
                } else if (p.bc_top >= MPCD_BC_THERMALIZE && p.bc_top <= MPCD_BC_SPECULAR) {
                    yi = 2.0 * Ly - yi;
                    if (p.bc_top == MPCD_BC_THERMALIZE) {
                        vx = p.Utop + s * rng.first(i, k, 3);
                        vy = -std::abs(s * rng.second(i, k, 3));
                    } else if (p.bc_top == MPCD_BC_BOUNCEBACK) {
                        vx = 2.0 * p.Utop - vx;
                        vy = -std::abs(vy);
                    } else {
                        vy = -std::abs(vy);
                    }
                    info.nTop += 1;
                    info.dEwall += 0.5 * ((vx * vx + vy * vy) - (ox * ox + oy * oy));
                    info.dPxTop += vx - ox;
                    info.dPyTop += std::abs(vy - oy);
                } else {
                    throw std::runtime_error("Unsupported boundary_top mode");
                }
            }
        }
    }
    /* final wrap / clamp :178-202 */
    for (int i = 0; i < p.n; ++i) {
        double& xi = x[2 * (size_t)i];
        if (per_x(p)) {
            xi = xi - std::floor(xi / p.Lx) * p.Lx;
        } else {
            if (xi < 0.0) xi = 0.0;
            if (xi > p.Lx) xi = p.Lx;
        }
    }
    for (int i = 0; i < p.n; ++i) {
        double& yi = x[2 * (size_t)i + 1];
        if (per_y(p)) {
            yi = yi - std::floor(yi / p.Ly) * p.Ly;
        } else {
            if (yi < 0.0) yi = 0.0;
            if (yi > p.Ly) yi = p.Ly;
        }
    }
}

template void apply_bc<StdRngBC>(const P&, std::vector<double>&, std::vector<double>&, mpcd_wall_info&, StdRngBC&);
template void apply_bc<CtrRngBC>(const P&, std::vector<double>&, std::vector<double>&, mpcd_wall_info&, CtrRngBC&);

/* ============ shifted binning: srd_collision.cpp:9-45 ============ */

void shift_cell_ids(const P& p, const std::vector<double>& x, double sx, double sy, std::vector<int>& cid) {
    cid.assign((size_t)p.n, 0);
    for (int i = 0; i < p.n; ++i) {
        double xs = x[2 * (size_t)i] + sx;
        double ys = x[2 * (size_t)i + 1] + sy;
        if (per_x(p)) {
            xs = xs - std::floor(xs / p.Lx) * p.Lx;
        } else {
            if (xs < 0.0) xs = 0.0;
            if (xs >= p.Lx) xs = std::nextafter(p.Lx, 0.0);
        }
        if (per_y(p)) {
            ys = ys - std::floor(ys / p.Ly) * p.Ly;
        } else {
            if (ys < 0.0) ys = 0.0;
            if (ys >= p.Ly) ys = std::nextafter(p.Ly, 0.0);
        }
        int ix = (int)std::floor(xs / p.a0);
        int iy = (int)std::floor(ys / p.a0);
        ix = std::min(std::max(ix, 0), p.Nx - 1);
        iy = std::min(std::max(iy, 0), p.Ny - 1);
        cid[(size_t)i] = iy + p.Ny * ix;
    }
}

/* ============ SRD collision: srd_collision.cpp:47-140 ============ */

template <class R>
void collide(const P& p, std::vector<double>& v, const std::vector<int>& cid, R& rng) {
    if ((int)cid.size() != p.n) throw std::runtime_error("srd_collision_step: unexpected cid size");
    const size_t Nc = (size_t)p.Nc;
    std::vector<int> cnt(Nc, 0);
    std::vector<double> sx(Nc, 0.0), sy(Nc, 0.0);
    for (int i = 0; i < p.n; ++i) {
        const size_t c = (size_t)cid[(size_t)i];
        cnt[c] += 1;
        sx[c] += v[2 * (size_t)i];
        sy[c] += v[2 * (size_t)i + 1];
    }
    std::vector<double> ux(Nc, 0.0), uy(Nc, 0.0), th(Nc, 0.0);
    for (size_t c = 0; c < Nc; ++c) {
        if (cnt[c] > 0) {
            ux[c] = sx[c] / (double)cnt[c];
            uy[c] = sy[c] / (double)cnt[c];
        }
    }
    for (size_t c = 0; c < Nc; ++c) {
        if (cnt[c] >= 2) th[c] = p.alpha * rng.sign((int)c);
    }
    std::vector<double> rx((size_t)p.n), ry((size_t)p.n), r2((size_t)p.n);
    for (int i = 0; i < p.n; ++i) {
        const size_t c = (size_t)cid[(size_t)i];
        const double vxr = v[2 * (size_t)i] - ux[c];
        const double vyr = v[2 * (size_t)i + 1] - uy[c];
        const double ct = std::cos(th[c]);
        const double st = std::sin(th[c]);
        rx[(size_t)i] = ct * vxr - st * vyr;
        ry[(size_t)i] = st * vxr + ct * vyr;
        r2[(size_t)i] = rx[(size_t)i] * rx[(size_t)i] + ry[(size_t)i] * ry[(size_t)i];
    }
    if (p.useThermostat) {
        std::vector<double> s2(Nc, 0.0), lam(Nc, 1.0);
        for (int i = 0; i < p.n; ++i) s2[(size_t)cid[(size_t)i]] += r2[(size_t)i];
        for (size_t c = 0; c < Nc; ++c) {
            const int dof = 2 * std::max(cnt[c] - 1, 0);
            if (cnt[c] > 1 && s2[c] > 0.0) lam[c] = std::sqrt((double)dof * p.kBT / s2[c]);
        }
        for (int i = 0; i < p.n; ++i) {
            const double l = lam[(size_t)cid[(size_t)i]];
            rx[(size_t)i] *= l;
            ry[(size_t)i] *= l;
        }
    }
    for (int i = 0; i < p.n; ++i) {
        const size_t c = (size_t)cid[(size_t)i];
        v[2 * (size_t)i] = ux[c] + rx[(size_t)i];
        v[2 * (size_t)i + 1] = uy[c] + ry[(size_t)i];
    }
    if (p.keepMeanFlow) {
        double mx = 0.0, my = 0.0;
        for (int i = 0; i < p.n; ++i) {
            mx += v[2 * (size_t)i];
            my += v[2 * (size_t)i + 1];
        }
        mx /= (double)p.n;
        my /= (double)p.n;
        for (int i = 0; i < p.n; ++i) {
            v[2 * (size_t)i] -= mx;
            v[2 * (size_t)i + 1] -= my;
        }
    }
}

template void collide<StdRngRot>(const P&, std::vector<double>&, const std::vector<int>&, StdRngRot&);
template void collide<CtrRngRot>(const P&, std::vector<double>&, const std::vector<int>&, CtrRngRot&);

/* ============ cell grid / fields: common_grid.cpp:9-100 ============ */

int cell_of(const P& p, double x, double y) {
    const double dx = p.Lx / (double)p.Nx;
    const double dy = p.Ly / (double)p.Ny;
    int ix = (int)std::floor(x / dx);
    int iy = (int)std::floor(y / dy);
    ix = std::min(std::max(ix, 0), p.Nx - 1);
    iy = std::min(std::max(iy, 0), p.Ny - 1);
    return iy + p.Ny * ix;
}

static std::vector<int> cell_ids(const P& p, const std::vector<double>& x) {
    std::vector<int> cid((size_t)p.n);
    for (int i = 0; i < p.n; ++i) cid[(size_t)i] = cell_of(p, x[2 * (size_t)i], x[2 * (size_t)i + 1]);
    return cid;
}

double rho_target(const P& p) {
    const double Vc = (p.Lx / (double)p.Nx) * (p.Ly / (double)p.Ny);
    return p.gamma / Vc;
}

Fields cell_fields(const P& p, const std::vector<double>& x, const std::vector<double>& v, double rhoT) {
    const size_t Nc = (size_t)p.Nc;
    const double Vc = (p.Lx / (double)p.Nx) * (p.Ly / (double)p.Ny);
    Fields G;
    G.Nx = p.Nx;
    G.Ny = p.Ny;
    G.N.assign(Nc, 0);
    G.Ux.assign(Nc, 0.0);
    G.Uy.assign(Nc, 0.0);
    G.Px.assign(Nc, 0.0);
    G.Py.assign(Nc, 0.0);
    G.T.assign(Nc, std::numeric_limits<double>::quiet_NaN());
    G.rho.assign(Nc, 0.0);
    G.Pkin.assign(Nc, 0.0);
    G.Pvir.assign(Nc, 0.0);
    G.P.assign(Nc, 0.0);
    G.rhoTarget.assign(Nc, rhoT);
    const auto cid = cell_ids(p, x);
    for (int i = 0; i < p.n; ++i) {
        const size_t c = (size_t)cid[(size_t)i];
        G.N[c] += 1;
        G.Px[c] += v[2 * (size_t)i];
        G.Py[c] += v[2 * (size_t)i + 1];
    }
    for (size_t c = 0; c < Nc; ++c) {
        if (G.N[c] > 0) {
            G.Ux[c] = G.Px[c] / (double)G.N[c];
            G.Uy[c] = G.Py[c] / (double)G.N[c];
        }
        G.rho[c] = (double)G.N[c] / Vc;
    }
    std::vector<double> s2(Nc, 0.0);
    for (int i = 0; i < p.n; ++i) {
        const size_t c = (size_t)cid[(size_t)i];
        const double dvx = v[2 * (size_t)i] - G.Ux[c];
        const double dvy = v[2 * (size_t)i + 1] - G.Uy[c];
        s2[c] += dvx * dvx + dvy * dvy;
    }
    for (size_t c = 0; c < Nc; ++c) {
        const int dof = 2 * std::max(G.N[c] - 1, 0);
        if (dof > 0) {
            G.T[c] = s2[c] / (double)dof;
            G.Pkin[c] = G.rho[c] * G.T[c];
        } else {
            G.T[c] = 0.0;
            G.Pkin[c] = 0.0;
        }
        G.Pvir[c] = p.Kvirial * (G.rho[c] - G.rhoTarget[c]);
        G.P[c] = G.Pkin[c] + G.Pvir[c];
    }
    return G;
}

/* ============ zone layout: zone_pass.cpp:160-221 ============ */

static std::vector<int> start_indices(int N, int tile, int offset) {
    std::vector<int> s{0};
    for (int k = offset; k < N; k += tile)
        if (k > 0 && k < N) s.push_back(k);
    std::sort(s.begin(), s.end());
    s.erase(std::unique(s.begin(), s.end()), s.end());
    return s;
}

Layout build_layout(const P& p, int mode) {
    Layout L;
    L.owner.assign((size_t)p.Nc, 0);
    if (mode == MPCD_LAYOUT_SINGLE_FULL || p.zoneSingleFullDomainMode) {
        Zone z;
        z.id = 1;
        z.x0 = 0;
        z.x1 = p.Nx;
        z.y0 = 0;
        z.y1 = p.Ny;
        for (int c = 0; c < p.Nc; ++c) {
            z.cells.push_back(c);
            L.owner[(size_t)c] = 1;
        }
        L.zones.push_back(std::move(z));
        return L;
    }
    const int offx = (mode == MPCD_LAYOUT_SHIFTED) ? p.zoneTileNx / 2 : 0;
    const int offy = (mode == MPCD_LAYOUT_SHIFTED) ? p.zoneTileNy / 2 : 0;
    const auto xs = start_indices(p.Nx, p.zoneTileNx, offx);
    const auto ys = start_indices(p.Ny, p.zoneTileNy, offy);
    int id = 0;
    for (int y0 : ys) {
        for (int x0 : xs) {
            const int x1 = std::min(x0 + p.zoneTileNx, p.Nx);
            const int y1 = std::min(y0 + p.zoneTileNy, p.Ny);
            Zone z;
            z.id = ++id;
            z.x0 = x0;
            z.x1 = x1;
            z.y0 = y0;
            z.y1 = y1;
            for (int ix = x0; ix < x1; ++ix)
                for (int iy = y0; iy < y1; ++iy) {
                    const int c = iy + p.Ny * ix;
                    z.cells.push_back(c); /* ascending cell id */
                    L.owner[(size_t)c] = z.id;
                }
            L.zones.push_back(std::move(z));
        }
    }
    return L;
}

/* ============ zone kernel helpers: zone_pass.cpp:62-667 ============ */

namespace {

inline int wrapi(int i, int n) {
    const int r = i % n;
    return (r < 0) ? r + n : r;
}

std::vector<int> unique_stable(const std::vector<int>& in) {
    std::unordered_set<int> seen;
    std::vector<int> out;
    for (int a : in)
        if (seen.insert(a).second) out.push_back(a);
    return out;
}

/* zone_pass.cpp:229-244 */
void minimal_step(const P& p, int src, int dst, int& ddx, int& ddy) {
    const int sx = src / p.Ny, sy = src - sx * p.Ny;
    const int tx = dst / p.Ny, ty = dst - tx * p.Ny;
    ddx = tx - sx;
    ddy = ty - sy;
    if (per_x(p)) {
        if (ddx > p.Nx / 2) ddx -= p.Nx;
        if (ddx < -p.Nx / 2) ddx += p.Nx;
    }
    if (per_y(p)) {
        if (ddy > p.Ny / 2) ddy -= p.Ny;
        if (ddy < -p.Ny / 2) ddy += p.Ny;
    }
}

/* zone_pass.cpp:246-261 */
bool shift_index(const P& p, int iy, int ix, int dy, int dx, int& jy, int& jx) {
    jy = iy + dy;
    if (per_y(p)) jy = wrapi(jy, p.Ny);
    else if (jy < 0 || jy >= p.Ny) return false;
    jx = ix + dx;
    if (per_x(p)) jx = wrapi(jx, p.Nx);
    else if (jx < 0 || jx >= p.Nx) return false;
    return true;
}

/* zone_pass.cpp:296-319 (no wrap; weights renormalised at the edges) */
std::vector<double> smooth_counts(const P& p, const std::vector<int>& cnt) {
    static const double K[9] = {1.0 / 16.0, 2.0 / 16.0, 1.0 / 16.0, 2.0 / 16.0, 4.0 / 16.0,
                                2.0 / 16.0, 1.0 / 16.0, 2.0 / 16.0, 1.0 / 16.0};
    std::vector<double> out((size_t)p.Nc, 0.0);
    for (int ix = 0; ix < p.Nx; ++ix)
        for (int iy = 0; iy < p.Ny; ++iy) {
            double acc = 0.0, ws = 0.0;
            int kk = 0;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx, ++kk) {
                    const int jy = iy + dy, jx = ix + dx;
                    if (jy < 0 || jy >= p.Ny || jx < 0 || jx >= p.Nx) continue;
                    acc += K[kk] * (double)cnt[(size_t)(jy + p.Ny * jx)];
                    ws += K[kk];
                }
            out[(size_t)(iy + p.Ny * ix)] = acc / std::max(ws, 1e-30);
        }
    return out;
}

/* zone_pass.cpp:321-345 */
void threshold_maps(const P& p, const std::vector<double>& Nsm, std::vector<double>& lowLoc,
                    std::vector<double>& highLoc) {
    const double lowFrac = 1.0 / (1.0 + p.coef);
    const double highFrac = 1.0 + p.coef;
    lowLoc.assign((size_t)p.Nc, 0.0);
    highLoc.assign((size_t)p.Nc, 0.0);
    for (size_t c = 0; c < (size_t)p.Nc; ++c) {
        double f = 1.0;
        if (p.useLocalFluidFractionThresholds) {
            const double fRaw = Nsm[c] / std::max(p.gamma, 1e-12);
            f = std::clamp(p.fluidFracGain * fRaw, 0.0, 1.0);
        }
        const double gl = p.gamma * f;
        lowLoc[c] = std::max(lowFrac * gl, p.lowThrFloor);
        highLoc[c] = std::max(highFrac * gl, p.highThrFloor);
    }
}

/* zone_pass.cpp:347-366 */
void occupancy_thresholds(const P& p, double& highThr, double& lowThrBulk) {
    highThr = p.high_mode_one_plus_coef ? p.gamma * (1.0 + p.coef) : p.gamma + p.gamma * p.coef;
    lowThrBulk = p.low_mode_one_plus_coef ? p.gamma / (1.0 + p.coef) : p.gamma - p.gamma * p.coef;
    if (std::isfinite(p.lowThrBulkOverride)) lowThrBulk = p.lowThrBulkOverride;
}

/* zone_pass.cpp:375-411 */
void inward_dir(const P& p, int cell, const std::vector<double>& Nsm, double& gx, double& gy) {
    const int ix = cell / p.Ny, iy = cell - ix * p.Ny;
    auto at = [&](int jy, int jx) { return Nsm[(size_t)(jy + p.Ny * jx)]; };
    if (per_x(p)) gx = 0.5 * (at(iy, wrapi(ix + 1, p.Nx)) - at(iy, wrapi(ix - 1, p.Nx)));
    else if (ix == 0) gx = at(iy, ix + 1) - at(iy, ix);
    else if (ix == p.Nx - 1) gx = at(iy, ix) - at(iy, ix - 1);
    else gx = 0.5 * (at(iy, ix + 1) - at(iy, ix - 1));
    if (per_y(p)) gy = 0.5 * (at(wrapi(iy + 1, p.Ny), ix) - at(wrapi(iy - 1, p.Ny), ix));
    else if (iy == 0) gy = at(iy + 1, ix) - at(iy, ix);
    else if (iy == p.Ny - 1) gy = at(iy, ix) - at(iy - 1, ix);
    else gy = 0.5 * (at(iy + 1, ix) - at(iy - 1, ix));
}

/* zone_pass.cpp:413-434: neighbours within Chebyshev range R whose count is
 * below the target map (dy outer, dx inner), de-duplicated. */
std::vector<int> admissible(const P& p, int cell, const std::vector<int>& cnt, const std::vector<double>& target,
                            int R) {
    const int ix = cell / p.Ny, iy = cell - ix * p.Ny;
    std::vector<int> lst;
    for (int dy = -R; dy <= R; ++dy)
        for (int dx = -R; dx <= R; ++dx) {
            if (dx == 0 && dy == 0) continue;
            int jy, jx;
            if (!shift_index(p, iy, ix, dy, dx, jy, jx)) continue;
            const int nb = jy + p.Ny * jx;
            if ((double)cnt[(size_t)nb] < target[(size_t)nb]) lst.push_back(nb);
        }
    return unique_stable(lst);
}

struct Cand {
    int nb;
    double deficitNeg, inwardNeg;
    int cnt, cheb, dist2;
    double tie;
};

/* zone_pass.cpp:436-490: 8 angular bins, per-bin sort, round robin. */
template <class R>
std::vector<int> rank_dense(const P& p, int cell, const std::vector<int>& adm, const std::vector<int>& cnt,
                            const std::vector<double>& target, int pass, int phase, R& rng) {
    if (adm.empty()) return {};
    constexpr int nB = 8;
    std::array<std::vector<Cand>, nB> bins;
    const int startBin = rng.start_bin(pass, phase, cell);
    for (int nb : adm) {
        int dx, dy;
        minimal_step(p, cell, nb, dx, dy);
        const double ang = std::atan2((double)dy, (double)dx);
        int ib = (int)std::floor((ang + M_PI) / (2.0 * M_PI / (double)nB));
        ib = ((ib % nB) + nB) % nB;
        const double deficit = std::max(0.0, target[(size_t)nb] - (double)cnt[(size_t)nb]);
        const double inwardWeight = 0.0, inward = 0.0; /* caller passes nHat = 0, weight 0 (:740-741) */
        bins[(size_t)ib].push_back(Cand{nb, -deficit, -inwardWeight * inward, cnt[(size_t)nb],
                                        std::max(std::abs(dx), std::abs(dy)), dx * dx + dy * dy,
                                        rng.tie(pass, phase, cell, nb)});
    }
    auto less = [](const Cand& a, const Cand& b) {
        if (a.deficitNeg != b.deficitNeg) return a.deficitNeg < b.deficitNeg;
        if (a.inwardNeg != b.inwardNeg) return a.inwardNeg < b.inwardNeg;
        if (a.cnt != b.cnt) return a.cnt < b.cnt;
        if (a.cheb != b.cheb) return a.cheb < b.cheb;
        if (a.dist2 != b.dist2) return a.dist2 < b.dist2;
        return a.tie < b.tie;
    };
    for (auto& b : bins) {
        if (R::stable_sorts) std::stable_sort(b.begin(), b.end(), less);
        else std::sort(b.begin(), b.end(), less);
    }
    std::vector<int> ordered;
    std::array<size_t, nB> head{};
    bool more = true;
    while (more) {
        more = false;
        for (int s = 0; s < nB; ++s) {
            const int ib = (startBin + s) % nB;
            if (head[(size_t)ib] >= bins[(size_t)ib].size()) continue;
            ordered.push_back(bins[(size_t)ib][head[(size_t)ib]++].nb);
            more = true;
        }
    }
    return ordered;
}

/* zone_pass.cpp:492-525 */
void transfer_masks(std::vector<char>& bb, std::vector<char>& skip, int src, const std::vector<int>& recv,
                    const std::vector<int>& bulkBefore) {
    std::vector<int> touched{src};
    touched.insert(touched.end(), recv.begin(), recv.end());
    touched = unique_stable(touched);
    bool allBulk = true;
    for (int c : touched)
        if (c < 0 || c >= (int)bulkBefore.size() || !bulkBefore[(size_t)c]) {
            allBulk = false;
            break;
        }
    for (int c : touched) {
        if (allBulk) {
            bb[(size_t)c] = 1;
        } else {
            skip[(size_t)c] = 1;
            bb[(size_t)c] = 0;
        }
    }
}

/* zone_pass.cpp:280-294: position uniform in the target cell. */
inline void place(const P& p, int cell, double ux, double uy, double& x, double& y) {
    const int ix = cell / p.Ny, iy = cell - ix * p.Ny;
    const double dx = p.Lx / (double)p.Nx;
    const double dy = p.Ly / (double)p.Ny;
    x = (double)ix * dx + dx * ux;
    y = (double)iy * dy + dy * uy;
}

/* zone_pass.cpp:527-589 */
template <class R>
std::vector<int> spread(const P& p, std::vector<double>& x, std::vector<std::vector<int>>& parts,
                        std::vector<int>& cnt, int src, const std::vector<int>& ordered,
                        const std::vector<double>& target, int needOut, int pass, R& rng, int& moved) {
    moved = 0;
    std::vector<int> recv;
    if (needOut <= 0 || ordered.empty()) return recv;
    std::vector<int> avail = parts[(size_t)src];
    if (avail.empty()) return recv;
    auto def_of = [&](int nb) {
        return std::max(0, (int)std::floor(target[(size_t)nb] - (double)cnt[(size_t)nb]));
    };
    while (needOut > 0 && !avail.empty()) {
        bool progressed = false;
        for (size_t kk = 0; kk < ordered.size(); ++kk) {
            const int nb = ordered[kk];
            const int deficit = def_of(nb);
            if (deficit <= 0) continue;
            int remaining = 0;
            for (size_t jj = kk; jj < ordered.size(); ++jj)
                if (def_of(ordered[jj]) > 0) ++remaining;
            const int quota = std::max(1, (int)std::ceil((double)needOut / (double)std::max(1, remaining)));
            const int nt = std::max(1, std::min({quota, deficit, needOut, (int)avail.size()}));
            for (int q = 0; q < nt; ++q) {
                const int pid = avail[(size_t)q];
                double ux, uy;
                rng.pos_dense(pass, pid, ux, uy);
                place(p, nb, ux, uy, x[2 * (size_t)pid], x[2 * (size_t)pid + 1]);
                parts[(size_t)nb].push_back(pid);
            }
            avail.erase(avail.begin(), avail.begin() + nt);
            cnt[(size_t)nb] += nt;
            recv.push_back(nb);
            moved += nt;
            needOut -= nt;
            progressed = true;
            if (needOut <= 0 || avail.empty()) break;
        }
        if (!progressed) break;
    }
    if (moved > 0) {
        parts[(size_t)src] = avail;
        cnt[(size_t)src] = (int)avail.size();
        recv = unique_stable(recv);
    } else {
        recv.clear();
    }
    return recv;
}

/* zone_pass.cpp:636-667 */
template <class R>
std::vector<int> rank_donors(const P& p, int cell, std::vector<int> donors, double gx, double gy,
                             const std::vector<double>& Nsm, R& rng) {
    if (donors.empty()) return {};
    rng.shuffle_donors(donors, cell);
    const double ndir = R::hyp(gx, gy);
    struct Sc {
        int c;
        double align, dense, noise;
    };
    std::vector<Sc> sc;
    rng.begin_noise();
    for (int cn : donors) {
        int dx, dy;
        minimal_step(p, cell, cn, dx, dy);
        const double nstep = R::hyp((double)dx, (double)dy);
        double align = 0.0;
        if (ndir >= 1e-14 && nstep >= 1e-14) align = ((double)dx * gx + (double)dy * gy) / (ndir * nstep);
        sc.push_back(Sc{cn, align, Nsm[(size_t)cn], 1e-9 * rng.noise(cell, cn)});
    }
    auto greater = [](const Sc& a, const Sc& b) {
        const double sa = 100.0 * a.align + a.dense + a.noise;
        const double sb = 100.0 * b.align + b.dense + b.noise;
        return sa > sb;
    };
    if (R::stable_sorts) std::stable_sort(sc.begin(), sc.end(), greater);
    else std::sort(sc.begin(), sc.end(), greater);
    std::vector<int> out;
    for (const auto& s : sc) out.push_back(s.c);
    return out;
}

}  // namespace

/* ============ zone kernel: zone_pass.cpp:669-973 ============ */

namespace {

/* 3x3 binomial smoothing of one cell, the terms of smooth_counts in its order */
inline double smooth_at(const P& p, const std::vector<int>& cnt, int ix, int iy) {
    static const double K[9] = {1.0 / 16.0, 2.0 / 16.0, 1.0 / 16.0, 2.0 / 16.0, 4.0 / 16.0,
                                2.0 / 16.0, 1.0 / 16.0, 2.0 / 16.0, 1.0 / 16.0};
    double acc = 0.0, ws = 0.0;
    int kk = 0;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx, ++kk) {
            const int jy = iy + dy, jx = ix + dx;
            if (jy < 0 || jy >= p.Ny || jx < 0 || jx >= p.Nx) continue;
            acc += K[kk] * (double)cnt[(size_t)(jy + p.Ny * jx)];
            ws += K[kk];
        }
    return acc / std::max(ws, 1e-30);
}

/* the empty-cell thresholds of one cell, the arithmetic of threshold_maps */
inline void map_at(const P& p, double Nsm, double& lowLoc, double& highLoc) {
    const double lowFrac = 1.0 / (1.0 + p.coef);
    const double highFrac = 1.0 + p.coef;
    double f = 1.0;
    if (p.useLocalFluidFractionThresholds) {
        const double fRaw = Nsm / std::max(p.gamma, 1e-12);
        f = std::clamp(p.fluidFracGain * fRaw, 0.0, 1.0);
    }
    const double gl = p.gamma * f;
    lowLoc = std::max(lowFrac * gl, p.lowThrFloor);
    highLoc = std::max(highFrac * gl, p.highThrFloor);
}

/* Working arrays of the zone kernel, indexed by global cell id.  The
 * full-domain form (zone_kernel) sizes a fresh set for one zone, as the
 * reference's kernel builds its per-cell lists over the whole domain
 * (zone_pass.cpp:680-685).  The tile-local form (zone_pass_snapshot) keeps
 * one set for a pass, holding the pass input, and restores after each zone
 * the entries that zone changed: a zone then costs what its cells and their
 * ring of width 2 hold, not the domain. */
struct ZoneWork {
    std::vector<int> cnt, bulk, bulkBefore;
    std::vector<std::vector<int>> parts;
    std::vector<double> PbX, PbY, Nsm, lowLoc, highLoc, lowA, highA;
    std::vector<char> recv, bb, skip;
    std::vector<double> dU2, pxE, pyE, uxN, uyN;
    void init(size_t Nc) {
        cnt.assign(Nc, 0);
        bulk.assign(Nc, 0);
        bulkBefore.assign(Nc, 0);
        parts.assign(Nc, {});
        for (auto* a : {&PbX, &PbY, &Nsm, &lowLoc, &highLoc, &lowA, &highA, &dU2, &pxE, &pyE, &uxN, &uyN})
            a->assign(Nc, 0.0);
        recv.assign(Nc, 0);
        bb.assign(Nc, 0);
        skip.assign(Nc, 0);
    }
};

/* every cell a zone kernel reads a count or a map of: the zone and the ring
 * of width 2 around it (wrapped on periodic axes, clipped at walls) */
std::vector<int> zone_region(const P& p, const Zone& z) {
    std::vector<int> r;
    for (int ix = z.x0 - 2; ix < z.x1 + 2; ++ix)
        for (int iy = z.y0 - 2; iy < z.y1 + 2; ++iy) {
            int jx = ix, jy = iy;
            if (per_x(p)) jx = wrapi(jx, p.Nx);
            else if (jx < 0 || jx >= p.Nx) continue;
            if (per_y(p)) jy = wrapi(jy, p.Ny);
            else if (jy < 0 || jy >= p.Ny) continue;
            r.push_back(jy + p.Ny * jx);
        }
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
    return r;
}

/* The decisions and the momentum correction of run_zone_kernel
 * (zone_pass.cpp:706-922) on W, which holds the counts and cell lists
 * (ascending particle id) of the kernel's input.  Maps are built over
 * `region`; the dense and low lists and the correction run over the zone's
 * cells in ascending id, which is the reference's order over the whole
 * domain with the zone mask (every receiving cell is a zone cell). */
template <class R>
void zone_core(const P& p, std::vector<double>& x, std::vector<double>& v, const Zone& z, R& rng, ZoneWork& W,
               const std::vector<int>& region, ZoneStats& out) {
    auto& cnt = W.cnt;
    auto& parts = W.parts;
    for (int c : z.cells) {
        double px = 0.0, py = 0.0;
        for (int id : parts[(size_t)c]) {
            px += v[2 * (size_t)id];
            py += v[2 * (size_t)id + 1];
        }
        W.PbX[(size_t)c] = px;
        W.PbY[(size_t)c] = py;
    }
    double highThr, lowThrBulk;
    occupancy_thresholds(p, highThr, lowThrBulk);
    auto& lowA = W.lowA;
    auto& highA = W.highA;
    auto& Nsm = W.Nsm;
    auto& recv = W.recv;
    auto& bb = W.bb;
    auto& skip = W.skip;
    auto build_active_maps = [&]() {
        for (int c : region) {
            const int ix = c / p.Ny, iy = c - ix * p.Ny;
            Nsm[(size_t)c] = smooth_at(p, cnt, ix, iy);
            map_at(p, Nsm[(size_t)c], W.lowLoc[(size_t)c], W.highLoc[(size_t)c]);
            W.bulk[(size_t)c] = cnt[(size_t)c] > 0 ? 1 : 0;
            lowA[(size_t)c] = W.bulk[(size_t)c] ? lowThrBulk : W.lowLoc[(size_t)c];
            highA[(size_t)c] = W.bulk[(size_t)c] ? highThr : W.highLoc[(size_t)c];
        }
    };
    auto inz = [&](int c) { return z.in(c, p.Ny); };

    for (int pass = 1; pass <= p.maxRedistribPasses; ++pass) { /* :708-781 */
        int nHigh = 0, moved = 0;
        build_active_maps();
        for (int c : z.cells) W.bulkBefore[(size_t)c] = W.bulk[(size_t)c];
        std::vector<int> dense;
        for (int c : z.cells)
            if ((double)cnt[(size_t)c] > highA[(size_t)c]) dense.push_back(c);
        if (!dense.empty()) rng.shuffle_dense(dense, pass);
        for (int c : dense) {
            const int nc = cnt[(size_t)c];
            if (nc <= 0) continue;
            const double hl = highA[(size_t)c];
            if ((double)nc <= hl) continue;
            ++nHigh;
            /* redistribute_dense_cell_at_range_zone_local :591-634 */
            std::vector<int> recvT;
            int movedCell = 0;
            const int needOut = std::max(0, (int)std::ceil((double)cnt[(size_t)c] - hl));
            if (needOut > 0) {
                auto poor = admissible(p, c, cnt, lowA, 2);
                poor.erase(std::remove_if(poor.begin(), poor.end(), [&](int q) { return !inz(q); }), poor.end());
                auto ordP = rank_dense(p, c, poor, cnt, lowA, pass, 0, rng);
                int mv = 0;
                auto rn = spread(p, x, parts, cnt, c, ordP, lowA, needOut, pass, rng, mv);
                if (!rn.empty()) {
                    recvT.insert(recvT.end(), rn.begin(), rn.end());
                    recvT = unique_stable(recvT);
                }
                movedCell += mv;
                const int remaining = std::max(0, (int)std::ceil((double)cnt[(size_t)c] - hl));
                if (remaining > 0) {
                    auto relief = admissible(p, c, cnt, highA, 2);
                    relief.erase(std::remove_if(relief.begin(), relief.end(), [&](int q) { return !inz(q); }),
                                 relief.end());
                    auto ordR = rank_dense(p, c, relief, cnt, highA, pass, 1, rng);
                    rn = spread(p, x, parts, cnt, c, ordR, highA, remaining, pass, rng, mv);
                    if (!rn.empty()) {
                        recvT.insert(recvT.end(), rn.begin(), rn.end());
                        recvT = unique_stable(recvT);
                    }
                    movedCell += mv;
                }
            }
            for (int r : recvT) recv[(size_t)r] = 1;
            transfer_masks(bb, skip, c, recvT, W.bulkBefore);
            moved += movedCell;
            out.movedDense += movedCell;
        }
        out.stats[0] += (double)nHigh;
        out.stats[2] += (double)moved;
        out.stats[3] += (double)moved;
        /* :755-779 recompute maps that the next pass overwrites before use */
    }

    build_active_maps(); /* :783-794 */
    for (int c : z.cells) W.bulkBefore[(size_t)c] = W.bulk[(size_t)c]; /* the sparse phase's bulk snapshot */
    std::vector<int> low;
    for (int c : z.cells)
        if (cnt[(size_t)c] > 0 && (double)cnt[(size_t)c] < lowA[(size_t)c]) low.push_back(c);
    if (!low.empty()) rng.shuffle_low(low);

    int nLow = 0, movedS = 0;
    const int lowInt = (int)std::ceil(lowThrBulk);
    for (int c : low) { /* :804-854 */
        const int nc = cnt[(size_t)c];
        if (nc <= 0) continue;
        ++nLow;
        if (parts[(size_t)c].empty()) continue;
        /* local_neighbors_in_zone :263-278 */
        std::vector<int> neigh;
        {
            const int ix = c / p.Ny, iy = c - ix * p.Ny;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if (dx == 0 && dy == 0) continue;
                    int jy, jx;
                    if (!shift_index(p, iy, ix, dy, dx, jy, jx)) continue;
                    const int q = jy + p.Ny * jx;
                    if (inz(q)) neigh.push_back(q);
                }
            neigh = unique_stable(neigh);
        }
        if (neigh.empty()) continue;
        int need = std::max(0, lowInt - nc);
        if (need <= 0) continue;
        std::vector<int> donors;
        for (int d : neigh)
            if (inz(d) && cnt[(size_t)d] > lowInt) donors.push_back(d);
        if (donors.empty()) continue;
        double gx, gy;
        inward_dir(p, c, Nsm, gx, gy);
        const auto ordered = rank_donors(p, c, donors, gx, gy, Nsm, rng);
        for (int d : ordered) {
            std::vector<int> ids = parts[(size_t)d];
            if (ids.empty()) continue;
            const int avail = cnt[(size_t)d] - lowInt;
            if (avail <= 0) continue;
            const int nt = std::min({need, avail, (int)ids.size()});
            if (nt <= 0) continue;
            rng.shuffle_ids(ids, c, d);
            for (int q = 0; q < nt; ++q) {
                const int pid = ids[(size_t)q];
                double ux, uy;
                rng.pos_sparse(pid, ux, uy);
                place(p, c, ux, uy, x[2 * (size_t)pid], x[2 * (size_t)pid + 1]);
                parts[(size_t)c].push_back(pid);
            }
            parts[(size_t)d].assign(ids.begin() + nt, ids.end());
            cnt[(size_t)d] = (int)parts[(size_t)d].size();
            cnt[(size_t)c] += nt;
            recv[(size_t)c] = 1;
            transfer_masks(bb, skip, d, std::vector<int>{c}, W.bulkBefore);
            movedS += nt;
            out.movedSparse += nt;
            need -= nt;
            if (need <= 0) break;
        }
    }
    out.stats[1] += (double)nLow;
    out.stats[2] += (double)movedS;
    out.stats[3] += (double)movedS;

    /* momentum correction :870-909 (over the zone-touched cells: the zone's,
     * which hold every receiving cell) */
    double dUmax = 0.0;
    for (int c : z.cells) {
        const auto& ids = parts[(size_t)c];
        const int nc = (int)ids.size();
        if (nc == 0) continue;
        double pxA = 0.0, pyA = 0.0;
        for (int id : ids) {
            pxA += v[2 * (size_t)id];
            pyA += v[2 * (size_t)id + 1];
        }
        W.uxN[(size_t)c] = pxA / (double)nc;
        W.uyN[(size_t)c] = pyA / (double)nc;
        const double dPx = W.PbX[(size_t)c] - pxA, dPy = W.PbY[(size_t)c] - pyA;
        const double dux = dPx / (double)nc, duy = dPy / (double)nc;
        W.pxE[(size_t)c] = dPx * dPx;
        W.pyE[(size_t)c] = dPy * dPy;
        W.dU2[(size_t)c] = dux * dux + duy * duy;
        dUmax = std::max(dUmax, std::sqrt(W.dU2[(size_t)c]));
        bool corr = p.enableMomentumCorrectionPostRedistribution != 0;
        if (corr && skip[(size_t)c]) corr = false;
        if (corr && recv[(size_t)c] && !bb[(size_t)c]) corr = false;
        if (corr && (dux != 0.0 || duy != 0.0))
            for (int id : ids) {
                v[2 * (size_t)id] += dux;
                v[2 * (size_t)id + 1] += duy;
            }
    }
    double acc = 0.0;
    int nDU = 0;
    for (int c : z.cells) {
        const double d = W.dU2[(size_t)c];
        if (d > 0.0) {
            acc += d;
            ++nDU;
        }
    }
    if (nDU > 0) {
        out.stats[4] = std::sqrt(acc / (double)nDU);
        out.stats[5] = dUmax;
    }
}

}  // namespace

/* run_zone_kernel (zone_pass.cpp:669-973) on the whole domain: cell lists
 * of every cell, maps over every cell, and the full-domain diagnostics
 * (:924-970). */
template <class R>
ZoneStats zone_kernel(const P& p, std::vector<double>& x, std::vector<double>& v, const Zone& z, R& rng) {
    ZoneStats out;
    const size_t Nc = (size_t)p.Nc;
    ZoneWork W;
    W.init(Nc);
    const auto cid0 = cell_ids(p, x);
    for (int i = 0; i < p.n; ++i) W.parts[(size_t)cid0[(size_t)i]].push_back(i);
    for (size_t c = 0; c < Nc; ++c) W.cnt[c] = (int)W.parts[c].size();
    double PgX0 = 0.0, PgY0 = 0.0, E0 = 0.0;
    for (int i = 0; i < p.n; ++i) {
        PgX0 += v[2 * (size_t)i];
        PgY0 += v[2 * (size_t)i + 1];
    }
    for (int i = 0; i < p.n; ++i)
        E0 += 0.5 * (v[2 * (size_t)i] * v[2 * (size_t)i] + v[2 * (size_t)i + 1] * v[2 * (size_t)i + 1]);
    std::vector<int> all(Nc);
    for (size_t c = 0; c < Nc; ++c) all[c] = (int)c;
    zone_core(p, x, v, z, rng, W, all, out);

    double highThr, lowThrBulk;
    occupancy_thresholds(p, highThr, lowThrBulk);
    { /* :861-868: final high map for the diagnostics (the low map stays) */
        const auto Nsm = smooth_counts(p, W.cnt);
        std::vector<double> lowLoc, highLoc;
        threshold_maps(p, Nsm, lowLoc, highLoc);
        for (size_t c = 0; c < Nc; ++c) W.highA[c] = (W.cnt[c] > 0) ? highThr : highLoc[c];
    }
    const auto& cnt = W.cnt;
    const auto& highA = W.highA;
    const auto& lowA = W.lowA;

    /* diagnostics :924-970 */
    double PgX1 = 0.0, PgY1 = 0.0, E1 = 0.0;
    for (int i = 0; i < p.n; ++i) {
        PgX1 += v[2 * (size_t)i];
        PgY1 += v[2 * (size_t)i + 1];
    }
    for (int i = 0; i < p.n; ++i)
        E1 += 0.5 * (v[2 * (size_t)i] * v[2 * (size_t)i] + v[2 * (size_t)i + 1] * v[2 * (size_t)i + 1]);
    {
        int nAct = 0, nOut = 0;
        double mean = 0.0;
        for (size_t c = 0; c < Nc; ++c) mean += (double)cnt[c];
        nAct = p.Nc;
        for (size_t c = 0; c < Nc; ++c)
            if ((double)cnt[c] > highA[c] || ((double)cnt[c] > 0.0 && (double)cnt[c] < lowA[c])) ++nOut;
        if (nAct > 0) {
            mean /= (double)nAct;
            double a2 = 0.0;
            for (size_t c = 0; c < Nc; ++c) {
                const double d = (double)cnt[c] - mean;
                a2 += d * d;
            }
            out.diag[0] = std::sqrt(a2 / (double)nAct) / std::max(p.gamma, 1e-30);
            out.diag[1] = (double)nOut / (double)nAct;
        }
        const double dX = PgX1 - PgX0, dY = PgY1 - PgY0;
        out.diag[2] = std::sqrt(dX * dX + dY * dY);
        double aX = 0.0, aY = 0.0;
        int nX = 0, nY = 0;
        for (double q : W.pxE)
            if (q > 0.0) { aX += q; ++nX; }
        for (double q : W.pyE)
            if (q > 0.0) { aY += q; ++nY; }
        out.diag[3] = nX > 0 ? std::sqrt(aX / (double)nX) : 0.0;
        out.diag[4] = nY > 0 ? std::sqrt(aY / (double)nY) : 0.0;
        out.diag[5] = E1 - E0;
        out.diag[6] = std::numeric_limits<double>::quiet_NaN();
        double sUx = 0.0, sUy = 0.0;
        int nV = 0;
        for (size_t c = 0; c < Nc; ++c)
            if (cnt[c] > 0) {
                sUx += W.uxN[c] * W.uxN[c];
                sUy += W.uyN[c] * W.uyN[c];
                ++nV;
            }
        out.diag[7] = 0.0;
        out.diag[8] = nV > 0 ? std::sqrt(sUx / (double)nV) : 0.0;
        out.diag[9] = nV > 0 ? std::sqrt(sUy / (double)nV) : 0.0;
    }
    return out;
}

template ZoneStats zone_kernel<StdRngZone>(const P&, std::vector<double>&, std::vector<double>&, const Zone&,
                                           StdRngZone&);
template ZoneStats zone_kernel<CtrRngZone>(const P&, std::vector<double>&, std::vector<double>&, const Zone&,
                                           CtrRngZone&);

static void check_pass_modes(const P& p) {
    if (!p.useIncompressibleRedistribution || !p.useZoneRedistribution)
        throw std::runtime_error("zone pass requires useIncompressibleRedistribution=true and useZoneRedistribution=true");
    if (p.redistributionEnableSurfaceTopology || p.redistributionWallWettingEnabled ||
        p.useInterfaceVelocityReorientation)
        throw std::runtime_error("zone pass is restricted to no surface / no wetting / no reorientation");
}

/* ============ snapshot zone pass: zone_pass_openmp.cpp:34-172 ============ */

static void add_zone_metrics(PassMetrics& m, const ZoneStats& st, double& acc2, int& nc, double& mx) {
    m.movedDense += st.movedDense;
    m.movedSparse += st.movedSparse;
    if (st.stats[4] > 0.0) {
        acc2 += st.stats[4] * st.stats[4];
        ++nc;
    }
    mx = std::max(mx, st.stats[5]);
}

/* The reference's structure: each zone's kernel on a private copy of the
 * whole pass input, then the serial merge (the last zone covering a
 * particle's input cell wins, :145-153).  Cost O(n) per zone. */
template <class R>
PassMetrics zone_pass_snapshot_full(const P& p, std::vector<double>& x, std::vector<double>& v, int mode,
                                    uint64_t seedBase, std::vector<int32_t>* owner) {
    check_pass_modes(p);
    const Layout L = build_layout(p, mode);
    if (owner) *owner = L.owner;
    const std::vector<double> x0 = x, v0 = v;
    const auto cid0 = cell_ids(p, x0);
    PassMetrics m;
    m.nZones = (int)L.zones.size();
    double acc2 = 0.0, mx = 0.0;
    int nc = 0;
    for (size_t iz = 0; iz < L.zones.size(); ++iz) {
        const Zone& z = L.zones[iz];
        std::vector<double> xl = x0, vl = v0;
        R rng(seedBase + 104729ULL * (uint64_t)(iz + 1));
        const ZoneStats st = zone_kernel(p, xl, vl, z, rng);
        for (int i = 0; i < p.n; ++i) {
            if (!z.in(cid0[(size_t)i], p.Ny)) continue;
            x[2 * (size_t)i] = xl[2 * (size_t)i];
            x[2 * (size_t)i + 1] = xl[2 * (size_t)i + 1];
            v[2 * (size_t)i] = vl[2 * (size_t)i];
            v[2 * (size_t)i + 1] = vl[2 * (size_t)i + 1];
        }
        add_zone_metrics(m, st, acc2, nc, mx);
    }
    m.corrRmsDU = nc > 0 ? std::sqrt(acc2 / (double)nc) : 0.0;
    m.corrMaxDU = mx;
    return m;
}

/* The same pass, tile-local: the cell lists and counts of the pass input
 * are built once; a zone's kernel reads its own cells and their ring of
 * width 2 and changes only its own cells and particles, which are merged
 * into the output and restored to the pass input before the next zone.
 * Bitwise equal to zone_pass_snapshot_full (tests/test_oracle_fast.py),
 * O(n) per pass: the form that checks the 1024 x 512 grid. */
template <class R>
PassMetrics zone_pass_snapshot(const P& p, std::vector<double>& x, std::vector<double>& v, int mode,
                               uint64_t seedBase, std::vector<int32_t>* owner) {
    check_pass_modes(p);
    const Layout L = build_layout(p, mode);
    if (owner) *owner = L.owner;
    const size_t Nc = (size_t)p.Nc;
    const std::vector<double> x0 = x, v0 = v;
    std::vector<double> xw = x0, vw = v0;
    const auto cid0 = cell_ids(p, x0);
    ZoneWork W;
    W.init(Nc);
    for (int i = 0; i < p.n; ++i) W.parts[(size_t)cid0[(size_t)i]].push_back(i);
    for (size_t c = 0; c < Nc; ++c) W.cnt[c] = (int)W.parts[c].size();
    const auto parts0 = W.parts;
    const auto cnt0 = W.cnt;
    PassMetrics m;
    m.nZones = (int)L.zones.size();
    double acc2 = 0.0, mx = 0.0;
    int nc = 0;
    for (size_t iz = 0; iz < L.zones.size(); ++iz) {
        const Zone& z = L.zones[iz];
        R rng(seedBase + 104729ULL * (uint64_t)(iz + 1));
        ZoneStats st;
        zone_core(p, xw, vw, z, rng, W, zone_region(p, z), st);
        for (int c : z.cells) {
            for (int i : parts0[(size_t)c]) {
                for (size_t k = 2 * (size_t)i; k < 2 * (size_t)i + 2; ++k) {
                    x[k] = xw[k];
                    v[k] = vw[k];
                    xw[k] = x0[k];
                    vw[k] = v0[k];
                }
            }
            W.cnt[(size_t)c] = cnt0[(size_t)c];
            W.parts[(size_t)c] = parts0[(size_t)c];
            W.recv[(size_t)c] = W.bb[(size_t)c] = W.skip[(size_t)c] = 0;
            W.dU2[(size_t)c] = W.pxE[(size_t)c] = W.pyE[(size_t)c] = W.uxN[(size_t)c] = W.uyN[(size_t)c] = 0.0;
        }
        add_zone_metrics(m, st, acc2, nc, mx);
    }
    m.corrRmsDU = nc > 0 ? std::sqrt(acc2 / (double)nc) : 0.0;
    m.corrMaxDU = mx;
    return m;
}

template PassMetrics zone_pass_snapshot<StdRngZone>(const P&, std::vector<double>&, std::vector<double>&, int,
                                                    uint64_t, std::vector<int32_t>*);
template PassMetrics zone_pass_snapshot<CtrRngZone>(const P&, std::vector<double>&, std::vector<double>&, int,
                                                    uint64_t, std::vector<int32_t>*);
template PassMetrics zone_pass_snapshot_full<StdRngZone>(const P&, std::vector<double>&, std::vector<double>&, int,
                                                         uint64_t, std::vector<int32_t>*);
template PassMetrics zone_pass_snapshot_full<CtrRngZone>(const P&, std::vector<double>&, std::vector<double>&, int,
                                                         uint64_t, std::vector<int32_t>*);

/* ============ Gauss-Seidel zone pass: zone_pass.cpp:1096-1163 ============ */

template <class R>
PassMetrics zone_pass_gauss_seidel(const P& p, std::vector<double>& x, std::vector<double>& v, int mode,
                                   uint64_t seedBase) {
    check_pass_modes(p);
    const Layout L = build_layout(p, mode);
    PassMetrics m;
    m.nZones = (int)L.zones.size();
    double acc2 = 0.0, mx = 0.0;
    int nc = 0;
    for (size_t iz = 0; iz < L.zones.size(); ++iz) {
        R rng(seedBase + 104729ULL * (uint64_t)(iz + 1));
        const ZoneStats st = zone_kernel(p, x, v, L.zones[iz], rng);
        m.movedDense += st.movedDense;
        m.movedSparse += st.movedSparse;
        if (st.stats[4] > 0.0) {
            acc2 += st.stats[4] * st.stats[4];
            ++nc;
        }
        mx = std::max(mx, st.stats[5]);
    }
    m.corrRmsDU = nc > 0 ? std::sqrt(acc2 / (double)nc) : 0.0;
    m.corrMaxDU = mx;
    return m;
}

template PassMetrics zone_pass_gauss_seidel<StdRngZone>(const P&, std::vector<double>&, std::vector<double>&, int,
                                                        uint64_t);
template PassMetrics zone_pass_gauss_seidel<CtrRngZone>(const P&, std::vector<double>&, std::vector<double>&, int,
                                                        uint64_t);

/* ============ liquid closure: liquid_closure.cpp:143-444 ============ */

static std::vector<int32_t> boundary_of(const P& p, const std::vector<int32_t>& own) {
    std::vector<int32_t> m((size_t)p.Nc, 0);
    for (int ix = 0; ix < p.Nx; ++ix)
        for (int iy = 0; iy < p.Ny; ++iy) {
            const int c = iy + p.Ny * ix;
            const int right = (ix + 1 < p.Nx) ? iy + p.Ny * (ix + 1) : (per_x(p) ? iy : -1);
            const int up = (iy + 1 < p.Ny) ? (iy + 1) + p.Ny * ix : (per_y(p) ? p.Ny * ix : -1);
            bool b = false;
            if (right >= 0 && own[(size_t)right] != own[(size_t)c]) b = true;
            if (up >= 0 && own[(size_t)up] != own[(size_t)c]) b = true;
            if (b) m[(size_t)c] = 1;
        }
    return m;
}

static std::vector<int32_t> dilate(const P& p, const std::vector<int32_t>& in, int w) {
    if (w <= 0) return in;
    std::vector<int32_t> out((size_t)p.Nc, 0);
    for (int ix = 0; ix < p.Nx; ++ix)
        for (int iy = 0; iy < p.Ny; ++iy) {
            bool on = false;
            for (int dx = -w; dx <= w && !on; ++dx)
                for (int dy = -w; dy <= w; ++dy) {
                    int jx = ix + dx, jy = iy + dy;
                    if (per_x(p)) jx = wrapi(jx, p.Nx);
                    else if (jx < 0 || jx >= p.Nx) continue;
                    if (per_y(p)) jy = wrapi(jy, p.Ny);
                    else if (jy < 0 || jy >= p.Ny) continue;
                    if (in[(size_t)(jy + p.Ny * jx)] != 0) {
                        on = true;
                        break;
                    }
                }
            if (on) out[(size_t)(iy + p.Ny * ix)] = 1;
        }
    return out;
}

std::vector<int32_t> interzone_mask(const P& p) {
    std::vector<int32_t> m((size_t)p.Nc, 0);
    if (!p.smoothPdriveAtInterzone || !p.useZoneRedistribution || p.zonePassthroughMode ||
        p.zoneSingleFullDomainMode)
        return m;
    const int w = p.smoothPdriveInterzoneWidthCells;
    m = dilate(p, boundary_of(p, build_layout(p, MPCD_LAYOUT_BASE).owner), w);
    if (p.smoothPdriveIncludeShiftedLayouts) {
        const auto s = dilate(p, boundary_of(p, build_layout(p, MPCD_LAYOUT_SHIFTED).owner), w);
        for (size_t c = 0; c < (size_t)p.Nc; ++c) m[c] = (m[c] || s[c]) ? 1 : 0;
    }
    return m;
}

static std::vector<double> cross_smooth(const P& p, const std::vector<double>& in) {
    std::vector<double> out((size_t)p.Nc, 0.0);
    for (int ix = 0; ix < p.Nx; ++ix)
        for (int iy = 0; iy < p.Ny; ++iy) {
            double acc = in[(size_t)(iy + p.Ny * ix)];
            int n = 1;
            auto add = [&](int jy, int jx) {
                if (per_x(p)) jx = wrapi(jx, p.Nx);
                else if (jx < 0 || jx >= p.Nx) return;
                if (per_y(p)) jy = wrapi(jy, p.Ny);
                else if (jy < 0 || jy >= p.Ny) return;
                acc += in[(size_t)(jy + p.Ny * jx)];
                ++n;
            };
            add(iy, ix - 1);
            add(iy, ix + 1);
            add(iy - 1, ix);
            add(iy + 1, ix);
            out[(size_t)(iy + p.Ny * ix)] = acc / (double)n;
        }
    return out;
}

static void gradient(const P& p, const std::vector<double>& f, std::vector<double>& gx, std::vector<double>& gy) {
    const double dx = p.Lx / (double)p.Nx, dy = p.Ly / (double)p.Ny;
    gx.assign((size_t)p.Nc, 0.0);
    gy.assign((size_t)p.Nc, 0.0);
    auto at = [&](int iy, int ix) { return f[(size_t)(iy + p.Ny * ix)]; };
    for (int ix = 0; ix < p.Nx; ++ix)
        for (int iy = 0; iy < p.Ny; ++iy) {
            const size_t c = (size_t)(iy + p.Ny * ix);
            if (per_x(p)) gx[c] = (at(iy, wrapi(ix + 1, p.Nx)) - at(iy, wrapi(ix - 1, p.Nx))) / (2.0 * dx);
            else if (ix == 0) gx[c] = (at(iy, ix + 1) - at(iy, ix)) / dx;
            else if (ix == p.Nx - 1) gx[c] = (at(iy, ix) - at(iy, ix - 1)) / dx;
            else gx[c] = (at(iy, ix + 1) - at(iy, ix - 1)) / (2.0 * dx);
            if (per_y(p)) gy[c] = (at(wrapi(iy + 1, p.Ny), ix) - at(wrapi(iy - 1, p.Ny), ix)) / (2.0 * dy);
            else if (iy == 0) gy[c] = (at(iy + 1, ix) - at(iy, ix)) / dy;
            else if (iy == p.Ny - 1) gy[c] = (at(iy, ix) - at(iy - 1, ix)) / dy;
            else gy[c] = (at(iy + 1, ix) - at(iy - 1, ix)) / (2.0 * dy);
        }
}

static void kick(std::vector<double>& v, const std::vector<int>& cid, const std::vector<double>& kx,
                 const std::vector<double>& ky, int n, double& rms, double& mx) {
    double acc = 0.0;
    mx = 0.0;
    for (int i = 0; i < n; ++i) {
        const size_t c = (size_t)cid[(size_t)i];
        v[2 * (size_t)i] += kx[c];
        v[2 * (size_t)i + 1] += ky[c];
        const double k = std::sqrt(kx[c] * kx[c] + ky[c] * ky[c]);
        acc += k * k;
        if (k > mx) mx = k;
    }
    rms = n > 0 ? std::sqrt(acc / (double)n) : 0.0;
    if (n <= 0) mx = 0.0;
}

static double mean_of(const std::vector<double>& a) {
    if (a.empty()) return 0.0;
    return std::accumulate(a.begin(), a.end(), 0.0) / (double)a.size();
}

/* ============ occupancy statistics of a pass: zone_pass.cpp:975-1044 ============
 * (every cell is active, zone_pass.cpp:128-130; the mean and the variance
 * in two passes in cell order, as compute_occ_std_from_counts sums them) */
mpcd_occupancy_stats occupancy_stats(const P& p, const std::vector<int>& cnt) {
    mpcd_occupancy_stats o{};
    const size_t Nc = (size_t)p.Nc;
    double mean = 0.0;
    for (size_t c = 0; c < Nc; ++c) mean += (double)cnt[c];
    mean /= (double)Nc;
    double acc = 0.0;
    for (size_t c = 0; c < Nc; ++c) {
        const double d = (double)cnt[c] - mean;
        acc += d * d;
    }
    o.occStd = std::sqrt(acc / (double)Nc) / std::max(p.gamma, 1e-30);
    const auto Nsm = smooth_counts(p, cnt);
    std::vector<double> lowLoc, highLoc;
    threshold_maps(p, Nsm, lowLoc, highLoc);
    double highThr, lowThrBulk;
    occupancy_thresholds(p, highThr, lowThrBulk);
    int nOut = 0;
    for (size_t c = 0; c < Nc; ++c) {
        const bool bulk = cnt[c] > 0;
        const double lo = bulk ? lowThrBulk : lowLoc[c], hi = bulk ? highThr : highLoc[c];
        if ((double)cnt[c] > hi || (cnt[c] > 0 && (double)cnt[c] < lo)) ++nOut;
        if ((double)cnt[c] > hi) ++o.nDense;
        if (cnt[c] > 0 && (double)cnt[c] < lo) ++o.nSparse;
        if (cnt[c] == 0) ++o.nEmptyActive;
    }
    o.outBand = (double)nOut / (double)Nc;
    return o;
}

ClosureOut liquid_closure(const P& p, const std::vector<double>& xRef, const std::vector<double>& vRef,
                          const std::vector<double>& xRed, const std::vector<double>& vRed,
                          std::vector<double>& xOut, std::vector<double>& vOut) {
    if (!p.useLiquidClosure) throw std::runtime_error("run_liquid_closure requires useLiquidClosure=true");
    if (xRef.size() != xRed.size() || vRef.size() != vRed.size())
        throw std::runtime_error("run_liquid_closure requires ref/red states with matching sizes");
    ClosureOut o;
    const size_t Nc = (size_t)p.Nc;
    const double rhoT = rho_target(p);
    o.ref = cell_fields(p, xRef, vRef, rhoT);
    o.red = cell_fields(p, xRed, vRed, rhoT);
    o.duRepairX.assign(Nc, 0.0);
    o.duRepairY.assign(Nc, 0.0);
    double num = 0.0, den = 0.0;
    for (size_t c = 0; c < Nc; ++c) { /* :349-360 */
        o.duRepairX[c] = (o.ref.Px[c] - o.red.Px[c]) / p.gamma;
        o.duRepairY[c] = (o.ref.Py[c] - o.red.Py[c]) / p.gamma;
        const double ex = o.red.Ux[c] - o.ref.Ux[c], ey = o.red.Uy[c] - o.ref.Uy[c];
        const double ax = o.duRepairX[c], ay = o.duRepairY[c];
        num += -(ex * ax + ey * ay);
        den += ax * ax + ay * ay;
    }
    o.m.betaRepairNum = num;
    o.m.betaRepairDen = den;
    double bopt = 0.0;
    if (std::isfinite(num) && std::isfinite(den) && den > 0.0) bopt = std::max(num / den, 0.0);
    o.m.betaRepairOpt = bopt;
    if (!p.useOptimalBetaRepair) o.m.betaRepairApplied = p.betaRepair;
    else if (!std::isfinite(bopt) || !std::isfinite(den) || den < 1e-20) o.m.betaRepairApplied = p.betaRepair;
    else o.m.betaRepairApplied = std::clamp(bopt, 0.0, 2.0);
    o.m.betaEOSApplied = p.betaEOS;

    std::vector<double> vRep = vRed;
    const auto cidRed = cell_ids(p, xRed);
    std::vector<double> ax(Nc), ay(Nc);
    for (size_t c = 0; c < Nc; ++c) {
        ax[c] = o.m.betaRepairApplied * o.duRepairX[c];
        ay[c] = o.m.betaRepairApplied * o.duRepairY[c];
    }
    kick(vRep, cidRed, ax, ay, p.n, o.m.rmsRepairDU, o.m.maxRepairDU);
    {
        double a = 0.0;
        for (size_t c = 0; c < Nc; ++c) a += std::hypot(ax[c], ay[c]);
        o.m.meanAbsRepairDU = a / (double)std::max(p.Nc, 1);
    }
    o.repaired = cell_fields(p, xRed, vRep, rhoT);
    o.pdriveRaw.assign(Nc, 0.0);
    for (size_t c = 0; c < Nc; ++c) o.pdriveRaw[c] = o.ref.Pkin[c] + o.ref.Pvir[c];
    o.m.meanPdriveRaw = mean_of(o.pdriveRaw);
    o.pdrive = o.pdriveRaw;
    o.mask = interzone_mask(p);
    o.m.nMaskedInterzoneCells = (int)std::count(o.mask.begin(), o.mask.end(), 1);
    if (o.m.nMaskedInterzoneCells > 0) {
        const auto ps = cross_smooth(p, o.pdriveRaw);
        for (size_t c = 0; c < Nc; ++c)
            if (o.mask[c] != 0)
                o.pdrive[c] = (1.0 - p.smoothPdriveInterzoneBlend) * o.pdriveRaw[c] +
                              p.smoothPdriveInterzoneBlend * ps[c];
    }
    o.m.meanPdrive = mean_of(o.pdrive);
    std::vector<double> gx, gy;
    gradient(p, o.pdrive, gx, gy);
    o.duEOSX.assign(Nc, 0.0);
    o.duEOSY.assign(Nc, 0.0);
    for (size_t c = 0; c < Nc; ++c) {
        o.duEOSX[c] = -(p.dt / rhoT) * gx[c] * p.betaEOS;
        o.duEOSY[c] = -(p.dt / rhoT) * gy[c] * p.betaEOS;
    }
    xOut = xRed;
    vOut = vRep;
    kick(vOut, cidRed, o.duEOSX, o.duEOSY, p.n, o.m.rmsEOSDU, o.m.maxEOSDU);
    o.out = cell_fields(p, xOut, vOut, rhoT);
    double pi0 = 0.0, pi1 = 0.0, po0 = 0.0, po1 = 0.0;
    for (int i = 0; i < p.n; ++i) {
        pi0 += vRed[2 * (size_t)i];
        pi1 += vRed[2 * (size_t)i + 1];
    }
    for (int i = 0; i < p.n; ++i) {
        po0 += vOut[2 * (size_t)i];
        po1 += vOut[2 * (size_t)i + 1];
    }
    o.m.momentumDeltaX = po0 - pi0;
    o.m.momentumDeltaY = po1 - pi1;
    o.vRepaired = std::move(vRep);
    return o;
}

}  // namespace orc
