/*
 * mpcd_core.h -- per-particle, per-cell and per-zone bodies of the gfx950
 * SRD/MPCD step.  Plain C++ with MPCD_HD (host+device) so the kernels in
 * mpcd_engine.hip and the host-side logic check in tests/ share one
 * definition.  FP64 throughout; every file compiling this must pass
 * -ffp-contract=off so products and sums round exactly as written.
 */
#pragma once

#include <stdint.h>

#include "mpcd_gpu.h"
#include "mpcd_rng.h"

#if defined(__HIP_DEVICE_COMPILE__)
#define MPCD_CEIL(x) __builtin_ceil(x)
#define MPCD_FABS(x) __builtin_fabs(x)
#else
#define MPCD_CEIL(x) std::ceil(x)
#define MPCD_FABS(x) std::fabs(x)
#endif

namespace mpcdk {

MPCD_HD int imin(int a, int b) { return a < b ? a : b; }
MPCD_HD int imax(int a, int b) { return a < b ? b : a; }
MPCD_HD double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max(a, b) */
MPCD_HD double dmin(double a, double b) { return (b < a) ? b : a; } /* std::min(a, b) */
MPCD_HD int wrapi(int i, int n) {
    const int r = i % n;
    return (r < 0) ? r + n : r;
}

/* Scalars every kernel needs, derived once on the host from mpcd_params. */
struct Geom {
    int Nx, Ny, Nc, n;
    int perx, pery;
    double Lx, Ly, a0, dxc, dyc; /* dxc = Lx/Nx, dyc = Ly/Ny */
};

/* ---------- force + stream + wall BC (driver :242-247 + boundary_conditions.cpp:19-203) ---------- */

struct BCParams {
    double dt, Fx, Fy;
    double Utop, Ubottom, sigma;
    int left, right, bottom, top; /* MPCD_BC_* */
};

/* Wall-hit accumulators (WallInfo). */
struct WallAcc {
    int nBot, nTop, nLeft, nRight;
    double dE, dPxBot, dPxTop, dPyBot, dPyTop, dPxLeft, dPxRight, dPyLeft, dPyRight;
};

MPCD_HD void wall_acc_zero(WallAcc& a) {
    a.nBot = a.nTop = a.nLeft = a.nRight = 0;
    a.dE = a.dPxBot = a.dPxTop = a.dPyBot = a.dPyTop = a.dPxLeft = a.dPxRight = a.dPyLeft = a.dPyRight = 0.0;
}

/* One particle: returns 0, or MPCD_E_BC_MODE when it hits a wall whose mode
 * is not one of the supported strings (the reference throws there). */
MPCD_HD int bc_particle(const Geom& g, const BCParams& b, uint64_t seed, uint32_t pid, int do_stream, double& x,
                        double& y, double& vx, double& vy, WallAcc& acc) {
    if (do_stream) {
        vx += b.Fx * b.dt;
        vy += b.Fy * b.dt;
        x += b.dt * vx;
        y += b.dt * vy;
    }
    const double Lx = g.Lx, Ly = g.Ly, s = b.sigma;
    for (int k = 0; k < 4; ++k) {
        if (x < 0.0) {
            const double ox = vx, oy = vy;
            if (b.left == MPCD_BC_PERIODIC) {
                x += Lx;
            } else if (b.left >= MPCD_BC_THERMALIZE && b.left <= MPCD_BC_SPECULAR) {
                x = -x;
                if (b.left == MPCD_BC_THERMALIZE) {
                    double n1, n2;
                    mpcd_draw_bc(seed, pid, k, 0, &n1, &n2);
                    vx = MPCD_FABS(s * n1);
                    vy = s * n2;
                } else if (b.left == MPCD_BC_BOUNCEBACK) {
                    vx = MPCD_FABS(vx);
                    vy = -vy;
                } else {
                    vx = MPCD_FABS(vx);
                }
                acc.nLeft += 1;
                acc.dE += 0.5 * ((vx * vx + vy * vy) - (ox * ox + oy * oy));
                acc.dPxLeft += MPCD_FABS(vx - ox);
                acc.dPyLeft += vy - oy;
            } else {
                return MPCD_E_BC_MODE;
            }
        }
        if (x > Lx) {
            const double ox = vx, oy = vy;
            if (b.right == MPCD_BC_PERIODIC) {
                x -= Lx;
            } else if (b.right >= MPCD_BC_THERMALIZE && b.right <= MPCD_BC_SPECULAR) {
                x = 2.0 * Lx - x;
                if (b.right == MPCD_BC_THERMALIZE) {
                    double n1, n2;
                    mpcd_draw_bc(seed, pid, k, 1, &n1, &n2);
                    vx = -MPCD_FABS(s * n1);
                    vy = s * n2;
                } else if (b.right == MPCD_BC_BOUNCEBACK) {
                    vx = -MPCD_FABS(vx);
                    vy = -vy;
                } else {
                    vx = -MPCD_FABS(vx);
                }
                acc.nRight += 1;
                acc.dE += 0.5 * ((vx * vx + vy * vy) - (ox * ox + oy * oy));
                acc.dPxRight += MPCD_FABS(vx - ox);
                acc.dPyRight += vy - oy;
            } else {
                return MPCD_E_BC_MODE;
            }
        }
        if (y < 0.0) {
            const double ox = vx, oy = vy;
            if (b.bottom == MPCD_BC_PERIODIC) {
                y += Ly;
            } else if (b.bottom >= MPCD_BC_THERMALIZE && b.bottom <= MPCD_BC_SPECULAR) {
                y = -y;
                if (b.bottom == MPCD_BC_THERMALIZE) {
                    double n1, n2;
                    mpcd_draw_bc(seed, pid, k, 2, &n1, &n2);
                    vx = b.Ubottom + s * n1;
                    vy = MPCD_FABS(s * n2);
                } else if (b.bottom == MPCD_BC_BOUNCEBACK) {
                    vx = 2.0 * b.Ubottom - vx;
                    vy = MPCD_FABS(vy);
                } else {
                    vy = MPCD_FABS(vy);
                }
                acc.nBot += 1;
                acc.dE += 0.5 * ((vx * vx + vy * vy) - (ox * ox + oy * oy));
                acc.dPxBot += vx - ox;
                acc.dPyBot += MPCD_FABS(vy - oy);
            } else {
                return MPCD_E_BC_MODE;
            }
        }
        if (y > Ly) {
            const double ox = vx, oy = vy;
            if (b.top == MPCD_BC_PERIODIC) {
                y -= Ly;
            } else if (b.top >= MPCD_BC_THERMALIZE && b.top <= MPCD_BC_SPECULAR) {
                y = 2.0 * Ly - y;
                if (b.top == MPCD_BC_THERMALIZE) {
                    double n1, n2;
                    mpcd_draw_bc(seed, pid, k, 3, &n1, &n2);
                    vx = b.Utop + s * n1;
                    vy = -MPCD_FABS(s * n2);
                } else if (b.top == MPCD_BC_BOUNCEBACK) {
                    vx = 2.0 * b.Utop - vx;
                    vy = -MPCD_FABS(vy);
                } else {
                    vy = -MPCD_FABS(vy);
                }
                acc.nTop += 1;
                acc.dE += 0.5 * ((vx * vx + vy * vy) - (ox * ox + oy * oy));
                acc.dPxTop += vx - ox;
                acc.dPyTop += MPCD_FABS(vy - oy);
            } else {
                return MPCD_E_BC_MODE;
            }
        }
    }
    if (g.perx) {
        x = x - MPCD_FLOOR(x / Lx) * Lx;
    } else {
        if (x < 0.0) x = 0.0;
        if (x > Lx) x = Lx;
    }
    if (g.pery) {
        y = y - MPCD_FLOOR(y / Ly) * Ly;
    } else {
        if (y < 0.0) y = 0.0;
        if (y > Ly) y = Ly;
    }
    return 0;
}

/* ---------- cell indices ---------- */

/* Unshifted cell (common_grid.cpp:9-19). */
MPCD_HD int cell_of(const Geom& g, double x, double y) {
    int ix = (int)MPCD_FLOOR(x / g.dxc);
    int iy = (int)MPCD_FLOOR(y / g.dyc);
    ix = imin(imax(ix, 0), g.Nx - 1);
    iy = imin(imax(iy, 0), g.Ny - 1);
    return iy + g.Ny * ix;
}

/* Random-shift collision cell (srd_collision.cpp:18-44); xmax/ymax are
 * nextafter(Lx, 0), nextafter(Ly, 0) computed on the host. */
MPCD_HD int shifted_cell_of(const Geom& g, double x, double y, double sx, double sy, double xmax, double ymax) {
    double xs = x + sx;
    double ys = y + sy;
    if (g.perx) {
        xs = xs - MPCD_FLOOR(xs / g.Lx) * g.Lx;
    } else {
        if (xs < 0.0) xs = 0.0;
        if (xs >= g.Lx) xs = xmax;
    }
    if (g.pery) {
        ys = ys - MPCD_FLOOR(ys / g.Ly) * g.Ly;
    } else {
        if (ys < 0.0) ys = 0.0;
        if (ys >= g.Ly) ys = ymax;
    }
    int ix = (int)MPCD_FLOOR(xs / g.a0);
    int iy = (int)MPCD_FLOOR(ys / g.a0);
    ix = imin(imax(ix, 0), g.Nx - 1);
    iy = imin(imax(iy, 0), g.Ny - 1);
    return iy + g.Ny * ix;
}

/* ---------- SRD collision of one cell (srd_collision.cpp:61-129) ----------
 * ids[0..m) are the particle slots of the cell in ascending particle id, so
 * every sum runs in the reference's order.  rot[4] = {cos(+a), sin(+a),
 * cos(-a), sin(-a)} from the host's libm. */
struct CollideParams {
    double kBT;
    int thermostat;
    double rot[4];
};

MPCD_HD void collide_cell(const CollideParams& cp, uint64_t seed, int cell, const int* ids, int m, double* vx,
                          double* vy) {
    /* ids == nullptr: the cell's particles are vx[0..m), vy[0..m) */
    if (m <= 0) return;
    double sx = 0.0, sy = 0.0;
    for (int k = 0; k < m; ++k) {
        const int i = ids ? ids[k] : k;
        sx += vx[i];
        sy += vy[i];
    }
    const double ux = sx / (double)m, uy = sy / (double)m;
    double ct = 1.0, st = 0.0; /* cos(0), sin(0) for cells with fewer than 2 particles */
    if (m >= 2) {
        const double sgn = mpcd_draw_rot_sign(seed, (uint32_t)cell);
        ct = sgn < 0.0 ? cp.rot[2] : cp.rot[0];
        st = sgn < 0.0 ? cp.rot[3] : cp.rot[1];
    }
    double lam = 1.0;
    if (cp.thermostat) {
        double s2 = 0.0;
        for (int k = 0; k < m; ++k) {
            const int i = ids ? ids[k] : k;
            const double vxr = vx[i] - ux, vyr = vy[i] - uy;
            const double rx = ct * vxr - st * vyr;
            const double ry = st * vxr + ct * vyr;
            s2 += rx * rx + ry * ry;
        }
        const int dof = 2 * imax(m - 1, 0);
        if (m > 1 && s2 > 0.0) lam = MPCD_SQRT((double)dof * cp.kBT / s2);
    }
    for (int k = 0; k < m; ++k) {
        const int i = ids ? ids[k] : k;
        const double vxr = vx[i] - ux, vyr = vy[i] - uy;
        double rx = ct * vxr - st * vyr;
        double ry = st * vxr + ct * vyr;
        if (cp.thermostat) {
            rx *= lam;
            ry *= lam;
        }
        vx[i] = ux + rx;
        vy[i] = uy + ry;
    }
}

/* ---------- cell fields of one cell (common_grid.cpp:62-97) ---------- */

struct CellStats {
    int N;
    double Px, Py, Ux, Uy, T, rho, Pkin, Pvir, P;
};

MPCD_HD CellStats cell_stats(const int* ids, int m, const double* vx, const double* vy, double Vc, double Kvirial,
                             double rhoTarget) {
    /* ids == nullptr: the cell's particles are vx[0..m), vy[0..m) */
    CellStats s;
    s.N = m;
    s.Px = 0.0;
    s.Py = 0.0;
    for (int k = 0; k < m; ++k) {
        const int i = ids ? ids[k] : k;
        s.Px += vx[i];
        s.Py += vy[i];
    }
    s.Ux = 0.0;
    s.Uy = 0.0;
    if (m > 0) {
        s.Ux = s.Px / (double)m;
        s.Uy = s.Py / (double)m;
    }
    s.rho = (double)m / Vc;
    double s2 = 0.0;
    for (int k = 0; k < m; ++k) {
        const int i = ids ? ids[k] : k;
        const double dvx = vx[i] - s.Ux, dvy = vy[i] - s.Uy;
        s2 += dvx * dvx + dvy * dvy;
    }
    const int dof = 2 * imax(m - 1, 0);
    if (dof > 0) {
        s.T = s2 / (double)dof;
        s.Pkin = s.rho * s.T;
    } else {
        s.T = 0.0;
        s.Pkin = 0.0;
    }
    s.Pvir = Kvirial * (s.rho - rhoTarget);
    s.P = s.Pkin + s.Pvir;
    return s;
}

}  // namespace mpcdk
