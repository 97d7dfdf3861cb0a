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

/* ---------- zone redistribution (zone_pass.cpp:62-973) ----------
 *
 * One zone is processed by one wavefront.  The reference's zone kernel is
 * sequential: a shuffled list of dense cells, each of which mutates the
 * counts its successors read.  That order is kept; inside each step the
 * wavefront works together wherever the work is independent by construction:
 * all counter-based draws of a step (shuffle partners, tie breaks, score
 * noise, new positions), candidate evaluation and ranking, order-preserving
 * compactions and prefix sums.  Only list bookkeeping runs on lane 0.
 * The same code runs on the host with one lane (nl = 1), where the wave
 * primitives below reduce to scalars and the barrier to nothing. */

#if defined(__HIP_DEVICE_COMPILE__)
#define ZBAR() __syncthreads()
MPCD_HD unsigned long long wave_ballot(bool f) { return __ballot(f); }
MPCD_HD int wave_popc(unsigned long long m) { return __popcll(m); }
MPCD_HD int wave_shfl_up_add(int v, int lane) {
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}
MPCD_HD int wave_bcast(int v, int src) { return __shfl(v, src, 64); }
#else
#define ZBAR() ((void)0)
MPCD_HD unsigned long long wave_ballot(bool f) { return f ? 1ull : 0ull; }
MPCD_HD int wave_popc(unsigned long long m) { return __builtin_popcountll(m); }
MPCD_HD int wave_shfl_up_add(int v, int) { return v; }
MPCD_HD int wave_bcast(int v, int) { return v; }
#endif
MPCD_HD unsigned long long lanes_below(int lane) { return (lane >= 64) ? ~0ull : ((1ull << lane) - 1ull); }

struct ZoneParams {
    double highThr, lowThrBulk;
    double lowFrac, highFrac; /* 1/(1+coef), 1+coef */
    double gamma, fluidFracGain, lowThrFloor, highThrFloor;
    double lowLoc1, highLoc1; /* empty-cell thresholds for fluid fraction 1 */
    int useLocal, maxPasses, corrEnabled;
    signed char binTab[25]; /* angular bin of (dx,dy) in [-2,2]^2: index (dy+2)*5+(dx+2) */
};

/* Candidate of a dense cell's target ranking. */
struct Cand {
    int lc, gc;
    double deficitNeg;
    int cnt, cheb, dist2, bin;
    double tie;
};

constexpr int ZONE_MAX_MOVE = 64; /* movers whose draws are staged at once */

/* Working set of one zone.  Cell arrays are indexed by the zone-local cell
 * lc = (ix-xs)*zh + (iy-ys) (ascending global cell id); particles by the
 * zone-local index k (the zone's cells' runs of the binning list, in lc
 * order).  Per-particle scratch lives at [off, off+P) of next and tmpg,
 * private to the zone: the shifted layout's zones overlap
 * (zone_pass.cpp:191-219 starts them at 0 and T/2 with width T), so two
 * zones may hold the same particle.  Every zone reads the pass input; a
 * zone writes only the particles it owns, i.e. whose input cell has this
 * zone as its owner-map entry (the last zone covering the cell, which is the
 * zone whose result the reference's merge keeps,
 * zone_pass_openmp.cpp:145-153); owner == nullptr: the zone owns all.
 * Positions are written in place (nobody reads them during a pass).
 * Velocities are written in place for particles whose input cell only one
 * zone covers; for cells several zones cover (gcover != 0) they go to
 * vsx/vsy and are copied back after the pass, so every zone reads inputs.
 * A cell's particle list stays implicit -- its input run of local indices
 * k0..k0+cnt-1 -- until the first change, when its links are written
 * (mat = 1); cells nothing touched cost nothing at the correction. */
struct ZoneCtx {
    Geom g;
    ZoneParams zp;
    int xs, xe, ys, ye, zw, zh, C;
    int zid, off, P;
    uint64_t seed;
    /* snapshot binning (read only) */
    const int* gcnt;
    const int* gstart;
    const int* glist;
    const int* gpid;
    const int* gcid;  /* input cell of each slot */
    const int* owner; /* owner map of the layout, or nullptr */
    /* state */
    double *x, *y, *vx, *vy;
    double *vsx, *vsy;
    const unsigned char* gcover; /* per cell: several zones cover it, or nullptr */
    int* next; /* [off, off+P): list links by local index */
    int* tmpg; /* [off, off+P): donor shuffles */
    /* zone cell arrays */
    int *cnt, *head, *tail, *dl, *cntS, *jd, *k0, *colL, *colG;
    double *lowA, *highA, *dux, *duy; /* lowA/highA only with local fluid fractions */
    unsigned char *bulk0, *recv, *bb, *skip, *mat, *corr;
    /* small scratch */
    Cand* cand;     /* 24 */
    int* ord;       /* 24 */
    int* nbGc;      /* 24 */
    int* rnk;       /* 24 */
    int* recvT;     /* 48 */
    double* dsc;    /* 8 */
    int* dord;      /* 8 */
    int* dord2;     /* 8 */
    double* nsm4;   /* 4 */
    double* mvx;    /* ZONE_MAX_MOVE */
    double* mvy;    /* ZONE_MAX_MOVE */
    int* mvk;       /* ZONE_MAX_MOVE */
    /* shared scalars */
    int m, nOrd, nr, nn, nd, startBin, need, stage;
    /* results */
    int nHigh, nLow;
    int movedDense, movedSparse;
    double stats4, stats5;
    int status;
};

MPCD_HD int zone_lc(const ZoneCtx& z, int gx, int gy) {
    if (gx >= z.xs && gx < z.xe && gy >= z.ys && gy < z.ye) return (gx - z.xs) * z.zh + (gy - z.ys);
    return -1;
}
MPCD_HD int lc_gx(const ZoneCtx& z, int lc) { return z.xs + lc / z.zh; }
MPCD_HD int lc_gy(const ZoneCtx& z, int lc) { return z.ys + lc % z.zh; }
MPCD_HD int lc_gc(const ZoneCtx& z, int lc) { return lc_gy(z, lc) + z.g.Ny * lc_gx(z, lc); }

/* count at a global cell: live zone count (or the sparse-phase copy), else
 * the pass-input snapshot. */
MPCD_HD int count_at(const ZoneCtx& z, int gx, int gy, int use_copy) {
    const int lc = zone_lc(z, gx, gy);
    if (lc >= 0) return use_copy ? z.cntS[lc] : z.cnt[lc];
    return z.gcnt[gy + z.g.Ny * gx];
}

/* 3x3 binomial smoothing without wrap (zone_pass.cpp:296-319). */
MPCD_HD double nsm_at(const ZoneCtx& z, int gx, int gy, int use_copy) {
    const double K[9] = {1.0 / 16.0, 2.0 / 16.0, 1.0 / 16.0, 2.0 / 16.0, 4.0 / 16.0,
                         2.0 / 16.0, 1.0 / 16.0, 2.0 / 16.0, 1.0 / 16.0};
    double acc = 0.0, ws = 0.0;
    int kk = 0;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx, ++kk) {
            const int jy = gy + dy, jx = gx + dx;
            if (jy < 0 || jy >= z.g.Ny || jx < 0 || jx >= z.g.Nx) continue;
            acc += K[kk] * (double)count_at(z, jx, jy, use_copy);
            ws += K[kk];
        }
    return acc / dmax(ws, 1e-30);
}

/* Active thresholds of a zone cell (zone_pass.cpp:321-345, :715-725): the
 * bulk flag is taken when the maps are built; with fluid fraction 1 the
 * empty-cell values are constants, otherwise they are stored per cell. */
MPCD_HD void zone_maps(ZoneCtx& z, int lc, int use_copy) {
    const int bulk = z.cnt[lc] > 0;
    z.bulk0[lc] = (unsigned char)bulk;
    if (!z.zp.useLocal) return;
    const double fRaw = nsm_at(z, lc_gx(z, lc), lc_gy(z, lc), use_copy) / dmax(z.zp.gamma, 1e-12);
    const double t = z.zp.fluidFracGain * fRaw;
    const double f = (t < 0.0) ? 0.0 : ((1.0 < t) ? 1.0 : t); /* std::clamp(t, 0, 1) */
    const double gl = z.zp.gamma * f;
    z.lowA[lc] = bulk ? z.zp.lowThrBulk : dmax(z.zp.lowFrac * gl, z.zp.lowThrFloor);
    z.highA[lc] = bulk ? z.zp.highThr : dmax(z.zp.highFrac * gl, z.zp.highThrFloor);
}
MPCD_HD double low_thr(const ZoneCtx& z, int lc) {
    if (z.zp.useLocal) return z.lowA[lc];
    return z.bulk0[lc] ? z.zp.lowThrBulk : z.zp.lowLoc1;
}
MPCD_HD double high_thr(const ZoneCtx& z, int lc) {
    if (z.zp.useLocal) return z.highA[lc];
    return z.bulk0[lc] ? z.zp.highThr : z.zp.highLoc1;
}
MPCD_HD double thr(const ZoneCtx& z, int lc, int high) { return high ? high_thr(z, lc) : low_thr(z, lc); }

/* zone_pass.cpp:229-244 */
MPCD_HD void min_step(const ZoneCtx& z, int sx, int sy, int tx, int ty, int& ddx, int& ddy) {
    ddx = tx - sx;
    ddy = ty - sy;
    if (z.g.perx) {
        if (ddx > z.g.Nx / 2) ddx -= z.g.Nx;
        if (ddx < -z.g.Nx / 2) ddx += z.g.Nx;
    }
    if (z.g.pery) {
        if (ddy > z.g.Ny / 2) ddy -= z.g.Ny;
        if (ddy < -z.g.Ny / 2) ddy += z.g.Ny;
    }
}

/* zone_pass.cpp:246-261 */
MPCD_HD int shift_idx(const ZoneCtx& z, int gx, int gy, int dx, int dy, int& jx, int& jy) {
    jy = gy + dy;
    if (z.g.pery) jy = wrapi(jy, z.g.Ny);
    else if (jy < 0 || jy >= z.g.Ny) return 0;
    jx = gx + dx;
    if (z.g.perx) jx = wrapi(jx, z.g.Nx);
    else if (jx < 0 || jx >= z.g.Nx) return 0;
    return 1;
}

/* slot of the zone's k-th particle: the zone's particles are its columns'
 * runs of the binning list, in column order */
MPCD_HD int slot_of(const ZoneCtx& z, int k) {
    int lo = 0, hi = z.zw - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (z.colL[mid] <= k) lo = mid;
        else hi = mid - 1;
    }
    return z.glist[z.colG[lo] + (k - z.colL[lo])];
}

/* successor of k in cell lc's list (implicit run or links) */
MPCD_HD int next_k(const ZoneCtx& z, int lc, int k) {
    return z.mat[lc] ? z.next[z.off + k] : (k < z.tail[lc] ? k + 1 : -1);
}

/* write the links of a still-implicit list before its first change */
MPCD_HD void ensure_mat(ZoneCtx& z, int lc) {
    if (z.mat[lc]) return;
    const int t = z.tail[lc];
    for (int k = z.head[lc]; k >= 0 && k <= t; ++k) z.next[z.off + k] = (k < t) ? k + 1 : -1;
    z.mat[lc] = 1;
}

/* requires a materialised list */
MPCD_HD void list_append(ZoneCtx& z, int lc, int k) {
    z.next[z.off + k] = -1;
    if (z.tail[lc] < 0) z.head[lc] = k;
    else z.next[z.off + z.tail[lc]] = k;
    z.tail[lc] = k;
}

MPCD_HD bool zone_owns(const ZoneCtx& z, int slot) { return !z.owner || z.owner[z.gcid[slot]] == z.zid; }

/* zone_pass.cpp:280-294: new position uniform in the target cell, kept
 * only where this zone owns the particle */
MPCD_HD void move_to(ZoneCtx& z, int lc, int slot, double ux, double uy) {
    const double nx = (double)(z.xs + lc / z.zh) * z.g.dxc + z.g.dxc * ux;
    const double ny = (double)(z.ys + lc % z.zh) * z.g.dyc + z.g.dyc * uy;
    if (zone_owns(z, slot)) {
        z.x[slot] = nx;
        z.y[slot] = ny;
    }
}

/* the bulk/skip masks of a transfer (zone_pass.cpp:492-525) */
MPCD_HD void transfer_masks(ZoneCtx& z, int src, const int* rv, int nr) {
    bool allBulk = z.bulk0[src] != 0;
    for (int k = 0; k < nr && allBulk; ++k)
        if (!z.bulk0[rv[k]]) allBulk = false;
    if (allBulk) {
        z.bb[src] = 1;
        for (int k = 0; k < nr; ++k) z.bb[rv[k]] = 1;
    } else {
        z.skip[src] = 1;
        z.bb[src] = 0;
        for (int k = 0; k < nr; ++k) {
            z.skip[rv[k]] = 1;
            z.bb[rv[k]] = 0;
        }
    }
}

MPCD_HD bool cand_less(const Cand& a, const Cand& b) {
    /* (-deficit, -inward (all equal: weight 0), cnt, cheb, dist2, tie) */
    if (a.deficitNeg != b.deficitNeg) return a.deficitNeg < b.deficitNeg;
    if (a.cnt != b.cnt) return a.cnt < b.cnt;
    if (a.cheb != b.cheb) return a.cheb < b.cheb;
    if (a.dist2 != b.dist2) return a.dist2 < b.dist2;
    return a.tie < b.tie;
}

/* Order-preserving compaction of the zone cells satisfying pred into z.dl;
 * all lanes; returns the count (same on every lane). */
template <class F>
MPCD_HD int zone_compact(ZoneCtx& z, int lane, int nl, F pred) {
    int base = 0;
    for (int b0 = 0; b0 < z.C; b0 += nl) {
        const int lc = b0 + lane;
        const bool f = lc < z.C && pred(lc);
        const unsigned long long msk = wave_ballot(f);
        if (f) z.dl[base + wave_popc(msk & lanes_below(lane))] = lc;
        base += wave_popc(msk);
    }
    return base;
}

/* Fisher-Yates of a[0..m) with partner draws computed in parallel into jd;
 * the swaps are applied in order by lane 0.  `draw(i)` gives j in [0, i]. */
template <class F>
MPCD_HD void zone_shuffle(ZoneCtx& z, int* a, int m, int lane, int nl, F draw) {
    for (int i = 1 + lane; i < m; i += nl) z.jd[i] = (int)draw((uint32_t)i);
    ZBAR();
    if (lane == 0)
        for (int i = m - 1; i >= 1; --i) {
            const int j = z.jd[i];
            const int t = a[i];
            a[i] = a[j];
            a[j] = t;
        }
    ZBAR();
}

/* Admissible targets of dense cell src within Chebyshev range 2 inside the
 * zone, ranked into the round-robin order of rank_dense_targets_unified
 * (zone_pass.cpp:413-490): 8 angular bins from a random start bin, each bin
 * sorted by the candidate key, then one per bin per round.  All lanes; the
 * order lands in z.ord[0..z.nOrd). */
MPCD_HD void dense_targets(ZoneCtx& z, int src, int high, int pass, int phase, int lane, int nl) {
    const int sgx = lc_gx(z, src), sgy = lc_gy(z, src), sgc = lc_gc(z, src);
    /* 24 neighbour offsets in (dy outer, dx inner) order, centre skipped */
    for (int o = lane; o < 24; o += nl) {
        const int f = o < 12 ? o : o + 1;
        const int dy = f / 5 - 2, dx = f % 5 - 2;
        int jx, jy, gc = -1;
        if (shift_idx(z, sgx, sgy, dx, dy, jx, jy)) {
            const int lc = zone_lc(z, jx, jy);
            if (lc >= 0 && (double)z.cnt[lc] < thr(z, lc, high)) gc = jy + z.g.Ny * jx;
        }
        z.nbGc[o] = gc;
    }
    ZBAR();
    int base = 0;
    for (int b0 = 0; b0 < 24; b0 += nl) {
        const int o = b0 + lane;
        bool keep = false;
        int gc = -1;
        if (o < 24) {
            gc = z.nbGc[o];
            keep = gc >= 0;
            for (int q = 0; q < o && keep; ++q) /* duplicates only on grids narrower than 5 cells */
                if (z.nbGc[q] == gc) keep = false;
        }
        const unsigned long long msk = wave_ballot(keep);
        if (keep) {
            Cand& c = z.cand[base + wave_popc(msk & lanes_below(lane))];
            const int jx = gc / z.g.Ny, jy = gc - jx * z.g.Ny;
            const int lc = zone_lc(z, jx, jy);
            int ddx, ddy;
            min_step(z, sgx, sgy, jx, jy, ddx, ddy);
            c.lc = lc;
            c.gc = gc;
            c.bin = z.zp.binTab[(ddy + 2) * 5 + (ddx + 2)] & 7;
            c.deficitNeg = -dmax(0.0, thr(z, lc, high) - (double)z.cnt[lc]);
            c.cnt = z.cnt[lc];
            c.cheb = imax(ddx < 0 ? -ddx : ddx, ddy < 0 ? -ddy : ddy);
            c.dist2 = ddx * ddx + ddy * ddy;
            c.tie = mpcd_draw_tie(z.seed, pass, phase, (uint32_t)sgc, (uint32_t)gc);
        }
        base += wave_popc(msk);
    }
    const int m = base;
    if (lane == 0) {
        z.nOrd = m;
        z.startBin = m > 0 ? mpcd_draw_start_bin(z.seed, pass, phase, (uint32_t)sgc) : 0;
    }
    ZBAR();
    if (m == 0) return;
    /* rank inside the bin: stable order of the candidate key */
    for (int i = lane; i < m; i += nl) {
        const Cand ci = z.cand[i];
        int r = 0;
        for (int j = 0; j < m; ++j) {
            if (j == i || z.cand[j].bin != ci.bin) continue;
            if (cand_less(z.cand[j], ci) || (j < i && !cand_less(ci, z.cand[j]))) ++r;
        }
        z.rnk[i] = r;
    }
    ZBAR();
    /* round robin position: earlier rounds, then earlier bins from startBin */
    for (int i = lane; i < m; i += nl) {
        const int ri = z.rnk[i];
        const int si = (z.cand[i].bin - z.startBin + 8) & 7;
        int pos = 0;
        for (int j = 0; j < m; ++j) {
            const int rj = z.rnk[j];
            if (rj < ri || (rj == ri && ((z.cand[j].bin - z.startBin + 8) & 7) < si)) ++pos;
        }
        z.ord[pos] = z.cand[i].lc;
    }
    ZBAR();
}

/* New-position draws for the first `cnt` particles of a cell's list (the
 * only ones a spread can move), staged in mvx/mvy; all lanes. */
MPCD_HD void stage_dense_draws(ZoneCtx& z, int src, int cnt, int pass, int lane, int nl) {
    if (lane == 0) {
        int k = z.head[src];
        for (int q = 0; q < cnt && k >= 0; ++q) {
            z.mvk[q] = k;
            k = next_k(z, src, k);
        }
    }
    ZBAR();
    for (int q = lane; q < cnt; q += nl) {
        const int slot = slot_of(z, z.mvk[q]);
        mpcd_draw_dense_pos(z.seed, pass, (uint32_t)z.gpid[slot], &z.mvx[q], &z.mvy[q]);
    }
    ZBAR();
}

/* spread_particles_from_dense_source_round_robin (zone_pass.cpp:527-589),
 * lane 0.  `used` counts the staged movers consumed so far. */
MPCD_HD int spread(ZoneCtx& z, int src, int high, int needOut, int pass, int* rv, int& nr, int& used) {
    int moved = 0;
    const int nOrd = z.nOrd;
    if (needOut <= 0 || nOrd == 0) return 0;
    int avail = z.cnt[src];
    if (avail <= 0) return 0;
    while (needOut > 0 && avail > 0) {
        bool progressed = false;
        for (int kk = 0; kk < nOrd; ++kk) {
            const int nb = z.ord[kk];
            const double tnb = thr(z, nb, high);
            const int deficit = imax(0, (int)MPCD_FLOOR(tnb - (double)z.cnt[nb]));
            if (deficit <= 0) continue;
            int remaining = 0;
            for (int jj = kk; jj < nOrd; ++jj) {
                const int q = z.ord[jj];
                if (imax(0, (int)MPCD_FLOOR(thr(z, q, high) - (double)z.cnt[q])) > 0) ++remaining;
            }
            const int quota = imax(1, (int)MPCD_CEIL((double)needOut / (double)imax(1, remaining)));
            const int nt = imax(1, imin(imin(quota, deficit), imin(needOut, avail)));
            ensure_mat(z, src);
            ensure_mat(z, nb);
            for (int q = 0; q < nt; ++q) {
                const int k = z.head[src];
                z.head[src] = z.next[z.off + k];
                if (z.head[src] < 0) z.tail[src] = -1;
                const int slot = slot_of(z, k);
                double ux, uy;
                if (used < ZONE_MAX_MOVE) {
                    ux = z.mvx[used];
                    uy = z.mvy[used];
                } else {
                    mpcd_draw_dense_pos(z.seed, pass, (uint32_t)z.gpid[slot], &ux, &uy);
                }
                move_to(z, nb, slot, ux, uy);
                ++used;
                list_append(z, nb, k);
            }
            avail -= nt;
            z.cnt[nb] += nt;
            bool seen = false;
            for (int k = 0; k < nr; ++k)
                if (rv[k] == nb) seen = true;
            if (!seen) rv[nr++] = nb;
            moved += nt;
            needOut -= nt;
            progressed = true;
            if (needOut <= 0 || avail <= 0) break;
        }
        if (!progressed) break;
    }
    if (moved > 0) z.cnt[src] = avail;
    return moved;
}

/* Dense passes (zone_pass.cpp:708-781), all lanes. */
MPCD_HD void zone_dense_passes(ZoneCtx& z, int lane, int nl) {
    for (int pass = 1; pass <= z.zp.maxPasses; ++pass) {
        for (int lc = lane; lc < z.C; lc += nl) zone_maps(z, lc, 0);
        ZBAR();
        const int m = zone_compact(z, lane, nl, [&](int lc) { return (double)z.cnt[lc] > high_thr(z, lc); });
        zone_shuffle(z, z.dl, m, lane, nl, [&](uint32_t i) { return mpcd_draw_dense_shuffle(z.seed, pass, i); });
        for (int k = 0; k < m; ++k) {
            const int c = z.dl[k];
            const int nc = z.cnt[c];
            if (nc <= 0) continue;
            const double hl = high_thr(z, c);
            if ((double)nc <= hl) continue;
            const int needOut = imax(0, (int)MPCD_CEIL((double)nc - hl));
            int moved = 0;
            if (lane == 0) {
                ++z.nHigh;
                z.nr = 0;
            }
            if (needOut > 0) {
                /* the spreads move at most needOut particles, all from the
                 * head of the list: stage their draws once */
                const int nStage = imin(imin(needOut, nc), ZONE_MAX_MOVE);
                stage_dense_draws(z, c, nStage, pass, lane, nl);
                int used = 0;
                dense_targets(z, c, 0, pass, 0, lane, nl);
                if (lane == 0) {
                    moved += spread(z, c, 0, needOut, pass, z.recvT, z.nr, used);
                    z.stage = used;
                }
                ZBAR();
                const int remaining = imax(0, (int)MPCD_CEIL((double)z.cnt[c] - hl));
                if (remaining > 0) {
                    used = z.stage;
                    dense_targets(z, c, 1, pass, 1, lane, nl);
                    if (lane == 0) moved += spread(z, c, 1, remaining, pass, z.recvT, z.nr, used);
                }
            }
            if (lane == 0) {
                for (int r = 0; r < z.nr; ++r) z.recv[z.recvT[r]] = 1;
                transfer_masks(z, c, z.recvT, z.nr);
                z.movedDense += moved;
            }
            ZBAR();
        }
    }
}

/* inward_direction_from_local_gradient (zone_pass.cpp:375-411) on the
 * sparse-phase count copy; the four smoothed counts in parallel. */
MPCD_HD void inward_dir(ZoneCtx& z, int gx, int gy, int lane, int nl, double& dxo, double& dyo) {
    const int Nx = z.g.Nx, Ny = z.g.Ny;
    /* x pair (xa - xb) * fx, y pair (ya - yb) * fy */
    int xa, xb, ya, yb;
    double fx, fy;
    if (z.g.perx) { xa = wrapi(gx + 1, Nx); xb = wrapi(gx - 1, Nx); fx = 0.5; }
    else if (gx == 0) { xa = gx + 1; xb = gx; fx = 1.0; }
    else if (gx == Nx - 1) { xa = gx; xb = gx - 1; fx = 1.0; }
    else { xa = gx + 1; xb = gx - 1; fx = 0.5; }
    if (z.g.pery) { ya = wrapi(gy + 1, Ny); yb = wrapi(gy - 1, Ny); fy = 0.5; }
    else if (gy == 0) { ya = gy + 1; yb = gy; fy = 1.0; }
    else if (gy == Ny - 1) { ya = gy; yb = gy - 1; fy = 1.0; }
    else { ya = gy + 1; yb = gy - 1; fy = 0.5; }
    for (int t = lane; t < 4; t += nl) {
        const int qx = (t == 0) ? xa : (t == 1) ? xb : gx;
        const int qy = (t == 2) ? ya : (t == 3) ? yb : gy;
        z.nsm4[t] = nsm_at(z, qx, qy, 1);
    }
    ZBAR();
    dxo = (fx == 0.5) ? 0.5 * (z.nsm4[0] - z.nsm4[1]) : (z.nsm4[0] - z.nsm4[1]);
    dyo = (fy == 0.5) ? 0.5 * (z.nsm4[2] - z.nsm4[3]) : (z.nsm4[2] - z.nsm4[3]);
}

/* Sparse pass (zone_pass.cpp:783-854), all lanes. */
MPCD_HD void zone_sparse_pass(ZoneCtx& z, int lane, int nl) {
    for (int lc = lane; lc < z.C; lc += nl) {
        zone_maps(z, lc, 0);
        z.cntS[lc] = z.cnt[lc];
    }
    ZBAR();
    const int m = zone_compact(z, lane, nl, [&](int lc) { return z.cnt[lc] > 0 && (double)z.cnt[lc] < low_thr(z, lc); });
    zone_shuffle(z, z.dl, m, lane, nl, [&](uint32_t i) { return mpcd_draw_low_shuffle(z.seed, i); });
    const int lowInt = (int)MPCD_CEIL(z.zp.lowThrBulk);
    for (int k = 0; k < m; ++k) {
        const int c = z.dl[k];
        const int nc = z.cnt[c];
        if (nc <= 0) continue;
        if (lane == 0) ++z.nLow;
        const int cgx = lc_gx(z, c), cgy = lc_gy(z, c), cgc = lc_gc(z, c);
        /* zone neighbours at range 1 (local_neighbors_in_zone :263-278) */
        for (int o = lane; o < 8; o += nl) {
            const int f = o < 4 ? o : o + 1;
            const int dy = f / 3 - 1, dx = f % 3 - 1;
            int jx, jy, q = -1;
            if (shift_idx(z, cgx, cgy, dx, dy, jx, jy)) q = zone_lc(z, jx, jy);
            z.nbGc[o] = q;
        }
        ZBAR();
        int nn = 0, nd = 0;
        for (int b0 = 0; b0 < 8; b0 += nl) {
            const int o = b0 + lane;
            bool keep = false;
            if (o < 8) {
                const int q = z.nbGc[o];
                keep = q >= 0;
                for (int t = 0; t < o && keep; ++t)
                    if (z.nbGc[t] == q) keep = false;
            }
            const unsigned long long mk = wave_ballot(keep);
            nn += wave_popc(mk);
            const bool don = keep && z.cnt[z.nbGc[o]] > lowInt;
            const unsigned long long md = wave_ballot(don);
            if (don) z.dord[nd + wave_popc(md & lanes_below(lane))] = z.nbGc[o];
            nd += wave_popc(md);
        }
        ZBAR();
        if (nn == 0) continue;
        int need = imax(0, lowInt - nc);
        if (need <= 0 || nd == 0) continue;
        double gxd, gyd;
        inward_dir(z, cgx, cgy, lane, nl, gxd, gyd);
        /* rank_admissible_neighbors_inward (zone_pass.cpp:636-667) */
        zone_shuffle(z, z.dord, nd, lane, nl,
                     [&](uint32_t i) { return mpcd_draw_donor_shuffle(z.seed, (uint32_t)cgc, i); });
        const double ndir = mpcd_hypot(gxd, gyd);
        for (int t = lane; t < nd; t += nl) {
            const int q = z.dord[t];
            int ddx, ddy;
            min_step(z, cgx, cgy, lc_gx(z, q), lc_gy(z, q), ddx, ddy);
            const double nstep = mpcd_hypot((double)ddx, (double)ddy);
            double align = 0.0;
            if (ndir >= 1e-14 && nstep >= 1e-14) align = ((double)ddx * gxd + (double)ddy * gyd) / (ndir * nstep);
            const double noise = 1e-9 * mpcd_draw_donor_noise(z.seed, (uint32_t)cgc, (uint32_t)lc_gc(z, q));
            z.dsc[t] = 100.0 * align + nsm_at(z, lc_gx(z, q), lc_gy(z, q), 1) + noise;
        }
        ZBAR();
        for (int t = lane; t < nd; t += nl) { /* stable, descending score */
            int r = 0;
            for (int u = 0; u < nd; ++u)
                if (z.dsc[u] > z.dsc[t] || (u < t && !(z.dsc[t] > z.dsc[u]))) ++r;
            z.dord2[r] = z.dord[t];
        }
        ZBAR();
        for (int t = 0; t < nd; ++t) {
            const int d = z.dord2[t];
            const int md = z.cnt[d];
            if (md <= 0) continue;
            const int avail = md - lowInt;
            if (avail <= 0) continue;
            const int nt = imin(imin(need, avail), md);
            if (nt <= 0) continue;
            /* copy the donor list, Fisher-Yates it, move the first nt */
            int* tmp = z.tmpg + z.off;
            if (lane == 0) {
                int kk = z.head[d];
                for (int q = 0; q < md; ++q) {
                    tmp[q] = kk;
                    kk = next_k(z, d, kk);
                }
            }
            const int dgc = lc_gc(z, d);
            zone_shuffle(z, tmp, md, lane, nl,
                         [&](uint32_t i) { return mpcd_draw_particle_shuffle(z.seed, (uint32_t)cgc, (uint32_t)dgc, i); });
            for (int q = lane; q < nt && q < ZONE_MAX_MOVE; q += nl) {
                const int slot = slot_of(z, tmp[q]);
                mpcd_draw_sparse_pos(z.seed, (uint32_t)z.gpid[slot], &z.mvx[q], &z.mvy[q]);
            }
            ZBAR();
            if (lane == 0) {
                ensure_mat(z, c);
                for (int q = 0; q < nt; ++q) {
                    const int slot = slot_of(z, tmp[q]);
                    double ux, uy;
                    if (q < ZONE_MAX_MOVE) {
                        ux = z.mvx[q];
                        uy = z.mvy[q];
                    } else {
                        mpcd_draw_sparse_pos(z.seed, (uint32_t)z.gpid[slot], &ux, &uy);
                    }
                    move_to(z, c, slot, ux, uy);
                    list_append(z, c, tmp[q]);
                }
                z.mat[d] = 1;
                z.head[d] = -1;
                z.tail[d] = -1;
                for (int q = nt; q < md; ++q) list_append(z, d, tmp[q]);
                z.cnt[d] = md - nt;
                z.cnt[c] += nt;
                z.recv[c] = 1;
                int one[1] = {c};
                transfer_masks(z, d, one, 1);
                z.movedSparse += nt;
            }
            ZBAR();
            need -= nt;
            if (need <= 0) break;
        }
    }
}

/* Zone setup (zone_pass.cpp:680-685), all lanes: counts, implicit list of
 * each cell (its run of local indices, from a wave prefix sum over the
 * zone's cells) and the column map. */
MPCD_HD void zone_setup(ZoneCtx& z, int lane, int nl) {
    int base = 0;
    for (int b0 = 0; b0 < z.C; b0 += nl) {
        const int lc = b0 + lane;
        const int m = lc < z.C ? z.gcnt[lc_gc(z, lc)] : 0;
        const int incl = wave_shfl_up_add(m, lane);
        if (lc < z.C) {
            const int k0 = base + incl - m;
            z.cnt[lc] = m;
            z.k0[lc] = k0;
            z.head[lc] = m > 0 ? k0 : -1;
            z.tail[lc] = m > 0 ? k0 + m - 1 : -1;
            z.recv[lc] = 0;
            z.bb[lc] = 0;
            z.skip[lc] = 0;
            z.mat[lc] = 0;
            z.corr[lc] = 0;
            z.dux[lc] = 0.0;
            z.duy[lc] = 0.0;
            if (lc % z.zh == 0) {
                z.colL[lc / z.zh] = k0;
                z.colG[lc / z.zh] = z.gstart[lc_gc(z, lc)];
            }
        }
        /* lanes past the zone add 0, so the last lane holds the chunk total */
        base += wave_bcast(incl, nl - 1);
    }
    if (lane == 0) {
        z.P = base;
        z.nHigh = 0;
        z.nLow = 0;
        z.movedDense = 0;
        z.movedSparse = 0;
        z.stats4 = 0.0;
        z.stats5 = 0.0;
        z.status = 0;
    }
}

/* Momentum correction (zone_pass.cpp:870-909), first half: the velocity
 * change of a changed cell, from the momentum of its input run (particle
 * order) and of its final list.  Reads only; a cell nothing changed has
 * the same list before and after, so its change is exactly zero. */
MPCD_HD void zone_correct_sums(ZoneCtx& z, int lc) {
    const int m = z.cnt[lc];
    if (!z.mat[lc] || m == 0) return;
    const int gc = lc_gc(z, lc);
    const int s0 = z.gstart[gc], m0 = z.gcnt[gc];
    double pbx = 0.0, pby = 0.0;
    for (int j = 0; j < m0; ++j) {
        const int slot = z.glist[s0 + j];
        pbx += z.vx[slot];
        pby += z.vy[slot];
    }
    double px = 0.0, py = 0.0;
    for (int k = z.head[lc]; k >= 0; k = z.next[z.off + k]) {
        const int slot = slot_of(z, k);
        px += z.vx[slot];
        py += z.vy[slot];
    }
    const double dPx = pbx - px, dPy = pby - py;
    const double dux = dPx / (double)m, duy = dPy / (double)m;
    z.dux[lc] = dux;
    z.duy[lc] = duy;
    bool corr = z.zp.corrEnabled != 0;
    if (corr && z.skip[lc]) corr = false;
    if (corr && z.recv[lc] && !z.bb[lc]) corr = false;
    if (corr && !(dux != 0.0 || duy != 0.0)) corr = false;
    z.corr[lc] = corr ? 1 : 0;
}

MPCD_HD void put_velocity(ZoneCtx& z, int slot, double vx, double vy) {
    if (z.gcover && z.gcover[z.gcid[slot]]) {
        z.vsx[slot] = vx;
        z.vsy[slot] = vy;
    } else {
        z.vx[slot] = vx;
        z.vy[slot] = vy;
    }
}

/* Second half: write the velocities of the owned particles of a changed
 * cell; for an unchanged cell of the overlap band, pass its owned particles'
 * input velocities through to the side buffer. */
MPCD_HD void zone_correct_write(ZoneCtx& z, int lc) {
    if (z.mat[lc]) {
        const bool corr = z.corr[lc] != 0;
        const double dux = z.dux[lc], duy = z.duy[lc];
        for (int k = z.head[lc]; k >= 0; k = z.next[z.off + k]) {
            const int slot = slot_of(z, k);
            if (!zone_owns(z, slot)) continue;
            double vx = z.vx[slot], vy = z.vy[slot];
            if (corr) {
                vx += dux;
                vy += duy;
            }
            put_velocity(z, slot, vx, vy);
        }
        return;
    }
    const int gc = lc_gc(z, lc);
    if (!z.gcover || !z.gcover[gc]) return;
    const int s0 = z.gstart[gc], m0 = z.gcnt[gc];
    for (int j = 0; j < m0; ++j) {
        const int slot = z.glist[s0 + j];
        if (!zone_owns(z, slot)) continue;
        z.vsx[slot] = z.vx[slot];
        z.vsy[slot] = z.vy[slot];
    }
}

/* stats[4], stats[5] of the zone (zone_pass.cpp:911-922), lane 0. */
MPCD_HD void zone_final_stats(ZoneCtx& z) {
    double acc = 0.0, mx = 0.0;
    int nDU = 0;
    for (int lc = 0; lc < z.C; ++lc) {
        const double d = z.dux[lc] * z.dux[lc] + z.duy[lc] * z.duy[lc];
        mx = dmax(mx, MPCD_SQRT(d));
        if (d > 0.0) {
            acc += d;
            ++nDU;
        }
    }
    if (nDU > 0) {
        z.stats4 = MPCD_SQRT(acc / (double)nDU);
        z.stats5 = mx;
    }
}

struct ZoneRect {
    int xs, xe, ys, ye;
};

struct ZoneOut {
    int nHigh, nLow, movedDense, movedSparse;
    double stats4, stats5;
};

/* bytes of the per-zone working set for C cells and W columns */
MPCD_HD size_t zone_bytes(int C, int W, int useLocal) {
    size_t b = 0;
    b += (size_t)C * sizeof(double) * (useLocal ? 4 : 2); /* dux duy (+ lowA highA) */
    b += (size_t)C * (7 * sizeof(int));                    /* cnt head tail dl cntS jd k0 */
    b += (size_t)W * (2 * sizeof(int));                    /* colL colG */
    b += (size_t)C * 6;                                    /* bulk0 recv bb skip mat corr */
    b = (b + 15) & ~(size_t)15;
    b += 24 * sizeof(Cand) + 24 * 4 * sizeof(int) + 48 * sizeof(int) + 8 * sizeof(double) + 3 * 8 * sizeof(int) +
         4 * sizeof(double) + ZONE_MAX_MOVE * (2 * sizeof(double) + sizeof(int));
    return (b + 15) & ~(size_t)15;
}

MPCD_HD void zone_carve(ZoneCtx& z, unsigned char* base, int C, int W, int useLocal) {
    unsigned char* p = base;
    z.dux = (double*)p; p += (size_t)C * sizeof(double);
    z.duy = (double*)p; p += (size_t)C * sizeof(double);
    if (useLocal) {
        z.lowA = (double*)p; p += (size_t)C * sizeof(double);
        z.highA = (double*)p; p += (size_t)C * sizeof(double);
    } else {
        z.lowA = nullptr;
        z.highA = nullptr;
    }
    z.cnt = (int*)p; p += (size_t)C * sizeof(int);
    z.head = (int*)p; p += (size_t)C * sizeof(int);
    z.tail = (int*)p; p += (size_t)C * sizeof(int);
    z.dl = (int*)p; p += (size_t)C * sizeof(int);
    z.cntS = (int*)p; p += (size_t)C * sizeof(int);
    z.jd = (int*)p; p += (size_t)C * sizeof(int);
    z.k0 = (int*)p; p += (size_t)C * sizeof(int);
    z.colL = (int*)p; p += (size_t)W * sizeof(int);
    z.colG = (int*)p; p += (size_t)W * sizeof(int);
    z.bulk0 = p; p += C;
    z.recv = p; p += C;
    z.bb = p; p += C;
    z.skip = p; p += C;
    z.mat = p; p += C;
    z.corr = p; p += C;
    p = base + ((size_t)(p - base + 15) & ~(size_t)15);
    z.cand = (Cand*)p; p += 24 * sizeof(Cand);
    z.ord = (int*)p; p += 24 * sizeof(int);
    z.nbGc = (int*)p; p += 24 * sizeof(int);
    z.rnk = (int*)p; p += 24 * sizeof(int);
    p += 24 * sizeof(int);
    z.recvT = (int*)p; p += 48 * sizeof(int);
    z.dsc = (double*)p; p += 8 * sizeof(double);
    z.dord = (int*)p; p += 8 * sizeof(int);
    z.dord2 = (int*)p; p += 8 * sizeof(int);
    p += 8 * sizeof(int);
    z.nsm4 = (double*)p; p += 4 * sizeof(double);
    z.mvx = (double*)p; p += ZONE_MAX_MOVE * sizeof(double);
    z.mvy = (double*)p; p += ZONE_MAX_MOVE * sizeof(double);
    z.mvk = (int*)p;
}

}  // namespace mpcdk
