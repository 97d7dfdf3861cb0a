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
    if (m <= 0) return;
    double sx = 0.0, sy = 0.0;
    for (int k = 0; k < m; ++k) {
        sx += vx[ids[k]];
        sy += vy[ids[k]];
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
            const double vxr = vx[ids[k]] - ux, vyr = vy[ids[k]] - uy;
            const double rx = ct * vxr - st * vyr;
            const double ry = st * vxr + ct * vyr;
            s2 += rx * rx + ry * ry;
        }
        const int dof = 2 * imax(m - 1, 0);
        if (m > 1 && s2 > 0.0) lam = MPCD_SQRT((double)dof * cp.kBT / s2);
    }
    for (int k = 0; k < m; ++k) {
        const int i = ids[k];
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
    CellStats s;
    s.N = m;
    s.Px = 0.0;
    s.Py = 0.0;
    for (int k = 0; k < m; ++k) {
        s.Px += vx[ids[k]];
        s.Py += vy[ids[k]];
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
        const double dvx = vx[ids[k]] - s.Ux, dvy = vy[ids[k]] - s.Uy;
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

/* ---------- zone redistribution (zone_pass.cpp:62-973) ---------- */

struct ZoneParams {
    double highThr, lowThrBulk;
    double lowFrac, highFrac; /* 1/(1+coef), 1+coef */
    double gamma, fluidFracGain, lowThrFloor, highThrFloor;
    int useLocal, maxPasses, corrEnabled;
    signed char binTab[25]; /* angular bin of (dx,dy) in [-2,2]^2: index (dy+2)*5+(dx+2) */
};

/* Candidate of a dense cell's target ranking. */
struct Cand {
    int lc, gc;
    double deficitNeg;
    int cnt, cheb, dist2;
    double tie;
};

/* Working set of one zone.  Cell arrays are indexed by the zone-local cell
 * lc = (ix-xs)*zh + (iy-ys) (ascending global cell id); particles by the
 * zone-local index k (the zone's cells' runs of the binning list, in lc
 * order).  Per-particle scratch lives at [off, off+P) of next/tmpg/slotOf,
 * private to the zone: the shifted layout's zones overlap
 * (zone_pass.cpp:191-219 starts them at 0 and T/2 with width T), so two
 * zones may hold the same particle.  Reads come from the pass-input
 * snapshot; a zone writes only the particles it owns, i.e. whose input cell
 * has this zone as its owner-map entry (the last zone covering the cell,
 * which is the zone whose result the reference's merge keeps,
 * zone_pass_openmp.cpp:145-153).  owner == nullptr: the zone owns all. */
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
    /* state: positions updated in place (nobody reads them during a pass),
     * velocities read from vxin/vyin and written to vxout/vyout */
    double *x, *y;
    const double *vxin, *vyin;
    double *vxout, *vyout;
    int* next;   /* [off, off+P): list links by local index */
    int* tmpg;   /* [off, off+P): donor shuffles */
    int* slotOf; /* [off, off+P): slot of each local index */
    /* zone cell arrays */
    int *cnt, *head, *tail, *dl, *cntS, *colL, *colG;
    double *lowA, *highA, *PbX, *PbY, *dU2;
    unsigned char *bulk0, *recv, *bb, *skip;
    /* candidate scratch */
    Cand* cand;   /* 24 */
    int* ord;     /* 24 */
    int* recvT;   /* 48 */
    double* dsc;  /* 8 donor scores */
    int* dord;    /* 8 */
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

/* Active thresholds of a zone cell from the current counts
 * (zone_pass.cpp:321-345 and :715-725). */
MPCD_HD void zone_maps(ZoneCtx& z, int lc, int use_copy) {
    double f = 1.0;
    if (z.zp.useLocal) {
        const double fRaw = nsm_at(z, lc_gx(z, lc), lc_gy(z, lc), use_copy) / dmax(z.zp.gamma, 1e-12);
        const double t = z.zp.fluidFracGain * fRaw;
        f = (t < 0.0) ? 0.0 : ((1.0 < t) ? 1.0 : t); /* std::clamp(t, 0, 1) */
    }
    const double gl = z.zp.gamma * f;
    const double lowLoc = dmax(z.zp.lowFrac * gl, z.zp.lowThrFloor);
    const double highLoc = dmax(z.zp.highFrac * gl, z.zp.highThrFloor);
    const int bulk = z.cnt[lc] > 0;
    z.bulk0[lc] = (unsigned char)bulk;
    z.lowA[lc] = bulk ? z.zp.lowThrBulk : lowLoc;
    z.highA[lc] = bulk ? z.zp.highThr : highLoc;
}

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

/* unique_stable of a touched list then the bulk/skip masks
 * (zone_pass.cpp:492-525). */
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

/* Admissible targets of a dense cell within Chebyshev range 2 inside the zone,
 * ranked into the round-robin order of rank_dense_targets_unified
 * (zone_pass.cpp:413-490).  Writes zone-local cells to z.ord, returns count. */
MPCD_HD int dense_targets(ZoneCtx& z, int src, const double* target, int pass, int phase) {
    const int sgx = lc_gx(z, src), sgy = lc_gy(z, src), sgc = lc_gc(z, src);
    int m = 0;
    for (int dy = -2; dy <= 2; ++dy)
        for (int dx = -2; dx <= 2; ++dx) {
            if (dx == 0 && dy == 0) continue;
            int jx, jy;
            if (!shift_idx(z, sgx, sgy, dx, dy, jx, jy)) continue;
            const int lc = zone_lc(z, jx, jy);
            if (lc < 0) continue;
            if (!((double)z.cnt[lc] < target[lc])) continue;
            bool dup = false;
            for (int k = 0; k < m; ++k)
                if (z.cand[k].lc == lc) dup = true;
            if (dup) continue;
            Cand& c = z.cand[m++];
            c.lc = lc;
            c.gc = jy + z.g.Ny * jx;
        }
    if (m == 0) return 0;
    const int startBin = mpcd_draw_start_bin(z.seed, pass, phase, (uint32_t)sgc);
    int bin[24];
    int binCount[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < m; ++k) {
        Cand& c = z.cand[k];
        int ddx, ddy;
        min_step(z, sgx, sgy, lc_gx(z, c.lc), lc_gy(z, c.lc), ddx, ddy);
        bin[k] = z.zp.binTab[(ddy + 2) * 5 + (ddx + 2)] & 7;
        binCount[bin[k]]++;
        c.deficitNeg = -dmax(0.0, target[c.lc] - (double)z.cnt[c.lc]);
        c.cnt = z.cnt[c.lc];
        c.cheb = imax(ddx < 0 ? -ddx : ddx, ddy < 0 ? -ddy : ddy);
        c.dist2 = ddx * ddx + ddy * ddy;
        c.tie = mpcd_draw_tie(z.seed, pass, phase, (uint32_t)sgc, (uint32_t)c.gc);
    }
    /* stable per-bin order: bucket in admissible order, then insertion sort */
    int off[8];
    int acc = 0;
    for (int b = 0; b < 8; ++b) {
        off[b] = acc;
        acc += binCount[b];
    }
    int fill[8];
    for (int b = 0; b < 8; ++b) fill[b] = off[b];
    int byBin[24];
    for (int k = 0; k < m; ++k) byBin[fill[bin[k]]++] = k;
    for (int b = 0; b < 8; ++b) {
        int* a = byBin + off[b];
        for (int i = 1; i < binCount[b]; ++i) {
            const int key = a[i];
            int j = i - 1;
            while (j >= 0 && cand_less(z.cand[key], z.cand[a[j]])) {
                a[j + 1] = a[j];
                --j;
            }
            a[j + 1] = key;
        }
    }
    int nOut = 0;
    for (int r = 0; nOut < m; ++r)
        for (int s = 0; s < 8; ++s) {
            const int b = (startBin + s) % 8;
            if (r < binCount[b]) z.ord[nOut++] = z.cand[byBin[off[b] + r]].lc;
        }
    return nOut;
}

/* spread_particles_from_dense_source_round_robin (zone_pass.cpp:527-589).
 * Appends receiving cells (unique, first-seen order) to rv. */
MPCD_HD int spread(ZoneCtx& z, int src, int nOrd, const double* target, int needOut, int pass, int* rv, int& nr) {
    int moved = 0;
    if (needOut <= 0 || nOrd == 0) return 0;
    int avail = z.cnt[src];
    if (avail <= 0) return 0;
    while (needOut > 0 && avail > 0) {
        bool progressed = false;
        for (int kk = 0; kk < nOrd; ++kk) {
            const int nb = z.ord[kk];
            const int deficit = imax(0, (int)MPCD_FLOOR(target[nb] - (double)z.cnt[nb]));
            if (deficit <= 0) continue;
            int remaining = 0;
            for (int jj = kk; jj < nOrd; ++jj) {
                const int q = z.ord[jj];
                if (imax(0, (int)MPCD_FLOOR(target[q] - (double)z.cnt[q])) > 0) ++remaining;
            }
            const int quota = imax(1, (int)MPCD_CEIL((double)needOut / (double)imax(1, remaining)));
            const int nt = imax(1, imin(imin(quota, deficit), imin(needOut, avail)));
            for (int q = 0; q < nt; ++q) {
                const int k = z.head[src];
                z.head[src] = z.next[z.off + k];
                if (z.head[src] < 0) z.tail[src] = -1;
                const int slot = z.slotOf[z.off + k];
                double ux, uy;
                mpcd_draw_dense_pos(z.seed, pass, (uint32_t)z.gpid[slot], &ux, &uy);
                move_to(z, nb, slot, ux, uy);
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

/* Dense passes (zone_pass.cpp:708-781), run by one lane. */
MPCD_HD void zone_dense_passes(ZoneCtx& z) {
    for (int pass = 1; pass <= z.zp.maxPasses; ++pass) {
        for (int lc = 0; lc < z.C; ++lc) zone_maps(z, lc, 0);
        int m = 0;
        for (int lc = 0; lc < z.C; ++lc)
            if ((double)z.cnt[lc] > z.highA[lc]) z.dl[m++] = lc;
        for (int i = m - 1; i >= 1; --i) {
            const int j = (int)mpcd_draw_dense_shuffle(z.seed, pass, (uint32_t)i);
            const int t = z.dl[i];
            z.dl[i] = z.dl[j];
            z.dl[j] = t;
        }
        for (int k = 0; k < m; ++k) {
            const int c = z.dl[k];
            const int nc = z.cnt[c];
            if (nc <= 0) continue;
            const double hl = z.highA[c];
            if ((double)nc <= hl) continue;
            ++z.nHigh;
            int nr = 0;
            int moved = 0;
            const int needOut = imax(0, (int)MPCD_CEIL((double)z.cnt[c] - hl));
            if (needOut > 0) {
                int no = dense_targets(z, c, z.lowA, pass, 0);
                moved += spread(z, c, no, z.lowA, needOut, pass, z.recvT, nr);
                const int remaining = imax(0, (int)MPCD_CEIL((double)z.cnt[c] - hl));
                if (remaining > 0) {
                    no = dense_targets(z, c, z.highA, pass, 1);
                    moved += spread(z, c, no, z.highA, remaining, pass, z.recvT, nr);
                }
            }
            for (int r = 0; r < nr; ++r) z.recv[z.recvT[r]] = 1;
            transfer_masks(z, c, z.recvT, nr);
            z.movedDense += moved;
        }
    }
}

/* inward_direction_from_local_gradient (zone_pass.cpp:375-411) on the
 * sparse-phase count copy. */
MPCD_HD void inward_dir(const ZoneCtx& z, int gx, int gy, double& dxo, double& dyo) {
    const int Nx = z.g.Nx, Ny = z.g.Ny;
    if (z.g.perx) dxo = 0.5 * (nsm_at(z, wrapi(gx + 1, Nx), gy, 1) - nsm_at(z, wrapi(gx - 1, Nx), gy, 1));
    else if (gx == 0) dxo = nsm_at(z, gx + 1, gy, 1) - nsm_at(z, gx, gy, 1);
    else if (gx == Nx - 1) dxo = nsm_at(z, gx, gy, 1) - nsm_at(z, gx - 1, gy, 1);
    else dxo = 0.5 * (nsm_at(z, gx + 1, gy, 1) - nsm_at(z, gx - 1, gy, 1));
    if (z.g.pery) dyo = 0.5 * (nsm_at(z, gx, wrapi(gy + 1, Ny), 1) - nsm_at(z, gx, wrapi(gy - 1, Ny), 1));
    else if (gy == 0) dyo = nsm_at(z, gx, gy + 1, 1) - nsm_at(z, gx, gy, 1);
    else if (gy == Ny - 1) dyo = nsm_at(z, gx, gy, 1) - nsm_at(z, gx, gy - 1, 1);
    else dyo = 0.5 * (nsm_at(z, gx, gy + 1, 1) - nsm_at(z, gx, gy - 1, 1));
}

/* Sparse pass (zone_pass.cpp:783-854), run by one lane. */
MPCD_HD void zone_sparse_pass(ZoneCtx& z) {
    for (int lc = 0; lc < z.C; ++lc) zone_maps(z, lc, 0);
    for (int lc = 0; lc < z.C; ++lc) z.cntS[lc] = z.cnt[lc];
    int m = 0;
    for (int lc = 0; lc < z.C; ++lc)
        if (z.cnt[lc] > 0 && (double)z.cnt[lc] < z.lowA[lc]) z.dl[m++] = lc;
    for (int i = m - 1; i >= 1; --i) {
        const int j = (int)mpcd_draw_low_shuffle(z.seed, (uint32_t)i);
        const int t = z.dl[i];
        z.dl[i] = z.dl[j];
        z.dl[j] = t;
    }
    const int lowInt = (int)MPCD_CEIL(z.zp.lowThrBulk);
    for (int k = 0; k < m; ++k) {
        const int c = z.dl[k];
        const int nc = z.cnt[c];
        if (nc <= 0) continue;
        ++z.nLow;
        const int cgx = lc_gx(z, c), cgy = lc_gy(z, c), cgc = lc_gc(z, c);
        int neigh[8];
        int nn = 0;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                if (dx == 0 && dy == 0) continue;
                int jx, jy;
                if (!shift_idx(z, cgx, cgy, dx, dy, jx, jy)) continue;
                const int q = zone_lc(z, jx, jy);
                if (q < 0) continue;
                bool dup = false;
                for (int t = 0; t < nn; ++t)
                    if (neigh[t] == q) dup = true;
                if (!dup) neigh[nn++] = q;
            }
        if (nn == 0) continue;
        int need = imax(0, lowInt - nc);
        if (need <= 0) continue;
        int nd = 0;
        for (int t = 0; t < nn; ++t)
            if (z.cnt[neigh[t]] > lowInt) z.dord[nd++] = neigh[t];
        if (nd == 0) continue;
        double gxd, gyd;
        inward_dir(z, cgx, cgy, gxd, gyd);
        /* rank_admissible_neighbors_inward (zone_pass.cpp:636-667) */
        for (int i = nd - 1; i >= 1; --i) {
            const int j = (int)mpcd_draw_donor_shuffle(z.seed, (uint32_t)cgc, (uint32_t)i);
            const int t = z.dord[i];
            z.dord[i] = z.dord[j];
            z.dord[j] = t;
        }
        const double ndir = mpcd_hypot(gxd, gyd);
        for (int t = 0; t < nd; ++t) {
            const int q = z.dord[t];
            int ddx, ddy;
            min_step(z, cgx, cgy, lc_gx(z, q), lc_gy(z, q), ddx, ddy);
            const double nstep = mpcd_hypot((double)ddx, (double)ddy);
            double align = 0.0;
            if (ndir >= 1e-14 && nstep >= 1e-14) align = ((double)ddx * gxd + (double)ddy * gyd) / (ndir * nstep);
            const double noise = 1e-9 * mpcd_draw_donor_noise(z.seed, (uint32_t)cgc, (uint32_t)lc_gc(z, q));
            z.dsc[t] = 100.0 * align + nsm_at(z, lc_gx(z, q), lc_gy(z, q), 1) + noise;
        }
        for (int i = 1; i < nd; ++i) { /* stable, descending score */
            const int kq = z.dord[i];
            const double ks = z.dsc[i];
            int j = i - 1;
            while (j >= 0 && ks > z.dsc[j]) {
                z.dord[j + 1] = z.dord[j];
                z.dsc[j + 1] = z.dsc[j];
                --j;
            }
            z.dord[j + 1] = kq;
            z.dsc[j + 1] = ks;
        }
        for (int t = 0; t < nd; ++t) {
            const int d = z.dord[t];
            const int md = z.cnt[d];
            if (md <= 0) continue;
            const int avail = md - lowInt;
            if (avail <= 0) continue;
            const int nt = imin(imin(need, avail), md);
            if (nt <= 0) continue;
            /* copy the donor list, Fisher-Yates it, move the first nt */
            int* tmp = z.tmpg + z.off;
            int k = z.head[d];
            for (int q = 0; q < md; ++q) {
                tmp[q] = k;
                k = z.next[z.off + k];
            }
            for (int i = md - 1; i >= 1; --i) {
                const int j = (int)mpcd_draw_particle_shuffle(z.seed, (uint32_t)cgc, (uint32_t)lc_gc(z, d), (uint32_t)i);
                const int tv = tmp[i];
                tmp[i] = tmp[j];
                tmp[j] = tv;
            }
            for (int q = 0; q < nt; ++q) {
                const int slot = z.slotOf[z.off + tmp[q]];
                double ux, uy;
                mpcd_draw_sparse_pos(z.seed, (uint32_t)z.gpid[slot], &ux, &uy);
                move_to(z, c, slot, ux, uy);
                list_append(z, c, tmp[q]);
            }
            z.head[d] = -1;
            z.tail[d] = -1;
            for (int q = nt; q < md; ++q) list_append(z, d, tmp[q]);
            z.cnt[d] = md - nt;
            z.cnt[c] += nt;
            z.recv[c] = 1;
            int one[1] = {c};
            transfer_masks(z, d, one, 1);
            z.movedSparse += nt;
            need -= nt;
            if (need <= 0) break;
        }
    }
}

/* Zone setup (zone_pass.cpp:680-685).  Lane 0 first: counts, local list
 * heads and the column map; then zone_setup_cell per cell on every lane:
 * slots, links and the momentum before the pass. */
MPCD_HD void zone_setup_serial(ZoneCtx& z) {
    int accl = 0;
    for (int col = 0; col < z.zw; ++col) {
        const int gx = z.xs + col;
        z.colL[col] = accl;
        z.colG[col] = z.gstart[z.ys + z.g.Ny * gx];
        for (int r = 0; r < z.zh; ++r) {
            const int lc = col * z.zh + r;
            const int m = z.gcnt[(z.ys + r) + z.g.Ny * gx];
            z.cnt[lc] = m;
            z.head[lc] = m > 0 ? accl : -1;
            z.tail[lc] = m > 0 ? accl + m - 1 : -1;
            accl += m;
        }
    }
    z.P = accl;
    z.nHigh = 0;
    z.nLow = 0;
    z.movedDense = 0;
    z.movedSparse = 0;
    z.stats4 = 0.0;
    z.stats5 = 0.0;
    z.status = 0;
}

MPCD_HD void zone_setup_cell(ZoneCtx& z, int lc) {
    const int col = lc / z.zh;
    const int gc = lc_gc(z, lc);
    const int s0 = z.gstart[gc], m = z.gcnt[gc];
    const int k0 = z.colL[col] + (s0 - z.colG[col]);
    double px = 0.0, py = 0.0;
    for (int j = 0; j < m; ++j) {
        const int slot = z.glist[s0 + j];
        px += z.vxin[slot];
        py += z.vyin[slot];
        z.slotOf[z.off + k0 + j] = slot;
        z.next[z.off + k0 + j] = (j + 1 < m) ? k0 + j + 1 : -1;
    }
    z.PbX[lc] = px;
    z.PbY[lc] = py;
    z.recv[lc] = 0;
    z.bb[lc] = 0;
    z.skip[lc] = 0;
    z.dU2[lc] = 0.0;
}

/* Momentum correction of one zone cell (zone_pass.cpp:879-909); writes the
 * output velocity of every particle of the cell this zone owns. */
MPCD_HD void zone_correct_cell(ZoneCtx& z, int lc) {
    const int m = z.cnt[lc];
    if (m == 0) return;
    double px = 0.0, py = 0.0;
    for (int k = z.head[lc]; k >= 0; k = z.next[z.off + k]) {
        const int slot = z.slotOf[z.off + k];
        px += z.vxin[slot];
        py += z.vyin[slot];
    }
    const double dPx = z.PbX[lc] - px, dPy = z.PbY[lc] - py;
    const double dux = dPx / (double)m, duy = dPy / (double)m;
    z.dU2[lc] = dux * dux + duy * duy;
    bool corr = z.zp.corrEnabled != 0;
    if (corr && z.skip[lc]) corr = false;
    if (corr && z.recv[lc] && !z.bb[lc]) corr = false;
    if (corr && !(dux != 0.0 || duy != 0.0)) corr = false;
    for (int k = z.head[lc]; k >= 0; k = z.next[z.off + k]) {
        const int slot = z.slotOf[z.off + k];
        if (!zone_owns(z, slot)) continue;
        double vx = z.vxin[slot], vy = z.vyin[slot];
        if (corr) {
            vx += dux;
            vy += duy;
        }
        z.vxout[slot] = vx;
        z.vyout[slot] = vy;
    }
}

/* stats[4], stats[5] of the zone (zone_pass.cpp:911-922), lane 0. */
MPCD_HD void zone_final_stats(ZoneCtx& z) {
    double acc = 0.0, mx = 0.0;
    int nDU = 0;
    for (int lc = 0; lc < z.C; ++lc) {
        const double d = z.dU2[lc];
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



/* bytes of the per-zone working set for `C` cells and `W` columns */
MPCD_HD size_t zone_bytes(int C, int W) {
    size_t b = 0;
    b += (size_t)C * (5 * sizeof(double));          /* lowA highA PbX PbY dU2 */
    b += (size_t)C * (5 * sizeof(int));             /* cnt head tail dl cntS */
    b += (size_t)W * (2 * sizeof(int));             /* colL colG */
    b += (size_t)C * 4;                             /* bulk0 recv bb skip */
    b = (b + 15) & ~(size_t)15;
    b += 24 * sizeof(Cand) + 24 * sizeof(int) + 48 * sizeof(int) + 8 * sizeof(double) + 8 * sizeof(int);
    return (b + 15) & ~(size_t)15;
}

MPCD_HD void zone_carve(ZoneCtx& z, unsigned char* base, int C, int W) {
    unsigned char* p = base;
    z.lowA = (double*)p; p += (size_t)C * sizeof(double);
    z.highA = (double*)p; p += (size_t)C * sizeof(double);
    z.PbX = (double*)p; p += (size_t)C * sizeof(double);
    z.PbY = (double*)p; p += (size_t)C * sizeof(double);
    z.dU2 = (double*)p; p += (size_t)C * sizeof(double);
    z.cnt = (int*)p; p += (size_t)C * sizeof(int);
    z.head = (int*)p; p += (size_t)C * sizeof(int);
    z.tail = (int*)p; p += (size_t)C * sizeof(int);
    z.dl = (int*)p; p += (size_t)C * sizeof(int);
    z.cntS = (int*)p; p += (size_t)C * sizeof(int);
    z.colL = (int*)p; p += (size_t)W * sizeof(int);
    z.colG = (int*)p; p += (size_t)W * sizeof(int);
    z.bulk0 = p; p += C;
    z.recv = p; p += C;
    z.bb = p; p += C;
    z.skip = p; p += C;
    p = base + ((size_t)(p - base + 15) & ~(size_t)15);
    z.cand = (Cand*)p; p += 24 * sizeof(Cand);
    z.ord = (int*)p; p += 24 * sizeof(int);
    z.recvT = (int*)p; p += 48 * sizeof(int);
    z.dsc = (double*)p; p += 8 * sizeof(double);
    z.dord = (int*)p;
}


}  // namespace mpcdk
