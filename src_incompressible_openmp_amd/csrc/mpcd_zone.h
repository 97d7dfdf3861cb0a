/*
 * mpcd_zone.h -- one zone of the snapshot zone redistribution
 * (zone_pass.cpp:62-973) on one wavefront, shared by the gfx950 kernel in
 * mpcd_engine.hip and the host logic check in tests/ (ZHD code).
 *
 * The reference's zone kernel is sequential: a shuffled list of dense cells,
 * each of which mutates the counts its successors read.  That order is kept;
 * inside each step the wavefront works together wherever the work is
 * independent by construction: all counter-based draws of a step (shuffle
 * partners, tie breaks, score noise), candidate evaluation and ranking,
 * order-preserving compactions and prefix sums.  On the device a dense cell's
 * candidates and movers are held in lanes and its spread is applied by the
 * lanes at once (dense_cell_wave); the one-lane form of the same functions
 * runs on the host (nl = 1), where the wave primitives reduce to scalars and
 * the barrier to nothing.  A move only records, per particle, the dense pass
 * or the sparse phase that moved it last; the momentum correction, which
 * visits every particle in its final cell, writes the new positions from
 * that record.  On the device the correction runs in a kernel of its own
 * (the zone's state handed over in global memory, see ZoneState).
 *
 * Memory.  Cell arrays are indexed by the zone-local cell
 * lc = (ix-xs)*zh + (iy-ys) (ascending global cell id); particles by the
 * zone-local index k (the zone's cells' runs of the pass-input binning list,
 * in lc order).  Index arrays use the type IT: uint16_t when the zone has
 * fewer than 65535 cells and particles (the cell arrays and most list links
 * then fit in LDS next to three other zones on a CU), int otherwise.  List
 * links of k < nL live in LDS, the rest in the zone's private global region.
 *
 * Snapshot semantics.  The shifted layout's zones overlap
 * (zone_pass.cpp:191-219 starts them at 0 and T/2 with width T), so two zones
 * may hold the same particle.  Every zone reads the pass input; a zone writes
 * only the particles it owns, i.e. whose input cell has this zone as its
 * owner-map entry (the last zone covering the cell, which is the zone whose
 * result the reference's merge keeps, zone_pass_openmp.cpp:145-153).
 * Positions are written in place (nobody reads them during a pass).
 * Velocities are written in place for particles whose input cell only one
 * zone covers; for cells several zones cover they go to vsx/vsy and are
 * copied back after the pass, so every zone reads inputs.
 *
 * Lists.  A cell's list stays implicit -- a run of consecutive local
 * indices head..tail, initially its input run -- while only its head is
 * popped; the first append or rewrite writes its links (ZF_MAT).  ZF_CHG
 * marks every cell whose list changed; the others cost nothing at the
 * correction.
 */
#pragma once

#include <stdint.h>

#include "mpcd_core.h"

/* every zone function is inlined into the kernel, so the per-lane zone
 * context stays in registers */
#if defined(__HIPCC__)
#define ZHD __host__ __device__ inline __attribute__((always_inline))
#else
#define ZHD inline
#endif

namespace mpcdk {

#if defined(__HIP_DEVICE_COMPILE__)
/* The zone code runs on one wavefront: its barrier orders the lanes' memory
 * operations (workgroup-scope fences) without a workgroup barrier, so a
 * larger workgroup can run it on one of its wavefronts. */
#define ZBAR()                                                 \
    do {                                                       \
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup"); \
        __builtin_amdgcn_wave_barrier();                       \
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup"); \
    } while (0)
/* compiler-level ordering of LDS accesses between lanes of the one
 * wavefront (the LDS executes a wavefront's accesses in issue order) */
#define ZWSYNC()                                               \
    do {                                                       \
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); \
        __builtin_amdgcn_wave_barrier();                       \
    } while (0)
ZHD unsigned long long wave_ballot(bool f) { return __ballot(f); }
ZHD int wave_popc(unsigned long long m) { return __popcll(m); }
ZHD int wave_shfl_up_add(int v, int lane) {
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}
ZHD int wave_bcast(int v, int src) { return __shfl(v, src, 64); }
/* value of lane l (l uniform across the wavefront) */
ZHD int wave_readlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
ZHD double wave_readlane_d(double v, int l) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
ZHD int wave_sum(int v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
ZHD double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) {
        const double t = __shfl_xor(v, o, 64);
        v = (v < t) ? t : v;
    }
    return v;
}
#else
#define ZBAR() ((void)0)
#define ZWSYNC() ((void)0)
ZHD unsigned long long wave_ballot(bool f) { return f ? 1ull : 0ull; }
ZHD int wave_popc(unsigned long long m) { return __builtin_popcountll(m); }
ZHD int wave_shfl_up_add(int v, int) { return v; }
ZHD int wave_bcast(int v, int) { return v; }
ZHD int wave_readlane(int v, int) { return v; }
ZHD double wave_readlane_d(double v, int) { return v; }
ZHD int wave_sum(int v) { return v; }
ZHD double wave_max(double v) { return v; }
#endif
ZHD unsigned long long lanes_below(int lane) { return (lane >= 64) ? ~0ull : ((1ull << lane) - 1ull); }

struct ZoneParams {
    double highThr, lowThrBulk;
    double lowFrac, highFrac; /* 1/(1+coef), 1+coef */
    double gamma, fluidFracGain, lowThrFloor, highThrFloor;
    double lowLoc1, highLoc1; /* empty-cell thresholds for fluid fraction 1 */
    int useLocal, maxPasses, corrEnabled;
    int dedupe; /* a periodic axis narrower than 5 cells: wrapped neighbours can repeat */
    /* angular bin (3 bits) of (dx,dy) in [-2,2]^2 at index (dy+2)*5+(dx+2):
     * indices 0..20 in binLo, 21..24 in binHi (registers, not an array that
     * a dynamic index would send to private memory) */
    unsigned long long binLo, binHi;
};

ZHD int zone_bin(const ZoneParams& zp, int idx) {
    return idx < 21 ? (int)((zp.binLo >> (3 * idx)) & 7ull) : (int)((zp.binHi >> (3 * (idx - 21))) & 7ull);
}

/* Candidate of a dense cell's target ranking. */
struct Cand {
    int lc, gc;
    double deficitNeg;
    int cnt, cheb, dist2, bin;
    double tie;
};

struct ZoneRect {
    int xs, xe, ys, ye;
};

struct ZoneOut {
    int nHigh, nLow, movedDense, movedSparse;
    double stats4, stats5;
};

/* (overridable so that a host build can exercise the paths past them) */
#ifndef MPCD_ZONE_MAX_MOVE
#define MPCD_ZONE_MAX_MOVE 64
#endif
#ifndef MPCD_ZONE_TMP_L
#define MPCD_ZONE_TMP_L 256
#endif
constexpr int ZONE_MAX_MOVE = MPCD_ZONE_MAX_MOVE; /* dense movers staged per source cell */
static_assert(ZONE_MAX_MOVE <= 64, "a dense cell's movers are the lanes of one wavefront");
constexpr int ZONE_TMP_L = MPCD_ZONE_TMP_L;       /* donor lists up to this length are shuffled in LDS */

/* per-cell flag bits */
enum : unsigned short {
    ZF_BULK0 = 1,   /* occupied when this pass's maps were built */
    ZF_RECV = 2,    /* received particles */
    ZF_BB = 4,      /* bulk-to-bulk transfer */
    ZF_SKIP = 8,    /* transfer touching a non-bulk cell */
    ZF_MAT = 16,    /* list links written */
    ZF_CORR = 32,   /* momentum correction applies */
    ZF_CHG = 64,    /* list changed */
    ZF_OWN = 128,   /* this zone owns the cell's input particles */
    ZF_COVER = 256, /* several zones cover the cell */
    ZF_OUT = 4096   /* in the zone's rectangle but not in its mask */
};

/* Inputs of a zone pass common to all its zones. */
struct ZoneIn {
    Geom g;
    ZoneParams zp;
    const int* gcnt;    /* pass-input binning: counts, run starts, slots */
    const int* gstart;
    const int* glist;
    const unsigned long long* gkey; /* (pid << 32 | slot) of each binning position */
    const int* owner;   /* owner map of the layout, or nullptr: each zone owns its cells */
    const unsigned char* gcover; /* per cell: several zones cover it, or nullptr */
    double *x, *y, *vx, *vy;
    double *vsx, *vsy;  /* side buffer for covered cells */
    /* per-cell velocity sums of the pass input (each cell's run summed in
     * order from zero), or nullptr: the correction sums the runs itself */
    const double *inPx = nullptr, *inPy = nullptr;
    /* non-null: each moved particle's cell change is recorded for the
     * binning update (cell of each slot, departures and arrivals per cell,
     * rank among the new cell's arrivals) */
    int *dcid = nullptr, *dout = nullptr, *dinc = nullptr, *drnk = nullptr;
    /* non-null (a single-zone call that reports ZoneKernelStats.diag): per
     * zone cell, the momentum defect and the final list's momentum before
     * the correction (4 doubles), and the count the sparse phase starts
     * from */
    double* dgCell = nullptr;
    int* dgCnt = nullptr;
    /* added to the recorded new cells (an ensemble replica's first cell in
     * the concatenated grid; 0 for a single simulation) */
    int cellOff = 0;
    /* per grid cell, nonzero in the zone (Zone.maskVec, zone_pass.cpp:20-24),
     * or nullptr: the zone is its whole rectangle.  A cell of the rectangle
     * outside the mask keeps its particles and counts but is never in the
     * dense or low lists, a target, a neighbour or a donor. */
    const unsigned char* zmask = nullptr;
};

/* Scalars the lanes share (LDS on the device). */
struct ZoneShared {
    int nOrd, startBin, nr, stage, P, allOwned;
    int mwCount;           /* the multi-wavefront phases' list length */
    int mwNext;            /* and the next list position a wavefront takes */
    int schedWait;         /* nonzero: a wait of the dependency-driven dense phase hit its bound */
    int nHigh, nLow, movedDense, movedSparse;
    int mixedW[3];         /* zone_setup_mw: wavefront 1..3 saw a cell not plainly owned */
    double stats4, stats5;
};

template <class IT>
struct ZoneCtxT {
    typedef IT idx_t;
    Geom g;
    ZoneParams zp;
    int xs, xe, ys, ye, zw, zh, C;
    int zid, P, nL;
    int allOwned; /* copied from the shared set-up result */
    uint64_t seed;
    const int *gcnt, *gstart, *glist;
    const unsigned long long* gkey;
    double *x, *y, *vx, *vy, *vsx, *vsy;
    const double *inPx, *inPy; /* input run sums per global cell, or nullptr */
    int *dcid, *dout, *dinc, *drnk; /* cell-change record, or nullptr */
    int cellOff;    /* added to the recorded new cells */
    double* dgCell; /* diagnostics' per-cell record, or nullptr */
    int* dgCnt;     /* diagnostics' sparse-phase start counts, or nullptr */
    IT* nxL;     /* links of k < nL */
    IT* nxG;     /* links of k >= nL (zone's global region, indexed by k) */
    IT* tmpG;    /* donor lists longer than ZONE_TMP_L (zone's global region) */
    double* duG; /* velocity change per cell (zone's global region, 2 per cell) */
    unsigned char* rec; /* last move of each particle by local index (zone's global region, 1 per particle) */
    /* cell arrays */
    IT *cnt, *head, *tail, *dl, *k0;
    IT* ext; /* sparse phase: the counts it starts from, zone and ring, (zw+4) x (zh+4) column-major */
    unsigned short* fl;
    int *colL, *colG; /* per column: first local index, start in the binning list */
    double *lowA, *highA; /* local fluid-fraction thresholds (only with useLocal) */
    IT* halo;             /* pass-input counts of the ring of width 2 around the zone */
    /* small scratch */
    Cand* cand;     /* 24 */
    int* ord;       /* 24 */
    double* ordThr; /* 24 */
    int* nbGc;      /* 24 */
    double* nbTie;  /* 24 */
    int* rnk;       /* 24 */
    int* recvT;     /* 24 */
    double* dsc;    /* 8 */
    int* dord;      /* 8 */
    int* dord2;     /* 8 */
    double* nsm4;   /* 4 */
    int* mvk;          /* ZONE_MAX_MOVE: local indices of a materialised list's head run */
    IT* tmpL;          /* tmpCap entries (ZONE_TMP_L, or a wavefront's share in the multi-wavefront sparse phase) */
    IT* lix;           /* per cell: its index in the multi-wavefront sparse phase's list, or the type's maximum */
    int tmpCap;
    ZoneShared* sh;
};

/* ---- index and flag access ---- */

template <class IT>
ZHD int iget(IT v) {
    return v == (IT)-1 ? -1 : (int)v;
}
template <class Z>
ZHD bool fget(const Z& z, int lc, unsigned short b) {
    return (z.fl[lc] & b) != 0;
}
template <class Z>
ZHD void fset(Z& z, int lc, unsigned short b) {
    z.fl[lc] = (unsigned short)(z.fl[lc] | b);
}
template <class Z>
ZHD void fclr(Z& z, int lc, unsigned short b) {
    z.fl[lc] = (unsigned short)(z.fl[lc] & ~b);
}
template <class Z>
ZHD int cnt_of(const Z& z, int lc) {
    return (int)z.cnt[lc];
}
template <class Z>
ZHD void cnt_set(Z& z, int lc, int v) {
    z.cnt[lc] = (typename Z::idx_t)v;
}
template <class Z>
ZHD int head_of(const Z& z, int lc) {
    return iget(z.head[lc]);
}
template <class Z>
ZHD int tail_of(const Z& z, int lc) {
    return iget(z.tail[lc]);
}
template <class Z>
ZHD void head_set(Z& z, int lc, int k) {
    z.head[lc] = (typename Z::idx_t)k;
}
template <class Z>
ZHD void tail_set(Z& z, int lc, int k) {
    z.tail[lc] = (typename Z::idx_t)k;
}
/* Accesses to the zone's global regions.  On the device they are
 * non-temporal operations: distinct from the working-set accesses, so the
 * compiler never merges the two into one generic (flat) access whose wait
 * would also stall on every outstanding global load. */
#if defined(__HIP_DEVICE_COMPILE__)
#define ZG_LOAD(p) __builtin_nontemporal_load(p)
#define ZG_STORE(p, v) __builtin_nontemporal_store((v), (p))
#else
#define ZG_LOAD(p) (*(p))
#define ZG_STORE(p, v) (*(p) = (v))
#endif

/* The link accessors read both base pointers before choosing: a choice
 * between two fields of the context (rather than two loaded values) would
 * keep the context out of registers. */
template <class Z>
ZHD int nx_get(const Z& z, int k) {
    const typename Z::idx_t* const L = z.nxL;
    const typename Z::idx_t* const G = z.nxG;
    if (k < z.nL) return iget(L[k]);
    return iget(ZG_LOAD(G + k));
}
template <class Z>
ZHD void nx_set(Z& z, int k, int v) {
    typename Z::idx_t* const L = z.nxL;
    typename Z::idx_t* const G = z.nxG;
    if (k < z.nL) L[k] = (typename Z::idx_t)v;
    else ZG_STORE(G + k, (typename Z::idx_t)v);
}

/* ---- geometry ---- */

template <class Z>
ZHD int zone_lc(const Z& z, int gx, int gy) {
    if (gx >= z.xs && gx < z.xe && gy >= z.ys && gy < z.ye) return (gx - z.xs) * z.zh + (gy - z.ys);
    return -1;
}
template <class Z>
ZHD int lc_gx(const Z& z, int lc) {
    return z.xs + lc / z.zh;
}
template <class Z>
ZHD int lc_gy(const Z& z, int lc) {
    return z.ys + lc % z.zh;
}
template <class Z>
ZHD int lc_gc(const Z& z, int lc) {
    return lc_gy(z, lc) + z.g.Ny * lc_gx(z, lc);
}

/* Slot in z.halo of the cell at zone-relative offset (dx, dy) of the ring of
 * width 2 around the zone: two rows below and above (all zw+4 columns), then
 * two columns left and right (the zh rows between); -1 inside the zone or
 * beyond the ring. */
template <class Z>
ZHD int halo_slot(const Z& z, int dx, int dy) {
    if (dx < -2 || dx >= z.zw + 2 || dy < -2 || dy >= z.zh + 2) return -1;
    const int hx = dx + 2, hy = dy + 2, W4 = z.zw + 4;
    if (hy < 2) return hx * 2 + hy;
    if (hy >= z.zh + 2) return 2 * W4 + hx * 2 + (hy - z.zh - 2);
    if (hx < 2) return 4 * W4 + hx * z.zh + (hy - 2);
    if (hx >= z.zw + 2) return 4 * W4 + 2 * z.zh + (hx - z.zw - 2) * z.zh + (hy - 2);
    return -1;
}
template <class Z>
ZHD int halo_size(const Z& z) {
    return 4 * (z.zw + 4) + 4 * z.zh;
}
/* the ring slot of global cell (gx, gy) outside the zone, periodic axes
 * taken to the image next to the zone */
template <class Z>
ZHD int halo_of(const Z& z, int gx, int gy) {
    int dx = gx - z.xs, dy = gy - z.ys;
    if (z.g.perx) {
        if (dx < -2) dx += z.g.Nx;
        else if (dx >= z.zw + 2) dx -= z.g.Nx;
    }
    if (z.g.pery) {
        if (dy < -2) dy += z.g.Ny;
        else if (dy >= z.zh + 2) dy -= z.g.Ny;
    }
    return halo_slot(z, dx, dy);
}

/* count at a global cell: live zone count (or the sparse-phase copy), else
 * the pass-input snapshot, from the ring in LDS (a count the index type
 * cannot hold is read from global memory). */
template <class Z>
ZHD int count_at(const Z& z, int gx, int gy, int use_copy) {
    const int lc = zone_lc(z, gx, gy);
    if (lc >= 0) {
        if (!use_copy) return (int)z.cnt[lc];
        const int ix = lc / z.zh;
        return (int)z.ext[(ix + 2) * (z.zh + 4) + (lc - ix * z.zh) + 2];
    }
    const int h = halo_of(z, gx, gy);
    if (h >= 0) {
        const int v = iget(z.halo[h]);
        if (v >= 0) return v;
    }
    return z.gcnt[gy + z.g.Ny * gx];
}

/* 3x3 binomial smoothing without wrap (zone_pass.cpp:296-319).  The
 * weights are multiples of 1/16 and the counts integers, so the weighted
 * sums are exact: they are taken in integers (scaled by 16) and divided once,
 * which is the same correctly rounded quotient. */
template <class Z>
ZHD double nsm_at(const Z& z, int gx, int gy, int use_copy) {
    long long num = 0;
    int den = 0;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            const int jy = gy + dy, jx = gx + dx;
            if (jy < 0 || jy >= z.g.Ny || jx < 0 || jx >= z.g.Nx) continue;
            const int w = (dx == 0 ? 2 : 1) * (dy == 0 ? 2 : 1); /* 1 2 1 / 2 4 2 / 1 2 1 */
            num += (long long)w * count_at(z, jx, jy, use_copy);
            den += w;
        }
    return (double)num / (double)den;
}

/* nsm_at(.., use_copy = 1) for a point within one cell of the zone: the
 * nine counts are direct reads of the extended array (no branch per
 * neighbour); the same integer sums and quotient. */
template <class Z>
ZHD double nsm_ext(const Z& z, int gx, int gy) {
    int rx = gx - z.xs, ry = gy - z.ys;
    if (z.g.perx) {
        if (rx < -2) rx += z.g.Nx;
        else if (rx >= z.zw + 2) rx -= z.g.Nx;
    }
    if (z.g.pery) {
        if (ry < -2) ry += z.g.Ny;
        else if (ry >= z.zh + 2) ry -= z.g.Ny;
    }
    if (rx < -1 || rx > z.zw || ry < -1 || ry > z.zh) return nsm_at(z, gx, gy, 1);
    const int H = z.zh + 4;
    const int base = (rx + 2) * H + (ry + 2);
    int v[9];
    bool ok[9];
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy) {
            const int t = (dx + 1) * 3 + (dy + 1);
            const int jx = gx + dx, jy = gy + dy;
            ok[t] = jx >= 0 && jx < z.g.Nx && jy >= 0 && jy < z.g.Ny;
            v[t] = ok[t] ? iget(z.ext[base + dx * H + dy]) : 0;
        }
    long long num = 0;
    int den = 0;
    bool far = false;
    for (int t = 0; t < 9; ++t) {
        const int w = (t / 3 == 1 ? 2 : 1) * (t % 3 == 1 ? 2 : 1);
        if (!ok[t]) continue;
        if (v[t] < 0) far = true;
        num += (long long)w * v[t];
        den += w;
    }
    if (far) return nsm_at(z, gx, gy, 1); /* a ring count the index type could not hold */
    return (double)num / (double)den;
}

/* Active thresholds of a zone cell (zone_pass.cpp:321-345, :715-725): the
 * bulk flag is taken when the maps are built; with fluid fraction 1 the
 * empty-cell values are constants, otherwise they are stored per cell. */
template <class Z>
ZHD void zone_maps(Z& z, int lc, int use_copy) {
    const int bulk = z.cnt[lc] > 0;
    if (bulk) fset(z, lc, ZF_BULK0);
    else fclr(z, lc, ZF_BULK0);
    if (!z.zp.useLocal) return;
    const double fRaw = nsm_at(z, lc_gx(z, lc), lc_gy(z, lc), use_copy) / dmax(z.zp.gamma, 1e-12);
    const double t = z.zp.fluidFracGain * fRaw;
    const double f = (t < 0.0) ? 0.0 : ((1.0 < t) ? 1.0 : t); /* std::clamp(t, 0, 1) */
    const double gl = z.zp.gamma * f;
    z.lowA[lc] = bulk ? z.zp.lowThrBulk : dmax(z.zp.lowFrac * gl, z.zp.lowThrFloor);
    z.highA[lc] = bulk ? z.zp.highThr : dmax(z.zp.highFrac * gl, z.zp.highThrFloor);
}
template <class Z>
ZHD double low_thr(const Z& z, int lc) {
    if (z.zp.useLocal) return z.lowA[lc];
    const double b = z.zp.lowThrBulk, e = z.zp.lowLoc1;
    return fget(z, lc, ZF_BULK0) ? b : e;
}
template <class Z>
ZHD double high_thr(const Z& z, int lc) {
    if (z.zp.useLocal) return z.highA[lc];
    const double b = z.zp.highThr, e = z.zp.highLoc1;
    return fget(z, lc, ZF_BULK0) ? b : e;
}
template <class Z>
ZHD double thr(const Z& z, int lc, int high) {
    return high ? high_thr(z, lc) : low_thr(z, lc);
}

/* zone_pass.cpp:229-244 */
template <class Z>
ZHD void min_step(const Z& z, int sx, int sy, int tx, int ty, int& ddx, int& ddy) {
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
template <class Z>
ZHD int shift_idx(const Z& z, int gx, int gy, int dx, int dy, int& jx, int& jy) {
    jy = gy + dy;
    if (z.g.pery) jy = wrapi(jy, z.g.Ny);
    else if (jy < 0 || jy >= z.g.Ny) return 0;
    jx = gx + dx;
    if (z.g.perx) jx = wrapi(jx, z.g.Nx);
    else if (jx < 0 || jx >= z.g.Nx) return 0;
    return 1;
}

/* ---- particles ---- */

/* column of the zone's k-th particle (the zone's particles are its columns'
 * runs of the binning list, in column order) */
template <class Z>
ZHD int col_of(const Z& z, int k) {
    int lo = 0, hi = z.zw - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (z.colL[mid] <= k) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
template <class Z>
ZHD int slot_in(const Z& z, int k, int col) {
    return z.glist[z.colG[col] + (k - z.colL[col])];
}
/* slot and particle id of the zone's k-th particle in one load */
template <class Z>
ZHD int slot_pid_in(const Z& z, int k, int col, uint32_t& pid) {
    const unsigned long long key = z.gkey[z.colG[col] + (k - z.colL[col])];
    pid = (uint32_t)(key >> 32);
    return (int)(uint32_t)key;
}
template <class Z>
ZHD int slot_of(const Z& z, int k) {
    return slot_in(z, k, col_of(z, k));
}
/* ZF_OWN / ZF_COVER of the input cell of k: the last cell of k's column
 * whose input run starts at or before k */
template <class Z>
ZHD unsigned short own_bits(const Z& z, int k, int col) {
    if (z.allOwned) return ZF_OWN;
    int lo = col * z.zh, hi = lo + z.zh - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((int)z.k0[mid] <= k) lo = mid;
        else hi = mid - 1;
    }
    return (unsigned short)(z.fl[lo] & (ZF_OWN | ZF_COVER));
}

/* successor of k in cell lc's list (implicit run or links) */
template <class Z>
ZHD int next_k(const Z& z, int lc, int k) {
    return fget(z, lc, ZF_MAT) ? nx_get(z, k) : (k < tail_of(z, lc) ? k + 1 : -1);
}

/* write the links of a still-implicit list before its first append */
template <class Z>
ZHD void ensure_mat(Z& z, int lc) {
    if (fget(z, lc, ZF_MAT)) return;
    const int t = tail_of(z, lc);
    for (int k = head_of(z, lc); k >= 0 && k <= t; ++k) nx_set(z, k, (k < t) ? k + 1 : -1);
    fset(z, lc, ZF_MAT);
}

/* ensure_mat by all lanes (lc the same on every lane) */
template <class Z>
ZHD void ensure_mat_lanes(Z& z, int lc, int lane, int nl) {
    if (fget(z, lc, ZF_MAT)) return;
    const int h = head_of(z, lc), t = tail_of(z, lc);
    if (h >= 0)
        for (int k = h + lane; k <= t; k += nl) nx_set(z, k, k < t ? k + 1 : -1);
    ZBAR();
    if (lane == 0) fset(z, lc, ZF_MAT);
    ZBAR();
}

/* requires a materialised list */
template <class Z>
ZHD void list_append(Z& z, int lc, int k) {
    typedef typename Z::idx_t IT;
    IT* const H = z.head;
    IT* const T = z.tail;
    fset(z, lc, ZF_CHG);
    nx_set(z, k, -1);
    const int t = iget(T[lc]);
    if (t < 0) H[lc] = (IT)k;
    else nx_set(z, t, k);
    T[lc] = (IT)k;
}

/* A moved particle's new position (zone_pass.cpp:280-294: uniform in its
 * target cell, drawn per dense pass or for the sparse phase) depends only on
 * its last move, so a move records what it was -- ZM_NONE, the dense pass
 * (1..254) or ZM_SPARSE -- and the correction sweep, which visits every
 * particle in its final cell, writes the positions (owned particles only).
 * The serial phases then touch no global memory for moves. */
enum { ZM_NONE = 0, ZM_SPARSE = 255 };
template <class Z>
ZHD void mark_moved(Z& z, int k, int tag) {
    z.rec[k] = (unsigned char)tag;
}
template <class Z>
ZHD void place_moved(Z& z, int lc, int slot, uint32_t pid, int tag) {
    double ux, uy;
    if (tag == ZM_SPARSE) mpcd_draw_sparse_pos(z.seed, pid, &ux, &uy);
    else mpcd_draw_dense_pos(z.seed, tag, pid, &ux, &uy);
    const int ix = lc / z.zh;
    const double nx = (double)(z.xs + ix) * z.g.dxc + z.g.dxc * ux;
    const double ny = (double)(z.ys + (lc - ix * z.zh)) * z.g.dyc + z.g.dyc * uy;
    z.x[slot] = nx;
    z.y[slot] = ny;
    if (z.dcid) {
        /* the cell the binning update will see (a particle is placed once
         * per pass); the pass's own counts stay untouched */
        const int cn = cell_of(z.g, nx, ny) + z.cellOff, co = z.dcid[slot];
        if (cn != co) {
            z.dcid[slot] = cn;
#if defined(__HIP_DEVICE_COMPILE__)
            atomicAdd(&z.dout[co], 1);
            z.drnk[slot] = atomicAdd(&z.dinc[cn], 1);
#else
            ++z.dout[co];
            z.drnk[slot] = z.dinc[cn]++;
#endif
        }
    }
}

/* the bulk/skip masks of a transfer (zone_pass.cpp:492-525) */
template <class Z>
ZHD void transfer_masks(Z& z, int src, const int* rv, int nr) {
    bool allBulk = fget(z, src, ZF_BULK0);
    for (int k = 0; k < nr && allBulk; ++k)
        if (!fget(z, rv[k], ZF_BULK0)) allBulk = false;
    if (allBulk) {
        fset(z, src, ZF_BB);
        for (int k = 0; k < nr; ++k) fset(z, rv[k], ZF_BB);
    } else {
        fset(z, src, ZF_SKIP);
        fclr(z, src, ZF_BB);
        for (int k = 0; k < nr; ++k) {
            fset(z, rv[k], ZF_SKIP);
            fclr(z, rv[k], ZF_BB);
        }
    }
}

ZHD bool cand_less(const Cand& a, const Cand& b) {
    /* (-deficit, -inward (all equal: weight 0), cnt, cheb, dist2, tie) */
    if (a.deficitNeg != b.deficitNeg) return a.deficitNeg < b.deficitNeg;
    if (a.cnt != b.cnt) return a.cnt < b.cnt;
    if (a.cheb != b.cheb) return a.cheb < b.cheb;
    if (a.dist2 != b.dist2) return a.dist2 < b.dist2;
    return a.tie < b.tie;
}

enum { ZC_DENSE = 0, ZC_LOW = 1 };

/* Order-preserving compaction into z.dl of the zone cells above their high
 * threshold (ZC_DENSE) or occupied and below their low threshold (ZC_LOW);
 * all lanes; returns the count (same on every lane). */
template <class Z>
ZHD int zone_compact(Z& z, int lane, int nl, int mode) {
    int base = 0;
    for (int b0 = 0; b0 < z.C; b0 += nl) {
        const int lc = b0 + lane;
        bool f = false;
        if (lc < z.C) {
            const double c = (double)z.cnt[lc];
            f = (mode == ZC_DENSE) ? c > high_thr(z, lc) : (z.cnt[lc] > 0 && c < low_thr(z, lc));
            f = f && !fget(z, lc, ZF_OUT);
        }
        const unsigned long long msk = wave_ballot(f);
        if (f) z.dl[base + wave_popc(msk & lanes_below(lane))] = (typename Z::idx_t)lc;
        base += wave_popc(msk);
    }
    return base;
}

/* The shuffles' partner draws, j in [0, i]. */
enum { ZS_DENSE = 0, ZS_LOW = 1, ZS_DONOR = 2, ZS_PARTICLE = 3 };
struct ShufDraw {
    uint64_t seed;
    int kind;
    uint32_t a, b;
    ZHD uint32_t operator()(uint32_t i) const {
        switch (kind) {
            case ZS_DENSE: return mpcd_draw_dense_shuffle(seed, (int)a, i);
            case ZS_LOW: return mpcd_draw_low_shuffle(seed, i);
            case ZS_DONOR: return mpcd_draw_donor_shuffle(seed, a, i);
            default: return mpcd_draw_particle_shuffle(seed, a, b, i);
        }
    }
};

/* Fisher-Yates of a[0..m), i = m-1 down to 1, swapping a[i] with a[j(i)],
 * `draw(i)` giving j in [0, i].  The partner draws of a chunk of nl indices
 * are computed by the lanes at once and kept in registers; lane 0 applies
 * the chunk's swaps in order, reading each partner from its lane. */
template <class T>
ZHD void zone_shuffle(T* a, int m, int lane, int nl, ShufDraw draw) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (m <= 64) {
        /* element q in lane q's register; each swap reads the two values
         * and writes them back by lane select */
        int v = lane < m ? (int)a[lane] : 0;
        const int j = (lane >= 1 && lane < m) ? (int)draw((uint32_t)lane) : 0;
        for (int ii = m - 1; ii >= 1; --ii) {
            const int jj = wave_readlane(j, ii);
            const int vi = wave_readlane(v, ii), vj = wave_readlane(v, jj);
            v = lane == ii ? vj : (lane == jj ? vi : v);
        }
        if (lane < m) a[lane] = (T)v;
        ZBAR();
        return;
    }
#endif
    for (int c0 = ((m - 1) / nl) * nl; c0 >= 0 && m > 1; c0 -= nl) {
        const int i = c0 + lane;
        const int j = (i >= 1 && i < m) ? (int)draw((uint32_t)i) : 0;
        const int hi = imin(m - 1, c0 + nl - 1);
        for (int ii = hi; ii >= imax(1, c0); --ii) {
            const int jj = wave_readlane(j, ii - c0);
            if (lane == 0) {
                const T t = a[ii];
                a[ii] = a[jj];
                a[jj] = t;
            }
        }
    }
    ZBAR();
}

/* ---- dense phase ---- */

/* Admissible targets of dense cell src within Chebyshev range 2 inside the
 * zone, ranked into the round-robin order of rank_dense_targets_unified
 * (zone_pass.cpp:413-490): 8 angular bins from a random start bin, each bin
 * sorted by the candidate key, then one per bin per round.  All lanes; the
 * order lands in z.ord[0..nOrd) with each target's threshold in z.ordThr.
 * Every draw is taken in one round: the tie breaks (t = 0..23, one per
 * neighbour offset) and the start bin (t = 24). */
template <class Z>
ZHD void dense_targets(Z& z, int src, int high, int pass, int phase, int lane, int nl) {
    const int sgx = lc_gx(z, src), sgy = lc_gy(z, src), sgc = lc_gc(z, src);
    for (int t = lane; t < 25; t += nl) {
        if (t < 24) {
            /* neighbour offsets in (dy outer, dx inner) order, centre skipped */
            const int f = t < 12 ? t : t + 1;
            const int dy = f / 5 - 2, dx = f % 5 - 2;
            int jx, jy, gc = -1;
            if (shift_idx(z, sgx, sgy, dx, dy, jx, jy)) {
                const int lc = zone_lc(z, jx, jy);
                if (lc >= 0 && !fget(z, lc, ZF_OUT) && (double)z.cnt[lc] < thr(z, lc, high)) gc = jy + z.g.Ny * jx;
            }
            z.nbGc[t] = gc;
            if (gc >= 0) z.nbTie[t] = mpcd_draw_tie(z.seed, pass, phase, (uint32_t)sgc, (uint32_t)gc);
        } else {
            z.sh->startBin = mpcd_draw_start_bin(z.seed, pass, phase, (uint32_t)sgc);
        }
    }
    ZBAR();
#if defined(__HIP_DEVICE_COMPILE__)
    {
        /* one wavefront: neighbour offset o = lane; each candidate stays in
         * its lane's registers and the ranking reads the others by readlane */
        bool keep = false;
        int gc = -1;
        if (lane < 24) {
            gc = z.nbGc[lane];
            keep = gc >= 0;
            if (z.zp.dedupe)
                for (int q = 0; q < lane && keep; ++q)
                    if (z.nbGc[q] == gc) keep = false;
        }
        const unsigned long long msk = wave_ballot(keep);
        const int m = wave_popc(msk);
        int lc = -1, bin = 0, cnt = 0, cheb = 0, dist2 = 0;
        double defn = 0.0, tie = 0.0, th = 0.0;
        if (keep) {
            const int jx = gc / z.g.Ny, jy = gc - jx * z.g.Ny;
            lc = zone_lc(z, jx, jy);
            int ddx, ddy;
            min_step(z, sgx, sgy, jx, jy, ddx, ddy);
            bin = zone_bin(z.zp, (ddy + 2) * 5 + (ddx + 2));
            th = thr(z, lc, high);
            cnt = z.cnt[lc];
            defn = -dmax(0.0, th - (double)cnt);
            cheb = imax(ddx < 0 ? -ddx : ddx, ddy < 0 ? -ddy : ddy);
            dist2 = ddx * ddx + ddy * ddy;
            tie = z.nbTie[lane];
        }
        if (lane == 0) z.sh->nOrd = m;
        if (m == 0) {
            ZBAR();
            return;
        }
        Cand me;
        me.deficitNeg = defn;
        me.cnt = cnt;
        me.cheb = cheb;
        me.dist2 = dist2;
        me.tie = tie;
        /* rank inside the bin: stable order of the candidate key */
        int r = 0;
        for (unsigned long long mm = msk; mm; mm &= mm - 1ull) {
            const int j = __builtin_ctzll(mm);
            Cand o;
            o.deficitNeg = wave_readlane_d(defn, j);
            o.cnt = wave_readlane(cnt, j);
            o.cheb = wave_readlane(cheb, j);
            o.dist2 = wave_readlane(dist2, j);
            o.tie = wave_readlane_d(tie, j);
            const int bj = wave_readlane(bin, j);
            if (keep && j != lane && bj == bin && (cand_less(o, me) || (j < lane && !cand_less(me, o)))) ++r;
        }
        /* round robin position: earlier rounds, then earlier bins from startBin */
        const int sb = z.sh->startBin;
        const int si = (bin - sb + 8) & 7;
        int pos = 0;
        for (unsigned long long mm = msk; mm; mm &= mm - 1ull) {
            const int j = __builtin_ctzll(mm);
            const int rj = wave_readlane(r, j);
            const int sj = (wave_readlane(bin, j) - sb + 8) & 7;
            if (rj < r || (rj == r && sj < si)) ++pos;
        }
        if (keep) {
            z.ord[pos] = lc;
            z.ordThr[pos] = th;
        }
        ZBAR();
    }
#else
    int base = 0;
    for (int b0 = 0; b0 < 24; b0 += nl) {
        const int o = b0 + lane;
        bool keep = false;
        int gc = -1;
        if (o < 24) {
            gc = z.nbGc[o];
            keep = gc >= 0;
            if (z.zp.dedupe)
                for (int q = 0; q < o && keep; ++q)
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
            c.bin = zone_bin(z.zp, (ddy + 2) * 5 + (ddx + 2));
            c.deficitNeg = -dmax(0.0, thr(z, lc, high) - (double)z.cnt[lc]);
            c.cnt = z.cnt[lc];
            c.cheb = imax(ddx < 0 ? -ddx : ddx, ddy < 0 ? -ddy : ddy);
            c.dist2 = ddx * ddx + ddy * ddy;
            c.tie = z.nbTie[o];
        }
        base += wave_popc(msk);
    }
    const int m = base;
    if (lane == 0) z.sh->nOrd = m;
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
    const int sb = z.sh->startBin;
    for (int i = lane; i < m; i += nl) {
        const int ri = z.rnk[i];
        const int si = (z.cand[i].bin - sb + 8) & 7;
        int pos = 0;
        for (int j = 0; j < m; ++j) {
            const int rj = z.rnk[j];
            if (rj < ri || (rj == ri && ((z.cand[j].bin - sb + 8) & 7) < si)) ++pos;
        }
        const int lc = z.cand[i].lc;
        z.ord[pos] = lc;
        z.ordThr[pos] = thr(z, lc, high);
    }
    ZBAR();
#endif
}

/* spread_particles_from_dense_source_round_robin (zone_pass.cpp:527-589),
 * lane 0.  The number of candidates at or after position kk
 * that still have a deficit is kept incrementally: within a sweep only the
 * candidate being served changes, and each target appears once. */
template <class Z>
ZHD int spread(Z& z, int src, int needOut, int pass, int* rv, int& nr) {
    int moved = 0;
    const int nOrd = z.sh->nOrd;
    if (needOut <= 0 || nOrd == 0) return 0;
    int avail = cnt_of(z, src);
    if (avail <= 0) return 0;
    while (needOut > 0 && avail > 0) {
        bool progressed = false;
        int R = 0;
        for (int jj = 0; jj < nOrd; ++jj)
            if (imax(0, (int)MPCD_FLOOR(z.ordThr[jj] - (double)cnt_of(z, z.ord[jj]))) > 0) ++R;
        for (int kk = 0; kk < nOrd; ++kk) {
            const int nb = z.ord[kk];
            const int deficit = imax(0, (int)MPCD_FLOOR(z.ordThr[kk] - (double)cnt_of(z, nb)));
            if (deficit <= 0) continue;
            const int remaining = R;
            --R;
            const int quota = imax(1, (int)MPCD_CEIL((double)needOut / (double)imax(1, remaining)));
            const int nt = imax(1, imin(imin(quota, deficit), imin(needOut, avail)));
            /* pops from the head leave an implicit run implicit */
            ensure_mat(z, nb);
            fset(z, src, ZF_CHG);
            for (int q = 0; q < nt; ++q) {
                const int k = head_of(z, src);
                const int nk = next_k(z, src, k);
                head_set(z, src, nk);
                if (nk < 0) tail_set(z, src, -1);
                mark_moved(z, k, pass);
                list_append(z, nb, k);
            }
            avail -= nt;
            cnt_set(z, nb, cnt_of(z, nb) + nt);
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
    if (moved > 0) cnt_set(z, src, avail);
    return moved;
}

#if defined(__HIP_DEVICE_COMPILE__)
/* ---- dense cells on the device: candidates and movers held in lanes ----
 *
 * Lane p holds candidate p of a ranking and lane q the q-th particle of the
 * source's list (a mover); everything the spread decides is computed by all
 * lanes from readlane values, so the wavefront follows one uniform control
 * path, and the moves and list links are written by the lanes at once.  The
 * results are those of dense_targets + spread + transfer_masks. */


/* value held by lane src (per-lane src) */
ZHD int lane_get(int v, int src) { return __shfl(v, src, 64); }
ZHD double lane_get_d(double v, int src) { return __shfl(v, src, 64); }
ZHD unsigned long long lane_get_u64(unsigned long long v, int src) {
    const unsigned lo = (unsigned)__shfl((int)(unsigned)v, src, 64);
    const unsigned hi = (unsigned)__shfl((int)(unsigned)(v >> 32), src, 64);
    return ((unsigned long long)hi << 32) | lo;
}
/* every lane's v to lane dst (dst a permutation of the lanes) */
ZHD int lane_put(int v, int dst) { return __builtin_amdgcn_ds_permute(dst << 2, v); }
ZHD double lane_put_d(double v, int dst) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)lane_put((int)(unsigned)b, dst);
    const unsigned hi = (unsigned)lane_put((int)(unsigned)(b >> 32), dst);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

/* The ranking of dense_targets (zone_pass.cpp:413-490).  Candidate o (lane
 * o < 24) is neighbour offset o; its key (deficit, count, Chebyshev and
 * squared distance, tie draw) is compared only with the other members of its
 * angular bin, fetched by lane permutes; the round-robin position follows
 * from the rank in the bin and the bins' sizes.  On return lane p < nOrd
 * (the return value, same on all lanes) holds the p-th target: zone cell,
 * threshold and count. */
template <class Z>
ZHD int dense_rank_wave(const Z& z, int sgx, int sgy, int sgc, int high, int pass, int phase, int lane, int& lcP,
                        double& thP, int& cP) {
    int gc = -1, lc = -1, cnt = 0, ddx = 0, ddy = 0;
    double th = 0.0;
    if (lane < 24) {
        const int f = lane < 12 ? lane : lane + 1;
        const int dy = f / 5 - 2, dx = f % 5 - 2;
        int jx, jy;
        if (shift_idx(z, sgx, sgy, dx, dy, jx, jy)) {
            const int l = zone_lc(z, jx, jy);
            if (l >= 0 && !fget(z, l, ZF_OUT)) {
                const int c = (int)z.cnt[l];
                const double t = thr(z, l, high);
                if ((double)c < t) {
                    gc = jy + z.g.Ny * jx;
                    lc = l;
                    cnt = c;
                    th = t;
                    min_step(z, sgx, sgy, jx, jy, ddx, ddy);
                }
            }
        }
    }
    bool keep = gc >= 0;
    if (z.zp.dedupe)
        for (int q = 0; q < 23; ++q) {
            const int gq = wave_readlane(gc, q);
            if (q < lane && gq == gc) keep = false;
        }
    const unsigned long long msk = wave_ballot(keep);
    const int m = wave_popc(msk);
    lcP = -1;
    cP = 0;
    thP = 0.0;
    if (m == 0) return 0;
    /* one Philox block per lane: tie break (lanes < 24), start bin (lane 24),
     * the counters of mpcd_draw_tie and mpcd_draw_start_bin */
    const mpcd_u32x4 rb = mpcd_philox(z.seed,
                                      mpcd_ctr0(lane < 24 ? MPCD_TAG_RTIE : MPCD_TAG_RBIN, (uint32_t)((pass << 1) | phase)),
                                      (uint32_t)sgc, lane < 24 ? (uint32_t)gc : 0u, 0u);
    const double tie = mpcd_u01(rb.w[0], rb.w[1]);
    const int sb = wave_readlane((int)mpcd_uint_below(rb.w[0], 8u), 24);
    const int bin = keep ? zone_bin(z.zp, (ddy + 2) * 5 + (ddx + 2)) : 8;
    const double defn = -dmax(0.0, th - (double)cnt);
    const int cheb = imax(ddx < 0 ? -ddx : ddx, ddy < 0 ? -ddy : ddy);
    /* (count, Chebyshev, squared distance) in one ordered word: cheb <= 2, dist2 <= 8 */
    const unsigned long long sm =
        ((unsigned long long)(unsigned)cnt << 8) | (unsigned long long)((cheb << 4) | (ddx * ddx + ddy * ddy));
    unsigned long long bm[8];
    unsigned long long mine = 0ull;
    int smax = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        bm[b] = wave_ballot(bin == b);
        smax = imax(smax, wave_popc(bm[b]));
        if (bin == b) mine = bm[b];
    }
    /* rank inside the bin: the stable order of the candidate key */
    unsigned long long mm = mine & ~(1ull << lane);
    int r = 0;
    for (int s = 1; s < smax; ++s) {
        const bool have = mm != 0ull;
        const int j = have ? (int)__builtin_ctzll(mm) : lane;
        if (have) mm &= mm - 1ull;
        const double od = lane_get_d(defn, j);
        const unsigned long long osm = lane_get_u64(sm, j);
        const double ot = lane_get_d(tie, j);
        if (have) {
            const bool oLess = od < defn || (od == defn && (osm < sm || (osm == sm && ot < tie)));
            const bool meLess = defn < od || (defn == od && (sm < osm || (sm == osm && tie < ot)));
            if (oLess || (j < lane && !meLess)) ++r;
        }
    }
    /* round-robin position: every bin's members of earlier rounds, then this
     * round's members of the bins before this one from the start bin */
    const int si = (bin - sb + 8) & 7;
    int pos = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const int nb = wave_popc(bm[b]);
        pos += imin(nb, r) + ((((b - sb + 8) & 7) < si && nb > r) ? 1 : 0);
    }
    const int dst = keep ? pos : m + wave_popc(~msk & lanes_below(lane));
    lcP = lane_put(lc, dst);
    cP = lane_put(cnt, dst);
    thP = lane_put_d(th, dst);
    return m;
}

/* One sweep series of spread (zone_pass.cpp:527-589) over the ranking in the
 * lanes: assigns movers used.. to candidates (myT), chains the movers of one
 * candidate in order (nextQ), and keeps each candidate's first and last mover
 * and new count.  All values uniform except the per-lane ones. */
ZHD int spread_plan_wave(int nOrd, double thP, int& cP, int need, int& avail, int& used, int cap, int lane,
                         int& myT, int& nextQ, int& firstQ, int& lastQ) {
    int moved = 0;
    myT = -1;
    firstQ = -1;
    lastQ = -1;
    if (need <= 0 || nOrd == 0 || avail <= 0) return 0;
    while (need > 0 && avail > 0) {
        const int defP = lane < nOrd ? imax(0, (int)MPCD_FLOOR(thP - (double)cP)) : 0;
        unsigned long long mk = wave_ballot(defP > 0);
        if (mk == 0ull) break;
        int R = wave_popc(mk);
        bool stop = false;
        for (; mk; mk &= mk - 1ull) {
            const int kk = (int)__builtin_ctzll(mk);
            const int d = wave_readlane(defP, kk);
            const int remaining = R--;
            const int quota = imax(1, (int)MPCD_CEIL((double)need / (double)imax(1, remaining)));
            /* cap (the staged movers) never binds: the callers stage
             * enough for both phases; it only keeps lanes in bounds */
            const int nt = imin(imax(1, imin(imin(quota, d), imin(need, avail))), cap - used);
            if (nt <= 0) {
                stop = true;
                break;
            }
            const int prev = wave_readlane(lastQ, kk);
            if (lane >= used && lane < used + nt) {
                myT = kk;
                nextQ = lane + 1 < used + nt ? lane + 1 : -1;
            }
            if (prev >= 0 && lane == prev) nextQ = used;
            if (lane == kk) {
                cP += nt;
                if (firstQ < 0) firstQ = used;
                lastQ = used + nt - 1;
            }
            used += nt;
            avail -= nt;
            need -= nt;
            moved += nt;
            if (need <= 0 || avail <= 0) {
                stop = true;
                break;
            }
        }
        if (stop) break;
    }
    return moved;
}

/* Applies a plan: pops the movers from src's head, records their moves,
 * appends them to their targets in mover order. */
template <class Z>
ZHD void spread_apply_wave(Z& z, int src, bool srcMat, int lane, int q0, int moved, int avail, int pass, int myK,
                           int myT, int nextQ, int lcP, int cP, int firstQ, int lastQ) {
    if (moved <= 0) return;
    /* src first: a materialised list's next head is read before the movers'
     * links are rewritten */
    const int kLast = wave_readlane(myK, q0 + moved - 1);
    const int nh = srcMat ? nx_get(z, kLast) : (kLast < tail_of(z, src) ? kLast + 1 : -1);
    const int nk = lane_get(myK, nextQ >= 0 ? nextQ : 0);
    const int fk = lane_get(myK, firstQ >= 0 ? firstQ : 0);
    const int lk = lane_get(myK, lastQ >= 0 ? lastQ : 0);
    ZWSYNC();
    if (lane == 0) {
        head_set(z, src, nh);
        if (nh < 0) tail_set(z, src, -1);
        fset(z, src, ZF_CHG);
        cnt_set(z, src, avail);
    }
    if (myT >= 0) {
        mark_moved(z, myK, pass);
        nx_set(z, myK, nextQ >= 0 ? nk : -1);
    }
    /* targets whose lists are still implicit runs: their links written by
     * all lanes at once, one target at a time */
    for (unsigned long long nm = wave_ballot(firstQ >= 0 && !fget(z, lcP, ZF_MAT)); nm; nm &= nm - 1ull) {
        const int lt = wave_readlane(lcP, (int)__builtin_ctzll(nm));
        const int h = head_of(z, lt), t = tail_of(z, lt);
        if (h >= 0)
            for (int k = h + lane; k <= t; k += 64) nx_set(z, k, k < t ? k + 1 : -1);
        ZWSYNC();
        if (lane == 0) fset(z, lt, ZF_MAT);
        ZWSYNC();
    }
    if (firstQ >= 0) {
        ensure_mat(z, lcP);
        const int t = tail_of(z, lcP);
        if (t < 0) head_set(z, lcP, fk);
        else nx_set(z, t, fk);
        tail_set(z, lcP, lk);
        fset(z, lcP, ZF_CHG);
        cnt_set(z, lcP, cP);
    }
    ZWSYNC();
}

/* One dense cell (zone_pass.cpp:738-779: both target phases, the spreads and
 * the transfer masks) when its movers fit the wavefront:
 * min(needOut + 1, count) <= 64 (phase 1 moves at most what phase 0 left of
 * needOut, up to one more by the rounding of its ceil).  Returns the number
 * of particles moved. */
template <class Z>
ZHD int dense_cell_wave(Z& z, int src, int nc, double hl, int needOut, int pass, int lane) {
    const int sgx = lc_gx(z, src), sgy = lc_gy(z, src), sgc = lc_gc(z, src);
    const int nStage = imin(needOut + 1, nc);
    const bool srcMat = fget(z, src, ZF_MAT);
    /* movers: lane q is the q-th particle of src's list */
    int myK = -1;
    if (srcMat) {
        if (lane == 0) {
            int k = head_of(z, src);
            for (int q = 0; q < nStage && k >= 0; ++q) {
                z.mvk[q] = k;
                k = nx_get(z, k);
            }
        }
        ZWSYNC();
        if (lane < nStage) myK = z.mvk[lane];
    } else if (lane < nStage) {
        myK = head_of(z, src) + lane;
    }
    int used = 0, avail = nc, moved = 0;
    int lcP, cP, myT, nextQ = -1, firstQ, lastQ;
    double thP;
    /* phase 0: low thresholds */
    const int n0 = dense_rank_wave(z, sgx, sgy, sgc, 0, pass, 0, lane, lcP, thP, cP);
    int q0 = used;
    int m0 = spread_plan_wave(n0, thP, cP, needOut, avail, used, nStage, lane, myT, nextQ, firstQ, lastQ);
    spread_apply_wave(z, src, srcMat, lane, q0, m0, avail, pass, myK, myT, nextQ, lcP, cP, firstQ, lastQ);
    moved += m0;
    const int rcv0 = firstQ >= 0 ? lcP : -1;
    /* phase 1: high thresholds, for what is left above the high threshold */
    int rcv1 = -1;
    const int remaining = imax(0, (int)MPCD_CEIL((double)avail - hl));
    if (remaining > 0) {
        const int n1 = dense_rank_wave(z, sgx, sgy, sgc, 1, pass, 1, lane, lcP, thP, cP);
        q0 = used;
        nextQ = -1;
        const int m1 = spread_plan_wave(n1, thP, cP, remaining, avail, used, nStage, lane, myT, nextQ, firstQ, lastQ);
        spread_apply_wave(z, src, srcMat, lane, q0, m1, avail, pass, myK, myT, nextQ, lcP, cP, firstQ, lastQ);
        moved += m1;
        rcv1 = firstQ >= 0 ? lcP : -1;
    }
    /* transfer masks (zone_pass.cpp:492-525) over both phases' receivers */
    const bool nonBulk = (rcv0 >= 0 && !fget(z, rcv0, ZF_BULK0)) || (rcv1 >= 0 && !fget(z, rcv1, ZF_BULK0));
    const bool allBulk = fget(z, src, ZF_BULK0) && wave_ballot(nonBulk) == 0ull;
    const unsigned short setb = (unsigned short)(ZF_RECV | (allBulk ? ZF_BB : ZF_SKIP));
    const unsigned short keepb = (unsigned short)(allBulk ? 0xffff : ~ZF_BB);
    if (rcv0 >= 0) z.fl[rcv0] = (unsigned short)((z.fl[rcv0] | setb) & keepb);
    ZWSYNC();
    if (rcv1 >= 0) z.fl[rcv1] = (unsigned short)((z.fl[rcv1] | setb) & keepb);
    if (lane == 0) {
        if (allBulk) fset(z, src, ZF_BB);
        else z.fl[src] = (unsigned short)((z.fl[src] | ZF_SKIP) & ~ZF_BB);
    }
    ZWSYNC();
    return moved;
}
#endif

/* Dense passes (zone_pass.cpp:708-781), all lanes. */
template <class Z>
ZHD void zone_dense_passes(Z& z, int lane, int nl) {
#if defined(__HIP_DEVICE_COMPILE__)
    int nHighW = 0, movedW = 0;
#endif
    for (int pass = 1; pass <= z.zp.maxPasses; ++pass) {
        for (int lc = lane; lc < z.C; lc += nl) zone_maps(z, lc, 0);
        ZBAR();
        const int m = zone_compact(z, lane, nl, ZC_DENSE);
        ZBAR();
        zone_shuffle(z.dl, m, lane, nl, ShufDraw{z.seed, ZS_DENSE, (uint32_t)pass, 0u});
        for (int k = 0; k < m; ++k) {
            const int c = (int)z.dl[k];
            const int nc = cnt_of(z, c);
            if (nc <= 0) continue;
            const double hl = high_thr(z, c);
            if ((double)nc <= hl) continue;
            const int needOut = imax(0, (int)MPCD_CEIL((double)nc - hl));
#if defined(__HIP_DEVICE_COMPILE__)
            if (needOut > 0 && imin(needOut + 1, nc) <= ZONE_MAX_MOVE) {
                ++nHighW;
                movedW += dense_cell_wave(z, c, nc, hl, needOut, pass, lane);
                continue;
            }
#endif
            int moved = 0;
            if (lane == 0) {
                ++z.sh->nHigh;
                z.sh->nr = 0;
            }
            if (needOut > 0) {
                dense_targets(z, c, 0, pass, 0, lane, nl);
                if (lane == 0) {
                    int nr = 0;
                    moved += spread(z, c, needOut, pass, z.recvT, nr);
                    z.sh->nr = nr;
                }
                ZBAR();
                const int remaining = imax(0, (int)MPCD_CEIL((double)cnt_of(z, c) - hl));
                if (remaining > 0) {
                    dense_targets(z, c, 1, pass, 1, lane, nl);
                    if (lane == 0) {
                        int nr = z.sh->nr;
                        moved += spread(z, c, remaining, pass, z.recvT, nr);
                        z.sh->nr = nr;
                    }
                }
            }
            if (lane == 0) {
                const int nr = z.sh->nr;
                for (int r = 0; r < nr; ++r) fset(z, z.recvT[r], ZF_RECV);
                transfer_masks(z, c, z.recvT, nr);
                z.sh->movedDense += moved;
            }
            ZBAR();
        }
    }
#if defined(__HIP_DEVICE_COMPILE__)
    if (lane == 0) {
        z.sh->nHigh += nHighW;
        z.sh->movedDense += movedW;
    }
    ZBAR();
#endif
}

#if defined(__HIP_DEVICE_COMPILE__)
/* The dense passes on the four wavefronts of a 256-thread workgroup
 * (device only).  A dense cell's spread reads counts within Chebyshev range 2
 * of it and writes the lists, counts and flags of itself and of targets
 * within range 2; its own count changes only by itself (a dense cell is above
 * its high threshold, a target below it).  Cells of the shuffled list more
 * than 4 cells apart therefore touch disjoint cells and commute.  A cell
 * whose movers exceed a wavefront takes the one-lane form, which uses the
 * zone's shared scratch.  Waves 1-3 stage their movers in the one-lane form's
 * candidate scratch, free while they run. */
constexpr int ZONE_WAVES = 4;
static_assert(24 * sizeof(Cand) >= (ZONE_WAVES - 1) * ZONE_MAX_MOVE * sizeof(int),
              "the candidate scratch holds the other wavefronts' movers");

/* The dense list on the four wavefronts in list order.  Each wavefront takes
 * the next cell from a counter in LDS; a cell starts once every earlier live
 * cell within Chebyshev range 4 is done.  Those are found through pos (the
 * list position of each zone cell, after the m working words in ext), and a
 * finished cell sets DL_DONE in its word.  Conflicting cells thus keep the
 * list order and the result is the serial loop's bit for bit, while a
 * wavefront waits only for the cells its own depends on.  Cells are taken in
 * increasing order and wait only on smaller positions, which are taken
 * already, so the smallest unfinished cell can always run; each wait is
 * bounded all the same (past DS_WAIT_MAX polls it records schedWait, which
 * the kernel reports as an error, and goes on).  A list with a cell that
 * needs the one-lane form (its movers exceed a wavefront), or one that leaves
 * ext no room for pos, runs on wavefront 0 alone, in list order. */
constexpr unsigned DL_LIVE = 0x8000u, DL_ONE = 0x4000u, DL_DONE = 0x2000u;
constexpr int DS_WAIT_MAX = 1 << 22;

template <class Z>
__device__ __forceinline__ void zone_dense_passes_mw(Z& z, int wave, int lane) {
    typedef typename Z::idx_t IT;
    constexpr int NT = 64 * ZONE_WAVES;
    constexpr unsigned NONE = (unsigned)(IT)~(IT)0;
    const int tid = wave * 64 + lane;
    int nHighW = 0, movedW = 0;
    if (wave > 0) z.mvk = reinterpret_cast<int*>(z.cand) + (wave - 1) * ZONE_MAX_MOVE;
    if (tid == 0) z.sh->schedWait = 0;
    const int nExt = (z.zw + 4) * (z.zh + 4);
    const bool wrapX = z.g.perx && z.zw == z.g.Nx, wrapY = z.g.pery && z.zh == z.g.Ny;
    volatile IT* vext = z.ext;
    for (int pass = 1; pass <= z.zp.maxPasses; ++pass) {
        for (int lc = tid; lc < z.C; lc += NT) zone_maps(z, lc, 0);
        __syncthreads();
        if (wave == 0) {
            const int m = zone_compact(z, lane, 64, ZC_DENSE);
            ZBAR();
            zone_shuffle(z.dl, m, lane, 64, ShufDraw{z.seed, ZS_DENSE, (uint32_t)pass, 0u});
            if (lane == 0) {
                z.sh->mwCount = m;
                z.sh->mwNext = 0;
            }
        }
        __syncthreads();
        const int m = z.sh->mwCount;
        bool one = false;
        for (int k = tid; k < m; k += NT) {
            const int c = (int)z.dl[k];
            const int nc = cnt_of(z, c);
            const double hl = high_thr(z, c);
            unsigned w = 0u;
            if (nc > 0 && (double)nc > hl) {
                const int needOut = imax(0, (int)MPCD_CEIL((double)nc - hl));
                const bool o = !(needOut > 0 && imin(needOut + 1, nc) <= ZONE_MAX_MOVE);
                one = one || o;
                w = DL_LIVE | (o ? DL_ONE : 0u);
            }
            z.ext[k] = (IT)w;
        }
        const bool serial = __syncthreads_or(one) || m + z.C > nExt;
        IT* pos = z.ext + m;
        if (!serial) {
            for (int lc = tid; lc < z.C; lc += NT) pos[lc] = (IT)NONE;
            __syncthreads();
            for (int k = tid; k < m; k += NT)
                if ((unsigned)z.ext[k] & DL_LIVE) pos[(int)z.dl[k]] = (IT)k;
            __syncthreads();
        }
        if (!serial || wave == 0)
            for (;;) {
                int k = 0;
                if (lane == 0) k = atomicAdd(&z.sh->mwNext, 1);
                k = __shfl(k, 0, 64);
                if (k >= m) break;
                const unsigned wk = (unsigned)vext[k];
                if (!(wk & DL_LIVE)) continue;
                const int c = (int)z.dl[k];
                if (!serial) {
                    /* the window's earlier live cells: lane t checks offsets t and t + 64 of the 81 */
                    const int cx = c / z.zh, cy = c - cx * z.zh;
                    int q0 = -1, q1 = -1;
                    for (int h = 0; h < 2; ++h) {
                        const int t = lane + 64 * h;
                        int qq = -1;
                        if (t < 81) {
                            int x = cx + t / 9 - 4, y = cy + t % 9 - 4;
                            bool in = true;
                            if (wrapX) x = x < 0 ? x + z.zw : (x >= z.zw ? x - z.zw : x);
                            else in = in && x >= 0 && x < z.zw;
                            if (wrapY) y = y < 0 ? y + z.zh : (y >= z.zh ? y - z.zh : y);
                            else in = in && y >= 0 && y < z.zh;
                            if (in) {
                                const unsigned pq = (unsigned)pos[x * z.zh + y];
                                if (pq < (unsigned)k) qq = (int)pq;
                            }
                        }
                        if (h == 0) q0 = qq;
                        else q1 = qq;
                    }
                    for (int polls = 0;; ++polls) {
                        const bool wait0 = q0 >= 0 && !((unsigned)vext[q0] & DL_DONE);
                        const bool wait1 = q1 >= 0 && !((unsigned)vext[q1] & DL_DONE);
                        if (wave_ballot(wait0 || wait1) == 0ull) break;
                        if (polls >= DS_WAIT_MAX) {
                            if (lane == 0) z.sh->schedWait = 1;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                }
                const int nc = cnt_of(z, c);
                const double hl = high_thr(z, c);
                const int needOut = imax(0, (int)MPCD_CEIL((double)nc - hl));
                if (!(wk & DL_ONE)) {
                    ++nHighW;
                    movedW += dense_cell_wave(z, c, nc, hl, needOut, pass, lane);
                } else {
                    /* the one-lane form (serial lists only: wavefront 0, the
                     * zone's shared scratch) */
                    int moved = 0;
                    if (lane == 0) {
                        ++z.sh->nHigh;
                        z.sh->nr = 0;
                    }
                    dense_targets(z, c, 0, pass, 0, lane, 64);
                    if (lane == 0) {
                        int nr = 0;
                        moved += spread(z, c, needOut, pass, z.recvT, nr);
                        z.sh->nr = nr;
                    }
                    ZBAR();
                    const int remaining = imax(0, (int)MPCD_CEIL((double)cnt_of(z, c) - hl));
                    if (remaining > 0) {
                        dense_targets(z, c, 1, pass, 1, lane, 64);
                        if (lane == 0) {
                            int nr = z.sh->nr;
                            moved += spread(z, c, remaining, pass, z.recvT, nr);
                            z.sh->nr = nr;
                        }
                    }
                    if (lane == 0) {
                        const int nr = z.sh->nr;
                        for (int r = 0; r < nr; ++r) fset(z, z.recvT[r], ZF_RECV);
                        transfer_masks(z, c, z.recvT, nr);
                        z.sh->movedDense += moved;
                    }
                    ZBAR();
                }
                if (!serial) {
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                    if (lane == 0) vext[k] = (IT)(wk | DL_DONE);
                    ZWSYNC();
                }
            }
        __syncthreads();
    }
    if (lane == 0) {
        atomicAdd(&z.sh->nHigh, nHighW);
        atomicAdd(&z.sh->movedDense, movedW);
    }
    __syncthreads();
}
#else
/* (the host passes of the compilers see a declaration only) */
template <class Z>
ZHD void zone_dense_passes_mw(Z&, int, int) {}
#endif

/* ---- sparse phase ---- */

/* inward_direction_from_local_gradient (zone_pass.cpp:375-411) on the
 * sparse-phase count copy; the four smoothed counts in parallel. */
template <class Z>
ZHD void inward_dir(Z& z, int gx, int gy, int lane, int nl, double& dxo, double& dyo) {
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
        z.nsm4[t] = nsm_ext(z, qx, qy);
    }
    ZBAR();
    dxo = (fx == 0.5) ? 0.5 * (z.nsm4[0] - z.nsm4[1]) : (z.nsm4[0] - z.nsm4[1]);
    dyo = (fy == 0.5) ? 0.5 * (z.nsm4[2] - z.nsm4[3]) : (z.nsm4[2] - z.nsm4[3]);
}

/* One donor transfer of the sparse phase (zone_pass.cpp:818-850): copy
 * donor d's list into tmp (LDS or the zone's global region), Fisher-Yates
 * it, move the first nt to low cell c; all lanes. */
template <class Z, class T>
ZHD void donor_transfer(Z& z, int c, int d, int md, int nt, int cgc, T* tmp, int lane, int nl) {
    typedef typename Z::idx_t IT;
    /* copy the donor list, Fisher-Yates it, move the first nt */
    if (!fget(z, d, ZF_MAT)) {
        const int h = head_of(z, d);
        for (int q = lane; q < md; q += nl) tmp[q] = (IT)(h + q);
    } else if (lane == 0) {
        int kk = head_of(z, d);
        for (int q = 0; q < md; ++q) {
            tmp[q] = (IT)kk;
            kk = nx_get(z, kk);
        }
    }
    ZBAR();
    const int dgc = lc_gc(z, d);
    zone_shuffle(tmp, md, lane, nl,
                 ShufDraw{z.seed, ZS_PARTICLE, (uint32_t)cgc, (uint32_t)dgc});
    /* the movers' records, links of the moved run and of the donor's
     * remaining run, all lanes */
    for (int q = lane; q < md; q += nl) {
        const int kq = (int)tmp[q];
        if (q < nt) {
            mark_moved(z, kq, ZM_SPARSE);
            nx_set(z, kq, q + 1 < nt ? (int)tmp[q + 1] : -1);
        } else {
            nx_set(z, kq, q + 1 < md ? (int)tmp[q + 1] : -1);
        }
    }
    ensure_mat_lanes(z, c, lane, nl);
    if (lane == 0) {
        const int tc = tail_of(z, c);
        if (tc < 0) head_set(z, c, (int)tmp[0]);
        else nx_set(z, tc, (int)tmp[0]);
        tail_set(z, c, (int)tmp[nt - 1]);
        fset(z, c, ZF_CHG);
        fset(z, d, ZF_MAT | ZF_CHG);
        head_set(z, d, md > nt ? (int)tmp[nt] : -1);
        tail_set(z, d, md > nt ? (int)tmp[md - 1] : -1);
        cnt_set(z, d, md - nt);
        cnt_set(z, c, cnt_of(z, c) + nt);
        fset(z, c, ZF_RECV);
        const int rv1[1] = {c};
        transfer_masks(z, d, rv1, 1);
    }
    ZBAR();
}

/* One low cell of the sparse phase (zone_pass.cpp:804-854): its donors among
 * the zone neighbours at range 1, ranked, and the transfers; all lanes of
 * one wavefront.  Returns the particles moved into c. */
template <class Z>
ZHD int sparse_cell(Z& z, int c, int nc, int lowInt, int lane, int nl) {
    const int cgx = lc_gx(z, c), cgy = lc_gy(z, c), cgc = lc_gc(z, c);
    /* zone neighbours at range 1 (local_neighbors_in_zone :263-278) */
    for (int o = lane; o < 8; o += nl) {
        const int f = o < 4 ? o : o + 1;
        const int dy = f / 3 - 1, dx = f % 3 - 1;
        int jx, jy, q = -1;
        if (shift_idx(z, cgx, cgy, dx, dy, jx, jy)) q = zone_lc(z, jx, jy);
        if (q >= 0 && fget(z, q, ZF_OUT)) q = -1;
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
            if (z.zp.dedupe)
                for (int t = 0; t < o && keep; ++t)
                    if (z.nbGc[t] == q) keep = false;
        }
        const unsigned long long mk = wave_ballot(keep);
        nn += wave_popc(mk);
        const bool don = keep && cnt_of(z, z.nbGc[o]) > lowInt;
        const unsigned long long mdn = wave_ballot(don);
        if (don) z.dord[nd + wave_popc(mdn & lanes_below(lane))] = z.nbGc[o];
        nd += wave_popc(mdn);
    }
    ZBAR();
    if (nn == 0) return 0;
    int need = imax(0, lowInt - nc);
    if (need <= 0 || nd == 0) return 0;
    double gxd, gyd;
    inward_dir(z, cgx, cgy, lane, nl, gxd, gyd);
    /* rank_admissible_neighbors_inward (zone_pass.cpp:636-667) */
    zone_shuffle(z.dord, nd, lane, nl, ShufDraw{z.seed, ZS_DONOR, (uint32_t)cgc, 0u});
    const double ndir = mpcd_hypot(gxd, gyd);
    for (int t = lane; t < nd; t += nl) {
        const int q = z.dord[t];
        int ddx, ddy;
        min_step(z, cgx, cgy, lc_gx(z, q), lc_gy(z, q), ddx, ddy);
        const double nstep = mpcd_hypot((double)ddx, (double)ddy);
        double align = 0.0;
        if (ndir >= 1e-14 && nstep >= 1e-14) align = ((double)ddx * gxd + (double)ddy * gyd) / (ndir * nstep);
        const double noise = 1e-9 * mpcd_draw_donor_noise(z.seed, (uint32_t)cgc, (uint32_t)lc_gc(z, q));
        z.dsc[t] = 100.0 * align + nsm_ext(z, lc_gx(z, q), lc_gy(z, q)) + noise;
    }
    ZBAR();
    for (int t = lane; t < nd; t += nl) { /* stable, descending score */
        int r = 0;
        for (int u = 0; u < nd; ++u)
            if (z.dsc[u] > z.dsc[t] || (u < t && !(z.dsc[t] > z.dsc[u]))) ++r;
        z.dord2[r] = z.dord[t];
    }
    ZBAR();
    int moved = 0;
    for (int t = 0; t < nd; ++t) {
        const int d = z.dord2[t];
        const int md = cnt_of(z, d);
        if (md <= 0) continue;
        const int avail = md - lowInt;
        if (avail <= 0) continue;
        const int nt = imin(imin(need, avail), md);
        if (nt <= 0) continue;
        if (md <= z.tmpCap) donor_transfer(z, c, d, md, nt, cgc, z.tmpL, lane, nl);
        else donor_transfer(z, c, d, md, nt, cgc, z.tmpG, lane, nl);
        moved += nt;
        need -= nt;
        if (need <= 0) break;
    }
    return moved;
}

/* the sparse phase's set-up: maps, and the counts it starts from (the zone
 * and its ring) in z.ext; threads tid .. of nt */
template <class Z>
ZHD void sparse_setup(Z& z, int tid, int nt) {
    for (int lc = tid; lc < z.C; lc += nt) zone_maps(z, lc, 0);
    const int H = z.zh + 4, ne = (z.zw + 4) * H;
    for (int e = tid; e < ne; e += nt) {
        const int rx = e / H - 2, ry = e % H - 2;
        const int h = halo_slot(z, rx, ry);
        z.ext[e] = h >= 0 ? z.halo[h] : z.cnt[rx * z.zh + ry];
    }
    if (z.dgCnt)
        for (int lc = tid; lc < z.C; lc += nt) z.dgCnt[lc] = (int)z.cnt[lc];
}

/* Sparse pass (zone_pass.cpp:783-854), all lanes. */
template <class Z>
ZHD void zone_sparse_pass(Z& z, int lane, int nl) {
    sparse_setup(z, lane, nl);
    ZBAR();
    const int m = zone_compact(z, lane, nl, ZC_LOW);
    ZBAR();
    zone_shuffle(z.dl, m, lane, nl, ShufDraw{z.seed, ZS_LOW, 0u, 0u});
    const int lowInt = (int)MPCD_CEIL(z.zp.lowThrBulk);
    for (int k = 0; k < m; ++k) {
        const int c = (int)z.dl[k];
        const int nc = cnt_of(z, c);
        if (nc <= 0) continue;
        if (lane == 0) ++z.sh->nLow;
        const int moved = sparse_cell(z, c, nc, lowInt, lane, nl);
        if (lane == 0) z.sh->movedSparse += moved;
    }
}

#if defined(__HIP_DEVICE_COMPILE__)
/* The sparse phase on the four wavefronts (device only).  A low cell's
 * transfers read and change only itself and its zone neighbours at range 1
 * (its own count changes only by itself: a low cell is never a donor), and
 * the smoothed counts they rank by are the phase's starting counts, so low
 * cells more than 2 apart commute.  The wavefronts take the cells in list
 * order from a counter in LDS, and a cell starts once no earlier unfinished
 * cell lies within range 2 of it (lix, each cell's list index, is cleared
 * when the cell is finished); the waits are bounded as in the dense phase.
 * Conflicting cells keep the list order, so the result is the serial loop's
 * bit for bit.  A list with a cell one of whose neighbours starts with more
 * particles than a wavefront's share of the transfer scratch runs on
 * wavefront 0 alone, in list order (it has the full scratch and the global
 * region). */
constexpr int ZONE_TMP_W = 112; /* transfer scratch of each of wavefronts 1..3 */

/* a low cell with a zone neighbour at range 1 that starts the phase with
 * more particles than a wavefront's transfer scratch holds (donor lists only
 * shrink in the phase) */
template <class Z>
ZHD bool sparse_alone(const Z& z, int c) {
    const int H = z.zh + 4;
    bool big = false;
    int jx, jy;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            if ((dx || dy) && shift_idx(z, lc_gx(z, c), lc_gy(z, c), dx, dy, jx, jy)) {
                const int q = zone_lc(z, jx, jy);
                if (q >= 0 && !fget(z, q, ZF_OUT) && iget(z.ext[(q / z.zh + 2) * H + q % z.zh + 2]) > ZONE_TMP_W)
                    big = true;
            }
    return big;
}

template <class Z>
ZHD bool sparse_conflict(const Z& z, int a, int b) {
    int dx = a / z.zh - b / z.zh, dy = a % z.zh - b % z.zh;
    dx = dx < 0 ? -dx : dx;
    dy = dy < 0 ? -dy : dy;
    if (z.g.perx && z.zw == z.g.Nx) dx = imin(dx, z.zw - dx);
    if (z.g.pery && z.zh == z.g.Ny) dy = imin(dy, z.zh - dy);
    return dx <= 2 && dy <= 2;
}

template <class Z>
__device__ __forceinline__ void zone_sparse_pass_mw(Z& z, int wave, int lane) {
    typedef typename Z::idx_t IT;
    static_assert(sizeof(IT) == 2, "the wavefronts' scratch shares are laid out for 16-bit indices");
    const int tid = wave * 64 + lane;
    if (wave > 0) {
        /* this wavefront's scratch: 448 bytes of the dense phase's one-lane
         * candidate and threshold arrays (cand, ordThr, nbTie, contiguous and
         * idle now) */
        unsigned char* b = reinterpret_cast<unsigned char*>(z.cand) + (wave - 1) * 448;
        z.dsc = reinterpret_cast<double*>(b);
        z.nsm4 = reinterpret_cast<double*>(b + 64);
        z.nbGc = reinterpret_cast<int*>(b + 96);
        z.dord = reinterpret_cast<int*>(b + 128);
        z.dord2 = reinterpret_cast<int*>(b + 160);
        z.tmpL = reinterpret_cast<IT*>(b + 224);
        z.tmpCap = ZONE_TMP_W;
    }
    const IT none = (IT)~(IT)0;
    sparse_setup(z, tid, 64 * ZONE_WAVES);
    for (int lc = tid; lc < z.C; lc += 64 * ZONE_WAVES) z.lix[lc] = none;
    __syncthreads();
    if (wave == 0) {
        const int m = zone_compact(z, lane, 64, ZC_LOW);
        ZBAR();
        zone_shuffle(z.dl, m, lane, 64, ShufDraw{z.seed, ZS_LOW, 0u, 0u});
        if (lane == 0) z.sh->mwCount = m;
    }
    __syncthreads();
    const int m = z.sh->mwCount;
    bool anyAlone = false;
    for (int k = tid; k < m; k += 64 * ZONE_WAVES) {
        const int c = (int)z.dl[k];
        z.lix[c] = (IT)k;
        anyAlone = anyAlone || sparse_alone(z, c);
    }
    anyAlone = __syncthreads_or(anyAlone);
    const int lowInt = (int)MPCD_CEIL(z.zp.lowThrBulk);
    const bool wx = z.g.perx && z.zw == z.g.Nx, wy = z.g.pery && z.zh == z.g.Ny;
    int nLowW = 0, movedW = 0;
    const bool serial = anyAlone;
    if (tid == 0) z.sh->mwNext = 0;
    __syncthreads();
    volatile IT* vlix = z.lix;
    if (!serial || wave == 0)
        for (;;) {
            int k = 0;
            if (lane == 0) k = atomicAdd(&z.sh->mwNext, 1);
            k = __shfl(k, 0, 64);
            if (k >= m) break;
            const int c = (int)z.dl[k];
            if (!serial) {
                int q = -1;
                if (lane < 25) {
                    const int rx = c / z.zh, ry = c - rx * z.zh;
                    int qx = rx + lane / 5 - 2, qy = ry + lane % 5 - 2;
                    if (wx) qx = wrapi(qx, z.zw);
                    if (wy) qy = wrapi(qy, z.zh);
                    if (qx >= 0 && qx < z.zw && qy >= 0 && qy < z.zh) q = qx * z.zh + qy;
                }
                for (int polls = 0;; ++polls) {
                    const bool wait = q >= 0 && (int)vlix[q] < k;
                    if (wave_ballot(wait) == 0ull) break;
                    if (polls >= DS_WAIT_MAX) {
                        if (lane == 0) z.sh->schedWait = 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            }
            const int nc = cnt_of(z, c);
            if (nc > 0) {
                ++nLowW;
                movedW += sparse_cell(z, c, nc, lowInt, lane, 64);
            }
            if (!serial) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                if (lane == 0) vlix[c] = none;
                ZWSYNC();
            }
        }
    __syncthreads();
    if (lane == 0) {
        atomicAdd(&z.sh->nLow, nLowW);
        atomicAdd(&z.sh->movedSparse, movedW);
    }
    __syncthreads();
}
#else
template <class Z>
ZHD void zone_sparse_pass_mw(Z&, int, int) {}
#endif

/* ---- set-up, correction, statistics ---- */

/* zone_setup's flags of the cell with global index gc */
template <class Z>
ZHD unsigned short setup_flags(const Z& z, const ZoneIn& in, int gc) {
    unsigned short f = 0;
    if (!in.owner || in.owner[gc] == z.zid) f |= ZF_OWN;
    if (in.gcover && in.gcover[gc]) f |= ZF_COVER;
    if (in.zmask && !in.zmask[gc]) f |= ZF_OUT;
    return f;
}

/* zone_setup's entry h of the ring of pass-input counts around the zone */
template <class Z>
ZHD void setup_halo(Z& z, int h) {
    const int W4 = z.zw + 4;
    int hx, hy;
    if (h < 2 * W4) {
        hx = h / 2;
        hy = h % 2;
    } else if (h < 4 * W4) {
        hx = (h - 2 * W4) / 2;
        hy = z.zh + 2 + (h - 2 * W4) % 2;
    } else if (h < 4 * W4 + 2 * z.zh) {
        hx = (h - 4 * W4) / z.zh;
        hy = 2 + (h - 4 * W4) % z.zh;
    } else {
        hx = z.zw + 2 + (h - 4 * W4 - 2 * z.zh) / z.zh;
        hy = 2 + (h - 4 * W4 - 2 * z.zh) % z.zh;
    }
    int gx = z.xs + hx - 2, gy = z.ys + hy - 2;
    if (z.g.perx) gx = wrapi(gx, z.g.Nx);
    if (z.g.pery) gy = wrapi(gy, z.g.Ny);
    int v = 0;
    if (gx >= 0 && gx < z.g.Nx && gy >= 0 && gy < z.g.Ny) v = z.gcnt[gy + z.g.Ny * gx];
    /* the unsigned 16-bit type keeps its top value for "read the global count" */
    if (sizeof(typename Z::idx_t) == 2 && v >= 0xFFFF) v = -1;
    z.halo[h] = (typename Z::idx_t)v;
}

/* Zone setup (zone_pass.cpp:680-685), all lanes: counts, implicit list of
 * each cell (its run of local indices, from a wave prefix sum over the
 * zone's cells), the column map and the ownership flags. */
template <class Z>
ZHD void zone_setup(Z& z, const ZoneIn& in, int lane, int nl) {
    int base = 0;
    bool mixed = false;
    for (int b0 = 0; b0 < z.C; b0 += nl) {
        const int lc = b0 + lane;
        const int gc = lc < z.C ? lc_gc(z, lc) : 0;
        const int m = lc < z.C ? z.gcnt[gc] : 0;
        const int incl = wave_shfl_up_add(m, lane);
        if (lc < z.C) {
            const int k0 = base + incl - m;
            cnt_set(z, lc, m);
            z.k0[lc] = (typename Z::idx_t)k0;
            head_set(z, lc, m > 0 ? k0 : -1);
            tail_set(z, lc, m > 0 ? k0 + m - 1 : -1);
            const unsigned short f = setup_flags(z, in, gc);
            if ((f & (ZF_OWN | ZF_COVER)) != ZF_OWN) mixed = true;
            z.fl[lc] = f;
            if (lc % z.zh == 0) {
                z.colL[lc / z.zh] = k0;
                z.colG[lc / z.zh] = z.gstart[gc];
            }
        }
        /* lanes past the zone add 0, so the last lane holds the chunk total */
        base += wave_bcast(incl, nl - 1);
    }
    const bool anyMixed = wave_ballot(mixed) != 0ull;
    /* the ring of pass-input counts around the zone */
    const int nh = halo_size(z);
    for (int h = lane; h < nh; h += nl) setup_halo(z, h);
    if (lane == 0) {
        z.sh->P = base;
        z.sh->allOwned = anyMixed ? 0 : 1;
        z.sh->nHigh = 0;
        z.sh->nLow = 0;
        z.sh->movedDense = 0;
        z.sh->movedSparse = 0;
        z.sh->stats4 = 0.0;
        z.sh->stats5 = 0.0;
    }
}

/* zone_setup on the four wavefronts of a 256-thread workgroup: wavefront 0
 * takes the counts, runs and column starts (its prefix sum over the cells),
 * wavefronts 1..3 the flags, the column map's global starts and the halo
 * ring; each of them records in sh->mixedW whether it saw a cell that is not
 * plainly owned.  After a workgroup barrier, zone_setup_all_owned gives the
 * zone's allOwned. */
template <class Z>
ZHD void zone_setup_mw(Z& z, const ZoneIn& in, int wave, int lane) {
    if (wave == 0) {
        int base = 0;
        for (int b0 = 0; b0 < z.C; b0 += 64) {
            const int lc = b0 + lane;
            const int m = lc < z.C ? z.gcnt[lc_gc(z, lc)] : 0;
            const int incl = wave_shfl_up_add(m, lane);
            if (lc < z.C) {
                const int k0 = base + incl - m;
                cnt_set(z, lc, m);
                z.k0[lc] = (typename Z::idx_t)k0;
                head_set(z, lc, m > 0 ? k0 : -1);
                tail_set(z, lc, m > 0 ? k0 + m - 1 : -1);
                if (lc % z.zh == 0) z.colL[lc / z.zh] = k0;
            }
            base += wave_bcast(incl, 63);
        }
        if (lane == 0) {
            z.sh->P = base;
            z.sh->nHigh = 0;
            z.sh->nLow = 0;
            z.sh->movedDense = 0;
            z.sh->movedSparse = 0;
            z.sh->stats4 = 0.0;
            z.sh->stats5 = 0.0;
        }
        return;
    }
    const int t = (wave - 1) * 64 + lane;
    bool mixed = false;
    for (int lc = t; lc < z.C; lc += 192) {
        const int gc = lc_gc(z, lc);
        const unsigned short f = setup_flags(z, in, gc);
        if ((f & (ZF_OWN | ZF_COVER)) != ZF_OWN) mixed = true;
        z.fl[lc] = f;
        if (lc % z.zh == 0) z.colG[lc / z.zh] = z.gstart[gc];
    }
    const int nh = halo_size(z);
    for (int h = t; h < nh; h += 192) setup_halo(z, h);
    const bool anyMixed = wave_ballot(mixed) != 0ull;
    if (lane == 0) z.sh->mixedW[wave - 1] = anyMixed ? 1 : 0;
}
template <class Z>
ZHD int zone_setup_all_owned(const Z& z) {
    return (z.sh->mixedW[0] | z.sh->mixedW[1] | z.sh->mixedW[2]) ? 0 : 1;
}

/* sum over the slots glist[s0 .. s0+n) in order, loads issued 8 at a time */
template <class Z>
ZHD void run_sum(const Z& z, int s0, int n, double& sx, double& sy) {
    for (int j0 = 0; j0 < n; j0 += 8) {
        int sl[8];
        double ax[8], ay[8];
        for (int u = 0; u < 8; ++u) sl[u] = (j0 + u < n) ? z.glist[s0 + j0 + u] : -1;
        for (int u = 0; u < 8; ++u) {
            ax[u] = sl[u] >= 0 ? z.vx[sl[u]] : 0.0;
            ay[u] = sl[u] >= 0 ? z.vy[sl[u]] : 0.0;
        }
        for (int u = 0; u < 8; ++u)
            if (sl[u] >= 0) {
                sx += ax[u];
                sy += ay[u];
            }
    }
}

/* The diagnostics' record of cell lc (zone_pass.cpp:885-897): the momentum
 * defect (input run minus final list) and the final list's momentum, both
 * before the correction. */
template <class Z>
ZHD void diag_cell(Z& z, int lc, double dPx, double dPy, double px, double py) {
    double* d = z.dgCell + 4 * (size_t)lc;
    d[0] = dPx;
    d[1] = dPy;
    d[2] = px;
    d[3] = py;
}

/* the same for an occupied cell nothing changed: its list is its input run */
template <class Z>
ZHD void diag_cell_unchanged(Z& z, int lc) {
    const int gc = lc_gc(z, lc);
    double px = 0.0, py = 0.0;
    run_sum(z, z.gstart[gc], z.gcnt[gc], px, py);
    diag_cell(z, lc, 0.0, 0.0, px, py);
}

/* Momentum correction (zone_pass.cpp:870-909), first half, one lane per
 * cell: the velocity change of a changed cell from the momentum of its
 * input run and of its final list (both in list order).  Reads only; a cell
 * nothing changed has the same list before and after, so its change is
 * exactly zero. */
template <class Z>
ZHD void zone_correct_sums(Z& z, int lc) {
    const int m = cnt_of(z, lc);
    if (!fget(z, lc, ZF_CHG) || m == 0) {
        if (z.dgCell && m > 0 && !fget(z, lc, ZF_OUT)) diag_cell_unchanged(z, lc);
        return;
    }
    const int gc = lc_gc(z, lc);
    double pbx = 0.0, pby = 0.0;
    run_sum(z, z.gstart[gc], z.gcnt[gc], pbx, pby);
    double px = 0.0, py = 0.0;
    int k = head_of(z, lc);
    while (k >= 0) {
        int sl[8];
        double ax[8], ay[8];
        for (int u = 0; u < 8; ++u) {
            sl[u] = -1;
            if (k >= 0) {
                sl[u] = slot_of(z, k);
                k = next_k(z, lc, k);
            }
        }
        for (int u = 0; u < 8; ++u) {
            ax[u] = sl[u] >= 0 ? z.vx[sl[u]] : 0.0;
            ay[u] = sl[u] >= 0 ? z.vy[sl[u]] : 0.0;
        }
        for (int u = 0; u < 8; ++u)
            if (sl[u] >= 0) {
                px += ax[u];
                py += ay[u];
            }
    }
    const double dPx = pbx - px, dPy = pby - py;
    if (z.dgCell) diag_cell(z, lc, dPx, dPy, px, py);
    const double dux = dPx / (double)m, duy = dPy / (double)m;
    z.duG[2 * lc] = dux;
    z.duG[2 * lc + 1] = duy;
    bool corr = z.zp.corrEnabled != 0;
    if (corr && fget(z, lc, ZF_SKIP)) corr = false;
    if (corr && fget(z, lc, ZF_RECV) && !fget(z, lc, ZF_BB)) corr = false;
    if (corr && !(dux != 0.0 || duy != 0.0)) corr = false;
    if (corr) fset(z, lc, ZF_CORR);
}

/* Second half, one lane per cell: write the velocities of the owned
 * particles of a changed cell; for an unchanged cell that several zones
 * cover, pass its owned particles' input velocities through to the side
 * buffer.  Returns the cell's |du|^2 (0 for an unchanged cell). */
template <class Z>
ZHD double zone_correct_write(Z& z, int lc) {
    if (fget(z, lc, ZF_CHG)) {
        if (cnt_of(z, lc) == 0) return 0.0;
        const bool corr = fget(z, lc, ZF_CORR);
        const double dux = z.duG[2 * lc], duy = z.duG[2 * lc + 1];
        int k = head_of(z, lc);
        while (k >= 0) {
            int sl[8], tg[8];
            uint32_t pd[8];
            unsigned short ow[8];
            double ax[8], ay[8];
            for (int u = 0; u < 8; ++u) {
                sl[u] = -1;
                ow[u] = 0;
                tg[u] = ZM_NONE;
                if (k >= 0) {
                    const int col = col_of(z, k);
                    sl[u] = slot_pid_in(z, k, col, pd[u]);
                    ow[u] = own_bits(z, k, col);
                    tg[u] = z.rec[k];
                    k = next_k(z, lc, k);
                }
            }
            for (int u = 0; u < 8; ++u) {
                const bool w = sl[u] >= 0 && (ow[u] & ZF_OWN);
                ax[u] = w ? z.vx[sl[u]] : 0.0;
                ay[u] = w ? z.vy[sl[u]] : 0.0;
            }
            for (int u = 0; u < 8; ++u) {
                if (sl[u] < 0 || !(ow[u] & ZF_OWN)) continue;
                if (tg[u] != ZM_NONE) place_moved(z, lc, sl[u], pd[u], tg[u]);
                double vx = ax[u], vy = ay[u];
                if (corr) {
                    vx += dux;
                    vy += duy;
                }
                if (ow[u] & ZF_COVER) {
                    z.vsx[sl[u]] = vx;
                    z.vsy[sl[u]] = vy;
                } else {
                    z.vx[sl[u]] = vx;
                    z.vy[sl[u]] = vy;
                }
            }
        }
        return dux * dux + duy * duy;
    }
    if (!fget(z, lc, ZF_COVER) || !fget(z, lc, ZF_OWN)) return 0.0;
    const int gc = lc_gc(z, lc);
    const int s0 = z.gstart[gc], m0 = z.gcnt[gc];
    for (int j0 = 0; j0 < m0; j0 += 8) {
        int sl[8];
        double ax[8], ay[8];
        for (int u = 0; u < 8; ++u) sl[u] = (j0 + u < m0) ? z.glist[s0 + j0 + u] : -1;
        for (int u = 0; u < 8; ++u) {
            ax[u] = sl[u] >= 0 ? z.vx[sl[u]] : 0.0;
            ay[u] = sl[u] >= 0 ? z.vy[sl[u]] : 0.0;
        }
        for (int u = 0; u < 8; ++u)
            if (sl[u] >= 0) {
                z.vsx[sl[u]] = ax[u];
                z.vsy[sl[u]] = ay[u];
            }
    }
    return 0.0;
}

/* Correction of all cells and stats[4], stats[5] of the zone
 * (zone_pass.cpp:870-922), all lanes.  The rms sums |du|^2 over the cells in
 * ascending order: each chunk of nl cells is added lane by lane. */
template <class Z>
ZHD void zone_correct(Z& z, int lane, int nl) {
    for (int lc = lane; lc < z.C; lc += nl) zone_correct_sums(z, lc);
    ZBAR();
    double acc = 0.0, mx = 0.0;
    int nDU = 0;
    for (int b0 = 0; b0 < z.C; b0 += nl) {
        const int lc = b0 + lane;
        const double d = lc < z.C ? zone_correct_write(z, lc) : 0.0;
        mx = dmax(mx, MPCD_SQRT(d));
        unsigned long long pos = wave_ballot(d > 0.0);
        nDU += wave_popc(pos);
        while (pos) {
            const int l = __builtin_ctzll(pos);
            pos &= pos - 1ull;
            acc += wave_readlane_d(d, l);
        }
    }
    mx = wave_max(mx);
    if (lane == 0 && nDU > 0) {
        z.sh->stats4 = MPCD_SQRT(acc / (double)nDU);
        z.sh->stats5 = mx;
    }
}

/* ---- working set ---- */

/* bytes of the per-zone cell arrays and scratch (links excluded) */
template <class IT>
ZHD size_t zone_fixed_bytes(int C, int W, int useLocal) {
    size_t b = 0;
    if (useLocal) b += (size_t)C * 2 * sizeof(double);         /* lowA highA */
    b += 24 * sizeof(Cand) + 2 * 24 * sizeof(double) + 8 * sizeof(double) + 4 * sizeof(double) +
         (size_t)ZONE_MAX_MOVE * sizeof(int); /* 8-byte items first */
    b += (size_t)W * 2 * sizeof(int);                           /* colL colG */
    b += (5 * 24 + 3 * 8) * sizeof(int);                        /* ord nbGc rnk recvT (+pad) dord dord2 (+pad) */
    b += (size_t)C * 5 * sizeof(IT);                            /* cnt head tail dl k0 */
    b += (size_t)(W + 4) * (size_t)(C / (W > 0 ? W : 1) + 4) * sizeof(IT); /* ext */
    b += (size_t)(4 * (W + 4) + 4 * (C / (W > 0 ? W : 1))) * sizeof(IT); /* halo */
    b += (size_t)C * sizeof(unsigned short); /* fl */
    b += (size_t)C * sizeof(IT);             /* lix */
    b += (size_t)ZONE_TMP_L * sizeof(IT);                       /* tmpL */
    b += 4 + 16;                                                /* alignment after fl and after colG */
    return (b + 15) & ~(size_t)15;
}

/* Carves the working set out of `base` (16-byte aligned); links of k < nL
 * follow the fixed part, nL = what the remaining `bytes` hold. */
template <class IT>
ZHD void zone_carve(ZoneCtxT<IT>& z, unsigned char* base, size_t bytes) {
    const int C = z.C, W = z.zw;
    unsigned char* p = base;
    /* first what the correction keeps (cell arrays, column map), at the
     * offsets of the correction's own LDS layout, so the correction that
     * follows the passes in one kernel finds them in place */
    z.cnt = (IT*)p; p += (size_t)C * sizeof(IT);
    z.head = (IT*)p; p += (size_t)C * sizeof(IT);
    z.tail = (IT*)p; p += (size_t)C * sizeof(IT);
    z.k0 = (IT*)p; p += (size_t)C * sizeof(IT);
    z.fl = (unsigned short*)p; p += (size_t)C * sizeof(unsigned short);
    p = base + (((size_t)(p - base) + 3) & ~(size_t)3);
    z.colL = (int*)p; p += (size_t)W * sizeof(int);
    z.colG = (int*)p; p += (size_t)W * sizeof(int);
    p = base + (((size_t)(p - base) + 15) & ~(size_t)15);
    /* then the passes' scratch */
    if (z.zp.useLocal) {
        z.lowA = (double*)p; p += (size_t)C * sizeof(double);
        z.highA = (double*)p; p += (size_t)C * sizeof(double);
    } else {
        z.lowA = nullptr;
        z.highA = nullptr;
    }
    z.cand = (Cand*)p; p += 24 * sizeof(Cand);
    z.ordThr = (double*)p; p += 24 * sizeof(double);
    z.nbTie = (double*)p; p += 24 * sizeof(double);
    z.dsc = (double*)p; p += 8 * sizeof(double);
    z.nsm4 = (double*)p; p += 4 * sizeof(double);
    z.mvk = (int*)p; p += ZONE_MAX_MOVE * sizeof(int);
    z.ord = (int*)p; p += 24 * sizeof(int);
    z.nbGc = (int*)p; p += 24 * sizeof(int);
    z.rnk = (int*)p; p += 24 * sizeof(int);
    z.recvT = (int*)p; p += 48 * sizeof(int);
    z.dord = (int*)p; p += 8 * sizeof(int);
    z.dord2 = (int*)p; p += 16 * sizeof(int);
    z.dl = (IT*)p; p += (size_t)C * sizeof(IT);
    z.ext = (IT*)p; p += (size_t)(z.zw + 4) * (size_t)(z.zh + 4) * sizeof(IT);
    z.halo = (IT*)p; p += (size_t)halo_size(z) * sizeof(IT);
    z.tmpL = (IT*)p; p += (size_t)ZONE_TMP_L * sizeof(IT);
    z.tmpCap = ZONE_TMP_L;
    z.lix = (IT*)p; p += (size_t)C * sizeof(IT);
    const size_t fixed = zone_fixed_bytes<IT>(C, W, z.zp.useLocal);
    p = base + fixed;
    z.nxL = (IT*)p;
    const size_t rest = bytes > fixed ? bytes - fixed : 0;
    const size_t nl = rest / sizeof(IT);
    z.nL = nl > (size_t)0x7fffffff ? 0x7fffffff : (int)nl;
}

/* Per-lane context of zone zid (1-based) with rectangle r. */
template <class IT>
ZHD void zone_init(ZoneCtxT<IT>& z, const ZoneIn& in, ZoneRect r, int zid, uint64_t seed, ZoneShared* sh) {
    z.g = in.g;
    z.zp = in.zp;
    z.xs = r.xs;
    z.xe = r.xe;
    z.ys = r.ys;
    z.ye = r.ye;
    z.zw = r.xe - r.xs;
    z.zh = r.ye - r.ys;
    z.C = z.zw * z.zh;
    z.zid = zid;
    z.seed = seed;
    z.gcnt = in.gcnt;
    z.gstart = in.gstart;
    z.glist = in.glist;
    z.gkey = in.gkey;
    z.inPx = in.inPx;
    z.inPy = in.inPy;
    z.dcid = in.dcid;
    z.dout = in.dout;
    z.dinc = in.dinc;
    z.drnk = in.drnk;
    z.cellOff = in.cellOff;
    z.dgCell = in.dgCell;
    z.dgCnt = in.dgCnt;
    z.x = in.x;
    z.y = in.y;
    z.vx = in.vx;
    z.vy = in.vy;
    z.vsx = in.vsx;
    z.vsy = in.vsy;
    z.sh = sh;
}

/* The zone after zone_setup and a barrier: both phases. */
template <class Z>
ZHD void zone_passes(Z& z, int lane, int nl) {
    z.P = z.sh->P;
    z.allOwned = z.sh->allOwned;
    for (int k = lane; k < z.P; k += nl) z.rec[k] = (unsigned char)ZM_NONE;
    zone_dense_passes(z, lane, nl);
    zone_sparse_pass(z, lane, nl);
    ZBAR();
}

/* Both phases, the correction and the stats, in one go. */
template <class Z>
ZHD void zone_run(Z& z, int lane, int nl) {
    zone_passes(z, lane, nl);
    zone_correct(z, lane, nl);
}

/* ---- the zone's state between the passes and the correction ----
 *
 * On the device the correction runs in a kernel of its own, with many
 * threads per zone: the pass kernel leaves the cell arrays, the column map
 * and the flags in a per-zone state region of global memory, and the links
 * all in the zone's global link region.  Layout (16-byte aligned):
 * |dsq C doubles|hdr 4 ints|colL W ints|colG W ints|cnt|head|tail|k0 C IT each|fl C u16| */
template <class IT>
struct ZoneState {
    double* dsq; /* |du|^2 of each cell, for the ordered stats */
    int* hdr;    /* P, allOwned */
    int *colL, *colG;
    IT *cnt, *head, *tail, *k0;
    unsigned short* fl;
};

template <class IT>
ZHD size_t zone_state_bytes(int C, int W) {
    const size_t b = (size_t)C * sizeof(double) + 4 * sizeof(int) + (size_t)W * 2 * sizeof(int) +
                     (size_t)C * 4 * sizeof(IT) + (size_t)C * sizeof(unsigned short);
    return (b + 15) & ~(size_t)15;
}

template <class IT>
ZHD ZoneState<IT> zone_state_at(unsigned char* base, int C, int W) {
    ZoneState<IT> st;
    unsigned char* p = base;
    st.dsq = (double*)p; p += (size_t)C * sizeof(double);
    st.hdr = (int*)p; p += 4 * sizeof(int);
    st.colL = (int*)p; p += (size_t)W * sizeof(int);
    st.colG = (int*)p; p += (size_t)W * sizeof(int);
    st.cnt = (IT*)p; p += (size_t)C * sizeof(IT);
    st.head = (IT*)p; p += (size_t)C * sizeof(IT);
    st.tail = (IT*)p; p += (size_t)C * sizeof(IT);
    st.k0 = (IT*)p; p += (size_t)C * sizeof(IT);
    st.fl = (unsigned short*)p;
    return st;
}

/* all lanes: the state after zone_passes, and the links held in LDS moved
 * to the zone's global link region */
template <class IT>
ZHD void zone_dump(ZoneCtxT<IT>& z, const ZoneState<IT>& st, int lane, int nl) {
    for (int lc = lane; lc < z.C; lc += nl) {
        st.cnt[lc] = z.cnt[lc];
        st.head[lc] = z.head[lc];
        st.tail[lc] = z.tail[lc];
        st.k0[lc] = z.k0[lc];
        st.fl[lc] = z.fl[lc];
    }
    for (int c = lane; c < z.zw; c += nl) {
        st.colL[c] = z.colL[c];
        st.colG[c] = z.colG[c];
    }
    if (lane == 0) {
        st.hdr[0] = z.P;
        st.hdr[1] = z.allOwned;
    }
    const int nk = z.P < z.nL ? z.P : z.nL;
    const IT* const L = z.nxL;
    IT* const G = z.nxG;
    for (int k = lane; k < nk; k += nl) ZG_STORE(G + k, L[k]);
}

/* the correction's view of a dumped zone: cell arrays and column map in the
 * state region, every link in the global link region */
template <class IT>
ZHD void zone_attach_state(ZoneCtxT<IT>& z, const ZoneState<IT>& st) {
    z.cnt = st.cnt;
    z.head = st.head;
    z.tail = st.tail;
    z.k0 = st.k0;
    z.fl = st.fl;
    z.colL = st.colL;
    z.colG = st.colG;
    z.P = st.hdr[0];
    z.allOwned = st.hdr[1];
    z.nxL = nullptr;
    z.nL = 0;
}

/* length of cell lc's input run (its local indices k0[lc] ..) */
template <class Z>
ZHD int in_cnt(const Z& z, int lc) {
    const int col = lc / z.zh;
    const int end = (lc - col * z.zh + 1 < z.zh) ? (int)z.k0[lc + 1] : (col + 1 < z.zw ? (int)z.colL[col + 1] : z.P);
    return end - (int)z.k0[lc];
}

/* ---- the correction kernel's per-cell form (a group of lanes per cell) ----
 * Phase 1 sums every changed cell's input run; after a barrier, phase 2
 * walks each cell's final list once: a list of up to FIN_G x FIN_L entries
 * stays in registers from the sums to the writes (a particle is in exactly
 * one final list, and every input-run read happened in phase 1), a longer
 * one is walked twice.  The moved particles' links are set to their final
 * cell for phase 3, which places them.  Same sums in the same order as
 * zone_correct_sums / zone_correct_write. */
constexpr int FIN_L = 8;
constexpr int FIN_G = 4;

/* column of particle k of cell lc's final list: its own input run lies in
 * lc's column */
template <class Z>
ZHD int fin_col(const Z& z, int lc, int kb, int ke, int k) {
    return (k >= kb && k < ke) ? lc / z.zh : col_of(z, k);
}

template <class Z>
ZHD void fin_write(Z& z, int sl, unsigned short ow, bool corr, double vx, double vy, double dux, double duy) {
    if (!(ow & ZF_OWN)) return;
    if (corr) {
        vx += dux;
        vy += duy;
    }
    if (ow & ZF_COVER) {
        z.vsx[sl] = vx;
        z.vsy[sl] = vy;
    } else {
        z.vx[sl] = vx;
        z.vy[sl] = vy;
    }
}

#if defined(__HIP_DEVICE_COMPILE__)
/* Phase 2 for changed, occupied cell lc on a group of FIN_G consecutive
 * lanes (r: the lane's rank; every lane of the group passes the same cell).
 * Every entry's link is set to lc, its final cell, once the walk has read
 * it: phase 3 reads the links of the moved particles, and nothing reads the
 * others.
 * The group takes the final list FIN_G * FIN_L entries at a time, lane r
 * entries [FIN_L r, FIN_L (r + 1)) of each chunk, so a list of up to
 * FIN_G * FIN_L entries has each slot and velocity loaded once instead of
 * once per pass.  The momentum sums keep the list order: lane r adds its
 * entries to the running sum lane r - 1 hands on.  A linked list is walked by
 * the whole group in step.  pbx, pby: the cell's input-run momentum.  Returns
 * |du|^2 (the same on every lane of the group). */
/* Phase 1 for changed, occupied cell lc on a group of FIN_G lanes: the
 * momentum of its input run (the particles binned into it, before the
 * passes), FIN_G * FIN_L entries at a time, summed in run order as
 * fin_cell_group sums the final list.  Every lane of the group gets the sums. */
template <class Z>
__device__ __forceinline__ void fin_input_sum_group(const Z& z, int lc, int r, int lane, double& pbx, double& pby) {
    constexpr int CH = FIN_G * FIN_L;
    const int gc = lc_gc(z, lc);
    const int s0 = z.gstart[gc], n = z.gcnt[gc];
    const int g0 = lane - r;
    double px = 0.0, py = 0.0;
    for (int c0 = 0; c0 < n; c0 += CH) {
        int sl[FIN_L];
        double ax[FIN_L], ay[FIN_L];
        for (int u = 0; u < FIN_L; ++u) {
            const int e = c0 + FIN_L * r + u;
            sl[u] = e < n ? z.glist[s0 + e] : -1;
        }
        for (int u = 0; u < FIN_L; ++u) {
            ax[u] = sl[u] >= 0 ? z.vx[sl[u]] : 0.0;
            ay[u] = sl[u] >= 0 ? z.vy[sl[u]] : 0.0;
        }
        for (int a = 0; a < FIN_G; ++a) {
            if (r == a)
                for (int u = 0; u < FIN_L; ++u)
                    if (sl[u] >= 0) {
                        px += ax[u];
                        py += ay[u];
                    }
            px = __shfl(px, g0 + a, 64);
            py = __shfl(py, g0 + a, 64);
        }
    }
    pbx = px;
    pby = py;
}

template <class Z>
__device__ __forceinline__ void fin_group_entries(const Z& z, bool mat, int h, int t, int m, int c0, int r, int& kw,
                                                  int* kk) {
    for (int u = 0; u < FIN_L; ++u) kk[u] = -1;
    if (!mat) {
        for (int u = 0; u < FIN_L; ++u) {
            const int e = c0 + FIN_L * r + u;
            kk[u] = (e < m && h + e <= t) ? h + e : -1;
        }
        return;
    }
    int k = kw;
    for (int a = 0; a < FIN_G; ++a)
        for (int u = 0; u < FIN_L; ++u) {
            if (a == r) kk[u] = k;
            if (k >= 0) k = nx_get(z, k);
        }
    kw = k;
}

template <class Z>
__device__ __forceinline__ double fin_cell_group(Z& z, int lc, int r, int lane, double pbx, double pby) {
    constexpr int CH = FIN_G * FIN_L;
    const int m = cnt_of(z, lc);
    const int kb = (int)z.k0[lc], ke = kb + in_cnt(z, lc);
    const bool mat = fget(z, lc, ZF_MAT);
    const int t = tail_of(z, lc), h = head_of(z, lc);
    const unsigned short ownLc = (unsigned short)(z.fl[lc] & (ZF_OWN | ZF_COVER));
    const int g0 = lane - r;
    int kk[FIN_L], sl[FIN_L];
    unsigned short ow[FIN_L];
    double ax[FIN_L], ay[FIN_L];
    double px = 0.0, py = 0.0;
    int kw = h;
    for (int c0 = 0; c0 < m; c0 += CH) {
        fin_group_entries(z, mat, h, t, m, c0, r, kw, kk);
        for (int u = 0; u < FIN_L; ++u) {
            sl[u] = -1;
            ow[u] = 0;
            if (kk[u] >= 0) {
                const int col = fin_col(z, lc, kb, ke, kk[u]);
                sl[u] = slot_in(z, kk[u], col);
                ow[u] = (kk[u] >= kb && kk[u] < ke) ? ownLc : own_bits(z, kk[u], col);
            }
        }
        for (int u = 0; u < FIN_L; ++u) {
            ax[u] = sl[u] >= 0 ? z.vx[sl[u]] : 0.0;
            ay[u] = sl[u] >= 0 ? z.vy[sl[u]] : 0.0;
        }
        for (int a = 0; a < FIN_G; ++a) {
            if (r == a)
                for (int u = 0; u < FIN_L; ++u)
                    if (sl[u] >= 0) {
                        px += ax[u];
                        py += ay[u];
                    }
            px = __shfl(px, g0 + a, 64);
            py = __shfl(py, g0 + a, 64);
        }
    }
    if (z.dgCell && r == 0) diag_cell(z, lc, pbx - px, pby - py, px, py);
    const double dux = (pbx - px) / (double)m, duy = (pby - py) / (double)m;
    bool corr = z.zp.corrEnabled != 0;
    if (corr && fget(z, lc, ZF_SKIP)) corr = false;
    if (corr && fget(z, lc, ZF_RECV) && !fget(z, lc, ZF_BB)) corr = false;
    if (corr && !(dux != 0.0 || duy != 0.0)) corr = false;
    if (m > CH) {
        /* a longer list: walk it again for the writes */
        kw = h;
        for (int c0 = 0; c0 < m; c0 += CH) {
            fin_group_entries(z, mat, h, t, m, c0, r, kw, kk);
            for (int u = 0; u < FIN_L; ++u) {
                sl[u] = -1;
                ow[u] = 0;
                if (kk[u] >= 0) {
                    const int col = fin_col(z, lc, kb, ke, kk[u]);
                    sl[u] = slot_in(z, kk[u], col);
                    ow[u] = (kk[u] >= kb && kk[u] < ke) ? ownLc : own_bits(z, kk[u], col);
                }
            }
            for (int u = 0; u < FIN_L; ++u) {
                ax[u] = sl[u] >= 0 ? z.vx[sl[u]] : 0.0;
                ay[u] = sl[u] >= 0 ? z.vy[sl[u]] : 0.0;
            }
            for (int u = 0; u < FIN_L; ++u)
                if (sl[u] >= 0) {
                    fin_write(z, sl[u], ow[u], corr, ax[u], ay[u], dux, duy);
                    nx_set(z, kk[u], lc);
                }
        }
    } else {
        for (int u = 0; u < FIN_L; ++u)
            if (sl[u] >= 0) {
                fin_write(z, sl[u], ow[u], corr, ax[u], ay[u], dux, duy);
                nx_set(z, kk[u], lc);
            }
    }
    return dux * dux + duy * duy;
}
#else
/* (the host passes of the compilers see a declaration only) */
template <class Z>
ZHD void fin_input_sum_group(const Z&, int, int, int, double&, double&) {}
template <class Z>
ZHD double fin_cell_group(Z&, int, int, int, double, double) { return 0.0; }
#endif

/* phase 3, all threads of the zone: the new positions of the moved
 * particles (links hold their final cells).  Each wavefront scans its share
 * of the local indices and gathers the moved ones into a queue of one entry
 * per lane, so each batch of draws runs once for up to 64 particles. */
template <class Z>
ZHD void fin_place_batch(Z& z, const int* queue, int n, int lane) {
    if (lane < n) {
        const int kq = queue[lane];
        const int col = col_of(z, kq);
        uint32_t pid;
        const int slot = slot_pid_in(z, kq, col, pid);
        if (own_bits(z, kq, col) & ZF_OWN) place_moved(z, nx_get(z, kq), slot, pid, z.rec[kq]);
    }
}
template <class Z>
ZHD void fin_place(Z& z, int tid, int nthr, int* queue) {
    const int lane = tid & 63;
    int nq = 0;
    for (int b = tid - lane; b < z.P; b += nthr) {
        const int k = b + lane;
        const bool mv = k < z.P && z.rec[k] != ZM_NONE;
        const unsigned long long msk = wave_ballot(mv);
        const int add = wave_popc(msk);
        if (add == 0) continue;
        if (nq + add > 64) {
            ZWSYNC();
            fin_place_batch(z, queue, nq, lane);
            ZWSYNC();
            nq = 0;
        }
        if (mv) queue[nq + wave_popc(msk & lanes_below(lane))] = k;
        nq += add;
    }
    ZWSYNC();
    fin_place_batch(z, queue, nq, lane);
}

/* The stats of zone_correct from the per-cell |du|^2 in cell order (one
 * wavefront: each chunk of nl cells added lane by lane; the loads of four
 * chunks are issued before their adds). */
template <class Z>
ZHD void zone_correct_stats(const Z& z, const double* dsq, int lane, int nl, double& rms, double& mxo) {
    constexpr int NB = 4;
    double acc = 0.0, mx = 0.0;
    int nDU = 0;
    for (int b0 = 0; b0 < z.C; b0 += NB * nl) {
        double dv[NB];
        for (int j = 0; j < NB; ++j) {
            const int lc = b0 + j * nl + lane;
            dv[j] = lc < z.C ? dsq[lc] : 0.0;
        }
        for (int j = 0; j < NB; ++j) {
            const double d = dv[j];
            mx = dmax(mx, MPCD_SQRT(d));
            unsigned long long pos = wave_ballot(d > 0.0);
            nDU += wave_popc(pos);
            while (pos) {
                const int l = __builtin_ctzll(pos);
                pos &= pos - 1ull;
                acc += wave_readlane_d(d, l);
            }
        }
    }
    mxo = wave_max(mx);
    rms = nDU > 0 ? MPCD_SQRT(acc / (double)nDU) : 0.0;
}

}  // namespace mpcdk
