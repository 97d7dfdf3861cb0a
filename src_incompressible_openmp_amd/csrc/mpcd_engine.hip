/*
 * mpcd_engine.hip -- gfx950 kernels and device orchestration of the SRD/MPCD
 * step, and the C ABI declared in include/mpcd_gpu.h.
 *
 * Device state is FP64 structure-of-arrays (x, y, vx, vy) indexed by slot,
 * with pid[slot] the reference particle index (slot == pid as uploaded).
 * Every per-cell stage reads the particles of a cell through a counting-sort
 * binning (cnt, start, list) whose per-cell runs are in ascending pid, so the
 * per-cell sums run in the reference's order.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <array>
#include <vector>

#include "mpcd_core.h"
#include "mpcd_hip_util.h"
#include "mpcd_host.h"

using namespace mpcdk;

namespace {

thread_local std::string g_last_error;

constexpr int TPB = 256;
constexpr int ZONE_TPB = 64;
constexpr int ZONE_LDS_CAP = 64 * 1024;
/* LDS of one zone workgroup on the uint16_t path: four fit a CU's 160 KB with
 * the static shared scalars and allocation rounding */
constexpr int ZONE_LDS_FAST = 38 * 1024;
constexpr int SCAN_TILE = 1024; /* 256 threads x 4 */

inline unsigned nblk(long long n, int tpb = TPB) { return (unsigned)((n + tpb - 1) / tpb); }

/* ===================================================================== */
/*                               kernels                                 */
/* ===================================================================== */

__global__ void k_deinterleave(const double* __restrict__ a, int n, double* __restrict__ x, double* __restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        x[i] = a[2 * (size_t)i];
        y[i] = a[2 * (size_t)i + 1];
    }
}

/* gather by pid: out[2*pid] = x[slot] */
__global__ void k_interleave(const double* __restrict__ x, const double* __restrict__ y, const int* __restrict__ pid,
                             int n, double* __restrict__ a) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) {
        const size_t p = (size_t)pid[s];
        a[2 * p] = x[s];
        a[2 * p + 1] = y[s];
    }
}

/* scatter by pid: ox[pid[s]] = x[s] (the slot order back to particle-id order) */
__global__ void k_scatter_pid(const double* __restrict__ x, const double* __restrict__ y, const int* __restrict__ pid,
                              int n, double* __restrict__ ox, double* __restrict__ oy) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) {
        const int q = pid[s];
        ox[q] = x[s];
        oy[q] = y[s];
    }
}

__global__ void k_iota(int* a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

__global__ void k_fill_double(double* a, int n, double v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_fill_int(int* a, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

/* Block reduction of a WallAcc into one partial per block (fixed tree, so the
 * result does not depend on scheduling). */
__device__ void wall_block_reduce(WallAcc a, WallAcc* out) {
    __shared__ double sd[13][TPB];
    const int t = threadIdx.x;
    sd[0][t] = (double)a.nBot;
    sd[1][t] = (double)a.nTop;
    sd[2][t] = (double)a.nLeft;
    sd[3][t] = (double)a.nRight;
    sd[4][t] = a.dE;
    sd[5][t] = a.dPxBot;
    sd[6][t] = a.dPxTop;
    sd[7][t] = a.dPyBot;
    sd[8][t] = a.dPyTop;
    sd[9][t] = a.dPxLeft;
    sd[10][t] = a.dPxRight;
    sd[11][t] = a.dPyLeft;
    sd[12][t] = a.dPyRight;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if (t < s)
            for (int k = 0; k < 13; ++k) sd[k][t] += sd[k][t + s];
        __syncthreads();
    }
    if (t == 0) {
        WallAcc r;
        r.nBot = (int)sd[0][0];
        r.nTop = (int)sd[1][0];
        r.nLeft = (int)sd[2][0];
        r.nRight = (int)sd[3][0];
        r.dE = sd[4][0];
        r.dPxBot = sd[5][0];
        r.dPxTop = sd[6][0];
        r.dPyBot = sd[7][0];
        r.dPyTop = sd[8][0];
        r.dPxLeft = sd[9][0];
        r.dPxRight = sd[10][0];
        r.dPyLeft = sd[11][0];
        r.dPyRight = sd[12][0];
        out[blockIdx.x] = r;
    }
}

/* K1: force + stream + wall BC (+ shifted collision cell). */
__global__ void __launch_bounds__(TPB) k_stream_bc(Geom g, BCParams b, uint64_t seed, int do_stream, double* __restrict__ x,
                                                   double* __restrict__ y, double* __restrict__ vx,
                                                   double* __restrict__ vy, const int* __restrict__ pid,
                                                   int* __restrict__ cid, int* __restrict__ cidU, double sx,
                                                   double sy, double xmax, double ymax, WallAcc* __restrict__ wall,
                                                   int* __restrict__ err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    WallAcc acc;
    wall_acc_zero(acc);
    if (i < g.n) {
        double px = x[i], py = y[i], pvx = vx[i], pvy = vy[i];
        const int rc = bc_particle(g, b, seed, (uint32_t)pid[i], do_stream, px, py, pvx, pvy, acc);
        if (rc != 0) atomicOr(err, 1);
        x[i] = px;
        y[i] = py;
        vx[i] = pvx;
        vy[i] = pvy;
        if (cid) cid[i] = shifted_cell_of(g, px, py, sx, sy, xmax, ymax);
        if (cidU) cidU[i] = cell_of(g, px, py);
    }
    if (wall) wall_block_reduce(acc, wall);
}

__global__ void k_shift_cells(Geom g, const double* __restrict__ x, const double* __restrict__ y, double sx,
                              double sy, double xmax, double ymax, int* __restrict__ cid) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < g.n) cid[i] = shifted_cell_of(g, x[i], y[i], sx, sy, xmax, ymax);
}

__global__ void k_cells(Geom g, const double* __restrict__ x, const double* __restrict__ y, int* __restrict__ cid) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < g.n) cid[i] = cell_of(g, x[i], y[i]);
}

/* ---- counting sort ---- */

/* Groups of lanes of a wavefront holding the same key: each active lane
 * gets its group's lowest lane (the leader), its rank in the group and the
 * group size.  ALU only (one readlane and one ballot per group); the state
 * is kept sorted by cell, so a wavefront holds a few keys. */
__device__ inline void wave_groups(int c, bool active, int lane, int& leader, int& rank, int& size) {
    unsigned long long act = __ballot(active);
    leader = -1;
    rank = 0;
    size = 0;
    while (act) {
        const int l = __ffsll((long long)act) - 1;
        const int lc = __builtin_amdgcn_readlane(c, l);
        const bool mine = active && c == lc;
        const unsigned long long grp = __ballot(mine);
        if (mine) {
            leader = l;
            rank = __popcll(grp & ((1ull << lane) - 1ull));
            size = __popcll(grp);
        }
        act &= ~grp;
    }
}

/* The binning kernels give each wavefront BIN_ITEMS consecutive chunks of 64
 * keys (lanes on consecutive keys, so a chunk holds a few cells) and issue
 * every chunk's loads before the first use. */
constexpr int BIN_ITEMS = 4;
__host__ __device__ inline int bin_blocks(int n) {
    const int per = TPB * BIN_ITEMS;
    return (n + per - 1) / per;
}

/* Cell histogram: one atomic per key group, all groups in one instruction. */
__global__ void __launch_bounds__(TPB) k_hist(const int* __restrict__ cid, int n, int* __restrict__ cnt) {
    const int lane = __lane_id();
    const int w0 = (blockIdx.x * TPB + (int)threadIdx.x - lane) * BIN_ITEMS; /* the wavefront's first key */
    int c[BIN_ITEMS];
    for (int r = 0; r < BIN_ITEMS; ++r) {
        const int i = w0 + r * 64 + lane;
        c[r] = (i < n) ? cid[i] : -1;
    }
    for (int r = 0; r < BIN_ITEMS; ++r) {
        int leader, rank, size;
        wave_groups(c[r], c[r] >= 0, lane, leader, rank, size);
        if (c[r] >= 0 && lane == leader) atomicAdd(&cnt[c[r]], size);
    }
}

/* exclusive scan, level 1: each block scans SCAN_TILE entries */
__global__ void __launch_bounds__(TPB) k_scan_tiles(const int* __restrict__ in, int m, int* __restrict__ out,
                                                    int* __restrict__ tileSum, const int* __restrict__ in2 = nullptr,
                                                    int* __restrict__ zero = nullptr,
                                                    const int* __restrict__ sub = nullptr) {
    /* in2 != nullptr: scans in + in2 (- sub); zero != nullptr: cleared over the range */
    __shared__ int s[TPB];
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * 4;
    int v[4];
    int tot = 0;
    for (int k = 0; k < 4; ++k) {
        v[k] = (base + k < m) ? in[base + k] + (in2 ? in2[base + k] : 0) - (sub ? sub[base + k] : 0) : 0;
        tot += v[k];
    }
    s[threadIdx.x] = tot;
    __syncthreads();
    for (int off = 1; off < TPB; off <<= 1) {
        const int t = (threadIdx.x >= (unsigned)off) ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    int run = s[threadIdx.x] - tot;
    for (int k = 0; k < 4; ++k) {
        if (base + k < m) {
            out[base + k] = run;
            if (zero) zero[base + k] = 0;
        }
        run += v[k];
    }
    if (threadIdx.x == TPB - 1) tileSum[blockIdx.x] = s[TPB - 1];
}

/* level 2: one block scans the tile sums (sequential chunks per thread) */
__global__ void __launch_bounds__(TPB) k_scan_top(int* __restrict__ tileSum, int nt) {
    __shared__ int s[TPB];
    const int per = (nt + TPB - 1) / TPB;
    const int b0 = threadIdx.x * per;
    int tot = 0;
    for (int k = 0; k < per; ++k)
        if (b0 + k < nt) tot += tileSum[b0 + k];
    s[threadIdx.x] = tot;
    __syncthreads();
    for (int off = 1; off < TPB; off <<= 1) {
        const int t = (threadIdx.x >= (unsigned)off) ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    int run = s[threadIdx.x] - tot;
    for (int k = 0; k < per; ++k)
        if (b0 + k < nt) {
            const int v = tileSum[b0 + k];
            tileSum[b0 + k] = run;
            run += v;
        }
}

__global__ void k_scan_add(int* __restrict__ out, int m, const int* __restrict__ tileSum) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) out[i] += tileSum[i / SCAN_TILE];
}

/* levels 2 and 3 in one: each block sums the totals of the tiles before
 * its own (integers, any order) and adds them (blocks of TPB elements lie
 * in one tile) */
__global__ void __launch_bounds__(TPB) k_scan_add_prefix(int* __restrict__ out, int m,
                                                         const int* __restrict__ tileSum) {
    static_assert(SCAN_TILE % TPB == 0, "a block lies in one tile");
    __shared__ int s[TPB];
    const int i = blockIdx.x * TPB + (int)threadIdx.x;
    const int t = (blockIdx.x * TPB) / SCAN_TILE;
    int acc = 0;
    for (int j = (int)threadIdx.x; j < t; j += TPB) acc += tileSum[j];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int st = TPB / 2; st > 0; st >>= 1) {
        if (threadIdx.x < (unsigned)st) s[threadIdx.x] += s[threadIdx.x + st];
        __syncthreads();
    }
    if (i < m) out[i] += s[0];
}

/* Scatter (pid, slot) keys into their cell's run: one returning atomic per
 * key group, all groups in one instruction, the group base passed back from
 * its leader. */
__global__ void __launch_bounds__(TPB) k_scatter(const int* __restrict__ cid, const int* __restrict__ pid, int n,
                                                 const int* __restrict__ start, int* __restrict__ cursor,
                                                 unsigned long long* __restrict__ key) {
    const int lane = __lane_id();
    const int w0 = (blockIdx.x * TPB + (int)threadIdx.x - lane) * BIN_ITEMS;
    int c[BIN_ITEMS], s0[BIN_ITEMS], leader[BIN_ITEMS], rank[BIN_ITEMS], old[BIN_ITEMS];
    unsigned p[BIN_ITEMS];
    for (int r = 0; r < BIN_ITEMS; ++r) {
        const int i = w0 + r * 64 + lane;
        c[r] = (i < n) ? cid[i] : -1;
        p[r] = (i < n) ? (unsigned)pid[i] : 0u;
    }
    for (int r = 0; r < BIN_ITEMS; ++r) s0[r] = (c[r] >= 0) ? start[c[r]] : 0;
    for (int r = 0; r < BIN_ITEMS; ++r) {
        int size;
        wave_groups(c[r], c[r] >= 0, lane, leader[r], rank[r], size);
        old[r] = 0;
        if (c[r] >= 0 && lane == leader[r]) old[r] = atomicAdd(&cursor[c[r]], size);
    }
    for (int r = 0; r < BIN_ITEMS; ++r) {
        const int base = __shfl(old[r], leader[r] < 0 ? lane : leader[r], 64);
        const int i = w0 + r * 64 + lane;
        if (c[r] >= 0) key[s0[r] + base + rank[r]] = ((unsigned long long)p[r] << 32) | (unsigned)i;
    }
}

/* Order each cell's run by pid; writes the sorted keys back and the slots
 * to the list.  A block takes SEG_CELLS consecutive cells, whose runs are one
 * contiguous range of the keys, and stages the range in LDS; a run of up to
 * SORT_NET keys is sorted by a bitonic network on registers (on 32-bit
 * values pid << 5 | index when every pid fits 27 bits, else on the keys), a
 * longer one by insertion; a range too large for LDS is sorted in place in
 * global memory. */
constexpr int STAGE_U = 16; /* entries per thread in flight in the staging loops */
constexpr int SEG_CELLS = 256;
constexpr int SORT_NET = 32; /* runs up to this length are sorted by a network on registers */
constexpr int SEG_CAP = 6144;
__global__ void __launch_bounds__(SEG_CELLS) k_segsort(const int* __restrict__ start, int Nc,
                                                       unsigned long long* __restrict__ key, int* __restrict__ list,
                                                       int narrow, int writeKey) {
    __shared__ unsigned long long sk[SEG_CAP];
    const int c0 = blockIdx.x * SEG_CELLS;
    const int c1 = min(c0 + SEG_CELLS, Nc);
    const int b0 = start[c0], b1 = start[c1];
    const int c = c0 + (int)threadIdx.x;
    if (b1 - b0 <= SEG_CAP) {
        for (int q0 = b0; q0 < b1; q0 += SEG_CELLS * STAGE_U) {
            unsigned long long kv[STAGE_U];
            for (int u = 0; u < STAGE_U; ++u) {
                const int q = q0 + u * SEG_CELLS + (int)threadIdx.x;
                kv[u] = q < b1 ? key[q] : 0ull;
            }
            for (int u = 0; u < STAGE_U; ++u) {
                const int q = q0 + u * SEG_CELLS + (int)threadIdx.x;
                if (q < b1) sk[q - b0] = kv[u];
            }
        }
        __syncthreads();
        if (c < c1) {
            const int b = start[c] - b0, e = start[c + 1] - b0;
            if (e - b <= SORT_NET && narrow) {
                /* 32-bit network values (pid << 5 | index in the run), the
                 * run padded with the largest value (larger than any real
                 * one, as pid < 2^27 - 1); the keys then gathered by index */
                const unsigned* hiw = reinterpret_cast<const unsigned*>(sk) + 1;
                unsigned v[SORT_NET];
#pragma unroll
                for (int u = 0; u < SORT_NET; ++u) v[u] = b + u < e ? (hiw[2 * (b + u)] << 5) | (unsigned)u : ~0u;
#pragma unroll
                for (int k = 2; k <= SORT_NET; k <<= 1)
#pragma unroll
                    for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
                        for (int i = 0; i < SORT_NET; ++i) {
                            const int l = i ^ j;
                            if (l > i) {
                                const unsigned lo = min(v[i], v[l]), hi = max(v[i], v[l]);
                                const bool up = (i & k) == 0;
                                v[i] = up ? lo : hi;
                                v[l] = up ? hi : lo;
                            }
                        }
                unsigned long long kv[SORT_NET];
#pragma unroll
                for (int u = 0; u < SORT_NET; ++u) kv[u] = b + u < e ? sk[b + (int)(v[u] & 31u)] : 0ull;
#pragma unroll
                for (int u = 0; u < SORT_NET; ++u)
                    if (b + u < e) sk[b + u] = kv[u];
            } else if (e - b <= SORT_NET) {
                /* a bitonic network on registers (static indices), the
                 * run padded with the largest key */
                unsigned long long v[SORT_NET];
#pragma unroll
                for (int u = 0; u < SORT_NET; ++u) v[u] = b + u < e ? sk[b + u] : ~0ull;
#pragma unroll
                for (int k = 2; k <= SORT_NET; k <<= 1)
#pragma unroll
                    for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
                        for (int i = 0; i < SORT_NET; ++i) {
                            const int l = i ^ j;
                            if (l > i) {
                                const unsigned long long a0 = v[i], a1 = v[l];
                                const bool up = (i & k) == 0;
                                const bool sw = up ? (a0 > a1) : (a0 < a1);
                                v[i] = sw ? a1 : a0;
                                v[l] = sw ? a0 : a1;
                            }
                        }
#pragma unroll
                for (int u = 0; u < SORT_NET; ++u)
                    if (b + u < e) sk[b + u] = v[u];
            } else {
                for (int i = b + 1; i < e; ++i) {
                    const unsigned long long kv = sk[i];
                    int j = i - 1;
                    while (j >= b && sk[j] > kv) {
                        sk[j + 1] = sk[j];
                        --j;
                    }
                    sk[j + 1] = kv;
                }
            }
        }
        __syncthreads();
        for (int q = b0 + (int)threadIdx.x; q < b1; q += SEG_CELLS) {
            const unsigned long long kv = sk[q - b0];
            if (writeKey) key[q] = kv; /* the sorted keys: slot and pid of each position in one load */
            list[q] = (int)(unsigned)(kv & 0xffffffffull);
        }
        return;
    }
    if (c >= c1) return;
    const int b = start[c], e = start[c + 1];
    for (int i = b + 1; i < e; ++i) {
        const unsigned long long kv = key[i];
        int j = i - 1;
        while (j >= b && key[j] > kv) {
            key[j + 1] = key[j];
            --j;
        }
        key[j + 1] = kv;
    }
    for (int i = b; i < e; ++i) list[i] = (int)(unsigned)(key[i] & 0xffffffffull);
}

/* ---- incremental unshifted binning ---- */

/* Arrivals go to the tail of their new cell's run (after the stayers), in
 * arrival order; the merge sorts them. */
constexpr int PLACE_PER = 4; /* slots per thread (one 16-byte load of ranks) */
__global__ void k_place_arrivals(const int* __restrict__ cid, const int* __restrict__ pid,
                                 const int* __restrict__ rnk, int n, const int* __restrict__ startB,
                                 const int* __restrict__ cnt, const int* __restrict__ dout,
                                 unsigned long long* __restrict__ keyB) {
    const int i0 = (blockIdx.x * blockDim.x + threadIdx.x) * PLACE_PER;
    if (i0 >= n) return;
    int r[PLACE_PER];
    if (i0 + PLACE_PER <= n) {
        const int4 q = *reinterpret_cast<const int4*>(rnk + i0);
        r[0] = q.x;
        r[1] = q.y;
        r[2] = q.z;
        r[3] = q.w;
    } else {
        for (int u = 0; u < PLACE_PER; ++u) r[u] = i0 + u < n ? rnk[i0 + u] : -1;
    }
    for (int u = 0; u < PLACE_PER; ++u) {
        if (r[u] < 0) continue;
        const int i = i0 + u;
        const int c = cid[i];
        keyB[startB[c] + cnt[c] - dout[c] + r[u]] = ((unsigned long long)(unsigned)pid[i] << 32) | (unsigned)i;
    }
}

constexpr int REB_CELLS = 256; /* threads per block of the merge */
constexpr int REB_CAP = 6144;  /* keys of a block's new runs staged in LDS */
constexpr int REB_REG = 32;    /* old-run keys loaded per batch */
constexpr int REB_IN = 8;      /* arrivals merged on registers (more are inserted) */

/* One cell's new run: its old run's stayers (still in this cell, in pid
 * order) merged with its arrivals (the tail of the new run in keyB).  The
 * old run is read in batches of REB_REG keys, their cells gathered at once;
 * up to REB_IN arrivals are sorted on registers and each stayer's position
 * counts the arrivals below it.  out[o..o+m) receives the run; the
 * arrivals are read before out is written (out may be keyB's run). */
__device__ __forceinline__ void rebin_cell(int c, int b, int mo, int nb, int ms, int ki,
                                           const unsigned long long* key, const int* cid,
                                           const unsigned long long* keyB, unsigned long long* out, int o,
                                           int* err) {
    const bool regIn = ki <= REB_IN;
    unsigned long long a[REB_IN];
#pragma unroll
    for (int j = 0; j < REB_IN; ++j) a[j] = (regIn && j < ki) ? keyB[nb + ms + j] : ~0ull;
#pragma unroll
    for (int k = 2; k <= REB_IN; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < REB_IN; ++i) {
                const int l = i ^ j;
                if (l > i) {
                    const unsigned long long a0 = a[i], a1 = a[l];
                    const bool up = (i & k) == 0;
                    const bool sw = up ? (a0 > a1) : (a0 < a1);
                    a[i] = sw ? a1 : a0;
                    a[l] = sw ? a0 : a1;
                }
            }
    int before = 0;
    int below[REB_IN]; /* stayers below each arrival */
#pragma unroll
    for (int j = 0; j < REB_IN; ++j) below[j] = 0;
    for (int u0 = 0; u0 < mo; u0 += REB_REG) {
        unsigned long long v[REB_REG];
        int sc[REB_REG];
#pragma unroll
        for (int u = 0; u < REB_REG; ++u) v[u] = u0 + u < mo ? key[b + u0 + u] : 0ull;
#pragma unroll
        for (int u = 0; u < REB_REG; ++u) sc[u] = u0 + u < mo ? cid[(int)(unsigned)v[u]] : -1;
#pragma unroll
        for (int u = 0; u < REB_REG; ++u) {
            if (sc[u] == c) {
                int pos = before;
#pragma unroll
                for (int j = 0; j < REB_IN; ++j) {
                    const bool lt = a[j] < v[u];
                    pos += lt ? 1 : 0;
                    below[j] += lt ? 0 : 1;
                }
                if (before < ms) out[o + pos] = v[u];
                ++before;
            }
        }
    }
    if (before != ms) {
        atomicOr(err, 4);
        return;
    }
    if (regIn) {
#pragma unroll
        for (int j = 0; j < REB_IN; ++j)
            if (j < ki) out[o + j + below[j]] = a[j];
        return;
    }
    /* many arrivals: appended after the stayers and inserted */
    if (out + o != keyB + nb)
        for (int j = 0; j < ki; ++j) out[o + ms + j] = keyB[nb + ms + j];
    for (int i = o + ms; i < o + ms + ki; ++i) {
        const unsigned long long k = out[i];
        int j = i - 1;
        while (j >= o && out[j] > k) {
            out[j + 1] = out[j];
            --j;
        }
        out[j + 1] = k;
    }
}

/* A block takes cpb (<= REB_CELLS) consecutive cells; their new runs are
 * one contiguous range, assembled in LDS and written out coalesced (keys
 * and slots), or written in place when the range exceeds REB_CAP.  Resets
 * cnt to the new counts. */
__global__ void __launch_bounds__(REB_CELLS) k_rebin_merge(const int* __restrict__ start,
                                                           const unsigned long long* __restrict__ key,
                                                           const int* __restrict__ cid, int Nc, int cpb,
                                                           const int* __restrict__ startB, int* __restrict__ cnt,
                                                           int* __restrict__ inc, int* __restrict__ dout,
                                                           int* __restrict__ rnk, unsigned long long* keyB,
                                                           int* __restrict__ listB, int* __restrict__ err) {
    __shared__ unsigned long long so[REB_CAP];
    const int c0 = blockIdx.x * cpb;
    const int c1 = min(c0 + cpb, Nc);
    const int q0 = startB[c0], q1 = startB[c1];
    const int c = c0 + (int)threadIdx.x;
    const bool mine = (int)threadIdx.x < cpb && c < c1;
    const bool lds = q1 - q0 <= REB_CAP; /* uniform over the block */
    int nb = 0, m = 0;
    if (mine) {
        const int b = start[c], mo = start[c + 1] - b;
        nb = startB[c];
        const int ms = cnt[c] - dout[c], ki = inc[c];
        inc[c] = 0; /* clear for the next update */
        dout[c] = 0;
        for (int j = 0; j < ki; ++j) rnk[(int)(unsigned)(keyB[nb + ms + j] & 0xffffffffull)] = -1;
        m = ms + ki;
        if (lds)
            rebin_cell(c, b, mo, nb, ms, ki, key, cid, keyB, so, nb - q0, err);
        else
            rebin_cell(c, b, mo, nb, ms, ki, key, cid, keyB, keyB, nb, err);
        cnt[c] = m;
    }
    if (lds) {
        __syncthreads();
        for (int q = q0 + (int)threadIdx.x; q < q1; q += REB_CELLS) {
            const unsigned long long kv = so[q - q0];
            keyB[q] = kv;
            listB[q] = (int)(unsigned)(kv & 0xffffffffull);
        }
    } else {
        for (int q = nb; q < nb + m; ++q) listB[q] = (int)(unsigned)(keyB[q] & 0xffffffffull);
    }
}

/* ---- collision ---- */

/* Per-cell kernels stage a block's particles in LDS: CELL_BLK consecutive
 * cells own one contiguous range of the binning list, read with coalesced
 * loads; each thread then runs its cell's sequential sums (particle order,
 * so the same rounding as one pass over global memory) on LDS.  A range
 * larger than `cap` entries falls back to the global-memory loop. */
constexpr int CELL_BLK = 64;

/* Staging of a block's particle range list[b0 .. b0+n) to and from LDS.
 * Each thread issues STAGE_U index loads, then STAGE_U gathers (or
 * scatters), so a thread's accesses are in flight together rather than one
 * dependent pair at a time. */
__device__ inline void stage_in(const int* __restrict__ list, int b0, int n, const double* __restrict__ vx,
                                const double* __restrict__ vy, double* svx, double* svy) {
    const int tid = (int)threadIdx.x, nthr = (int)blockDim.x;
    for (int q0 = 0; q0 < n; q0 += nthr * STAGE_U) {
        int sl[STAGE_U];
        for (int u = 0; u < STAGE_U; ++u) {
            const int q = q0 + u * nthr + tid;
            sl[u] = q < n ? list[b0 + q] : -1;
        }
        double ax[STAGE_U], ay[STAGE_U];
        for (int u = 0; u < STAGE_U; ++u) {
            ax[u] = sl[u] >= 0 ? vx[sl[u]] : 0.0;
            ay[u] = sl[u] >= 0 ? vy[sl[u]] : 0.0;
        }
        for (int u = 0; u < STAGE_U; ++u)
            if (sl[u] >= 0) {
                const int q = q0 + u * nthr + tid;
                svx[q] = ax[u];
                svy[q] = ay[u];
            }
    }
}
__device__ inline void stage_out(const int* __restrict__ list, int b0, int n, double* __restrict__ vx,
                                 double* __restrict__ vy, const double* svx, const double* svy) {
    const int tid = (int)threadIdx.x, nthr = (int)blockDim.x;
    for (int q0 = 0; q0 < n; q0 += nthr * STAGE_U) {
        int sl[STAGE_U];
        for (int u = 0; u < STAGE_U; ++u) {
            const int q = q0 + u * nthr + tid;
            sl[u] = q < n ? list[b0 + q] : -1;
        }
        for (int u = 0; u < STAGE_U; ++u)
            if (sl[u] >= 0) {
                const int q = q0 + u * nthr + tid;
                vx[sl[u]] = svx[q];
                vy[sl[u]] = svy[q];
            }
    }
}

__global__ void __launch_bounds__(CELL_BLK) k_collide(CollideParams cp, uint64_t seed, const int* __restrict__ start,
                                                      const int* __restrict__ list, int Nc, double* __restrict__ vx,
                                                      double* __restrict__ vy, int cap) {
    extern __shared__ double sv[];
    const int c0 = blockIdx.x * CELL_BLK;
    const int c1 = min(c0 + CELL_BLK, Nc);
    const int b0 = start[c0], b1 = start[c1];
    const int c = c0 + (int)threadIdx.x;
    if (b1 - b0 <= cap) {
        double* svx = sv;
        double* svy = sv + cap;
        stage_in(list, b0, b1 - b0, vx, vy, svx, svy);
        __syncthreads();
        if (c < c1) {
            const int o = start[c] - b0;
            collide_cell(cp, seed, c, nullptr, start[c + 1] - start[c], svx + o, svy + o);
        }
        __syncthreads();
        stage_out(list, b0, b1 - b0, vx, vy, svx, svy);
        return;
    }
    if (c >= c1) return;
    collide_cell(cp, seed, c, list + start[c], start[c + 1] - start[c], vx, vy);
}

/* deterministic block partial sums of (vx, vy) */
__global__ void __launch_bounds__(TPB) k_sum2(const double* __restrict__ a, const double* __restrict__ b, int n,
                                              double* __restrict__ part) {
    __shared__ double s0[TPB], s1[TPB];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    s0[threadIdx.x] = (i < n) ? a[i] : 0.0;
    s1[threadIdx.x] = (i < n) ? b[i] : 0.0;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < (unsigned)s) {
            s0[threadIdx.x] += s0[threadIdx.x + s];
            s1[threadIdx.x] += s1[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s0[0];
        part[2 * blockIdx.x + 1] = s1[0];
    }
}

/* single-block reduction of `np` partial tuples of width w into out[0..w) */
/* column sums of part[np][w] by one block (fixed order: strided partial
 * sums, then a tree), written by thread 0 */
__device__ inline void reduce_parts_block(const double* __restrict__ part, int np, int w, double* __restrict__ out) {
    __shared__ double s[TPB];
    for (int k = 0; k < w; ++k) {
        double acc = 0.0;
        for (int j = threadIdx.x; j < np; j += TPB) acc += part[(size_t)j * w + k];
        s[threadIdx.x] = acc;
        __syncthreads();
        for (int st = TPB / 2; st > 0; st >>= 1) {
            if (threadIdx.x < (unsigned)st) s[threadIdx.x] += s[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[k] = s[0];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(TPB) k_reduce_parts(const double* __restrict__ part, int np, int w,
                                                      double* __restrict__ out) {
    reduce_parts_block(part, np, w, out);
}

__global__ void k_sub_mean(double* __restrict__ vx, double* __restrict__ vy, int n, const double* __restrict__ sums) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double mx = sums[0] / (double)n, my = sums[1] / (double)n;
        vx[i] -= mx;
        vy[i] -= my;
    }
}

/* ---- diagnostics: occupancy statistics and state means ---- */

/* Per-block partials of the occupancy statistics of a count grid
 * (compute_occ_std_from_counts, compute_out_band_from_counts,
 * count_dense_from_counts, count_sparse_from_counts, count_empty_active:
 * zone_pass.cpp:975-1044; every cell is active, zone_pass.cpp:128-130).
 * part[b][6] = sum (N - gamma), sum (N - gamma)^2, out-of-band, dense,
 * sparse, empty.  An occupied cell takes the bulk thresholds, an empty one
 * the fluid-fraction map of its 3x3 binomial smoothing (no wrap): an empty
 * cell is never out of band, dense or sparse, so the map only has to exist
 * for the comparison to read like the reference's. */
constexpr int OCC_W = 6;
/* slots of a simulation's diagnostics buffer */
constexpr int DIAG_OCC_BASE = 0, DIAG_OCC_SHIFTED = 5, DIAG_PSUM = 10, DIAG_USUM = 13, DIAG_SUR_OCC = 16,
              DIAG_SUR_DIV = 21, DIAG_ZK_KE0 = 24, DIAG_ZK_KE1 = 27, DIAG_ZK_CELLS = 30, DIAG_ZK_OCC = 36,
              DIAG_ZK_OUT = 41, DIAG_SLOTS = 51;
/* gamma x the clamped fluid fraction of cell c's 3x3 binomial smoothing of
 * the count grid (no wrap) */
__device__ __forceinline__ double occ_local_gl(const int* __restrict__ cnt, int c, int Nx, int Ny,
                                               const ZoneParams& zp) {
    const int ix = c / Ny, iy = c - ix * Ny;
    long long num = 0;
    int den = 0;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy) {
            const int jx = ix + dx, jy = iy + dy;
            if (jx < 0 || jx >= Nx || jy < 0 || jy >= Ny) continue;
            const int w = (dx == 0 ? 2 : 1) * (dy == 0 ? 2 : 1);
            num += (long long)w * cnt[jy + Ny * jx];
            den += w;
        }
    const double t = zp.fluidFracGain * (((double)num / (double)den) / dmax(zp.gamma, 1e-12));
    return zp.gamma * ((t < 0.0) ? 0.0 : ((1.0 < t) ? 1.0 : t));
}

/* cntLow: the grid the low thresholds come from (cnt itself for the pass
 * statistics; for a zone kernel's diagnostics the counts its sparse phase
 * started from, zone_pass.cpp:783-795, which the final maps keep) */
__global__ void __launch_bounds__(TPB) k_occupancy(const int* __restrict__ cnt, const int* __restrict__ cntLow,
                                                   int Nx, int Ny, ZoneParams zp, double* __restrict__ part) {
    __shared__ double sh[OCC_W][TPB];
    const int Nc = Nx * Ny;
    double a[OCC_W] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int c = blockIdx.x * TPB + (int)threadIdx.x; c < Nc; c += gridDim.x * TPB) {
        const int m = cnt[c], mL = cntLow[c];
        const double d = (double)m - zp.gamma;
        a[0] += d;
        a[1] += d * d;
        double lo = zp.lowThrBulk, hi = zp.highThr;
        if (m == 0) hi = zp.useLocal ? dmax(zp.highFrac * occ_local_gl(cnt, c, Nx, Ny, zp), zp.highThrFloor)
                                     : zp.highLoc1;
        if (mL == 0) lo = zp.useLocal ? dmax(zp.lowFrac * occ_local_gl(cntLow, c, Nx, Ny, zp), zp.lowThrFloor)
                                      : zp.lowLoc1;
        const bool dense = (double)m > hi, sparse = m > 0 && (double)m < lo;
        a[2] += (dense || sparse) ? 1.0 : 0.0;
        a[3] += dense ? 1.0 : 0.0;
        a[4] += sparse ? 1.0 : 0.0;
        a[5] += m == 0 ? 1.0 : 0.0;
    }
    for (int k = 0; k < OCC_W; ++k) sh[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int st = TPB / 2; st > 0; st >>= 1) {
        if (threadIdx.x < (unsigned)st)
            for (int k = 0; k < OCC_W; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < OCC_W) part[(size_t)blockIdx.x * OCC_W + threadIdx.x] = sh[threadIdx.x][0];
}

/* out[0..5) = occStd, outBand, nDense, nSparse, nEmptyActive from the
 * partials of k_occupancy.  With an integer gamma every partial is an exact
 * integer, so the sums are exact and the variance is
 * sum d^2 / Nc - (sum d / Nc)^2 with d = N - gamma. */
__global__ void __launch_bounds__(TPB) k_occupancy_final(const double* __restrict__ part, int np, int Nc, double gamma,
                                                         double* __restrict__ out) {
    __shared__ double tot[OCC_W];
    reduce_parts_block(part, np, OCC_W, tot);
    __syncthreads();
    if (threadIdx.x == 0) {
        const double m = tot[0] / (double)Nc;
        double var = tot[1] / (double)Nc - m * m;
        if (var < 0.0) var = 0.0;
        out[0] = sqrt(var) / dmax(gamma, 1e-30);
        out[1] = tot[2] / (double)Nc;
        out[2] = tot[3];
        out[3] = tot[4];
        out[4] = tot[5];
    }
}

/* deterministic block partials of (vx, vy, 0.5 |v|^2) */
__global__ void __launch_bounds__(TPB) k_sum_ke(const double* __restrict__ vx, const double* __restrict__ vy, int n,
                                                double* __restrict__ part) {
    __shared__ double s0[TPB], s1[TPB], s2[TPB];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const double a = (i < n) ? vx[i] : 0.0, b = (i < n) ? vy[i] : 0.0;
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = b;
    s2[threadIdx.x] = 0.5 * (a * a + b * b);
    __syncthreads();
    for (int st = TPB / 2; st > 0; st >>= 1) {
        if (threadIdx.x < (unsigned)st) {
            s0[threadIdx.x] += s0[threadIdx.x + st];
            s1[threadIdx.x] += s1[threadIdx.x + st];
            s2[threadIdx.x] += s2[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * (size_t)blockIdx.x] = s0[0];
        part[3 * (size_t)blockIdx.x + 1] = s1[0];
        part[3 * (size_t)blockIdx.x + 2] = s2[0];
    }
}

/* ---- ZoneKernelStats.diag of a single-zone call (zone_pass.cpp:924-970) ---- */

/* the final counts, with the zone's cells replaced by the counts its sparse
 * phase started from: the grid of the final low thresholds */
__global__ void __launch_bounds__(TPB) k_zone_low_counts(const int* __restrict__ cnt, int Nc, int Ny, ZoneRect r,
                                                         const int* __restrict__ dgCnt, int* __restrict__ out) {
    const int c = blockIdx.x * TPB + (int)threadIdx.x;
    if (c >= Nc) return;
    const int ix = c / Ny, iy = c - ix * Ny;
    const bool in = ix >= r.xs && ix < r.xe && iy >= r.ys && iy < r.ye;
    out[c] = in ? dgCnt[(ix - r.xs) * (r.ye - r.ys) + (iy - r.ys)] : cnt[c];
}

/* Sums over the zone's cells of the kernel's per-cell record, one block:
 * out[0..6) = sum and count of the nonzero squared momentum defects in x,
 * the same in y, and the sums of the squared final velocities
 * (momentum / count) of the occupied cells in x and y. */
__global__ void __launch_bounds__(TPB) k_zone_diag_cells(const double* __restrict__ dg, const int* __restrict__ cnt,
                                                         int Ny, ZoneRect r, double* __restrict__ out) {
    __shared__ double sh[6][TPB];
    const int zh = r.ye - r.ys, C = (r.xe - r.xs) * zh;
    double a[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int lc = (int)threadIdx.x; lc < C; lc += TPB) {
        const int gc = (r.ys + lc % zh) + Ny * (r.xs + lc / zh);
        const double ex = dg[4 * (size_t)lc] * dg[4 * (size_t)lc], ey = dg[4 * (size_t)lc + 1] * dg[4 * (size_t)lc + 1];
        if (ex > 0.0) {
            a[0] += ex;
            a[1] += 1.0;
        }
        if (ey > 0.0) {
            a[2] += ey;
            a[3] += 1.0;
        }
        const int m = cnt[gc];
        if (m > 0) {
            const double ux = dg[4 * (size_t)lc + 2] / (double)m, uy = dg[4 * (size_t)lc + 3] / (double)m;
            a[4] += ux * ux;
            a[5] += uy * uy;
        }
    }
    for (int k = 0; k < 6; ++k) sh[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int st = TPB / 2; st > 0; st >>= 1) {
        if (threadIdx.x < (unsigned)st)
            for (int k = 0; k < 6; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x] = sh[threadIdx.x][0];
}

/* diag[0..10) from the occupancy statistics of the final counts (occ:
 * occStd, outBand, nDense, nSparse, nEmpty), the particle sums (vx, vy,
 * kinetic energy) before and after the kernel, and the cell sums above;
 * every cell is active and the cells with a velocity are the occupied ones */
__global__ void k_zone_diag_final(const double* __restrict__ occ, const double* __restrict__ ke0,
                                  const double* __restrict__ ke1, const double* __restrict__ cs, int Nc,
                                  double* __restrict__ diag) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    diag[0] = occ[0];
    diag[1] = occ[1];
    const double dX = ke1[0] - ke0[0], dY = ke1[1] - ke0[1];
    diag[2] = sqrt(dX * dX + dY * dY);
    diag[3] = cs[1] > 0.0 ? sqrt(cs[0] / cs[1]) : 0.0;
    diag[4] = cs[3] > 0.0 ? sqrt(cs[2] / cs[3]) : 0.0;
    diag[5] = ke1[2] - ke0[2];
    diag[6] = __builtin_nan("");
    diag[7] = 0.0;
    const double nV = (double)Nc - occ[4];
    diag[8] = nV > 0.0 ? sqrt(cs[4] / nV) : 0.0;
    diag[9] = nV > 0.0 ? sqrt(cs[5] / nV) : 0.0;
}

/* ---- cell fields ---- */

struct FieldsOut {
    int* N;
    double *Ux, *Uy, *Px, *Py, *T, *rho, *Pkin, *Pvir, *P;
};

__global__ void __launch_bounds__(CELL_BLK) k_cell_fields(const int* __restrict__ start, const int* __restrict__ list,
                                                          int Nc, const double* __restrict__ vx,
                                                          const double* __restrict__ vy, double Vc, double K,
                                                          double rhoT, FieldsOut f, int cap) {
    extern __shared__ double sv[];
    const int c0 = blockIdx.x * CELL_BLK;
    const int c1 = min(c0 + CELL_BLK, Nc);
    const int b0 = start[c0], b1 = start[c1];
    const int c = c0 + (int)threadIdx.x;
    CellStats s;
    if (b1 - b0 <= cap) {
        double* svx = sv;
        double* svy = sv + cap;
        stage_in(list, b0, b1 - b0, vx, vy, svx, svy);
        __syncthreads();
        if (c >= c1) return;
        const int o = start[c] - b0;
        s = cell_stats(nullptr, start[c + 1] - start[c], svx + o, svy + o, Vc, K, rhoT);
    } else {
        if (c >= c1) return;
        s = cell_stats(list + start[c], start[c + 1] - start[c], vx, vy, Vc, K, rhoT);
    }
    if (f.N) f.N[c] = s.N;
    if (f.Ux) f.Ux[c] = s.Ux;
    if (f.Uy) f.Uy[c] = s.Uy;
    if (f.Px) f.Px[c] = s.Px;
    if (f.Py) f.Py[c] = s.Py;
    if (f.T) f.T[c] = s.T;
    if (f.rho) f.rho[c] = s.rho;
    if (f.Pkin) f.Pkin[c] = s.Pkin;
    if (f.Pvir) f.Pvir[c] = s.Pvir;
    if (f.P) f.P[c] = s.P;
}

/* ---- zone pass: one workgroup (one wavefront) per zone ---- */

struct ZoneArgs {
    ZoneIn in;
    const ZoneRect* rects;
    int nZones;
    uint64_t seedBase;
    int singleSeed; /* 1: seedBase is the zone's own seed (single-zone call) */
    const int* zsize; /* particles of each zone */
    const int* zoff;  /* per-zone offsets into the particle regions (exclusive scan of zsize) */
    long long scratchCap;
    void* nextG;    /* links beyond LDS, IT per particle */
    void* tmpG;     /* long donor lists, IT per particle */
    double* duG;    /* 2 per cell per zone */
    unsigned char* rec; /* last move of each particle, 1 per particle slot of the zone regions */
    unsigned char* zstate; /* per-zone state between the passes and the correction */
    size_t zsStride;
    int duStride;   /* doubles per zone */
    int pMin, pMax; /* zones with pMin <= P <= pMax run in this launch */
    size_t ldsBytes;
    unsigned char* gscratch; /* working sets in global memory when they do not fit LDS */
    size_t gstride;
    ZoneOut* out;
    int* err;
    /* an ensemble's pass (the kEns kernels): zone -> replica, the replicas'
     * own inputs (binning views, owner and cover maps, geometry, thresholds),
     * their zone seed bases and their first zone index in this pass */
    const int* ensRep = nullptr;
    const ZoneIn* ensIn = nullptr;
    const uint64_t* ensSeed = nullptr;
    const int* ensZoneOff = nullptr;
};

/* The inputs, zone number (1-based, within its own layout) and seed of zone
 * zi of a pass: the pass's own, or (kEns) those of the zone's replica. */
template <bool kEns>
__device__ __forceinline__ const ZoneIn& zone_in_of(const ZoneArgs& a, int zi, int& zid, uint64_t& seed) {
    if constexpr (kEns) {
        const int r = a.ensRep[zi];
        zid = zi - a.ensZoneOff[r] + 1;
        seed = a.ensSeed[r] + 104729ULL * (uint64_t)zid;
        return a.ensIn[r];
    } else {
        zid = zi + 1;
        seed = a.singleSeed ? a.seedBase : a.seedBase + 104729ULL * (uint64_t)(zi + 1);
        return a.in;
    }
}

/* particles of each zone: its columns' runs of the binning list */
/* One wavefront per zone (block nZ writes the terminating zero): each lane
 * takes columns of the zone, whose cells are one run of the binning. */
__global__ void __launch_bounds__(64) k_zone_sizes(const ZoneRect* __restrict__ rects, int nZ,
                                                   const int* __restrict__ start, int Ny, int* __restrict__ sizes) {
    const int zi = blockIdx.x, lane = threadIdx.x;
    if (zi > nZ) return;
    if (zi == nZ) {
        if (lane == 0) sizes[zi] = 0;
        return;
    }
    const ZoneRect r = rects[zi];
    int P = 0;
    for (int gx = r.xs + lane; gx < r.xe; gx += 64) P += start[(r.ye - 1 + Ny * gx) + 1] - start[r.ys + Ny * gx];
    for (int o = 32; o > 0; o >>= 1) P += __shfl_xor(P, o, 64);
    if (lane == 0) sizes[zi] = P;
}

/* One zone per workgroup of one wavefront.  IT = uint16_t keeps the working
 * set in LDS at four zones per CU; IT = int serves zones of 65535 particles
 * or more (and the single-zone layout), in LDS or in global scratch. */
/* The passes of zone zi by one wavefront (lanes 0..63): set-up, the dense
 * and sparse phases, the state hand-over to the correction.  The zone code
 * synchronises its lanes with wavefront-level barriers only. */
template <class IT, bool kLds>
__device__ __forceinline__ void zone_pass_body(const ZoneArgs& a, int zi, long long off, int lane,
                                               unsigned char* lds, ZoneShared* sh) {
    ZoneCtxT<IT> z;
    const uint64_t seed = a.singleSeed ? a.seedBase : a.seedBase + 104729ULL * (uint64_t)(zi + 1);
    zone_init(z, a.in, a.rects[zi], zi + 1, seed, sh);
    z.nxG = (IT*)a.nextG + off;
    z.tmpG = (IT*)a.tmpG + off;
    z.duG = a.duG + (size_t)zi * (size_t)a.duStride;
    z.rec = a.rec + off;
    if (kLds) zone_carve(z, lds, a.ldsBytes);
    else zone_carve(z, a.gscratch + (size_t)zi * a.gstride, a.gstride);
    zone_setup(z, a.in, lane, ZONE_TPB);
    ZBAR();
    zone_passes(z, lane, ZONE_TPB);
    zone_dump(z, zone_state_at<IT>(a.zstate + (size_t)zi * a.zsStride, z.C, z.zw), lane, ZONE_TPB);
    if (lane == 0) {
        ZoneOut o;
        o.nHigh = sh->nHigh;
        o.nLow = sh->nLow;
        o.movedDense = sh->movedDense;
        o.movedSparse = sh->movedSparse;
        o.stats4 = 0.0;
        o.stats5 = 0.0;
        a.out[zi] = o;
    }
}

/* A workgroup takes zones blockIdx.x, + gridDim.x, ... (the grid is
 * smaller than the zone count when few zones need this form). */
/* The zones block b takes, b + gridDim.x * k, 64 at a time: each lane reads
 * one zone's size and a ballot names the zones in this launch's range (the
 * others go to another launch), so a small grid finds its few zones in one
 * pass of loads.  Returns false when the block has no zone left. */
__device__ inline bool next_zone(const ZoneArgs& a, int& k0, unsigned long long& todo, int& zi) {
    while (!todo) {
        if ((long long)blockIdx.x + (long long)k0 * gridDim.x >= a.nZones) return false;
        const int ln = (int)(threadIdx.x & 63);
        const long long zl = (long long)blockIdx.x + (long long)(k0 + ln) * gridDim.x;
        const int pl = zl < a.nZones ? a.zsize[zl] : -1;
        todo = __ballot(zl < a.nZones && pl >= a.pMin && pl <= a.pMax);
        k0 += 64;
    }
    const int j = __builtin_ctzll(todo);
    todo &= todo - 1ull;
    zi = (int)((long long)blockIdx.x + (long long)(k0 - 64 + j) * gridDim.x);
    return true;
}

template <class IT, bool kLds>
__global__ void __launch_bounds__(ZONE_TPB) k_zone(ZoneArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ ZoneShared sh;
    const int lane = threadIdx.x;
    int k0 = 0, zi = 0;
    unsigned long long todo = 0ull;
    while (next_zone(a, k0, todo, zi)) {
        const int P = a.zsize[zi];
        const long long off = a.zoff[zi];
        if (off + (long long)P > a.scratchCap) {
            if (lane == 0) atomicOr(a.err, 2);
            continue;
        }
        zone_pass_body<IT, kLds>(a, zi, off, lane, lds, &sh);
        ZBAR(); /* the next zone reuses the working set */
    }
}

/* The momentum correction of the zones k_zone left (zone_pass.cpp:870-922),
 * FINISH_TPB threads per zone: a group of FIN_G lanes per cell for the
 * input-run sums and for the final lists' sums and writes (the sums in list
 * order), then one wavefront for the ordered stats. */
constexpr int FINISH_TPB = 256;
constexpr int FINISH_LDS = 40 * 1024; /* cell arrays, column map and as many links as fit */
/* The passes' zone, still in the workgroup's LDS (zone_carve's layout), for
 * the correction in the same kernel: its particle count, ownership flag and
 * the number of links the passes held in LDS. */
struct ZoneHandoff {
    int P, allOwned, nL;
};

/* The correction of zone zi by TPB threads (zones of more than CPT x TPB
 * cells take the two-sweep form); lds/ldsBytes: the staging space (kLds).
 * hand (kLds only): the passes left the zone in lds, so the cell arrays and
 * column map are already in place and the links move down from behind the
 * passes' scratch; otherwise the zone state comes from global memory. */
template <class IT, bool kLds, int TPB, int CPT>
__device__ __forceinline__ void zone_finish_body(const ZoneArgs& a, const ZoneIn& in, int zid, uint64_t seed, int zi,
                                                 long long off, int tid, unsigned char* lds, long long ldsBytes,
                                                 int* queue, const ZoneHandoff* hand = nullptr) {
    ZoneCtxT<IT> z;
    zone_init(z, in, a.rects[zi], zid, seed, nullptr);
    z.nxG = (IT*)a.nextG + off;
    z.tmpG = (IT*)a.tmpG + off;
    z.duG = a.duG + (size_t)zi * (size_t)a.duStride;
    z.rec = a.rec + off;
    const ZoneState<IT> st = zone_state_at<IT>(a.zstate + (size_t)zi * a.zsStride, z.C, z.zw);
    if (kLds && hand) {
        z.P = hand->P;
        z.allOwned = hand->allOwned;
    } else {
        zone_attach_state(z, st);
    }
    if (kLds && hand) {
        unsigned char* p = lds;
        IT* cnt = (IT*)p; p += (size_t)z.C * sizeof(IT);
        IT* head = (IT*)p; p += (size_t)z.C * sizeof(IT);
        IT* tail = (IT*)p; p += (size_t)z.C * sizeof(IT);
        IT* k0 = (IT*)p; p += (size_t)z.C * sizeof(IT);
        unsigned short* fl = (unsigned short*)p; p += (size_t)z.C * sizeof(unsigned short);
        p = lds + (((size_t)(p - lds) + 3) & ~(size_t)3);
        int* colL = (int*)p; p += (size_t)z.zw * sizeof(int);
        int* colG = (int*)p; p += (size_t)z.zw * sizeof(int);
        IT* links = (IT*)p;
        const long long room = (ldsBytes - (long long)(p - lds)) / (long long)sizeof(IT);
        const int nL = (int)(room < (long long)z.P ? (room > 0 ? room : 0) : z.P);
        /* the passes' links k < hand->nL, behind their scratch, move down in
         * chunks no longer than the distance they move, so a chunk overwrites
         * only links earlier chunks have copied; the rest come from the
         * zone's global link region */
        const IT* src = (const IT*)(lds + zone_fixed_bytes<IT>(z.C, z.zw, z.zp.useLocal));
        const int gap = (int)(src - links);
        const int nP = hand->nL < nL ? hand->nL : nL;
        for (int c0 = 0; c0 < nP; c0 += gap) {
            const int c1 = c0 + gap < nP ? c0 + gap : nP;
            for (int k = c0 + tid; k < c1; k += TPB) links[k] = src[k];
            __syncthreads();
        }
        for (int k = nP + tid; k < nL; k += TPB) links[k] = z.nxG[k];
        z.cnt = cnt;
        z.head = head;
        z.tail = tail;
        z.k0 = k0;
        z.fl = fl;
        z.colL = colL;
        z.colG = colG;
        z.nxL = links;
        z.nL = nL;
        __syncthreads();
    } else if (kLds) {
        /* the state and the first links to LDS */
        unsigned char* p = lds;
        IT* cnt = (IT*)p; p += (size_t)z.C * sizeof(IT);
        IT* head = (IT*)p; p += (size_t)z.C * sizeof(IT);
        IT* tail = (IT*)p; p += (size_t)z.C * sizeof(IT);
        IT* k0 = (IT*)p; p += (size_t)z.C * sizeof(IT);
        unsigned short* fl = (unsigned short*)p; p += (size_t)z.C * sizeof(unsigned short);
        p = lds + (((size_t)(p - lds) + 3) & ~(size_t)3);
        int* colL = (int*)p; p += (size_t)z.zw * sizeof(int);
        int* colG = (int*)p; p += (size_t)z.zw * sizeof(int);
        IT* links = (IT*)p;
        const long long room = (ldsBytes - (long long)(p - lds)) / (long long)sizeof(IT);
        const int nL = (int)(room < (long long)z.P ? (room > 0 ? room : 0) : z.P);
        for (int lc = tid; lc < z.C; lc += TPB) {
            cnt[lc] = st.cnt[lc];
            head[lc] = st.head[lc];
            tail[lc] = st.tail[lc];
            k0[lc] = st.k0[lc];
            fl[lc] = st.fl[lc];
        }
        for (int c = tid; c < z.zw; c += TPB) {
            colL[c] = st.colL[c];
            colG[c] = st.colG[c];
        }
        for (int k = tid; k < nL; k += TPB) links[k] = z.nxG[k];
        z.cnt = cnt;
        z.head = head;
        z.tail = tail;
        z.k0 = k0;
        z.fl = fl;
        z.colL = colL;
        z.colG = colG;
        z.nxL = links;
        z.nL = nL;
        __syncthreads();
    }
    /* phase 1: input-run sums of the changed, occupied cells, into the zone's
     * per-cell region (the base pass has them from the reference fields) */
    static_assert(TPB % FIN_G == 0, "whole groups");
    const int r = tid % FIN_G, gi = tid / FIN_G;
    const bool regs = z.C <= CPT * TPB;
    if (regs && !z.inPx)
        for (int lc0 = 0; lc0 < z.C; lc0 += TPB / FIN_G) {
            const int lc = lc0 + gi;
            if (lc >= z.C || !fget(z, lc, ZF_CHG) || cnt_of(z, lc) == 0) continue;
            double bx, by;
            fin_input_sum_group(z, lc, r, tid & 63, bx, by);
            if (r == 0) {
                z.duG[2 * lc] = bx;
                z.duG[2 * lc + 1] = by;
            }
        }
    __syncthreads();
    /* phase 2: final lists, a group of FIN_G lanes per changed cell */
    if (regs) {
        for (int lc0 = 0; lc0 < z.C; lc0 += TPB / FIN_G) {
            const int lc = lc0 + gi;
            if (lc >= z.C) continue;
            if (!fget(z, lc, ZF_CHG) || cnt_of(z, lc) == 0) {
                if (r == 0) {
                    /* nothing to correct: the diagnostics' record, the side-buffer copy of a covered cell */
                    if (z.dgCell && cnt_of(z, lc) > 0 && !fget(z, lc, ZF_CHG) && !fget(z, lc, ZF_OUT))
                        diag_cell_unchanged(z, lc);
                    if (!fget(z, lc, ZF_CHG)) zone_correct_write(z, lc);
                    st.dsq[lc] = 0.0;
                }
                continue;
            }
            double bx, by;
            if (z.inPx) {
                const int gc = lc_gc(z, lc);
                bx = z.inPx[gc];
                by = z.inPy[gc];
            } else {
                bx = z.duG[2 * lc];
                by = z.duG[2 * lc + 1];
            }
            const double d = fin_cell_group(z, lc, r, tid & 63, bx, by);
            if (r == 0) st.dsq[lc] = d;
        }
    } else {
        /* more cells than threads x CPT: the two-sweep form, positions included */
        for (int lc = tid; lc < z.C; lc += TPB) zone_correct_sums(z, lc);
        __syncthreads();
        for (int lc = tid; lc < z.C; lc += TPB) st.dsq[lc] = zone_correct_write(z, lc);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    /* phase 3: moved particles' positions (the two-sweep form placed them) on
     * all wavefronts but the last, which meanwhile takes the ordered stats of
     * the per-cell |du|^2 (phase 3 reads none of them) */
    static_assert(TPB >= 128 && TPB % 64 == 0, "one wavefront for the stats, at least one for phase 3");
    constexpr int TS = TPB - 64;
    if (regs && tid < TS) fin_place(z, tid, TS, queue + (tid & ~63));
    if (tid >= TS) {
        double rms, mx;
        zone_correct_stats(z, st.dsq, tid - TS, 64, rms, mx);
        if (tid == TS) {
            a.out[zi].stats4 = rms;
            a.out[zi].stats5 = mx;
        }
    }
}

template <class IT, bool kLds>
__global__ void __launch_bounds__(FINISH_TPB) k_zone_finish(ZoneArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int queue[FINISH_TPB];
    int k0 = 0, zi = 0;
    unsigned long long todo = 0ull; /* the same in every wavefront of the block */
    while (next_zone(a, k0, todo, zi)) {
        const int P = a.zsize[zi];
        const long long off = a.zoff[zi];
        if (off + (long long)P > a.scratchCap) continue;
        int zid;
        uint64_t seed;
        const ZoneIn& in = zone_in_of<false>(a, zi, zid, seed);
        zone_finish_body<IT, kLds, FINISH_TPB, 4>(a, in, zid, seed, zi, off, (int)threadIdx.x, lds, FINISH_LDS, queue);
        __syncthreads(); /* the next zone reuses the staging */
    }
}

/* The passes and the correction of a zone in one workgroup of FUSED_TPB
 * threads (four wavefronts of at most 128 registers, so four zones share a
 * CU): all four wavefronts run the set-up, the scheduled dense and sparse
 * phases, then the correction, which takes the zone over in LDS where the
 * passes left it.  A zone's correction starts as soon as its own passes
 * end, beside the passes of slower zones. */
constexpr int FUSED_TPB = 256;
template <class IT, bool kEns = false>
__global__ void __launch_bounds__(FUSED_TPB, 4) k_zone_fused(ZoneArgs a) { /* 4 waves per SIMD: four zones per CU */
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ ZoneShared sh;
    __shared__ int queue[FUSED_TPB];
    const int zi = blockIdx.x;
    if (zi >= a.nZones) return;
    const int P = a.zsize[zi];
    if (P < a.pMin || P > a.pMax) return;
    const long long off = a.zoff[zi];
    if (off + (long long)P > a.scratchCap) {
        if (threadIdx.x == 0) atomicOr(a.err, 2);
        return;
    }
    /* the passes are the critical path: their wavefronts issue ahead of the
     * corrections of zones that are done */
    __builtin_amdgcn_s_setprio(3);
    const int wave = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    ZoneCtxT<IT> z;
    int zid;
    uint64_t seed;
    const ZoneIn& inSrc = zone_in_of<kEns>(a, zi, zid, seed);
    /* an ensemble zone's inputs (its replica's) are staged in LDS, read as
     * cheaply as the kernel arguments */
    const ZoneIn* inp = &inSrc;
    if constexpr (kEns) {
        __shared__ __align__(16) uint32_t sIn[sizeof(ZoneIn) / sizeof(uint32_t)];
        constexpr int W = (int)(sizeof(ZoneIn) / sizeof(uint32_t));
        static_assert(sizeof(ZoneIn) % sizeof(uint32_t) == 0, "ZoneIn is copied as 32-bit words");
        const uint32_t* src = reinterpret_cast<const uint32_t*>(&inSrc);
        for (int k = (int)threadIdx.x; k < W; k += FUSED_TPB) sIn[k] = src[k];
        __syncthreads();
        inp = reinterpret_cast<const ZoneIn*>(sIn);
    }
    const ZoneIn& in = *inp;
    zone_init(z, in, a.rects[zi], zid, seed, &sh);
    z.nxG = (IT*)a.nextG + off;
    z.tmpG = (IT*)a.tmpG + off;
    z.duG = a.duG + (size_t)zi * (size_t)a.duStride;
    z.rec = a.rec + off;
    zone_carve(z, lds, a.ldsBytes);
    static_assert(FUSED_TPB == 256, "zone_setup_mw and the phases take four wavefronts");
    zone_setup_mw(z, in, wave, lane);
    __syncthreads();
    z.P = sh.P;
    z.allOwned = zone_setup_all_owned(z);
    for (int k = (int)threadIdx.x; k < z.P; k += FUSED_TPB) z.rec[k] = (unsigned char)ZM_NONE;
    __syncthreads();
    /* the dense passes and the sparse phase on all four wavefronts */
    zone_dense_passes_mw(z, wave, lane);
    if (threadIdx.x == 0 && sh.schedWait) atomicOr(a.err, 16);
    zone_sparse_pass_mw(z, wave, lane);
    if (threadIdx.x == 0 && sh.schedWait) atomicOr(a.err, 16);
    if (threadIdx.x == 0) {
        ZoneOut o;
        o.nHigh = sh.nHigh;
        o.nLow = sh.nLow;
        o.movedDense = sh.movedDense;
        o.movedSparse = sh.movedSparse;
        o.stats4 = 0.0;
        o.stats5 = 0.0;
        a.out[zi] = o;
    }
    __builtin_amdgcn_s_setprio(0);
    /* the correction takes the zone over where the passes left it in LDS */
    const ZoneHandoff hand{z.P, z.allOwned, z.nL};
    __syncthreads();
    zone_finish_body<IT, true, FUSED_TPB, 3>(a, in, zid, seed, zi, off, (int)threadIdx.x, lds, a.ldsBytes, queue,
                                            &hand);
}

/* velocities of the particles of cells several zones cover, from the side
 * buffer (their input runs in the pass-input binning) */
__global__ void k_overlap_copyback(const int* __restrict__ cells, int ncells, const int* __restrict__ start,
                                   const int* __restrict__ list, const double* __restrict__ vsx,
                                   const double* __restrict__ vsy, double* __restrict__ vx, double* __restrict__ vy) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncells) return;
    const int c = cells[t];
    for (int j = start[c]; j < start[c + 1]; ++j) {
        const int slot = list[j];
        vx[slot] = vsx[slot];
        vy[slot] = vsy[slot];
    }
}

/* a plain closure's grids and applied beta into a snapshot: out = 6 grids of
 * Nc (pdriveRaw, pdrive, duRepairX, duRepairY, duEOSX, duEOSY), then beta */
__global__ void __launch_bounds__(TPB) k_snap_grids(int Nc, const double* __restrict__ a, const double* __restrict__ b,
                                                    const double* __restrict__ c, const double* __restrict__ d,
                                                    const double* __restrict__ e, const double* __restrict__ f,
                                                    const double* __restrict__ scal, double* __restrict__ out) {
    const int i = blockIdx.x * TPB + (int)threadIdx.x;
    if (i < Nc) {
        const size_t C = (size_t)Nc;
        out[i] = a[i];
        out[C + i] = b[i];
        out[2 * C + i] = c[i];
        out[3 * C + i] = d[i];
        out[4 * C + i] = e[i];
        out[5 * C + i] = f[i];
    }
    if (i == 0) out[6 * (size_t)Nc] = scal[3];
}

/* ---- closure (liquid_closure.cpp:340-436) ---- */

struct ClosureArgs {
    int Nc;
    double gamma;
    /* ref / red cell sums */
    const double *refPx, *refPy, *refUx, *refUy;
    const double *redPx, *redPy, *redUx, *redUy;
    double *duRX, *duRY;
    double* part; /* per block: num, den */
    const int* gate; /* nullptr, or the kernel runs only when *gate != 0 (the surrogate's fallback) */
};

/* the surrogate's steps launch the plain closure predicated on the residual
 * check's flag: with the flag clear every block leaves at once */
__device__ inline bool gated_off(const int* gate) { return gate && *gate == 0; }

__global__ void __launch_bounds__(TPB) k_closure_repair(ClosureArgs a) {
    if (gated_off(a.gate)) return;
    __shared__ double s0[TPB], s1[TPB];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    double num = 0.0, den = 0.0;
    if (c < a.Nc) {
        const double ax = (a.refPx[c] - a.redPx[c]) / a.gamma;
        const double ay = (a.refPy[c] - a.redPy[c]) / a.gamma;
        a.duRX[c] = ax;
        a.duRY[c] = ay;
        const double ex = a.redUx[c] - a.refUx[c], ey = a.redUy[c] - a.refUy[c];
        num = -(ex * ax + ey * ay);
        den = ax * ax + ay * ay;
    }
    s0[threadIdx.x] = num;
    s1[threadIdx.x] = den;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < (unsigned)s) {
            s0[threadIdx.x] += s0[threadIdx.x + s];
            s1[threadIdx.x] += s1[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.part[2 * blockIdx.x] = s0[0];
        a.part[2 * blockIdx.x + 1] = s1[0];
    }
}

/* beta from (num, den) (liquid_closure.cpp:364-377): scal[0..1] = num, den;
 * writes scal[2] = betaOpt, scal[3] = betaApplied. */
/* beta from the repair sums scal[0] (num), scal[1] (den): scal[2] the
 * optimum, scal[3] the applied value */
__device__ inline void beta_from_sums(double* scal, int useOptimal, double betaRepair) {
    const double num = scal[0], den = scal[1];
    double bopt = 0.0;
    if (isfinite(num) && isfinite(den) && den > 0.0) bopt = dmax(num / den, 0.0);
    double app;
    if (!useOptimal) app = betaRepair;
    else if (!isfinite(bopt) || !isfinite(den) || den < 1e-20) app = betaRepair;
    else app = (bopt < 0.0) ? 0.0 : ((2.0 < bopt) ? 2.0 : bopt);
    scal[2] = bopt;
    scal[3] = app;
}

/* the repair sums and beta in one block (thread 0 wrote the sums) */
__global__ void __launch_bounds__(TPB) k_reduce_beta(const double* __restrict__ part, int np, double* scal,
                                                     int useOptimal, double betaRepair, const int* gate = nullptr) {
    if (gated_off(gate)) return;
    reduce_parts_block(part, np, 2, scal);
    if (threadIdx.x == 0) beta_from_sums(scal, useOptimal, betaRepair);
}

/* pdrive with the interzone cross smoothing (liquid_closure.cpp:401-419) */
__global__ void k_pdrive(Geom g, const double* __restrict__ raw, const int* __restrict__ mask, double blend,
                         int anyMask, double* __restrict__ pd, const int* gate = nullptr) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.Nc || gated_off(gate)) return;
    double v = raw[c];
    if (anyMask && mask[c] != 0) {
        const int ix = c / g.Ny, iy = c - ix * g.Ny;
        double acc = raw[c];
        int n = 1;
        /* left, right, down, up (liquid_closure.cpp:233-236) */
        const int ddx[4] = {-1, 1, 0, 0}, ddy[4] = {0, 0, -1, 1};
        for (int k = 0; k < 4; ++k) {
            int jx = ix + ddx[k], jy = iy + ddy[k];
            if (g.perx) jx = wrapi(jx, g.Nx);
            else if (jx < 0 || jx >= g.Nx) continue;
            if (g.pery) jy = wrapi(jy, g.Ny);
            else if (jy < 0 || jy >= g.Ny) continue;
            acc += raw[jy + g.Ny * jx];
            ++n;
        }
        const double ps = acc / (double)n;
        v = (1.0 - blend) * raw[c] + blend * ps;
    }
    pd[c] = v;
}

/* duEOS = -(dt/rhoT) * grad(pdrive) * betaEOS (liquid_closure.cpp:243-295, :424-430) */
__global__ void k_eos(Geom g, const double* __restrict__ p, double dtOverRho, double betaEOS, double* __restrict__ ex,
                      double* __restrict__ ey, const int* gate = nullptr) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.Nc || gated_off(gate)) return;
    const int ix = c / g.Ny, iy = c - ix * g.Ny;
    const double dx = g.dxc, dy = g.dyc;
    double gx, gy;
    if (g.perx) gx = (p[iy + g.Ny * wrapi(ix + 1, g.Nx)] - p[iy + g.Ny * wrapi(ix - 1, g.Nx)]) / (2.0 * dx);
    else if (ix == 0) gx = (p[iy + g.Ny * (ix + 1)] - p[c]) / dx;
    else if (ix == g.Nx - 1) gx = (p[c] - p[iy + g.Ny * (ix - 1)]) / dx;
    else gx = (p[iy + g.Ny * (ix + 1)] - p[iy + g.Ny * (ix - 1)]) / (2.0 * dx);
    if (g.pery) gy = (p[wrapi(iy + 1, g.Ny) + g.Ny * ix] - p[wrapi(iy - 1, g.Ny) + g.Ny * ix]) / (2.0 * dy);
    else if (iy == 0) gy = (p[(iy + 1) + g.Ny * ix] - p[c]) / dy;
    else if (iy == g.Ny - 1) gy = (p[c] - p[(iy - 1) + g.Ny * ix]) / dy;
    else gy = (p[(iy + 1) + g.Ny * ix] - p[(iy - 1) + g.Ny * ix]) / (2.0 * dy);
    ex[c] = -dtOverRho * gx * betaEOS;
    ey[c] = -dtOverRho * gy * betaEOS;
}

/* The closure's kick: repair (scaled by beta) then EOS correction. */
__device__ __forceinline__ void kick_velocity(double v0x, double v0y, double kx, double ky, double ex, double ey,
                                              double& nx, double& ny) {
    nx = v0x + kx;
    ny = v0y + ky;
    nx += ex;
    ny += ey;
}

/* The repaired velocities alone, v + beta*duRepair[c] of the particle's
 * (red) cell (stateRepaired, liquid_closure.cpp:379-387), interleaved in
 * particle-id order: the state of slot s is particle pid[s]. */
__global__ void __launch_bounds__(TPB) k_repaired_velocity(Geom g, const double* __restrict__ x,
                                                           const double* __restrict__ y, const double* __restrict__ vx,
                                                           const double* __restrict__ vy, const int* __restrict__ pid,
                                                           const double* __restrict__ duRX,
                                                           const double* __restrict__ duRY,
                                                           const double* __restrict__ scal, double* __restrict__ a) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.n) return;
    const double beta = scal[3];
    const int c = cell_of(g, x[s], y[s]);
    const size_t q = (size_t)pid[s];
    a[2 * q] = vx[s] + beta * duRX[c];
    a[2 * q + 1] = vy[s] + beta * duRY[c];
}

/* v = (v + beta*duRepair[c]) + duEOS[c] for the particle's (red) cell
 * (apply_cell_kick twice, liquid_closure.cpp:379-436), fused with a reorder
 * of the state into the order of the current (red) binning: position s of
 * the list becomes slot s, so the next step starts with particles sorted by
 * cell and every per-cell gather is local.  Optional per-block partials:
 * sum |kR|^2, max |kR|, sum |kE|^2, max |kE|, sum dvx, sum dvy. */
__global__ void __launch_bounds__(TPB) k_kick_reorder(Geom g, const int* __restrict__ list,
                                                      const double* __restrict__ x, const double* __restrict__ y,
                                                      const double* __restrict__ vx, const double* __restrict__ vy,
                                                      const int* __restrict__ pid, double* __restrict__ xo,
                                                      double* __restrict__ yo, double* __restrict__ vxo,
                                                      double* __restrict__ vyo, int* __restrict__ pido,
                                                      const double* __restrict__ duRX, const double* __restrict__ duRY,
                                                      const double* __restrict__ duEX, const double* __restrict__ duEY,
                                                      const double* __restrict__ scal, double* __restrict__ part) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const double beta = scal[3];
    double r2 = 0.0, rm = 0.0, e2 = 0.0, em = 0.0, dvx = 0.0, dvy = 0.0;
    if (s < g.n) {
        const int i = list[s];
        const double px = x[i], py = y[i];
        const int c = cell_of(g, px, py);
        const double kx = beta * duRX[c], ky = beta * duRY[c];
        const double v0x = vx[i], v0y = vy[i];
        const double ex = duEX[c], ey = duEY[c];
        double nx, ny;
        kick_velocity(v0x, v0y, kx, ky, ex, ey, nx, ny);
        xo[s] = px;
        yo[s] = py;
        vxo[s] = nx;
        vyo[s] = ny;
        pido[s] = pid[i];
        if (part) {
            const double kr = MPCD_SQRT(kx * kx + ky * ky);
            const double ke = MPCD_SQRT(ex * ex + ey * ey);
            r2 = kr * kr;
            rm = kr;
            e2 = ke * ke;
            em = ke;
            dvx = nx - v0x;
            dvy = ny - v0y;
        }
    }
    if (!part) return;
    __shared__ double sh[6][TPB];
    sh[0][threadIdx.x] = r2;
    sh[1][threadIdx.x] = rm;
    sh[2][threadIdx.x] = e2;
    sh[3][threadIdx.x] = em;
    sh[4][threadIdx.x] = dvx;
    sh[5][threadIdx.x] = dvy;
    __syncthreads();
    for (int st = TPB / 2; st > 0; st >>= 1) {
        if (threadIdx.x < (unsigned)st) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + st];
            sh[1][threadIdx.x] = dmax(sh[1][threadIdx.x], sh[1][threadIdx.x + st]);
            sh[2][threadIdx.x] += sh[2][threadIdx.x + st];
            sh[3][threadIdx.x] = dmax(sh[3][threadIdx.x], sh[3][threadIdx.x + st]);
            sh[4][threadIdx.x] += sh[4][threadIdx.x + st];
            sh[5][threadIdx.x] += sh[5][threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; ++k) part[6 * (size_t)blockIdx.x + k] = sh[k][0];
}

/* The kick and reorder of one step fused with the next step's force,
 * stream and wall conditions (k_stream_bc's work on the same values, in
 * registers): the state is written once, already streamed, with both cell
 * ids of the next step and their histograms (cntS, cntU, cleared before).
 * Used between steps of one call that report no metrics. */
__global__ void __launch_bounds__(TPB) k_kick_reorder_stream(
    Geom g, BCParams b, uint64_t seed, const int* __restrict__ list, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ vx, const double* __restrict__ vy,
    const int* __restrict__ pid, double* __restrict__ xo, double* __restrict__ yo, double* __restrict__ vxo,
    double* __restrict__ vyo, int* __restrict__ pido, const double* __restrict__ duRX,
    const double* __restrict__ duRY, const double* __restrict__ duEX, const double* __restrict__ duEY,
    const double* __restrict__ scal, int* __restrict__ cidS, int* __restrict__ cidU, double sx, double sy,
    double xmax, double ymax, int* __restrict__ cntS, int* __restrict__ cntU, int* __restrict__ err) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.n) return;
    const int lane = __lane_id();
    const double beta = scal[3];
    const int i = list[s];
    double px = x[i], py = y[i];
    const int c = cell_of(g, px, py);
    const double kx = beta * duRX[c], ky = beta * duRY[c];
    const double ex = duEX[c], ey = duEY[c];
    double nx, ny;
    kick_velocity(vx[i], vy[i], kx, ky, ex, ey, nx, ny);
    const int pi = pid[i];
    WallAcc acc;
    wall_acc_zero(acc);
    const int rc = bc_particle(g, b, seed, (uint32_t)pi, 1, px, py, nx, ny, acc);
    if (rc != 0) atomicOr(err, 1);
    xo[s] = px;
    yo[s] = py;
    vxo[s] = nx;
    vyo[s] = ny;
    pido[s] = pi;
    const int cs = shifted_cell_of(g, px, py, sx, sy, xmax, ymax);
    const int cu = cell_of(g, px, py);
    cidS[s] = cs;
    cidU[s] = cu;
    /* both histograms of the next step's binnings, as k_hist counts them */
    int leader, rank, size;
    wave_groups(cs, true, lane, leader, rank, size);
    if (lane == leader) atomicAdd(&cntS[cs], size);
    wave_groups(cu, true, lane, leader, rank, size);
    if (lane == leader) atomicAdd(&cntU[cu], size);
}

/* per-cell closure metrics partials: sum hypot(beta*duR), sum pdriveRaw, sum pdrive */
/* ---- closure surrogate ---- */

/* The surrogate's input features of each cell: occupancy deviation
 * (N - gamma)/gamma, the cell velocity's deviation from the mean of its 3x3
 * neighbourhood (periodic axes wrap, walls clip; the local mean flow is
 * removed, so the features do not depend on the flow's speed, only on its
 * local structure: a uniform or linearly sheared flow gives zero), and
 * T/kBT - 1 (0 where the temperature is undefined, N < 2). */
__global__ void k_features(Geom g, const int* __restrict__ N, const double* __restrict__ Ux,
                           const double* __restrict__ Uy, const double* __restrict__ T, double gamma, double kBT,
                           float* __restrict__ feat) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.Nc) return;
    const int ix = c / g.Ny, iy = c - ix * g.Ny;
    double sx = 0.0, sy = 0.0;
    int m = 0;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy) {
            int jx = ix + dx, jy = iy + dy;
            if (g.perx) jx = (jx + g.Nx) % g.Nx;
            else if (jx < 0 || jx >= g.Nx) continue;
            if (g.pery) jy = (jy + g.Ny) % g.Ny;
            else if (jy < 0 || jy >= g.Ny) continue;
            const int q = jy + g.Ny * jx;
            sx += Ux[q];
            sy += Uy[q];
            ++m;
        }
    const double t = T[c];
    feat[4 * (size_t)c] = (float)(((double)N[c] - gamma) / gamma);
    feat[4 * (size_t)c + 1] = (float)(Ux[c] - sx / (double)m);
    feat[4 * (size_t)c + 2] = (float)(Uy[c] - sy / (double)m);
    feat[4 * (size_t)c + 3] = (N[c] >= 2 && isfinite(t)) ? (float)(t / kBT - 1.0) : 0.0f;
}

/* the plain closure's applied correction per cell, beta duRepair + duEOS
 * (the surrogate's training target, scaled by 1/outScale) */
__global__ void k_closure_target(int Nc, const double* __restrict__ scal, const double* __restrict__ duRX,
                                 const double* __restrict__ duRY, const double* __restrict__ duEX,
                                 const double* __restrict__ duEY, double invScale, float* __restrict__ tgt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= Nc) return;
    const double beta = scal[3];
    tgt[2 * (size_t)c] = (float)((beta * duRX[c] + duEX[c]) * invScale);
    tgt[2 * (size_t)c + 1] = (float)((beta * duRY[c] + duEY[c]) * invScale);
}

/* the surrogate's correction as the kick's grids (duEOS = prediction,
 * duRepair = 0, beta = 0); flag = 1 when a value is not finite or exceeds
 * the cap (the residual check that sends the step to the plain closure);
 * per-block partials of the momentum it would add, sum of N du */
__global__ void __launch_bounds__(TPB) k_learned_du(int Nc, const float* __restrict__ pred, const int* __restrict__ N,
                                                    double scale, double cap, double* __restrict__ duRX,
                                                    double* __restrict__ duRY, double* __restrict__ duEX,
                                                    double* __restrict__ duEY, double* __restrict__ scal,
                                                    int* __restrict__ flag, double* __restrict__ part) {
    __shared__ double sx[TPB], sy[TPB];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c == 0) scal[3] = 0.0;
    double mx = 0.0, my = 0.0;
    if (c < Nc) {
        const double ux = (double)pred[2 * (size_t)c] * scale, uy = (double)pred[2 * (size_t)c + 1] * scale;
        if (!(fabs(ux) <= cap) || !(fabs(uy) <= cap)) atomicOr(flag, 1);
        duRX[c] = 0.0;
        duRY[c] = 0.0;
        duEX[c] = ux;
        duEY[c] = uy;
        mx = (double)N[c] * ux;
        my = (double)N[c] * uy;
    }
    sx[threadIdx.x] = mx;
    sy[threadIdx.x] = my;
    __syncthreads();
    for (int st = TPB / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            sx[threadIdx.x] += sx[threadIdx.x + st];
            sy[threadIdx.x] += sy[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * (size_t)blockIdx.x] = sx[0];
        part[2 * (size_t)blockIdx.x + 1] = sy[0];
    }
}

/* removes the predicted correction's net momentum: du -= (sum N du) / n,
 * so the surrogate's kick, like a closure kick, leaves the total momentum
 * to the physics (sums in sums[0..1]) */
__global__ void k_learned_center(int Nc, const double* __restrict__ sums, double invN, double* __restrict__ duEX,
                                 double* __restrict__ duEY) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= Nc) return;
    duEX[c] -= sums[0] * invN;
    duEY[c] -= sums[1] * invN;
}

/* Removes the predicted correction's momentum line by line along a
 * periodic axis: with axis 0 (x periodic) each cell row iy gets
 * du_x -= (sum over the row of N du_x) / (particles of the row), with axis
 * 1 (y periodic) each column gets the same for du_y.  The plain closure's
 * pressure kick, -grad P along a periodic axis, moves no momentum from one
 * line to another on average; a learned correction constrained the same way
 * cannot carry flow momentum across the channel (which would act as a
 * spurious viscosity on the flow profile).  One block per line. */
__global__ void __launch_bounds__(TPB) k_learned_center_lines(Geom g, int axis, const int* __restrict__ N,
                                                              double* __restrict__ du) {
    __shared__ double s0[TPB], s1[TPB];
    const int line = blockIdx.x;
    const int len = axis == 0 ? g.Nx : g.Ny;
    auto cell = [&](int k) { return axis == 0 ? line + g.Ny * k : k + g.Ny * line; };
    double a = 0.0, m = 0.0;
    for (int k = threadIdx.x; k < len; k += TPB) {
        const int c = cell(k);
        a += (double)N[c] * du[c];
        m += (double)N[c];
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = m;
    __syncthreads();
    for (int st = TPB / 2; st > 0; st >>= 1) {
        if (threadIdx.x < (unsigned)st) {
            s0[threadIdx.x] += s0[threadIdx.x + st];
            s1[threadIdx.x] += s1[threadIdx.x + st];
        }
        __syncthreads();
    }
    const double shift = s1[0] > 0.0 ? s0[0] / s1[0] : 0.0;
    for (int k = threadIdx.x; k < len; k += TPB) du[cell(k)] -= shift;
}

/* Divergence residual of a closure's correction: per cell the central
 * (one-sided at walls) divergence of the corrected cell velocity
 * U + beta duRepair + duEOS; per-block partial sums of its square. */
__global__ void __launch_bounds__(TPB) k_div_residual(Geom g, const double* __restrict__ Ux,
                                                      const double* __restrict__ Uy, const double* __restrict__ duRX,
                                                      const double* __restrict__ duRY, const double* __restrict__ duEX,
                                                      const double* __restrict__ duEY, const double* __restrict__ scal,
                                                      double* __restrict__ part) {
    __shared__ double sh[TPB];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    double d2 = 0.0;
    if (c < g.Nc) {
        const double beta = scal[3];
        auto ux = [&](int q) { return Ux[q] + beta * duRX[q] + duEX[q]; };
        auto uy = [&](int q) { return Uy[q] + beta * duRY[q] + duEY[q]; };
        const int ix = c / g.Ny, iy = c - ix * g.Ny;
        double gx, gy;
        if (g.perx) {
            const int a = iy + g.Ny * ((ix + 1) % g.Nx), b = iy + g.Ny * ((ix - 1 + g.Nx) % g.Nx);
            gx = (ux(a) - ux(b)) / (2.0 * g.dxc);
        } else if (g.Nx < 2) {
            gx = 0.0;
        } else {
            const int a = iy + g.Ny * min(ix + 1, g.Nx - 1), b = iy + g.Ny * max(ix - 1, 0);
            gx = (ux(a) - ux(b)) / ((double)(min(ix + 1, g.Nx - 1) - max(ix - 1, 0)) * g.dxc);
        }
        if (g.pery) {
            const int a = (iy + 1) % g.Ny + g.Ny * ix, b = (iy - 1 + g.Ny) % g.Ny + g.Ny * ix;
            gy = (uy(a) - uy(b)) / (2.0 * g.dyc);
        } else if (g.Ny < 2) {
            gy = 0.0;
        } else {
            const int a = min(iy + 1, g.Ny - 1) + g.Ny * ix, b = max(iy - 1, 0) + g.Ny * ix;
            gy = (uy(a) - uy(b)) / ((double)(min(iy + 1, g.Ny - 1) - max(iy - 1, 0)) * g.dyc);
        }
        d2 = (gx + gy) * (gx + gy);
    }
    sh[threadIdx.x] = d2;
    __syncthreads();
    for (int st = TPB / 2; st > 0; st >>= 1) {
        if (threadIdx.x < (unsigned)st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

/* slots of the surrogate's device state */
enum { SUR_OCC_REF = 0, SUR_DIV_REF = 1, SUR_HAVE_REF = 2, SUR_OCC = 3, SUR_DIV = 4, SUR_N_LEARNED = 5,
       SUR_N_FALLBACK = 6, SUR_SLOTS = 8 };

/* The residual check of a surrogate step, on the device (one thread).
 * res[0] = occStd of the red state's counts, res[1] = sum of the squared
 * divergence of its corrected velocity.  A plain step (plain = 1) records
 * them as the reference; a learned step fails when its prediction was not
 * finite or beyond the cap (flag already 1), when the red state's
 * occupancy spread exceeds occTol times the reference's, or when the
 * divergence residual of the predicted correction exceeds divTol times the
 * plain closure's; the flag then sends the step to the plain closure (its
 * kernels run predicated on it). */
__global__ void k_sur_decide(const double* __restrict__ occ, const double* __restrict__ div, double* __restrict__ sur,
                             int* __restrict__ flag, int plain, double occTol, double divTol) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double o = occ[0], d = sqrt(div[0]);
    sur[SUR_OCC] = o;
    sur[SUR_DIV] = d;
    if (plain) {
        sur[SUR_OCC_REF] = o;
        sur[SUR_DIV_REF] = d;
        sur[SUR_HAVE_REF] = 1.0;
        return;
    }
    bool bad = *flag != 0;
    if (sur[SUR_HAVE_REF] != 0.0) {
        if (o > occTol * sur[SUR_OCC_REF]) bad = true;
        if (d > divTol * sur[SUR_DIV_REF]) bad = true;
    }
    *flag = bad ? 1 : 0;
    sur[bad ? SUR_N_FALLBACK : SUR_N_LEARNED] += 1.0;
}

/* Row profiles accumulated over steps: sum of vx, vx^2 + vy^2 and the
 * particle count of each cell row (cell iy of the particle's y).  One block
 * strides over the particles with the rows in LDS. */
__global__ void __launch_bounds__(TPB) k_row_profile(Geom g, const double* __restrict__ y, const double* __restrict__ vx,
                                                     const double* __restrict__ vy, double* __restrict__ acc) {
    extern __shared__ double rows[]; /* 3 * Ny */
    for (int r = threadIdx.x; r < 3 * g.Ny; r += TPB) rows[r] = 0.0;
    __syncthreads();
    for (int i = blockIdx.x * TPB + (int)threadIdx.x; i < g.n; i += gridDim.x * TPB) {
        int iy = (int)floor(y[i] / g.dyc);
        iy = iy < 0 ? 0 : (iy >= g.Ny ? g.Ny - 1 : iy);
        const double a = vx[i], b = vy[i];
        atomicAdd(&rows[iy], a);
        atomicAdd(&rows[g.Ny + iy], a * a + b * b);
        atomicAdd(&rows[2 * g.Ny + iy], 1.0);
    }
    __syncthreads();
    for (int r = threadIdx.x; r < 3 * g.Ny; r += TPB)
        if (rows[r] != 0.0) atomicAdd(&acc[r], rows[r]);
}

__global__ void __launch_bounds__(TPB) k_closure_cell_metrics(int Nc, const double* __restrict__ duRX,
                                                              const double* __restrict__ duRY,
                                                              const double* __restrict__ raw,
                                                              const double* __restrict__ pd,
                                                              const double* __restrict__ scal, double* __restrict__ part) {
    __shared__ double s[3][TPB];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const double beta = scal[3];
    double a = 0.0, b = 0.0, d = 0.0;
    if (c < Nc) {
        a = mpcd_hypot(beta * duRX[c], beta * duRY[c]);
        b = raw[c];
        d = pd[c];
    }
    s[0][threadIdx.x] = a;
    s[1][threadIdx.x] = b;
    s[2][threadIdx.x] = d;
    __syncthreads();
    for (int st = TPB / 2; st > 0; st >>= 1) {
        if (threadIdx.x < (unsigned)st)
            for (int k = 0; k < 3; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) part[3 * (size_t)blockIdx.x + k] = s[k][0];
}

/* ===================================================================== */
/*                        host-side geometry                             */
/* ===================================================================== */

/* ===================================================================== */
/*                               engine                                  */
/* ===================================================================== */

struct Binning {
    int *cnt = nullptr, *start = nullptr, *cursor = nullptr, *list = nullptr;
    unsigned long long* key = nullptr; /* (pid << 32 | slot) in (cell, pid) order */
    /* the incremental update's second buffers (swapped with the above),
     * per-cell incoming counts and per-slot rank among a cell's incoming */
    int *startB = nullptr, *listB = nullptr, *inc = nullptr, *rnk = nullptr;
    unsigned long long* keyB = nullptr;
    int* cntS = nullptr; /* counts of the shifted collision cells */
    int* dout = nullptr; /* departures per cell recorded by a zone pass (zero between updates) */
};

}  // namespace

struct mpcd_gpu_sim {
    mpcd_params p{};
    Geom g{};
    int n = 0, Nc = 0;
    hipStream_t stream = nullptr;
    double *x = nullptr, *y = nullptr, *vx = nullptr, *vy = nullptr;
    double *vx2 = nullptr, *vy2 = nullptr; /* zone-pass output velocities (swapped after each pass) */
    double *x2 = nullptr, *y2 = nullptr;   /* reorder targets */
    int* pid = nullptr;
    int* pid2 = nullptr;
    int* cid = nullptr;  /* unshifted cell of each slot (key of the current binning) */
    int* cidS = nullptr; /* shifted collision cell of each slot */
    Binning bin;
    int* tileSum = nullptr;
    int nTiles = 0;
    int *next = nullptr, *tmpg = nullptr;
    long long scratchCap = 0;
    int *zsize = nullptr, *zoff = nullptr;
    double* stage = nullptr; /* 2n doubles, host<->device transfers */
    double *refPx = nullptr, *refPy = nullptr, *refUx = nullptr, *refUy = nullptr, *pdriveRaw = nullptr;
    double *redPx = nullptr, *redPy = nullptr, *redUx = nullptr, *redUy = nullptr;
    double *duRX = nullptr, *duRY = nullptr, *duEX = nullptr, *duEY = nullptr, *pdrive = nullptr;
    double* cellTmp = nullptr; /* Nc doubles for Pkin scratch */
    int* mask = nullptr;
    int anyMask = 0, nMasked = 0;
    double* part = nullptr;
    size_t partCap = 0;
    double* scal = nullptr; /* device scalars */
    /* closure surrogate */
    mpcd_gpu_net* sur = nullptr;
    int surK = 0;
    double surScale = 1.0, surCap = 0.0;
    bool wantSample = false;
    /* added to every base seed of a step (0: the benchmark driver's seeds);
     * a replica of an ensemble with seed key k uses k * SEED_KEY_MIX */
    uint64_t seedOff = 0;
    /* closure snapshots (training data in the reference's formats): the
     * reference and red CellFields (N, then 9 fields each), the six grids
     * and beta of the last plain closure since they were enabled */
    bool wantSnap = false;
    int* snapN = nullptr;     /* 2 Nc: ref, red */
    double* snapF = nullptr;  /* 9 Nc ref, 9 Nc red, 6 Nc grids, beta */
    int snapStep = -1, curStep = 0;
    int* redN = nullptr;
    double* redT = nullptr;
    int* refN = nullptr;    /* the reference state's counts and temperatures (the surrogate's features) */
    double* refT = nullptr;
    float *featBuf = nullptr, *tgtBuf = nullptr;
    int* surFlag = nullptr;
    double* surDev = nullptr; /* SUR_* slots */
    double surOccTol = 1.1, surDivTol = 2.0;
    int nPlain = 0;
    double* rowAcc = nullptr; /* 3 x Ny row profiles, accumulated while rowSteps >= 0 */
    int rowSteps = -1;
    int* err = nullptr;
    /* diagnostics: device results of the last diagnosed step, and the cell
     * fields of the current state (allocated on first use) */
    bool wantDiag = false;
    double* diagDev = nullptr; /* DIAG_* slots */
    /* a single-zone call's diagnostics: the zone kernel's per-cell record
     * and sparse-phase start counts (non-null only during such a call) */
    double* zdCell = nullptr;
    int* zdCnt = nullptr;
    int* outN = nullptr;
    double* outF = nullptr; /* 9 x Nc: Ux, Uy, Px, Py, T, rho, Pkin, Pvir, P */
    WallAcc* wallPart = nullptr;
    ZoneRect* rects[2] = {nullptr, nullptr};
    int* owner[2] = {nullptr, nullptr};
    unsigned char* cover[2] = {nullptr, nullptr}; /* per cell: several zones cover it (nullptr: none do) */
    int* overCells[2] = {nullptr, nullptr};
    int nOver[2] = {0, 0};
    int nz[2] = {0, 0};
    int maxCells[2] = {0, 0}, maxCols[2] = {0, 0};
    ZoneOut* zout = nullptr;
    int zoutCap = 0;
    double* duG = nullptr;
    unsigned char* recG = nullptr; /* move records, 1 byte per particle slot of the zone regions */
    unsigned char* zstate = nullptr; /* per-zone pass state for the correction kernel */
    size_t zstateCap = 0;
    size_t duCap = 0;
    unsigned char* zscratch = nullptr;
    size_t zscratchCap = 0;
    ZoneParams zp{};
    CollideParams cp{};
    BCParams bc{};
    double xmax = 0, ymax = 0;
    int fuseNext = -1; /* step whose stream the current kick performs (-1: none) */
    bool recDirty = false; /* a recording pass was launched and its update has not run */
    int streamed = -1; /* step whose stream has been performed by a kick (-1: none) */
    int cellCap = 0; /* LDS entries of a per-cell block (k_collide, k_cell_fields) */

    /* per-class kernel timing with events on this stream (MPCD_PROF_*) */
    bool prof = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[MPCD_PROF_CLASSES];
    std::vector<hipEvent_t> evPool;
    size_t evUsed = 0;
    hipEvent_t take_event() {
        if (evUsed == evPool.size()) {
            hipEvent_t e;
            HIP_OK(hipEventCreate(&e));
            evPool.push_back(e);
        }
        return evPool[evUsed++];
    }
    void pbeg(int c) {
        if (!prof) return;
        hipEvent_t a = take_event();
        HIP_OK(hipEventRecord(a, stream));
        ev[c].push_back({a, nullptr});
    }
    void pend(int c) {
        if (!prof) return;
        hipEvent_t b = take_event();
        HIP_OK(hipEventRecord(b, stream));
        ev[c].back().second = b;
    }
    /* phase boundaries of each profiled step: start, after the reference
     * build (stream, collision, reference fields), after the base pass and
     * its binning update, after the shifted pass and its update, after the
     * closure (the closure phase holds the next step's stream when the kick
     * performs it) */
    std::vector<std::array<hipEvent_t, 5>> phaseEv;
    bool phases = false;
    double phaseAcc[4] = {0, 0, 0, 0};
    void phase_mark(int k) {
        if (!phases) return;
        if (k == 0) phaseEv.push_back({});
        hipEvent_t e = take_event();
        HIP_OK(hipEventRecord(e, stream));
        phaseEv.back()[(size_t)k] = e;
    }
    void prof_reset() {
        for (auto& v : ev) v.clear();
        phaseEv.clear();
        for (double& a : phaseAcc) a = 0.0;
        evUsed = 0;
    }
    /* totals since the last reset; the events read are returned to the pool
     * when no kernel-class timing holds events */
    void phase_read(double* ms4) {
        HIP_OK(hipStreamSynchronize(stream));
        for (const auto& a : phaseEv)
            for (int k = 0; k < 4; ++k) {
                float f = 0.f;
                HIP_OK(hipEventElapsedTime(&f, a[(size_t)k], a[(size_t)k + 1]));
                phaseAcc[k] += (double)f;
            }
        phaseEv.clear();
        if (!prof) evUsed = 0;
        for (int k = 0; k < 4; ++k) ms4[k] = phaseAcc[k];
    }
    void prof_read(double* ms, int* count) {
        HIP_OK(hipStreamSynchronize(stream));
        for (int c = 0; c < MPCD_PROF_CLASSES; ++c) {
            double t = 0.0;
            for (const auto& pr : ev[c]) {
                float f = 0.f;
                HIP_OK(hipEventElapsedTime(&f, pr.first, pr.second));
                t += (double)f;
            }
            ms[c] = t;
            count[c] = (int)ev[c].size();
        }
    }

    explicit mpcd_gpu_sim(const mpcd_params& prm) : p(prm) {
        validate(p);
        if (p.n == 0) throw ApiError(MPCD_E_INVALID, "no particles: the GPU path needs n >= 1");
        g = make_geom(p);
        n = p.n;
        Nc = p.Nc;
        HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        x = dalloc<double>(n);
        y = dalloc<double>(n);
        vx = dalloc<double>(n);
        vy = dalloc<double>(n);
        vx2 = dalloc<double>(n);
        vy2 = dalloc<double>(n);
        x2 = dalloc<double>(n);
        y2 = dalloc<double>(n);
        pid = dalloc<int>(n);
        pid2 = dalloc<int>(n);
        cid = dalloc<int>(n);
        cidS = dalloc<int>(n);
        bin.cnt = dalloc<int>((size_t)Nc + 1);
        bin.start = dalloc<int>((size_t)Nc + 1);
        bin.cursor = dalloc<int>((size_t)Nc + 1);
        bin.list = dalloc<int>(n);
        bin.key = dalloc<unsigned long long>(n);
        bin.startB = dalloc<int>((size_t)Nc + 1);
        bin.listB = dalloc<int>(n);
        bin.inc = dalloc<int>((size_t)Nc + 1);
        bin.rnk = dalloc<int>(n);
        bin.keyB = dalloc<unsigned long long>(n);
        bin.cntS = dalloc<int>((size_t)Nc + 1);
        bin.dout = dalloc<int>((size_t)Nc + 1);
        nTiles = (int)((Nc + 1 + SCAN_TILE - 1) / SCAN_TILE);
        tileSum = dalloc<int>(nTiles);
        stage = dalloc<double>(2 * (size_t)n);
        double** cells[] = {&refPx, &refPy, &refUx, &refUy, &pdriveRaw, &redPx, &redPy, &redUx, &redUy,
                            &duRX, &duRY, &duEX, &duEY, &pdrive, &cellTmp};
        for (double** c : cells) *c = dalloc<double>(Nc);
        mask = dalloc<int>(Nc);
        redN = dalloc<int>(Nc);
        redT = dalloc<double>(Nc);
        refN = dalloc<int>(Nc);
        refT = dalloc<double>(Nc);
        featBuf = dalloc<float>(4 * (size_t)Nc);
        tgtBuf = dalloc<float>(2 * (size_t)Nc);
        surFlag = dalloc<int>(1);
        surDev = dalloc<double>(SUR_SLOTS);
        HIP_OK(hipMemset(surDev, 0, SUR_SLOTS * sizeof(double)));
        scal = dalloc<double>(64);
        diagDev = dalloc<double>(DIAG_SLOTS);
        err = dalloc<int>(1);
        ensure_parts(std::max(std::max((size_t)nblk(n) * 6, (size_t)nblk(Nc) * 3), std::min<size_t>(nblk(Nc), 512) * OCC_W) + 64);
        wallPart = dalloc<WallAcc>(nblk(n));
        hipLaunchKernelGGL(k_iota, dim3(nblk(n)), dim3(TPB), 0, stream, pid, n);
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, stream, err, 1, 0); /* no error flagged yet */
        hipLaunchKernelGGL(k_fill_int, dim3(nblk(Nc + 1)), dim3(TPB), 0, stream, bin.inc, Nc + 1, 0);
        hipLaunchKernelGGL(k_fill_int, dim3(nblk(Nc + 1)), dim3(TPB), 0, stream, bin.dout, Nc + 1, 0);
        hipLaunchKernelGGL(k_fill_int, dim3(nblk(n)), dim3(TPB), 0, stream, bin.rnk, n, -1);
        check_launch();
        /* constants */
        bc = make_bc_params(p);
        xmax = std::nextafter(p.Lx, 0.0);
        ymax = std::nextafter(p.Ly, 0.0);
        cp = make_collide_params(p);
        zp = make_zone_params(p);
        {
            /* 1.15x the mean particles of a block plus slack, at most 64 KB: the
             * LDS per block sets how many blocks a CU holds; a block above
             * the cap takes the global-memory path */
            const double mean = (double)n / (double)std::max(Nc, 1) * CELL_BLK;
            cellCap = (int)std::min(4096.0, std::max(256.0, 1.15 * mean + 2.0 * CELL_BLK));
        }
        /* layouts (zones of the shifted layout overlap: a cell is covered by
         * up to four zones, so per-zone particle scratch is sized by the
         * largest coverage) and the interzone mask */
        int coverMax = 1;
        for (int L = 0; L < 2; ++L) {
            std::vector<int32_t> own;
            const auto r = layout_rects(p, L == 0 ? MPCD_LAYOUT_BASE : MPCD_LAYOUT_SHIFTED, &own);
            nz[L] = (int)r.size();
            std::vector<int> cov((size_t)Nc, 0);
            for (const auto& q : r) {
                maxCells[L] = std::max(maxCells[L], (q.xe - q.xs) * (q.ye - q.ys));
                maxCols[L] = std::max(maxCols[L], q.xe - q.xs);
                for (int ix = q.xs; ix < q.xe; ++ix)
                    for (int iy = q.ys; iy < q.ye; ++iy) coverMax = std::max(coverMax, ++cov[(size_t)(iy + p.Ny * ix)]);
            }
            std::vector<unsigned char> ov((size_t)Nc, 0);
            std::vector<int> oc;
            for (int c = 0; c < Nc; ++c)
                if (cov[(size_t)c] > 1) {
                    ov[(size_t)c] = 1;
                    oc.push_back(c);
                }
            nOver[L] = (int)oc.size();
            if (!oc.empty()) {
                cover[L] = dalloc<unsigned char>((size_t)Nc);
                HIP_OK(hipMemcpy(cover[L], ov.data(), ov.size(), hipMemcpyHostToDevice));
                overCells[L] = dalloc<int>(oc.size());
                HIP_OK(hipMemcpy(overCells[L], oc.data(), oc.size() * sizeof(int), hipMemcpyHostToDevice));
            }
            rects[L] = dalloc<ZoneRect>(r.size());
            HIP_OK(hipMemcpy(rects[L], r.data(), r.size() * sizeof(ZoneRect), hipMemcpyHostToDevice));
            owner[L] = dalloc<int>((size_t)Nc);
            HIP_OK(hipMemcpy(owner[L], own.data(), own.size() * sizeof(int), hipMemcpyHostToDevice));
        }
        scratchCap = (long long)coverMax * (long long)n;
        next = dalloc<int>((size_t)scratchCap + 1);
        tmpg = dalloc<int>((size_t)scratchCap + 1);
        recG = dalloc<unsigned char>((size_t)scratchCap + 1);
        zsize = dalloc<int>((size_t)std::max(nz[0], nz[1]) + 2);
        zoff = dalloc<int>((size_t)std::max(nz[0], nz[1]) + 2);
        ensure_zout(std::max(nz[0], nz[1]));
        ensure_du(2 * (size_t)std::max((size_t)nz[0] * (size_t)maxCells[0], (size_t)nz[1] * (size_t)maxCells[1]));
        const auto m = interzone_mask_host(p);
        nMasked = (int)std::count(m.begin(), m.end(), 1);
        anyMask = nMasked > 0;
        HIP_OK(hipMemcpy(mask, m.data(), m.size() * sizeof(int), hipMemcpyHostToDevice));
        HIP_OK(hipStreamSynchronize(stream));
    }

    ~mpcd_gpu_sim() {
        void* ptrs[] = {x, y, vx, vy, vx2, vy2, x2, y2, pid, pid2, cid, cidS, bin.cnt, bin.start, bin.cursor, bin.list, bin.key, bin.startB, bin.listB, bin.inc, bin.rnk, bin.keyB, bin.cntS, bin.dout, tileSum, next,
                        tmpg, zsize, zoff, owner[0], owner[1], cover[0], cover[1], overCells[0], overCells[1], stage,
                        refPx, refPy, refUx, refUy, pdriveRaw, redPx, redPy, redUx, redUy, duRX, duRY, duEX, duEY,
                        pdrive, cellTmp, mask, part, scal, err, wallPart, rects[0], rects[1], zout, zscratch, duG, recG, zstate,
                        redN, redT, featBuf, tgtBuf, surFlag, diagDev, outN, outF, surDev, rowAcc, refN, refT, snapN, snapF};
        for (void* q : ptrs)
            if (q) (void)hipFree(q);
        for (hipEvent_t e : evPool) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void ensure_parts(size_t m) {
        if (m <= partCap) return;
        if (part) HIP_OK(hipFree(part));
        part = dalloc<double>(m);
        partCap = m;
    }
    void ensure_du(size_t m) {
        if (m <= duCap) return;
        if (duG) HIP_OK(hipFree(duG));
        duG = dalloc<double>(m);
        duCap = m;
    }
    void ensure_zstate(size_t m) {
        if (m <= zstateCap) return;
        if (zstate) HIP_OK(hipFree(zstate));
        zstate = dalloc<unsigned char>(m);
        zstateCap = m;
    }
    void ensure_zout(int m) {
        if (m <= zoutCap) return;
        if (zout) HIP_OK(hipFree(zout));
        zout = dalloc<ZoneOut>(m);
        zoutCap = m;
    }

    void check_launch() { HIP_OK(hipGetLastError()); }

    /* ---- transfers ---- */
    void upload_xy(const double* h, double* dx_, double* dy_) {
        HIP_OK(hipMemcpyAsync(stage, h, 2 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(k_deinterleave, dim3(nblk(n)), dim3(TPB), 0, stream, stage, n, dx_, dy_);
        check_launch();
    }
    void download_xy(const double* dx_, const double* dy_, double* h) {
        hipLaunchKernelGGL(k_interleave, dim3(nblk(n)), dim3(TPB), 0, stream, dx_, dy_, pid, n, stage);
        check_launch();
        HIP_OK(hipMemcpyAsync(h, stage, 2 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
    }
    /* Slots are reset to particle-id order (pid = iota).  An array the
     * caller does not upload (null) is kept: it is first scattered from the
     * current slot order into particle-id order, so every particle keeps its
     * own position or velocity. */
    void set_state(const double* hx, const double* hv) {
        if (!hx || !hv) {
            double* kx = hx ? vx : x;
            double* ky = hx ? vy : y;
            double* tx = hx ? vx2 : x2;
            double* ty = hx ? vy2 : y2;
            hipLaunchKernelGGL(k_scatter_pid, dim3(nblk(n)), dim3(TPB), 0, stream, kx, ky, pid, n, tx, ty);
            check_launch();
            if (hx) {
                std::swap(vx, vx2);
                std::swap(vy, vy2);
            } else {
                std::swap(x, x2);
                std::swap(y, y2);
            }
        }
        hipLaunchKernelGGL(k_iota, dim3(nblk(n)), dim3(TPB), 0, stream, pid, n);
        check_launch();
        if (hx) upload_xy(hx, x, y);
        if (hv) upload_xy(hv, vx, vy);
        streamed = -1;
        HIP_OK(hipStreamSynchronize(stream));
    }
    void get_state(double* hx, double* hv) {
        if (hx) download_xy(x, y, hx);
        if (hv) download_xy(vx, vy, hv);
    }

    /* ---- binning: cnt/start/list of the cell ids in `key` ---- */
    void bin_cells(const int* key) { bin_cells(key, bin.cnt, false, true); }
    /* cnt: the counts (the unshifted binning keeps them in bin.cnt);
     * counted: cnt already holds the histogram of key; keepKeys: the sorted
     * keys are written back (the zone passes and the update read them) */
    void bin_cells(const int* key, int* cnt, bool counted, bool keepKeys) {
        /* fills by kernel, not memset, so the sequence can be graph-captured */
        if (!counted) {
            hipLaunchKernelGGL(k_fill_int, dim3(nblk(Nc + 1)), dim3(TPB), 0, stream, cnt, Nc + 1, 0);
            hipLaunchKernelGGL(k_hist, dim3(bin_blocks(n)), dim3(TPB), 0, stream, key, n, cnt);
        }
        /* the scan clears the scatter's cursors */
        hipLaunchKernelGGL(k_scan_tiles, dim3(nTiles), dim3(TPB), 0, stream, cnt, Nc + 1, bin.start, tileSum,
                           (const int*)nullptr, bin.cursor);
        hipLaunchKernelGGL(k_scan_add_prefix, dim3(nblk(Nc + 1)), dim3(TPB), 0, stream, bin.start, Nc + 1, tileSum);
        hipLaunchKernelGGL(k_scatter, dim3(bin_blocks(n)), dim3(TPB), 0, stream, key, pid, n, bin.start, bin.cursor,
                           bin.key);
        hipLaunchKernelGGL(k_segsort, dim3(nblk(Nc, SEG_CELLS)), dim3(SEG_CELLS), 0, stream, bin.start, Nc, bin.key,
                           bin.list, (int)(n < (1 << 27)), (int)keepKeys);
        check_launch();
    }
    void bin_unshifted() {
        hipLaunchKernelGGL(k_cells, dim3(nblk(n)), dim3(TPB), 0, stream, g, x, y, cid);
        check_launch();
        bin_cells(cid);
    }
    /* The unshifted binning after a zone pass that recorded its cell changes
     * (bin.rnk -1 for every slot before it, then each moved particle's new cell in
     * cid, its departure in dout, its rank among its new cell's arrivals):
     * the runs are the old runs' stayers merged with the arrivals, in pid
     * order, the same binning a counting sort of the new cells gives. */
    void rebin_recorded() {
        hipLaunchKernelGGL(k_scan_tiles, dim3(nTiles), dim3(TPB), 0, stream, bin.cnt, Nc + 1, bin.startB, tileSum,
                           bin.inc, (int*)nullptr, bin.dout);
        hipLaunchKernelGGL(k_scan_add_prefix, dim3(nblk(Nc + 1)), dim3(TPB), 0, stream, bin.startB, Nc + 1, tileSum);
        hipLaunchKernelGGL(k_place_arrivals, dim3(nblk((n + PLACE_PER - 1) / PLACE_PER)), dim3(TPB), 0, stream, cid, pid,
                           bin.rnk, n, bin.startB, bin.cnt, bin.dout, bin.keyB);
        /* cells per block: the block's runs fit the LDS staging at 1.25x the mean density */
        int cpb = REB_CELLS;
        while (cpb > 64 && (double)cpb * (double)n / (double)Nc * 1.25 > (double)REB_CAP) cpb >>= 1;
        hipLaunchKernelGGL(k_rebin_merge, dim3(nblk(Nc, cpb)), dim3(REB_CELLS), 0, stream, bin.start, bin.key, cid,
                           Nc, cpb, bin.startB, bin.cnt, bin.inc, bin.dout, bin.rnk, bin.keyB, bin.listB, err);
        check_launch();
        recDirty = false;
        std::swap(bin.start, bin.startB);
        std::swap(bin.key, bin.keyB);
        std::swap(bin.list, bin.listB);
    }

    void clear_err() {
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, stream, err, 1, 0);
        check_launch();
    }
    /* reads the error flag and clears it (in stream order), so a reported
     * error is reported once */
    void check_err() {
        int h = 0;
        HIP_OK(hipMemcpyAsync(&h, err, sizeof(int), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        if (h) clear_err();
        if (h & 1) throw ApiError(MPCD_E_BC_MODE, "Unsupported boundary mode at a wall a particle crossed");
        if (h & 2) throw ApiError(MPCD_E_CAPACITY, "zone scratch capacity exceeded");
        if (h & 4) throw ApiError(MPCD_E_INVALID, "binning update: a cell's stayers do not match its count");
        if (h & 16) throw ApiError(MPCD_E_INVALID, "dense phase: a wait for earlier cells hit its bound");
    }

    /* ---- stages ---- */
    /* shift != nullptr: also write the shifted cells (cidS) and the
     * unshifted cells (cid) of the new positions */
    void stream_bc(uint64_t seed, int do_stream, const double* shift, mpcd_wall_info* wi) {
        hipLaunchKernelGGL(k_stream_bc, dim3(nblk(n)), dim3(TPB), 0, stream, g, bc, seed, do_stream, x, y, vx, vy,
                           pid, shift ? cidS : (int*)nullptr, shift ? cid : (int*)nullptr, shift ? shift[0] : 0.0,
                           shift ? shift[1] : 0.0, xmax, ymax, wi ? wallPart : (WallAcc*)nullptr, err);
        check_launch();
        if (wi) {
            const int nb = (int)nblk(n);
            std::vector<WallAcc> h((size_t)nb);
            HIP_OK(hipMemcpyAsync(h.data(), wallPart, (size_t)nb * sizeof(WallAcc), hipMemcpyDeviceToHost, stream));
            HIP_OK(hipStreamSynchronize(stream));
            std::memset(wi, 0, sizeof(*wi));
            for (const auto& a : h) {
                wi->nBot += a.nBot;
                wi->nTop += a.nTop;
                wi->nLeft += a.nLeft;
                wi->nRight += a.nRight;
                wi->dEwall += a.dE;
                wi->dPxBot += a.dPxBot;
                wi->dPxTop += a.dPxTop;
                wi->dPyBot += a.dPyBot;
                wi->dPyTop += a.dPyTop;
                wi->dPxLeft += a.dPxLeft;
                wi->dPxRight += a.dPxRight;
                wi->dPyLeft += a.dPyLeft;
                wi->dPyRight += a.dPyRight;
            }
        }
    }

    void collide(uint64_t seed) {
        hipLaunchKernelGGL(k_collide, dim3(nblk(Nc, CELL_BLK)), dim3(CELL_BLK), cellCap * 2 * sizeof(double), stream, cp,
                           seed, bin.start, bin.list, Nc, vx, vy, cellCap);
        check_launch();
        if (p.keepMeanFlow) {
            const int nb = (int)nblk(n);
            hipLaunchKernelGGL(k_sum2, dim3(nb), dim3(TPB), 0, stream, vx, vy, n, part);
            hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(TPB), 0, stream, part, nb, 2, scal + 8);
            hipLaunchKernelGGL(k_sub_mean, dim3(nb), dim3(TPB), 0, stream, vx, vy, n, scal + 8);
            check_launch();
        }
    }

    double vc() const { return (p.Lx / (double)p.Nx) * (p.Ly / (double)p.Ny); }
    double rho_target() const { return p.gamma / vc(); }

    /* ref cell sums for the closure, from the current (unshifted) binning */
    void ref_fields() {
        FieldsOut f{};
        f.Ux = refUx;
        f.Uy = refUy;
        f.Px = refPx;
        f.Py = refPy;
        f.P = pdriveRaw; /* P = Pkin + Pvir = pdriveRaw (liquid_closure.cpp:401-405) */
        if (sur || wantSample) { /* the surrogate's features are the reference state's */
            f.N = refN;
            f.T = refT;
        }
        hipLaunchKernelGGL(k_cell_fields, dim3(nblk(Nc, CELL_BLK)), dim3(CELL_BLK), cellCap * 2 * sizeof(double), stream,
                           bin.start, bin.list, Nc, vx, vy, vc(), p.Kvirial, rho_target(), f, cellCap);
        check_launch();
        if (wantSnap) snap_fields(0);
    }

    /* every CellFields field of the current binning into snapshot k (0: the
     * reference state, 1: the red state) */
    void snap_fields(int k) {
        const size_t C = (size_t)Nc;
        double* b = snapF + (size_t)k * 9 * C;
        FieldsOut f{snapN + (size_t)k * C, b, b + C, b + 2 * C, b + 3 * C, b + 4 * C, b + 5 * C, b + 6 * C, b + 7 * C,
                    b + 8 * C};
        hipLaunchKernelGGL(k_cell_fields, dim3(nblk(Nc, CELL_BLK)), dim3(CELL_BLK), cellCap * 2 * sizeof(double), stream,
                           bin.start, bin.list, Nc, vx, vy, vc(), p.Kvirial, rho_target(), f, cellCap);
        check_launch();
    }

    void check_pass_modes() const {
        if (!p.useIncompressibleRedistribution || !p.useZoneRedistribution)
            throw ApiError(MPCD_E_INVALID,
                           "zone pass requires useIncompressibleRedistribution=true and useZoneRedistribution=true");
        if (p.redistributionEnableSurfaceTopology || p.redistributionWallWettingEnabled ||
            p.useInterfaceVelocityReorientation)
            throw ApiError(MPCD_E_INVALID, "zone pass is restricted to no surface / no wetting / no reorientation");
    }

    /* One snapshot zone pass over the current unshifted binning (whose cid
     * holds the input cells).  coverDev/overDev: cells several zones cover
     * (nullptr when none do).  single = 1: one zone that owns all its
     * particles, seeded with seedBase itself (run_zone_kernel_inplace). */
    void zone_pass(uint64_t seedBase, const ZoneRect* rectsDev, int nZ, int maxC, int maxW, const int* ownerDev,
                   const unsigned char* coverDev, const int* overDev, int nOverCells, int single,
                   const double* inPx = nullptr, const double* inPy = nullptr, bool record = false,
                   const unsigned char* zmask = nullptr) {
        hipLaunchKernelGGL(k_zone_sizes, dim3(nZ + 1), dim3(64), 0, stream, rectsDev, nZ, bin.start, p.Ny, zsize);
        const int nt = (nZ + 1 + SCAN_TILE - 1) / SCAN_TILE;
        hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(TPB), 0, stream, zsize, nZ + 1, zoff, tileSum,
                           (const int*)nullptr, (int*)nullptr);
        if (nt > 1) { /* one tile is already the whole scan */
            hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(TPB), 0, stream, tileSum, nt);
            hipLaunchKernelGGL(k_scan_add, dim3(nblk(nZ + 1)), dim3(TPB), 0, stream, zoff, nZ + 1, tileSum);
        }
        check_launch();
        ensure_du((size_t)nZ * 2 * (size_t)maxC);
        ZoneArgs a{};
        a.in.g = g;
        a.in.zp = zp;
        a.in.gcnt = bin.cnt;
        a.in.gstart = bin.start;
        a.in.glist = bin.list;
        a.in.gkey = bin.key;
        a.in.owner = single ? nullptr : ownerDev;
        a.in.gcover = single ? nullptr : coverDev;
        a.in.x = x;
        a.in.y = y;
        a.in.vx = vx;
        a.in.vy = vy;
        a.in.vsx = vx2; /* the side buffer: free during a pass */
        a.in.vsy = vy2;
        a.in.inPx = inPx;
        a.in.inPy = inPy;
        a.in.dgCell = zdCell;
        a.in.dgCnt = zdCnt;
        a.in.zmask = zmask;
        if (record) {
            /* moved particles record their cell changes for rebin_recorded()
             * (bin.rnk is -1 for every slot here: set at creation, and each
             * merge resets the ranks of the arrivals it took; after an
             * interrupted pass-and-update the record is cleared again) */
            if (recDirty) {
                hipLaunchKernelGGL(k_fill_int, dim3(nblk(n)), dim3(TPB), 0, stream, bin.rnk, n, -1);
                hipLaunchKernelGGL(k_fill_int, dim3(nblk(Nc + 1)), dim3(TPB), 0, stream, bin.inc, Nc + 1, 0);
                hipLaunchKernelGGL(k_fill_int, dim3(nblk(Nc + 1)), dim3(TPB), 0, stream, bin.dout, Nc + 1, 0);
            }
            recDirty = true;
            a.in.dcid = cid;
            a.in.dout = bin.dout;
            a.in.dinc = bin.inc;
            a.in.drnk = bin.rnk;
        }
        a.rects = rectsDev;
        a.nZones = nZ;
        a.seedBase = seedBase;
        a.singleSeed = single;
        a.zsize = zsize;
        a.zoff = zoff;
        a.scratchCap = scratchCap;
        a.nextG = next;
        a.tmpG = tmpg;
        a.duG = duG;
        a.rec = recG;
        a.zsStride = zone_state_bytes<int>(maxC, maxW);
        ensure_zstate(a.zsStride * (size_t)nZ);
        a.zstate = zstate;
        a.duStride = 2 * maxC;
        a.out = zout;
        a.err = err;
        pbeg(MPCD_PROF_ZONE_KERNEL);
        const size_t fix16 = zone_fixed_bytes<uint16_t>(maxC, maxW, zp.useLocal);
        const bool fast = maxC < 65535 && fix16 + 2048 <= (size_t)ZONE_LDS_FAST;
        /* LDS staging of the correction: the cell arrays and the column map leave room for links */
        const size_t cellB16 = (size_t)maxC * (4 * sizeof(uint16_t) + 2) + 8 * (size_t)maxW + 16;
        const size_t cellB32 = (size_t)maxC * (4 * sizeof(int) + 2) + 8 * (size_t)maxW + 16;
        /* passes and correction in one kernel when the correction fits the passes' LDS */
        const bool fused = fast && cellB16 + 4096 <= (size_t)ZONE_LDS_FAST;
        if (fast) {
            a.pMin = 0;
            a.pMax = 65534;
            a.ldsBytes = ZONE_LDS_FAST;
            if (fused)
                hipLaunchKernelGGL((k_zone_fused<uint16_t>), dim3(nZ), dim3(FUSED_TPB), ZONE_LDS_FAST, stream, a);
            else
                hipLaunchKernelGGL((k_zone<uint16_t, true>), dim3(nZ), dim3(ZONE_TPB), ZONE_LDS_FAST, stream, a);
            check_launch();
        }
        /* zones the uint16_t path does not take (none, as a rule: a small grid
         * strides over the zones then; all zones when the 16-bit form is out) */
        const int nBig = fast ? std::min(nZ, 64) : nZ;
        a.pMin = fast ? 65535 : 0;
        a.pMax = 0x7fffffff;
        const size_t fix32 = zone_fixed_bytes<int>(maxC, maxW, zp.useLocal);
        if (fix32 <= (size_t)ZONE_LDS_CAP) {
            a.ldsBytes = ZONE_LDS_CAP;
            hipLaunchKernelGGL((k_zone<int, true>), dim3(nBig), dim3(ZONE_TPB), ZONE_LDS_CAP, stream, a);
        } else {
            const size_t need = fix32 * (size_t)nZ;
            if (need > zscratchCap) {
                if (zscratch) HIP_OK(hipFree(zscratch));
                zscratch = dalloc<unsigned char>(need);
                zscratchCap = need;
            }
            a.gscratch = zscratch;
            a.gstride = fix32;
            hipLaunchKernelGGL((k_zone<int, false>), dim3(nBig), dim3(ZONE_TPB), 0, stream, a);
        }
        pend(MPCD_PROF_ZONE_KERNEL);
        check_launch();
        /* the correction, with the same split of the zones by index type */
        pbeg(MPCD_PROF_ZONE_FINISH);
        if (fast && !fused) {
            a.pMin = 0;
            a.pMax = 65534;
            if (cellB16 + 4096 <= (size_t)FINISH_LDS)
                hipLaunchKernelGGL((k_zone_finish<uint16_t, true>), dim3(nZ), dim3(FINISH_TPB), FINISH_LDS, stream, a);
            else
                hipLaunchKernelGGL((k_zone_finish<uint16_t, false>), dim3(nZ), dim3(FINISH_TPB), 0, stream, a);
        }
        a.pMin = fast ? 65535 : 0;
        a.pMax = 0x7fffffff;
        if (cellB32 + 4096 <= (size_t)FINISH_LDS)
            hipLaunchKernelGGL((k_zone_finish<int, true>), dim3(nBig), dim3(FINISH_TPB), FINISH_LDS, stream, a);
        else
            hipLaunchKernelGGL((k_zone_finish<int, false>), dim3(nBig), dim3(FINISH_TPB), 0, stream, a);
        pend(MPCD_PROF_ZONE_FINISH);
        check_launch();
        if (!single && coverDev && nOverCells > 0) {
            hipLaunchKernelGGL(k_overlap_copyback, dim3(nblk(nOverCells)), dim3(TPB), 0, stream, overDev, nOverCells,
                               bin.start, bin.list, vx2, vy2, vx, vy);
            check_launch();
        }
    }

    void layout_pass(int L, uint64_t seedBase, const double* inPx = nullptr, const double* inPy = nullptr,
                     bool record = false) {
        zone_pass(seedBase, rects[L], nz[L], maxCells[L], maxCols[L], owner[L], cover[L], overCells[L], nOver[L], 0,
                  inPx, inPy, record);
    }

    void pass_metrics(int nZ, mpcd_zone_pass_metrics* m) {
        std::vector<ZoneOut> h((size_t)nZ);
        HIP_OK(hipMemcpyAsync(h.data(), zout, (size_t)nZ * sizeof(ZoneOut), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        std::memset(m, 0, sizeof(*m));
        m->nZones = nZ;
        double acc = 0.0, mx = 0.0;
        int cntz = 0;
        for (const auto& o : h) {
            m->nParticlesMovedDense += o.movedDense;
            m->nParticlesMovedSparse += o.movedSparse;
            if (o.stats4 > 0.0) {
                acc += o.stats4 * o.stats4;
                ++cntz;
            }
            mx = std::max(mx, o.stats5);
        }
        m->corrRmsDU = cntz > 0 ? std::sqrt(acc / (double)cntz) : 0.0;
        m->corrMaxDU = mx;
    }

    /* The surrogate's closure: red fields -> features -> network -> kick.
     * A prediction that fails the residual check (not finite, or beyond the
     * cap) sends the step to the plain closure instead.  ref sums and the
     * red binning as for closure(). */
    /* The kick and the reorder into red-cell order.  fuseNext >= 0 (set by
     * step() between steps of one call) and no metrics: the next step's
     * force, stream and wall conditions run in the same kernel, and that
     * step skips its own (streamed records which step is done). */
    void kick_reorder(double* partOut) {
        const bool fuse = fuseNext >= 0 && !partOut;
        const int cls = fuse ? MPCD_PROF_KICK_STREAM : MPCD_PROF_KICK;
        pbeg(cls);
        if (fuse) {
            const uint64_t un = (uint64_t)fuseNext;
            double sh[2];
            mpcd_draw_shift(130363ULL + 10007ULL * un + seedOff, p.a0, &sh[0], &sh[1]);
            hipLaunchKernelGGL(k_fill_int, dim3(nblk(Nc + 1)), dim3(TPB), 0, stream, bin.cntS, Nc + 1, 0);
            hipLaunchKernelGGL(k_fill_int, dim3(nblk(Nc + 1)), dim3(TPB), 0, stream, bin.cnt, Nc + 1, 0);
            hipLaunchKernelGGL(k_kick_reorder_stream, dim3(nblk(n)), dim3(TPB), 0, stream, g, bc,
                               104729ULL + 10007ULL * un + seedOff, bin.list, x, y, vx, vy, pid, x2, y2, vx2, vy2, pid2, duRX,
                               duRY, duEX, duEY, scal, cidS, cid, sh[0], sh[1], xmax, ymax, bin.cntS, bin.cnt, err);
            streamed = fuseNext;
        } else {
            hipLaunchKernelGGL(k_kick_reorder, dim3(nblk(n)), dim3(TPB), 0, stream, g, bin.list, x, y, vx, vy, pid,
                               x2, y2, vx2, vy2, pid2, duRX, duRY, duEX, duEY, scal, partOut);
        }
        pend(cls);
        std::swap(x, x2);
        std::swap(y, y2);
        std::swap(vx, vx2);
        std::swap(vy, vy2);
        std::swap(pid, pid2); /* the binning now describes the old order: stale until rebuilt */
        check_launch();
    }

    /* After closure(): the repaired velocities (host, interleaved, particle
     * order).  The kick swapped the buffers, so the second set holds the red
     * state in its pre-kick slot order with its particle ids. */
    void download_repaired(double* hv) {
        if (n == 0) return;
        hipLaunchKernelGGL(k_repaired_velocity, dim3(nblk(n)), dim3(TPB), 0, stream, g, x2, y2, vx2, vy2, pid2, duRX,
                           duRY, scal, stage);
        check_launch();
        HIP_OK(hipMemcpyAsync(hv, stage, 2 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
    }

    ClosureArgs closure_args() {
        ClosureArgs ca{};
        ca.Nc = Nc;
        ca.gamma = p.gamma;
        ca.refPx = refPx;
        ca.refPy = refPy;
        ca.refUx = refUx;
        ca.refUy = refUy;
        ca.redPx = redPx;
        ca.redPy = redPy;
        ca.redUx = redUx;
        ca.redUy = redUy;
        ca.duRX = duRX;
        ca.duRY = duRY;
        ca.part = part;
        return ca;
    }

    /* the plain closure's grids (repair, beta, pressure drive, EOS
     * gradient) from the reference and red fields; with a gate, predicated
     * on it (the kernels leave at once when it is 0) */
    void closure_grids(const int* gate) {
        ClosureArgs ca = closure_args();
        ca.gate = gate;
        const int nbc = (int)nblk(Nc);
        hipLaunchKernelGGL(k_closure_repair, dim3(nbc), dim3(TPB), 0, stream, ca);
        hipLaunchKernelGGL(k_reduce_beta, dim3(1), dim3(TPB), 0, stream, part, nbc, scal, p.useOptimalBetaRepair,
                           p.betaRepair, gate);
        hipLaunchKernelGGL(k_pdrive, dim3(nbc), dim3(TPB), 0, stream, g, pdriveRaw, mask, p.smoothPdriveInterzoneBlend,
                           anyMask, pdrive, gate);
        const double dtr = p.dt / rho_target();
        hipLaunchKernelGGL(k_eos, dim3(nbc), dim3(TPB), 0, stream, g, pdrive, dtr, p.betaEOS, duEX, duEY, gate);
        check_launch();
    }

    /* the residual check's inputs for the grids in place (occupancy of the
     * red counts, divergence of the corrected red velocity) and the
     * decision on the device (plain = 1: the plain closure's values become
     * the reference) */
    void sur_residual(int plain) {
        occupancy(redN, DIAG_SUR_OCC);
        const int nbc = (int)nblk(Nc);
        hipLaunchKernelGGL(k_div_residual, dim3(nbc), dim3(TPB), 0, stream, g, redUx, redUy, duRX, duRY, duEX, duEY,
                           scal, part);
        hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(TPB), 0, stream, part, nbc, 1, diagDev + DIAG_SUR_DIV);
        hipLaunchKernelGGL(k_sur_decide, dim3(1), dim3(64), 0, stream, diagDev + DIAG_SUR_OCC, diagDev + DIAG_SUR_DIV,
                           surDev, surFlag, plain, surOccTol, surDivTol);
        check_launch();
    }

    /* A surrogate step: red fields -> features -> network -> the kick's
     * grids, then the residual check on the device; the plain closure's
     * kernels are launched predicated on its flag, so a failed check makes
     * this step a plain one with no host round trip. */
    void closure_learned() {
        FieldsOut f{};
        f.N = redN;
        f.Ux = redUx;
        f.Uy = redUy;
        f.Px = redPx;
        f.Py = redPy;
        f.T = redT;
        hipLaunchKernelGGL(k_cell_fields, dim3(nblk(Nc, CELL_BLK)), dim3(CELL_BLK), cellCap * 2 * sizeof(double), stream,
                           bin.start, bin.list, Nc, vx, vy, vc(), p.Kvirial, rho_target(), f, cellCap);
        const int nbc = (int)nblk(Nc);
        hipLaunchKernelGGL(k_features, dim3(nbc), dim3(TPB), 0, stream, g, refN, refUx, refUy, refT, p.gamma, p.kBT,
                           featBuf);
        const float* pred = net_forward_device(sur, featBuf, stream);
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, stream, surFlag, 1, 0);
        hipLaunchKernelGGL(k_learned_du, dim3(nbc), dim3(TPB), 0, stream, Nc, pred, redN, surScale, surCap, duRX, duRY,
                           duEX, duEY, scal, surFlag, part);
        hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(TPB), 0, stream, part, nbc, 2, scal + 8);
        hipLaunchKernelGGL(k_learned_center, dim3(nbc), dim3(TPB), 0, stream, Nc, scal + 8, 1.0 / (double)n, duEX,
                           duEY);
        /* along a periodic axis the momentum is removed line by line */
        if (g.perx) hipLaunchKernelGGL(k_learned_center_lines, dim3(p.Ny), dim3(TPB), 0, stream, g, 0, redN, duEX);
        if (g.pery) hipLaunchKernelGGL(k_learned_center_lines, dim3(p.Nx), dim3(TPB), 0, stream, g, 1, redN, duEY);
        check_launch();
        sur_residual(0);
        closure_grids(surFlag);
        kick_reorder(nullptr);
    }

    /* closure: ref sums already in ref*; red binning current */
    void closure(mpcd_closure_metrics* cm) {
        if (!p.useLiquidClosure) throw ApiError(MPCD_E_INVALID, "run_liquid_closure requires useLiquidClosure=true");
        FieldsOut f{};
        f.Ux = redUx;
        f.Uy = redUy;
        f.Px = redPx;
        f.Py = redPy;
        if (wantSample || sur) {
            f.N = redN;
            f.T = redT;
        }
        hipLaunchKernelGGL(k_cell_fields, dim3(nblk(Nc, CELL_BLK)), dim3(CELL_BLK), cellCap * 2 * sizeof(double), stream,
                           bin.start, bin.list, Nc, vx, vy, vc(), p.Kvirial, rho_target(), f, cellCap);
        closure_grids(nullptr);
        const int nbc = (int)nblk(Nc);
        if (wantSnap) {
            snap_fields(1);
            double* gr = snapF + 18 * (size_t)Nc;
            hipLaunchKernelGGL(k_snap_grids, dim3(nbc), dim3(TPB), 0, stream, Nc, pdriveRaw, pdrive, duRX, duRY, duEX, duEY,
                               scal, gr);
            check_launch();
            snapStep = curStep;
        }
        if (sur) sur_residual(1); /* the plain closure's residuals become the surrogate's reference */
        if (wantSample) {
            hipLaunchKernelGGL(k_features, dim3(nbc), dim3(TPB), 0, stream, g, refN, refUx, refUy, refT, p.gamma, p.kBT,
                               featBuf);
            hipLaunchKernelGGL(k_closure_target, dim3(nbc), dim3(TPB), 0, stream, Nc, scal, duRX, duRY, duEX, duEY,
                               1.0 / surScale, tgtBuf);
        }
        check_launch();
        const int nbp = (int)nblk(n);
        if (cm) {
            hipLaunchKernelGGL(k_closure_cell_metrics, dim3(nbc), dim3(TPB), 0, stream, Nc, duRX, duRY, pdriveRaw,
                               pdrive, scal, part);
            hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(TPB), 0, stream, part, nbc, 3, scal + 16);
        }
        kick_reorder(cm ? part : (double*)nullptr);
        if (cm) {
            std::vector<double> h((size_t)nbp * 6);
            HIP_OK(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
            double sc[24];
            HIP_OK(hipMemcpyAsync(sc, scal, sizeof(sc), hipMemcpyDeviceToHost, stream));
            HIP_OK(hipStreamSynchronize(stream));
            double r2 = 0, rm = 0, e2 = 0, em = 0, dvx = 0, dvy = 0;
            for (int b = 0; b < nbp; ++b) {
                r2 += h[6 * (size_t)b];
                rm = std::max(rm, h[6 * (size_t)b + 1]);
                e2 += h[6 * (size_t)b + 2];
                em = std::max(em, h[6 * (size_t)b + 3]);
                dvx += h[6 * (size_t)b + 4];
                dvy += h[6 * (size_t)b + 5];
            }
            std::memset(cm, 0, sizeof(*cm));
            cm->betaRepairNum = sc[0];
            cm->betaRepairDen = sc[1];
            cm->betaRepairOpt = sc[2];
            cm->betaRepairApplied = sc[3];
            cm->betaEOSApplied = p.betaEOS;
            cm->meanAbsRepairDU = sc[16] / (double)std::max(Nc, 1);
            cm->meanPdriveRaw = sc[17] / (double)Nc;
            cm->meanPdrive = sc[18] / (double)Nc;
            cm->rmsRepairDU = n > 0 ? std::sqrt(r2 / (double)n) : 0.0;
            cm->maxRepairDU = n > 0 ? rm : 0.0;
            cm->rmsEOSDU = n > 0 ? std::sqrt(e2 / (double)n) : 0.0;
            cm->maxEOSDU = n > 0 ? em : 0.0;
            cm->momentumDeltaX = dvx;
            cm->momentumDeltaY = dvy;
            cm->nMaskedInterzoneCells = nMasked;
        }
    }

    /* ---- diagnostics (device reductions; the host reads a few scalars) ---- */

    /* occupancy statistics of the count grid `cnt` into diagDev[slot..slot+5)
     * (cntLow: the grid of the low thresholds, if not cnt) */
    void occupancy(const int* cnt, int slot, const int* cntLow = nullptr) {
        const int nb = (int)std::min<size_t>(nblk(Nc), 512);
        ensure_parts((size_t)nb * OCC_W);
        hipLaunchKernelGGL(k_occupancy, dim3(nb), dim3(TPB), 0, stream, cnt, cntLow ? cntLow : cnt, p.Nx, p.Ny,
                           zp, part);
        hipLaunchKernelGGL(k_occupancy_final, dim3(1), dim3(TPB), 0, stream, part, nb, Nc, p.gamma, diagDev + slot);
        check_launch();
    }

    /* all cell fields of the current state (common_grid.cpp:32-100) into
     * outN / outF, from a fresh unshifted binning of the current positions */
    void fields_current() {
        if (!outN) {
            outN = dalloc<int>((size_t)Nc);
            outF = dalloc<double>(9 * (size_t)Nc);
        }
        bin_unshifted();
        const size_t C = (size_t)Nc;
        FieldsOut f{outN, outF, outF + C, outF + 2 * C, outF + 3 * C, outF + 4 * C, outF + 5 * C, outF + 6 * C,
                    outF + 7 * C, outF + 8 * C};
        hipLaunchKernelGGL(k_cell_fields, dim3(nblk(Nc, CELL_BLK)), dim3(CELL_BLK), cellCap * 2 * sizeof(double), stream,
                           bin.start, bin.list, Nc, vx, vy, vc(), p.Kvirial, rho_target(), f, cellCap);
        check_launch();
    }

    /* sums over the particles of the current state (vx, vy, kinetic energy)
     * and over the cells of its fields (Ux, Uy) into diagDev */
    void state_diag() {
        const int nb = (int)nblk(n);
        hipLaunchKernelGGL(k_sum_ke, dim3(nb), dim3(TPB), 0, stream, vx, vy, n, part);
        hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(TPB), 0, stream, part, nb, 3, diagDev + DIAG_PSUM);
        fields_current();
        const int nbc = (int)nblk(Nc);
        hipLaunchKernelGGL(k_sum2, dim3(nbc), dim3(TPB), 0, stream, outF, outF + Nc, Nc, part);
        hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(TPB), 0, stream, part, nbc, 2, diagDev + DIAG_USUM);
        check_launch();
    }

    void read_diag(mpcd_step_diag* d) {
        double h[DIAG_SLOTS];
        HIP_OK(hipMemcpyAsync(h, diagDev, sizeof(h), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        mpcd_occupancy_stats* o[2] = {&d->base, &d->shifted};
        const int slot[2] = {DIAG_OCC_BASE, DIAG_OCC_SHIFTED};
        for (int k = 0; k < 2; ++k) {
            const double* q = h + slot[k];
            o[k]->occStd = q[0];
            o[k]->outBand = q[1];
            o[k]->nDense = (int32_t)q[2];
            o[k]->nSparse = (int32_t)q[3];
            o[k]->nEmptyActive = (int32_t)q[4];
            o[k]->_pad = 0;
        }
        const double inv = 1.0 / (double)std::max(n, 1);
        d->meanVelocityX = h[DIAG_PSUM] * inv;
        d->meanVelocityY = h[DIAG_PSUM + 1] * inv;
        d->meanKinetic = h[DIAG_PSUM + 2] * inv;
        d->meanUx = h[DIAG_USUM] / (double)Nc;
        d->meanUy = h[DIAG_USUM + 1] / (double)Nc;
        d->Qx = d->meanUx * p.Ly;
    }

    /* one full time step with the benchmark driver's seeds
     * (main_benchmark_poiseuille_openmp_profiled_shifted_profiled.cpp:240-279) */
    /* next >= 0: the step that follows in the same call (its stream is
     * fused into this step's kick when no metrics are asked) */
    void step(int s, mpcd_step_metrics* sm, double* const* stages, int next = -1) {
        check_pass_modes();
        curStep = s;
        const uint64_t us = (uint64_t)s;
        const bool pre = streamed == s && !sm;
        phase_mark(0);
        if (pre) {
            streamed = -1; /* the previous step's kick streamed this one */
        } else {
            if (streamed >= 0) throw ApiError(MPCD_E_INVALID, "internal: a fused stream was left unused");
            double sh[2];
            mpcd_draw_shift(130363ULL + 10007ULL * us + seedOff, p.a0, &sh[0], &sh[1]);
            pbeg(MPCD_PROF_STREAM_BC);
            stream_bc(104729ULL + 10007ULL * us + seedOff, 1, sh, sm ? &sm->wall : nullptr);
            pend(MPCD_PROF_STREAM_BC);
        }
        /* after a fused stream both histograms are done (the previous kick) */
        pbeg(MPCD_PROF_BIN);
        bin_cells(cidS, bin.cntS, pre, false);
        pend(MPCD_PROF_BIN);
        pbeg(MPCD_PROF_COLLIDE);
        collide(433494437ULL + 10007ULL * us + seedOff);
        pend(MPCD_PROF_COLLIDE);
        pbeg(MPCD_PROF_BIN);
        bin_cells(cid, bin.cnt, pre, true); /* collision does not move particles */
        pend(MPCD_PROF_BIN);
        if (stages && stages[0]) get_state(stages[0], stages[1]);
        pbeg(MPCD_PROF_CELL_FIELDS);
        ref_fields();
        pend(MPCD_PROF_CELL_FIELDS);
        phase_mark(1);
        const uint64_t zs = 88172645463325252ULL + 10007ULL * us + seedOff;
        pbeg(MPCD_PROF_ZONE_PASS);
        /* the base pass's input is the state the reference fields were
         * summed over (same runs, same order): the correction takes its
         * input sums from them */
        layout_pass(0, zs, refPx, refPy, true);
        pend(MPCD_PROF_ZONE_PASS);
        if (sm) pass_metrics(nz[0], &sm->base);
        /* the passes move a few particles: the binning is updated */
        pbeg(MPCD_PROF_BIN);
        rebin_recorded();
        pend(MPCD_PROF_BIN);
        phase_mark(2);
        if (wantDiag) occupancy(bin.cnt, DIAG_OCC_BASE); /* the counts after the base pass */
        if (stages && stages[2]) get_state(stages[2], stages[3]);
        pbeg(MPCD_PROF_ZONE_PASS);
        layout_pass(1, zs + 4099ULL, nullptr, nullptr, true);
        pend(MPCD_PROF_ZONE_PASS);
        if (sm) pass_metrics(nz[1], &sm->shifted);
        pbeg(MPCD_PROF_BIN);
        rebin_recorded();
        pend(MPCD_PROF_BIN);
        phase_mark(3);
        if (wantDiag) occupancy(bin.cnt, DIAG_OCC_SHIFTED);
        if (stages && stages[4]) get_state(stages[4], stages[5]);
        pbeg(MPCD_PROF_CLOSURE);
        fuseNext = (stages || sm || rowSteps >= 0) ? -1 : next;
        if (sur && surK > 0 && s % surK != 0) {
            closure_learned();
        } else {
            ++nPlain;
            closure(sm ? &sm->closure : nullptr);
        }
        fuseNext = -1;
        pend(MPCD_PROF_CLOSURE);
        phase_mark(4);
        if (wantDiag) state_diag();
        if (rowSteps >= 0) { /* row profiles of the state after the step */
            hipLaunchKernelGGL(k_row_profile, dim3(std::min<unsigned>(nblk(n), 1024u)), dim3(TPB),
                               3 * (size_t)p.Ny * sizeof(double), stream, g, y, vx, vy, rowAcc);
            check_launch();
            ++rowSteps;
        }
    }
};

/* ===================================================================== */
/*          compute_cell_fields without allocation (static scratch)      */
/* ===================================================================== */

namespace {

/* Scratch of mpcd_gpu_compute_cell_fields_noalloc: a file-scope device
 * array (each GPU the library runs on holds its own copy until the process
 * ends).  Regions, each 256-byte aligned: vx, vy (f64), cid, pid, list
 * (i32) and the sort keys (u64) per particle; counts, starts, cursors (i32)
 * per cell + 1; the scan's tile totals. */
constexpr size_t NA_N = MPCD_NOALLOC_MAX_PARTICLES, NA_C = (size_t)MPCD_NOALLOC_MAX_CELLS + 1;
constexpr size_t na_al(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t NA_VX = 0, NA_VY = NA_VX + na_al(8 * NA_N), NA_CID = NA_VY + na_al(8 * NA_N),
                 NA_PID = NA_CID + na_al(4 * NA_N), NA_LIST = NA_PID + na_al(4 * NA_N),
                 NA_KEY = NA_LIST + na_al(4 * NA_N), NA_CNT = NA_KEY + na_al(8 * NA_N),
                 NA_START = NA_CNT + na_al(4 * NA_C), NA_CUR = NA_START + na_al(4 * NA_C),
                 NA_TILE = NA_CUR + na_al(4 * NA_C),
                 NA_BYTES = NA_TILE + na_al(4 * ((NA_C + SCAN_TILE - 1) / SCAN_TILE));

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

/* per-device address of `scratch` (the host-side name is not a device
 * address), taken once per device with that device current */
constexpr int NA_MAX_DEV = 64;
std::mutex g_na_mu;
unsigned char* g_na_ptr[NA_MAX_DEV] = {};

unsigned char* noalloc_scratch() {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_MAX_DEV) throw ApiError(MPCD_E_INVALID, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[dev]) {
        void* ptr = nullptr;
        HIP_OK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw ApiError(MPCD_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_na_ptr[dev] = (unsigned char*)ptr;
    }
    return g_na_ptr[dev];
}

/* unshifted cells of interleaved positions (common_grid.cpp:9-19) */
__global__ void k_cells_xy(Geom g, const double* __restrict__ xy, int* __restrict__ cid) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < g.n) cid[i] = cell_of(g, xy[2 * (size_t)i], xy[2 * (size_t)i + 1]);
}

}  // namespace

/* ===================================================================== */
/*                                C ABI                                  */
/* ===================================================================== */

void mpcdk::set_last_error(const std::string& m) { g_last_error = m; }

namespace {

/* Scratch simulations reused across stateless calls of the same shape. */
std::mutex g_cache_mu;
mpcd_gpu_sim* g_cache = nullptr;

mpcd_gpu_sim& scratch_sim(const mpcd_params& p) {
    if (g_cache && std::memcmp(&g_cache->p, &p, sizeof(p)) == 0) return *g_cache;
    delete g_cache;
    g_cache = nullptr;
    g_cache = new mpcd_gpu_sim(p);
    return *g_cache;
}

}  // namespace

extern "C" {

const char* mpcd_gpu_last_error(void) { return g_last_error.c_str(); }

int mpcd_gpu_device_count(int* count) {
    return guarded([&] {
        int c = 0;
        HIP_OK(hipGetDeviceCount(&c));
        *count = c;
    });
}

int mpcd_gpu_sim_create(const mpcd_params* p, mpcd_gpu_sim** out) {
    return guarded([&] {
        if (!p || !out) throw ApiError(MPCD_E_INVALID, "null argument");
        *out = new mpcd_gpu_sim(*p);
    });
}

int mpcd_gpu_sim_destroy(mpcd_gpu_sim* s) {
    return guarded([&] { delete s; });
}

int mpcd_gpu_sim_set_state(mpcd_gpu_sim* s, const double* x, const double* v) {
    return guarded([&] { s->set_state(x, v); });
}

int mpcd_gpu_sim_get_state(mpcd_gpu_sim* s, double* x, double* v) {
    return guarded([&] { s->get_state(x, v); });
}

int mpcd_gpu_sim_step(mpcd_gpu_sim* s, int first_step, int n_steps, mpcd_step_metrics* last) {
    return guarded([&] {
        s->clear_err();
        s->streamed = -1;
        for (int k = 0; k < n_steps; ++k) {
            /* the step before the one that reports metrics streams nothing ahead */
            const bool fuse = k + 1 < n_steps && !(last && k + 2 == n_steps);
            s->step(first_step + k, (k == n_steps - 1) ? last : nullptr, nullptr, fuse ? first_step + k + 1 : -1);
        }
        s->check_err();
    });
}

/* Enqueue steps without waiting or checking the wall-mode flag (timed loops). */
int mpcd_gpu_sim_step_async(mpcd_gpu_sim* s, int first_step, int n_steps) {
    return guarded([&] {
        s->streamed = -1;
        for (int k = 0; k < n_steps; ++k) s->step(first_step + k, nullptr, nullptr, k + 1 < n_steps ? first_step + k + 1 : -1);
    });
}

int mpcd_gpu_sim_check(mpcd_gpu_sim* s) {
    return guarded([&] { s->check_err(); });
}

int mpcd_gpu_sim_step_stages(mpcd_gpu_sim* s, int step, double* x_ref, double* v_ref, double* x_base, double* v_base,
                             double* x_red, double* v_red, mpcd_step_metrics* m) {
    return guarded([&] {
        double* st[6] = {x_ref, v_ref, x_base, v_base, x_red, v_red};
        s->clear_err();
        s->streamed = -1;
        s->step(step, m, st);
        s->check_err();
    });
}

int mpcd_gpu_sim_step_diag(mpcd_gpu_sim* s, int step, mpcd_step_metrics* m, mpcd_step_diag* d) {
    return guarded([&] {
        if (!d) throw ApiError(MPCD_E_INVALID, "null diagnostics");
        mpcd_step_metrics local{};
        s->clear_err();
        s->streamed = -1;
        s->wantDiag = true;
        try {
            s->step(step, m ? m : &local, nullptr);
        } catch (...) {
            s->wantDiag = false;
            throw;
        }
        s->wantDiag = false;
        s->read_diag(d);
        s->check_err();
    });
}

int mpcd_gpu_sim_cell_fields(mpcd_gpu_sim* s, double rhoTarget, mpcd_cell_fields* out) {
    return guarded([&] {
        if (!out) throw ApiError(MPCD_E_INVALID, "null cell fields");
        if (s->streamed >= 0) throw ApiError(MPCD_E_INVALID, "internal: a fused stream was left unused");
        s->fields_current();
        const size_t C = (size_t)s->Nc;
        if (out->N) HIP_OK(hipMemcpyAsync(out->N, s->outN, C * sizeof(int), hipMemcpyDeviceToHost, s->stream));
        double* hs[9] = {out->Ux, out->Uy, out->Px, out->Py, out->T, out->rho, out->Pkin, out->Pvir, out->P};
        for (int k = 0; k < 9; ++k)
            if (hs[k])
                HIP_OK(hipMemcpyAsync(hs[k], s->outF + (size_t)k * C, C * sizeof(double), hipMemcpyDeviceToHost,
                                      s->stream));
        HIP_OK(hipStreamSynchronize(s->stream));
        if (out->rhoTarget)
            for (size_t c = 0; c < C; ++c) out->rhoTarget[c] = rhoTarget;
    });
}

int mpcd_gpu_occupancy_stats(const mpcd_params* p, const int32_t* counts, mpcd_occupancy_stats* out) {
    return guarded([&] {
        if (!p || !counts || !out) throw ApiError(MPCD_E_INVALID, "null argument");
        if (p->Nx <= 0 || p->Ny <= 0 || p->Nc != p->Nx * p->Ny) throw ApiError(MPCD_E_INVALID, "bad grid size");
        const ZoneParams zp = make_zone_params(*p);
        const int Nc = p->Nc;
        const int nb = (int)std::min<size_t>(nblk(Nc), 512);
        DevBuf<int> dcb((size_t)Nc);
        DevBuf<double> dpartb((size_t)nb * OCC_W + 8);
        int* dc = dcb.p;
        double* dpart = dpartb.p;
        hipStream_t st = nullptr;
        HIP_OK(hipMemcpy(dc, counts, (size_t)Nc * sizeof(int), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_occupancy, dim3(nb), dim3(TPB), 0, st, dc, dc, p->Nx, p->Ny, zp, dpart);
        hipLaunchKernelGGL(k_occupancy_final, dim3(1), dim3(TPB), 0, st, dpart, nb, Nc, p->gamma,
                           dpart + (size_t)nb * OCC_W);
        HIP_OK(hipGetLastError());
        double h[5];
        HIP_OK(hipMemcpy(h, dpart + (size_t)nb * OCC_W, sizeof(h), hipMemcpyDeviceToHost));
        out->occStd = h[0];
        out->outBand = h[1];
        out->nDense = (int32_t)h[2];
        out->nSparse = (int32_t)h[3];
        out->nEmptyActive = (int32_t)h[4];
        out->_pad = 0;
    });
}

int mpcd_gpu_device_synchronize(void) {
    return guarded([&] { HIP_OK(hipDeviceSynchronize()); });
}

int mpcd_gpu_sim_profile(mpcd_gpu_sim* s, int enable) {
    return guarded([&] {
        HIP_OK(hipStreamSynchronize(s->stream));
        s->prof_reset();
        s->prof = (enable & MPCD_PROFILE_KERNELS) != 0;
        s->phases = (enable & MPCD_PROFILE_PHASES) != 0;
    });
}

int mpcd_gpu_sim_profile_read(mpcd_gpu_sim* s, double* total_ms, int32_t* launches) {
    return guarded([&] { s->prof_read(total_ms, launches); });
}

int mpcd_gpu_sim_phase_read(mpcd_gpu_sim* s, double* ms4) {
    return guarded([&] {
        if (!ms4) throw ApiError(MPCD_E_INVALID, "null argument");
        s->phase_read(ms4);
    });
}

int mpcd_gpu_sim_synchronize(mpcd_gpu_sim* s) {
    return guarded([&] { HIP_OK(hipStreamSynchronize(s->stream)); });
}

int mpcd_gpu_sim_stream(mpcd_gpu_sim* s, void** stream) {
    return guarded([&] { *stream = (void*)s->stream; });
}

int mpcd_gpu_force_stream_bc(const mpcd_params* p, double* x, double* v, uint64_t rng_seed, int do_stream,
                             mpcd_wall_info* info) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto& s = scratch_sim(*p);
        s.set_state(x, v);
        mpcd_wall_info wi{};
        s.clear_err();
        s.stream_bc(rng_seed, do_stream, nullptr, &wi);
        s.check_err();
        s.get_state(x, v);
        if (info) *info = wi;
    });
}

int mpcd_gpu_apply_bc_general(const mpcd_params* p, double* x, double* v, uint64_t rng_seed, mpcd_wall_info* info) {
    return mpcd_gpu_force_stream_bc(p, x, v, rng_seed, 0, info);
}

int mpcd_gpu_srd_cell_id_with_random_shift(const mpcd_params* p, const double* x, uint64_t rng_seed, int32_t* cid) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto& s = scratch_sim(*p);
        s.set_state(x, nullptr);
        double sx, sy;
        mpcd_draw_shift(rng_seed, p->a0, &sx, &sy);
        hipLaunchKernelGGL(k_shift_cells, dim3(nblk(s.n)), dim3(TPB), 0, s.stream, s.g, s.x, s.y, sx, sy, s.xmax,
                           s.ymax, s.cid);
        s.check_launch();
        HIP_OK(hipMemcpyAsync(cid, s.cid, (size_t)s.n * sizeof(int), hipMemcpyDeviceToHost, s.stream));
        HIP_OK(hipStreamSynchronize(s.stream));
    });
}

int mpcd_gpu_srd_collision_step(const mpcd_params* p, double* v, const int32_t* cid, uint64_t rng_seed) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        for (int i = 0; i < p->n; ++i)
            if (cid[i] < 0 || cid[i] >= p->Nc) throw ApiError(MPCD_E_INVALID, "srd_collision_step: cell id out of range");
        auto& s = scratch_sim(*p);
        s.set_state(nullptr, v);
        HIP_OK(hipMemcpyAsync(s.cid, cid, (size_t)s.n * sizeof(int), hipMemcpyHostToDevice, s.stream));
        s.bin_cells(s.cid);
        s.collide(rng_seed);
        s.get_state(nullptr, v);
    });
}

int mpcd_gpu_compute_cell_fields(const mpcd_params* p, const double* x, const double* v, double rhoTarget,
                                 mpcd_cell_fields* out) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto& s = scratch_sim(*p);
        s.set_state(x, v);
        s.bin_unshifted();
        DevBuf<double> db[9];
        double* d[9];
        for (int k = 0; k < 9; ++k) d[k] = db[k].p = dalloc<double>((size_t)s.Nc);
        DevBuf<int> dNb((size_t)s.Nc);
        int* dN = dNb.p;
        FieldsOut f{dN, d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8]};
        hipLaunchKernelGGL(k_cell_fields, dim3(nblk(s.Nc, CELL_BLK)), dim3(CELL_BLK), s.cellCap * 2 * sizeof(double),
                           s.stream, s.bin.start, s.bin.list, s.Nc, s.vx, s.vy, s.vc(), p->Kvirial, rhoTarget, f,
                           s.cellCap);
        s.check_launch();
        const size_t nb = (size_t)s.Nc * sizeof(double);
        if (out->N) HIP_OK(hipMemcpyAsync(out->N, dN, (size_t)s.Nc * sizeof(int), hipMemcpyDeviceToHost, s.stream));
        double* hs[9] = {out->Ux, out->Uy, out->Px, out->Py, out->T, out->rho, out->Pkin, out->Pvir, out->P};
        for (int k = 0; k < 9; ++k)
            if (hs[k]) HIP_OK(hipMemcpyAsync(hs[k], d[k], nb, hipMemcpyDeviceToHost, s.stream));
        HIP_OK(hipStreamSynchronize(s.stream));
        if (out->rhoTarget)
            for (int c = 0; c < s.Nc; ++c) out->rhoTarget[c] = rhoTarget;
    });
}

int mpcd_gpu_compute_cell_fields_noalloc(const mpcd_params* p, const double* x, const double* v, double rhoTarget,
                                         const mpcd_cell_fields* out, void* stream) {
    return guarded([&] {
        if (!p || !x || !v || !out) throw ApiError(MPCD_E_INVALID, "null argument");
        validate(*p);
        if (p->n == 0) throw ApiError(MPCD_E_INVALID, "no particles: the GPU path needs n >= 1");
        if (p->n > MPCD_NOALLOC_MAX_PARTICLES || p->Nc > MPCD_NOALLOC_MAX_CELLS)
            throw ApiError(MPCD_E_INVALID, "compute_cell_fields_noalloc: at most " +
                                               std::to_string(MPCD_NOALLOC_MAX_PARTICLES) + " particles and " +
                                               std::to_string(MPCD_NOALLOC_MAX_CELLS) + " cells (got n=" +
                                               std::to_string(p->n) + ", Nc=" + std::to_string(p->Nc) + ")");
        unsigned char* b = noalloc_scratch();
        hipStream_t st = (hipStream_t)stream;
        const int n = p->n, Nc = p->Nc;
        const Geom g = make_geom(*p);
        double* vx = (double*)(b + NA_VX);
        double* vy = (double*)(b + NA_VY);
        int* cid = (int*)(b + NA_CID);
        int* pid = (int*)(b + NA_PID);
        int* list = (int*)(b + NA_LIST);
        unsigned long long* key = (unsigned long long*)(b + NA_KEY);
        int* cnt = (int*)(b + NA_CNT);
        int* start = (int*)(b + NA_START);
        int* cursor = (int*)(b + NA_CUR);
        int* tileSum = (int*)(b + NA_TILE);
        const int nTiles = (int)((Nc + 1 + SCAN_TILE - 1) / SCAN_TILE);
        /* the scratch keeps the last call's contents: everything read is
         * written first (counts cleared by a fill kernel, not a memset) */
        if (n > 0) {
            hipLaunchKernelGGL(k_deinterleave, dim3(nblk(n)), dim3(TPB), 0, st, v, n, vx, vy);
            hipLaunchKernelGGL(k_cells_xy, dim3(nblk(n)), dim3(TPB), 0, st, g, x, cid);
            hipLaunchKernelGGL(k_iota, dim3(nblk(n)), dim3(TPB), 0, st, pid, n);
        }
        hipLaunchKernelGGL(k_fill_int, dim3(nblk(Nc + 1)), dim3(TPB), 0, st, cnt, Nc + 1, 0);
        if (n > 0) hipLaunchKernelGGL(k_hist, dim3(bin_blocks(n)), dim3(TPB), 0, st, cid, n, cnt);
        hipLaunchKernelGGL(k_scan_tiles, dim3(nTiles), dim3(TPB), 0, st, cnt, Nc + 1, start, tileSum,
                           (const int*)nullptr, cursor);
        hipLaunchKernelGGL(k_scan_add_prefix, dim3(nblk(Nc + 1)), dim3(TPB), 0, st, start, Nc + 1, tileSum);
        if (n > 0) {
            hipLaunchKernelGGL(k_scatter, dim3(bin_blocks(n)), dim3(TPB), 0, st, cid, pid, n, start, cursor, key);
            hipLaunchKernelGGL(k_segsort, dim3(nblk(Nc, SEG_CELLS)), dim3(SEG_CELLS), 0, st, start, Nc, key, list,
                               (int)(n < (1 << 27)), 0);
        }
        const double mean = (double)n / (double)std::max(Nc, 1) * CELL_BLK;
        const int cap = (int)std::min(4096.0, std::max(256.0, 1.15 * mean + 2.0 * CELL_BLK));
        FieldsOut f{out->N, out->Ux, out->Uy, out->Px, out->Py, out->T, out->rho, out->Pkin, out->Pvir, out->P};
        const double vc = (p->Lx / (double)p->Nx) * (p->Ly / (double)p->Ny);
        hipLaunchKernelGGL(k_cell_fields, dim3(nblk(Nc, CELL_BLK)), dim3(CELL_BLK), cap * 2 * sizeof(double), st,
                           start, list, Nc, vx, vy, vc, p->Kvirial, rhoTarget, f, cap);
        if (out->rhoTarget)
            hipLaunchKernelGGL(k_fill_double, dim3(nblk(Nc)), dim3(TPB), 0, st, out->rhoTarget, Nc, rhoTarget);
        HIP_OK(hipGetLastError());
    });
}

int mpcd_gpu_build_zone_layout(const mpcd_params* p, int mode, int32_t* ownerMap, int32_t* nZones, int32_t* rects4) {
    return guarded([&] {
        validate(*p);
        std::vector<int32_t> own;
        const auto r = layout_rects(*p, mode, ownerMap ? &own : nullptr);
        if (ownerMap) std::memcpy(ownerMap, own.data(), own.size() * sizeof(int32_t));
        if (nZones) *nZones = (int32_t)r.size();
        if (rects4)
            for (size_t k = 0; k < r.size(); ++k) {
                rects4[4 * k] = r[k].xs;
                rects4[4 * k + 1] = r[k].xe;
                rects4[4 * k + 2] = r[k].ys;
                rects4[4 * k + 3] = r[k].ye;
            }
    });
}

/* One zone's kernel on the scratch simulation: the rectangle [xs,xe) x
 * [ys,ye), restricted to the cells of mask (Nc bytes, host) when mask is
 * given.  The caller holds g_cache_mu. */
static void run_zone_kernel_on(const mpcd_params* p, double* x, double* v, int xs, int xe, int ys, int ye,
                               const unsigned char* mask, uint64_t rng_seed, double* stats6, int32_t* moved2,
                               double* diag10) {
    {
        auto& s = scratch_sim(*p);
        s.set_state(x, v);
        s.bin_unshifted();
        /* the temporaries are freed, and the simulation's diagnostics record
         * detached, however the call ends */
        struct Detach {
            mpcd_gpu_sim& s;
            ~Detach() {
                s.zdCell = nullptr;
                s.zdCnt = nullptr;
            }
        } detach{s};
        DevBuf<unsigned char> zm;
        DevBuf<ZoneRect> dr(1);
        DevBuf<double> zdc;
        DevBuf<int> zdn, low;
        if (mask) {
            zm.p = dalloc<unsigned char>((size_t)s.Nc);
            HIP_OK(hipMemcpyAsync(zm.p, mask, (size_t)s.Nc, hipMemcpyHostToDevice, s.stream));
        }
        const unsigned char* zmask = zm.p;
        ZoneRect r{xs, xe, ys, ye};
        const int C = (xe - xs) * (ye - ys);
        ZoneRect* d = dr.p;
        HIP_OK(hipMemcpyAsync(d, &r, sizeof(r), hipMemcpyHostToDevice, s.stream));
        s.clear_err();
        if (diag10) {
            /* particle sums before; the kernel's per-cell record starts at zero */
            const int nb = (int)nblk(s.n);
            hipLaunchKernelGGL(k_sum_ke, dim3(nb), dim3(TPB), 0, s.stream, s.vx, s.vy, s.n, s.part);
            hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(TPB), 0, s.stream, s.part, nb, 3, s.diagDev + DIAG_ZK_KE0);
            zdc.p = dalloc<double>(4 * (size_t)C);
            zdn.p = dalloc<int>((size_t)C);
            s.zdCell = zdc.p;
            s.zdCnt = zdn.p;
            hipLaunchKernelGGL(k_fill_double, dim3(nblk(4 * (size_t)C)), dim3(TPB), 0, s.stream, s.zdCell, 4 * C, 0.0);
            s.check_launch();
        }
        s.zone_pass(rng_seed, d, 1, C, xe - xs, nullptr, nullptr, nullptr, 0, 1, nullptr, nullptr, false, zmask);
        if (diag10) {
            /* ZoneKernelStats.diag: occupancy of the final counts (low maps from
             * the sparse phase's start counts), particle sums after, cell sums */
            s.bin_unshifted();
            const int nb = (int)nblk(s.n);
            hipLaunchKernelGGL(k_sum_ke, dim3(nb), dim3(TPB), 0, s.stream, s.vx, s.vy, s.n, s.part);
            hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(TPB), 0, s.stream, s.part, nb, 3, s.diagDev + DIAG_ZK_KE1);
            low.p = dalloc<int>((size_t)s.Nc);
            hipLaunchKernelGGL(k_zone_low_counts, dim3(nblk(s.Nc)), dim3(TPB), 0, s.stream, s.bin.cnt, s.Nc, p->Ny, r,
                               s.zdCnt, low.p);
            s.check_launch();
            s.occupancy(s.bin.cnt, DIAG_ZK_OCC, low.p);
            hipLaunchKernelGGL(k_zone_diag_cells, dim3(1), dim3(TPB), 0, s.stream, s.zdCell, s.bin.cnt, p->Ny, r,
                               s.diagDev + DIAG_ZK_CELLS);
            hipLaunchKernelGGL(k_zone_diag_final, dim3(1), dim3(64), 0, s.stream, s.diagDev + DIAG_ZK_OCC,
                               s.diagDev + DIAG_ZK_KE0, s.diagDev + DIAG_ZK_KE1, s.diagDev + DIAG_ZK_CELLS, s.Nc,
                               s.diagDev + DIAG_ZK_OUT);
            s.check_launch();
            HIP_OK(hipMemcpyAsync(diag10, s.diagDev + DIAG_ZK_OUT, 10 * sizeof(double), hipMemcpyDeviceToHost,
                                  s.stream));
            HIP_OK(hipStreamSynchronize(s.stream));
        }
        ZoneOut o{};
        HIP_OK(hipMemcpyAsync(&o, s.zout, sizeof(o), hipMemcpyDeviceToHost, s.stream));
        s.get_state(x, v);
        s.check_err();
        if (stats6) {
            stats6[0] = (double)o.nHigh;
            stats6[1] = (double)o.nLow;
            stats6[2] = (double)(o.movedDense + o.movedSparse);
            stats6[3] = (double)(o.movedDense + o.movedSparse);
            stats6[4] = o.stats4;
            stats6[5] = o.stats5;
        }
        if (moved2) {
            moved2[0] = o.movedDense;
            moved2[1] = o.movedSparse;
        }
    }
}

int mpcd_gpu_run_zone_kernel(const mpcd_params* p, double* x, double* v, int xs, int xe, int ys, int ye,
                             uint64_t rng_seed, double* stats6, int32_t* moved2, double* diag10) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        if (xs < 0 || ys < 0 || xe > p->Nx || ye > p->Ny || xs >= xe || ys >= ye)
            throw ApiError(MPCD_E_INVALID, "zone rectangle out of range");
        run_zone_kernel_on(p, x, v, xs, xe, ys, ye, nullptr, rng_seed, stats6, moved2, diag10);
    });
}

int mpcd_gpu_run_zone_kernel_mask(const mpcd_params* p, double* x, double* v, const uint8_t* zone_mask,
                                  uint64_t rng_seed, double* stats6, int32_t* moved2, double* diag10) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        validate(*p);
        if (!zone_mask) throw ApiError(MPCD_E_INVALID, "zone mask is NULL");
        int xs = p->Nx, xe = -1, ys = p->Ny, ye = -1;
        long long nIn = 0;
        for (int c = 0; c < p->Nc; ++c) {
            if (!zone_mask[c]) continue;
            const int ix = c / p->Ny, iy = c - ix * p->Ny;
            xs = std::min(xs, ix);
            xe = std::max(xe, ix + 1);
            ys = std::min(ys, iy);
            ye = std::max(ye, iy + 1);
            ++nIn;
        }
        if (nIn == 0) throw ApiError(MPCD_E_INVALID, "the zone mask selects no cell");
        /* a mask that fills its bounding rectangle is that rectangle */
        const bool rect = nIn == (long long)(xe - xs) * (long long)(ye - ys);
        run_zone_kernel_on(p, x, v, xs, xe, ys, ye, rect ? nullptr : zone_mask, rng_seed, stats6, moved2, diag10);
    });
}

int mpcd_gpu_zone_pass(const mpcd_params* p, double* x, double* v, int layoutMode, uint64_t rngSeedBase,
                       mpcd_zone_pass_metrics* m, int32_t* ownerMap) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto& s = scratch_sim(*p);
        s.check_pass_modes();
        s.set_state(x, v);
        s.bin_unshifted();
        std::vector<int32_t> own;
        const auto r = layout_rects(*p, layoutMode, &own);
        if (layoutMode != MPCD_LAYOUT_BASE && layoutMode != MPCD_LAYOUT_SHIFTED && layoutMode != MPCD_LAYOUT_SINGLE_FULL)
            throw ApiError(MPCD_E_INVALID, "unknown layout mode");
        int maxC = 0, maxW = 0;
        for (const auto& q : r) {
            maxC = std::max(maxC, (q.xe - q.xs) * (q.ye - q.ys));
            maxW = std::max(maxW, q.xe - q.xs);
        }
        std::vector<int> cov((size_t)p->Nc, 0);
        for (const auto& q : r)
            for (int ix = q.xs; ix < q.xe; ++ix)
                for (int iy = q.ys; iy < q.ye; ++iy) cov[(size_t)(iy + p->Ny * ix)]++;
        std::vector<unsigned char> ov((size_t)p->Nc, 0);
        std::vector<int> oc;
        for (int c = 0; c < p->Nc; ++c)
            if (cov[(size_t)c] > 1) {
                ov[(size_t)c] = 1;
                oc.push_back(c);
            }
        DevBuf<ZoneRect> db(r.size());
        DevBuf<int> dOwnb(own.size()), dOverb;
        DevBuf<unsigned char> dCovb;
        if (!oc.empty()) {
            dCovb.p = dalloc<unsigned char>(ov.size());
            dOverb.p = dalloc<int>(oc.size());
        }
        ZoneRect* d = db.p;
        int* dOwn = dOwnb.p;
        unsigned char* dCov = dCovb.p;
        int* dOver = dOverb.p;
        HIP_OK(hipMemcpyAsync(d, r.data(), r.size() * sizeof(ZoneRect), hipMemcpyHostToDevice, s.stream));
        HIP_OK(hipMemcpyAsync(dOwn, own.data(), own.size() * sizeof(int), hipMemcpyHostToDevice, s.stream));
        if (dCov) {
            HIP_OK(hipMemcpyAsync(dCov, ov.data(), ov.size(), hipMemcpyHostToDevice, s.stream));
            HIP_OK(hipMemcpyAsync(dOver, oc.data(), oc.size() * sizeof(int), hipMemcpyHostToDevice, s.stream));
        }
        s.ensure_zout((int)r.size());
        s.clear_err();
        s.zone_pass(rngSeedBase, d, (int)r.size(), maxC, maxW, dOwn, dCov, dOver, (int)oc.size(), 0);
        s.check_err();
        mpcd_zone_pass_metrics mm{};
        s.pass_metrics((int)r.size(), &mm);
        s.get_state(x, v);
        if (m) *m = mm;
        if (ownerMap) std::memcpy(ownerMap, own.data(), own.size() * sizeof(int32_t));
    });
}

int mpcd_gpu_sim_set_surrogate(mpcd_gpu_sim* s, mpcd_gpu_net* net, int K, double outScale, double duCap) {
    return guarded([&] {
        if (!s) throw ApiError(MPCD_E_INVALID, "null simulation");
        if (net && K > 0) {
            int nx, ny, px, py;
            net_grid(net, &nx, &ny, &px, &py);
            const int sx = s->p.bc_left == MPCD_BC_PERIODIC && s->p.bc_right == MPCD_BC_PERIODIC;
            const int sy = s->p.bc_bottom == MPCD_BC_PERIODIC && s->p.bc_top == MPCD_BC_PERIODIC;
            if (nx != s->p.Nx || ny != s->p.Ny || px != sx || py != sy)
                throw ApiError(MPCD_E_INVALID, "surrogate grid does not match the simulation");
            if (!(outScale > 0.0) || !(duCap > 0.0)) throw ApiError(MPCD_E_INVALID, "outScale and duCap must be > 0");
        }
        s->sur = (K > 0) ? net : nullptr;
        s->surK = K;
        if (outScale > 0.0) s->surScale = outScale;
        s->surCap = duCap;
    });
}

int mpcd_gpu_sim_set_surrogate_check(mpcd_gpu_sim* s, double occTol, double divTol) {
    return guarded([&] {
        if (!s) throw ApiError(MPCD_E_INVALID, "null simulation");
        if (!(occTol > 0.0) || !(divTol > 0.0)) throw ApiError(MPCD_E_INVALID, "tolerances must be > 0");
        s->surOccTol = occTol;
        s->surDivTol = divTol;
    });
}

int mpcd_gpu_sim_row_profile(mpcd_gpu_sim* s, int enable, double* rows3, int32_t* steps) {
    return guarded([&] {
        if (!s) throw ApiError(MPCD_E_INVALID, "null simulation");
        if (3 * (size_t)s->p.Ny * sizeof(double) > 60 * 1024)
            throw ApiError(MPCD_E_INVALID, "row profiles support at most 2560 rows");
        if (rows3 || steps) {
            HIP_OK(hipStreamSynchronize(s->stream));
            if (rows3) {
                if (s->rowAcc)
                    HIP_OK(hipMemcpy(rows3, s->rowAcc, 3 * (size_t)s->p.Ny * sizeof(double), hipMemcpyDeviceToHost));
                else
                    std::fill(rows3, rows3 + 3 * (size_t)s->p.Ny, 0.0);
            }
            if (steps) *steps = s->rowSteps < 0 ? 0 : s->rowSteps;
        }
        if (enable > 0) {
            if (!s->rowAcc) s->rowAcc = dalloc<double>(3 * (size_t)s->p.Ny);
            hipLaunchKernelGGL(k_fill_double, dim3(nblk(3 * s->p.Ny)), dim3(TPB), 0, s->stream, s->rowAcc, 3 * s->p.Ny,
                               0.0);
            HIP_OK(hipGetLastError());
            s->rowSteps = 0;
        } else if (enable == 0) {
            s->rowSteps = -1;
        }
    });
}

int mpcd_gpu_sim_surrogate_counts(mpcd_gpu_sim* s, int32_t* counts3) {
    return guarded([&] {
        if (!s || !counts3) throw ApiError(MPCD_E_INVALID, "null argument");
        double h[SUR_SLOTS];
        HIP_OK(hipMemcpyAsync(h, s->surDev, sizeof(h), hipMemcpyDeviceToHost, s->stream));
        HIP_OK(hipStreamSynchronize(s->stream));
        counts3[0] = (int32_t)h[SUR_N_LEARNED];
        counts3[1] = s->nPlain;
        counts3[2] = (int32_t)h[SUR_N_FALLBACK];
    });
}

int mpcd_gpu_sim_closure_sample(mpcd_gpu_sim* s, int enable, double outScale, float* feat, float* target) {
    return guarded([&] {
        if (!s) throw ApiError(MPCD_E_INVALID, "null simulation");
        if (feat || target) {
            HIP_OK(hipStreamSynchronize(s->stream));
            if (feat)
                HIP_OK(hipMemcpy(feat, s->featBuf, 4 * (size_t)s->Nc * sizeof(float), hipMemcpyDeviceToHost));
            if (target)
                HIP_OK(hipMemcpy(target, s->tgtBuf, 2 * (size_t)s->Nc * sizeof(float), hipMemcpyDeviceToHost));
        }
        s->wantSample = enable != 0;
        if (outScale > 0.0) s->surScale = outScale;
    });
}

int mpcd_gpu_sim_closure_snapshot(mpcd_gpu_sim* s, int enable, const mpcd_cell_fields* ref,
                                  const mpcd_cell_fields* red, const mpcd_closure_grids* grids, double* beta,
                                  int32_t* step) {
    return guarded([&] {
        if (!s) throw ApiError(MPCD_E_INVALID, "null simulation");
        const size_t C = (size_t)s->Nc;
        if (enable && !s->snapF) {
            s->snapN = dalloc<int>(2 * C);
            s->snapF = dalloc<double>(24 * C + 8);
        }
        if (ref || red || grids || beta || step) {
            HIP_OK(hipStreamSynchronize(s->stream));
            if (step) *step = s->snapStep;
            if (s->snapStep >= 0) {
                const mpcd_cell_fields* cf[2] = {ref, red};
                for (int k = 0; k < 2; ++k) {
                    if (!cf[k]) continue;
                    if (cf[k]->N)
                        HIP_OK(hipMemcpy(cf[k]->N, s->snapN + k * C, C * sizeof(int), hipMemcpyDeviceToHost));
                    double* hs[9] = {cf[k]->Ux,  cf[k]->Uy,   cf[k]->Px,   cf[k]->Py, cf[k]->T,
                                     cf[k]->rho, cf[k]->Pkin, cf[k]->Pvir, cf[k]->P};
                    for (int j = 0; j < 9; ++j)
                        if (hs[j])
                            HIP_OK(hipMemcpy(hs[j], s->snapF + (size_t)(9 * k + j) * C, C * sizeof(double),
                                             hipMemcpyDeviceToHost));
                    if (cf[k]->rhoTarget)
                        for (size_t c = 0; c < C; ++c) cf[k]->rhoTarget[c] = s->rho_target();
                }
                if (grids) {
                    double* hg[6] = {grids->pdriveRaw, grids->pdrive, grids->duRepairX,
                                     grids->duRepairY, grids->duEOSX, grids->duEOSY};
                    for (int j = 0; j < 6; ++j)
                        if (hg[j])
                            HIP_OK(hipMemcpy(hg[j], s->snapF + (size_t)(18 + j) * C, C * sizeof(double),
                                             hipMemcpyDeviceToHost));
                    if (grids->interzoneMask) /* a constant of the layouts */
                        HIP_OK(hipMemcpy(grids->interzoneMask, s->mask, C * sizeof(int), hipMemcpyDeviceToHost));
                }
                if (beta) HIP_OK(hipMemcpy(beta, s->snapF + 24 * C, sizeof(double), hipMemcpyDeviceToHost));
            }
        }
        s->wantSnap = enable != 0;
    });
}

int mpcd_gpu_closure_training_pair(const mpcd_params* p, const int32_t* refN, const double* refUx,
                                   const double* refUy, const double* refT, const double* duRepairX,
                                   const double* duRepairY, const double* duEOSX, const double* duEOSY, double beta,
                                   double outScale, float* feat, float* target) {
    return guarded([&] {
        if (!p || !refN || !refUx || !refUy || !refT || !duRepairX || !duRepairY || !duEOSX || !duEOSY)
            throw ApiError(MPCD_E_INVALID, "null argument");
        if (p->Nx <= 0 || p->Ny <= 0 || p->Nc != p->Nx * p->Ny) throw ApiError(MPCD_E_INVALID, "bad grid size");
        if (!(outScale > 0.0)) throw ApiError(MPCD_E_INVALID, "outScale must be positive");
        const Geom g = make_geom(*p);
        const size_t C = (size_t)p->Nc;
        DevBuf<int> dNb(C);
        DevBuf<double> db(8 * C + 8);
        DevBuf<float> fb(6 * C);
        int* dN = dNb.p;
        double* d = db.p;
        float* f = fb.p;
        hipStream_t st = nullptr;
        HIP_OK(hipMemcpy(dN, refN, C * sizeof(int), hipMemcpyHostToDevice));
        const double* hs[7] = {refUx, refUy, refT, duRepairX, duRepairY, duEOSX, duEOSY};
        for (int k = 0; k < 7; ++k) HIP_OK(hipMemcpy(d + k * C, hs[k], C * sizeof(double), hipMemcpyHostToDevice));
        const double sc[4] = {0.0, 0.0, 0.0, beta};
        HIP_OK(hipMemcpy(d + 8 * C, sc, sizeof(sc), hipMemcpyHostToDevice));
        const int nb = (int)nblk(C);
        hipLaunchKernelGGL(k_features, dim3(nb), dim3(TPB), 0, st, g, dN, d, d + C, d + 2 * C, p->gamma, p->kBT, f);
        hipLaunchKernelGGL(k_closure_target, dim3(nb), dim3(TPB), 0, st, p->Nc, d + 8 * C, d + 3 * C, d + 4 * C,
                           d + 5 * C, d + 6 * C, 1.0 / outScale, f + 4 * C);
        HIP_OK(hipGetLastError());
        if (feat) HIP_OK(hipMemcpy(feat, f, 4 * C * sizeof(float), hipMemcpyDeviceToHost));
        if (target) HIP_OK(hipMemcpy(target, f + 4 * C, 2 * C * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int mpcd_gpu_zone_pass_gauss_seidel(const mpcd_params* p, double* x, double* v, int layoutMode,
                                    uint64_t rngSeedBase, mpcd_zone_pass_metrics* m, int32_t* ownerMap) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        if (layoutMode != MPCD_LAYOUT_BASE && layoutMode != MPCD_LAYOUT_SHIFTED && layoutMode != MPCD_LAYOUT_SINGLE_FULL)
            throw ApiError(MPCD_E_INVALID, "unknown layout mode");
        auto& s = scratch_sim(*p);
        s.check_pass_modes();
        s.set_state(x, v);
        std::vector<int32_t> own;
        const auto r = layout_rects(*p, layoutMode, &own);
        DevBuf<ZoneRect> db(r.size());
        ZoneRect* d = db.p;
        HIP_OK(hipMemcpyAsync(d, r.data(), r.size() * sizeof(ZoneRect), hipMemcpyHostToDevice, s.stream));
        s.ensure_zout(1);
        s.clear_err();
        /* zones one after another, each on the state its predecessors left
         * (zone_pass.cpp:1131-1141): fresh binning, one zone owning all
         * its particles, seed rngSeedBase + 104729 (iz+1) */
        mpcd_zone_pass_metrics mm{};
        mm.nZones = (int)r.size();
        double acc = 0.0, mx = 0.0;
        int cnt = 0;
        for (size_t iz = 0; iz < r.size(); ++iz) {
            const ZoneRect q = r[iz];
            s.bin_unshifted();
            s.zone_pass(rngSeedBase + 104729ULL * (uint64_t)(iz + 1), d + iz, 1, (q.xe - q.xs) * (q.ye - q.ys),
                        q.xe - q.xs, nullptr, nullptr, nullptr, 0, 1);
            ZoneOut o{};
            HIP_OK(hipMemcpyAsync(&o, s.zout, sizeof(o), hipMemcpyDeviceToHost, s.stream));
            HIP_OK(hipStreamSynchronize(s.stream));
            mm.nParticlesMovedDense += o.movedDense;
            mm.nParticlesMovedSparse += o.movedSparse;
            if (o.stats4 > 0.0) {
                acc += o.stats4 * o.stats4;
                ++cnt;
            }
            mx = std::max(mx, o.stats5);
        }
        s.check_err();
        mm.corrRmsDU = cnt > 0 ? std::sqrt(acc / (double)cnt) : 0.0;
        mm.corrMaxDU = mx;
        s.get_state(x, v);
        if (m) *m = mm;
        if (ownerMap) std::memcpy(ownerMap, own.data(), own.size() * sizeof(int32_t));
    });
}

int mpcd_gpu_liquid_closure(const mpcd_params* p, const double* xRef, const double* vRef, const double* xRed,
                            const double* vRed, double* xOut, double* vOut, mpcd_closure_metrics* m,
                            mpcd_closure_grids* grids) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto& s = scratch_sim(*p);
        if (!p->useLiquidClosure) throw ApiError(MPCD_E_INVALID, "run_liquid_closure requires useLiquidClosure=true");
        s.set_state(xRef, vRef);
        s.bin_unshifted();
        s.ref_fields();
        s.set_state(xRed, vRed);
        s.bin_unshifted();
        mpcd_closure_metrics cm{};
        s.closure(&cm);
        if (grids) {
            const size_t nb = (size_t)s.Nc * sizeof(double);
            const double* src[6] = {s.pdriveRaw, s.pdrive, s.duRX, s.duRY, s.duEX, s.duEY};
            double* dst[6] = {grids->pdriveRaw, grids->pdrive, grids->duRepairX, grids->duRepairY, grids->duEOSX,
                              grids->duEOSY};
            for (int k = 0; k < 6; ++k)
                if (dst[k]) HIP_OK(hipMemcpyAsync(dst[k], src[k], nb, hipMemcpyDeviceToHost, s.stream));
            if (grids->interzoneMask)
                HIP_OK(hipMemcpyAsync(grids->interzoneMask, s.mask, (size_t)s.Nc * sizeof(int), hipMemcpyDeviceToHost,
                                      s.stream));
            HIP_OK(hipStreamSynchronize(s.stream));
        }
        if (grids && grids->vRepaired) s.download_repaired(grids->vRepaired);
        s.get_state(xOut, vOut);
        if (m) *m = cm;
    });
}

}  // extern "C"

/* batched ensembles (kernels and the C ABI) */
#include "mpcd_ensemble.h"
