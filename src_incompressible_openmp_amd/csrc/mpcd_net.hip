/*
 * mpcd_net.hip -- the closure surrogate: a 4-layer 3x3 convolutional network
 * over the cell grid (4 -> 32 -> 32 -> 32 -> 2 channels, ReLU after the first
 * three layers) that maps the red state's cell features to the liquid
 * closure's per-cell velocity correction.  It has no counterpart in the
 * reference; it stands in for the run_liquid_closure call of the benchmark
 * driver (main_benchmark_poiseuille_openmp_profiled_shifted_profiled.cpp:277)
 * on the steps the plain closure skips.
 *
 * Every convolution is an implicit GEMM on gfx950's bf16 matrix cores
 * (v_mfma_f32_16x16x32_bf16): M = 16 cells per wavefront, N = output
 * channels, K = 3x3 taps x input channels; weights and activations are
 * bfloat16, sums are float32.  The input-gradient pass is the same kernel
 * on tap-flipped, transposed weights; the weight gradient is a GEMM over
 * cells (K = cells) from LDS-transposed tiles, reduced in a fixed order so
 * training is deterministic.
 *
 * Layouts: grids of B samples, cell index iy + Ny*ix; activations
 * [cell][pitch] bfloat16 (pitch 8 for the 4 features and the 2-channel
 * output gradient, else 32) plus a transposed copy [channel][cell] for the
 * weight gradient; weights W[co][tap][ci], tap = 3*(dx+1) + (dy+1).
 * The x axis pads periodically when both x walls are periodic, else with
 * zeros; likewise y.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "mpcd_gpu.h"
#include "mpcd_hip_util.h"
#include "mpcd_rng.h"

using namespace mpcdk;

namespace {

typedef unsigned short u16; /* raw bfloat16 bits */
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NTPB = 256;          /* 4 wavefronts */
constexpr int NLAYERS = 4;
constexpr int CH[5] = {4, 32, 32, 32, 2};
constexpr int WG_PITCH = 128;      /* cells per weight-gradient block */

__host__ __device__ inline u16 f2bf(float f) {
    /* round to nearest even, as torch's float -> bfloat16 */
    unsigned u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (u16)0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (u16)(u >> 16);
}
__host__ __device__ inline float bf2f(u16 h) {
    const unsigned u = (unsigned)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

__device__ inline bf16x8 ld8(const u16* p) {
    const uint4 v = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, v);
}
__device__ inline bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0u, 0u, 0u, 0u)); }

struct Grid {
    int B, Nx, Ny, Nc, perx, pery;
    long long total; /* B * Nc */
    long long tpitch; /* pitch of the transposed copies: total rounded up to 16 */
};

/* neighbour (dx, dy) of cell index q, or -1 outside the padded domain */
__device__ inline long long neighbour(const Grid& g, long long q, int dx, int dy) {
    const long long s = q / g.Nc;
    const int c = (int)(q - s * g.Nc);
    int jx = c / g.Ny + dx, jy = c % g.Ny + dy;
    if (jx < 0 || jx >= g.Nx) {
        if (!g.perx) return -1;
        jx = (jx + g.Nx) % g.Nx;
    }
    if (jy < 0 || jy >= g.Ny) {
        if (!g.pery) return -1;
        jy = (jy + g.Ny) % g.Ny;
    }
    return s * g.Nc + jy + (long long)g.Ny * jx;
}

enum { EPI_RELU = 0, EPI_LINEAR = 1, EPI_MASK = 2 };

struct ConvArgs {
    Grid g;
    const u16* in;     /* [cell][CIN] */
    const u16* w;      /* [16*NB][9][CIN] */
    const float* bias; /* [coReal] (RELU, LINEAR) */
    int coReal;
    u16* out;          /* [cell][32] (RELU, MASK) */
    u16* outT;         /* [32][tpitch] or nullptr */
    const u16* mask;   /* [cell][32] (MASK) */
    float* outf;       /* [cell][2] (LINEAR) */
};

/* One 3x3 convolution as an implicit GEMM.  Each wavefront owns 16 cells
 * (rows) x 16*NB output channels; K runs over (tap, 8-channel block) pairs,
 * 4 per MFMA: with CIN = 32 one tap per step, with CIN = 8 four taps.  The
 * weight fragments are loaded once into registers. */
template <int CIN, int NB, int EPI>
__global__ void __launch_bounds__(NTPB) k_conv(ConvArgs a) {
    constexpr int BPT = CIN / 8;                /* 8-channel blocks per tap */
    constexpr int NS = (9 * BPT + 3) / 4;       /* MFMA steps */
    __shared__ __align__(16) u16 tile[4][16][32];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int row = lane & 15, kb = lane >> 4;
    bf16x8 bw[NS][NB];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int gblk = 4 * s + kb, tap = gblk / BPT, chb = gblk % BPT;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
            bw[s][nb] = (tap < 9) ? ld8(a.w + ((size_t)(16 * nb + row) * 9 + tap) * CIN + chb * 8) : zero8();
    }
    const Grid& g = a.g;
    for (long long t0 = (long long)blockIdx.x * 64; t0 < g.total; t0 += (long long)gridDim.x * 64) {
        const long long cell0 = t0 + 16 * wv;
        const long long q = cell0 + row;
        const bool valid = q < g.total;
        f32x4 acc[NB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int gblk = 4 * s + kb, tap = gblk / BPT, chb = gblk % BPT;
            bf16x8 av = zero8();
            if (valid && tap < 9) {
                const long long nbq = neighbour(g, q, tap / 3 - 1, tap % 3 - 1);
                if (nbq >= 0) av = ld8(a.in + nbq * CIN + chb * 8);
            }
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bw[s][nb], acc[nb], 0, 0, 0);
        }
        /* D[i][n] for i = 4*kb + r, n = row (+16 nb) */
        if (EPI == EPI_LINEAR) {
            if (row < a.coReal) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const long long qi = cell0 + 4 * kb + r;
                    if (qi < g.total) a.outf[qi * 2 + row] = acc[0][r] + a.bias[row];
                }
            }
            continue;
        }
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int n = 16 * nb + row;
                float v = acc[nb][r];
                if (EPI == EPI_RELU) v = fmaxf(v + a.bias[n], 0.0f);
                tile[wv][4 * kb + r][n] = f2bf(v);
            }
        __syncthreads();
        {
            /* row-major store: lane -> cell lane/4, channels 8*(lane%4).. */
            const int ci = lane >> 2, cb = (lane & 3) * 8;
            const long long qi = cell0 + ci;
            if (qi < g.total) {
                uint4 v = *reinterpret_cast<const uint4*>(&tile[wv][ci][cb]);
                if (EPI == EPI_MASK) {
                    const uint4 m = *reinterpret_cast<const uint4*>(a.mask + qi * 32 + cb);
                    const unsigned mw[4] = {m.x, m.y, m.z, m.w};
                    unsigned vw[4] = {v.x, v.y, v.z, v.w};
                    for (int k = 0; k < 4; ++k) {
                        /* keep each half where the forward activation is > 0 */
                        const unsigned lo = (bf2f((u16)(mw[k] & 0xffffu)) > 0.0f) ? 0x0000ffffu : 0u;
                        const unsigned hi = (bf2f((u16)(mw[k] >> 16)) > 0.0f) ? 0xffff0000u : 0u;
                        vw[k] &= (lo | hi);
                    }
                    v = make_uint4(vw[0], vw[1], vw[2], vw[3]);
                    *reinterpret_cast<uint4*>(&tile[wv][ci][cb]) = v; /* masked values for the transposed copy */
                }
                *reinterpret_cast<uint4*>(a.out + qi * 32 + cb) = v;
            }
        }
        __syncthreads();
        if (a.outT) {
            /* transposed store: lane -> channel lane/2, cells 8*(lane%2).. */
            const int ch = lane >> 1, c8 = (lane & 1) * 8;
            u16 h[8];
            for (int k = 0; k < 8; ++k) h[k] = tile[wv][c8 + k][ch];
            if (cell0 + c8 + 8 <= g.tpitch) {
                uint4 v;
                v.x = (unsigned)h[0] | ((unsigned)h[1] << 16);
                v.y = (unsigned)h[2] | ((unsigned)h[3] << 16);
                v.z = (unsigned)h[4] | ((unsigned)h[5] << 16);
                v.w = (unsigned)h[6] | ((unsigned)h[7] << 16);
                *reinterpret_cast<uint4*>(a.outT + (size_t)ch * g.tpitch + cell0 + c8) = v;
            }
        }
        __syncthreads();
    }
}

/* Weight and bias gradient of one layer: dW[co][tap][ci] = sum over cells of
 * g[cell][co] * in[neighbour(cell, tap)][ci], db[co] = sum of g[cell][co].
 * A block takes WG_PITCH cells at a time: the output gradient (transposed)
 * once, then per tap the gathered input (transposed) into LDS; wavefront w
 * owns the 16x16 output block (ci block w>>1, co block w&1) for all taps.
 * Per-block partials go to part[block][32][9][32] and bpart[block][32]. */
template <int CIN, int NB>
__global__ void __launch_bounds__(NTPB) k_conv_wgrad(Grid g, const u16* __restrict__ inT, const u16* __restrict__ gT,
                                                     int gRows, float* __restrict__ part, float* __restrict__ bpart) {
    __shared__ __align__(16) u16 sg[32][WG_PITCH];
    __shared__ __align__(16) u16 si[32][WG_PITCH];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int row = lane & 15, kb = lane >> 4;
    const int mb = wv >> 1, nb = wv & 1;
    const bool active = (16 * mb < CIN) && (nb < NB);
    f32x4 acc[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    float bsum = 0.0f; /* threads 0..31: co */
    const long long nblkTot = (g.total + WG_PITCH - 1) / WG_PITCH;
    for (long long bk = blockIdx.x; bk < nblkTot; bk += gridDim.x) {
        const long long c0 = bk * WG_PITCH;
        for (int e = threadIdx.x; e < 32 * WG_PITCH; e += NTPB) {
            const int co = e / WG_PITCH, k = e % WG_PITCH;
            const long long q = c0 + k;
            sg[co][k] = (co < gRows && q < g.total) ? gT[(size_t)co * g.tpitch + q] : (u16)0;
        }
        __syncthreads();
        if (threadIdx.x < 32) {
            float s = 0.0f;
            for (int k = 0; k < WG_PITCH; ++k) s += bf2f(sg[threadIdx.x][k]);
            bsum += s;
        }
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int dx = tap / 3 - 1, dy = tap % 3 - 1;
            for (int e = threadIdx.x; e < CIN * WG_PITCH; e += NTPB) {
                const int ci = e / WG_PITCH, k = e % WG_PITCH;
                const long long q = c0 + k;
                u16 v = 0;
                if (q < g.total) {
                    const long long nq = neighbour(g, q, dx, dy);
                    if (nq >= 0) v = inT[(size_t)ci * g.tpitch + nq];
                }
                si[ci][k] = v;
            }
            __syncthreads();
            if (active) {
#pragma unroll
                for (int ks = 0; ks < WG_PITCH / 32; ++ks) {
                    const int kk = 32 * ks + 8 * kb;
                    const int ci = 16 * mb + row;
                    const bf16x8 av = (ci < CIN) ? __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&si[ci][kk]))
                                                 : zero8();
                    const bf16x8 bv = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&sg[16 * nb + row][kk]));
                    acc[tap] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[tap], 0, 0, 0);
                }
            }
            __syncthreads();
        }
    }
    float* pb = part + (size_t)blockIdx.x * (32 * 9 * 32);
    if (active) {
#pragma unroll
        for (int tap = 0; tap < 9; ++tap)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int ci = 16 * mb + 4 * kb + r, co = 16 * nb + row;
                pb[((size_t)co * 9 + tap) * 32 + ci] = acc[tap][r];
            }
    }
    if (threadIdx.x < 32) bpart[(size_t)blockIdx.x * 32 + threadIdx.x] = bsum;
}

/* Sums the per-block partials in block order into the flat gradient of
 * one layer (W[co][tap][ci] for the real channels, then b[co]). */
__global__ void k_wgrad_reduce(const float* __restrict__ part, const float* __restrict__ bpart, int nparts, int co,
                               int ci, double* __restrict__ grad) {
    const int nW = co * 9 * ci;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nW + co) return;
    double s = 0.0;
    if (e < nW) {
        const int o = e / (9 * ci), t = (e / ci) % 9, i = e % ci;
        for (int b = 0; b < nparts; ++b) s += (double)part[(size_t)b * (32 * 9 * 32) + ((size_t)o * 9 + t) * 32 + i];
    } else {
        for (int b = 0; b < nparts; ++b) s += (double)bpart[(size_t)b * 32 + (e - nW)];
    }
    grad[e] = s;
}

/* features (float [cell][4]) -> bfloat16 [cell][8] and its transpose */
__global__ void k_pack_features(Grid g, const float* __restrict__ f, u16* __restrict__ x, u16* __restrict__ xT) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.tpitch) return;
    for (int k = 0; k < 8; ++k) {
        const u16 h = (q < g.total && k < 4) ? f2bf(f[q * 4 + k]) : (u16)0;
        if (q < g.total) x[q * 8 + k] = h;
        xT[(size_t)k * g.tpitch + q] = h;
    }
}

/* squared error, its gradient 2 (pred - target) / count in bfloat16
 * ([cell][8] and transposed), and per-block loss partials */
__global__ void __launch_bounds__(NTPB) k_loss(Grid g, const float* __restrict__ pred, const float* __restrict__ tgt,
                                               double inv, u16* __restrict__ gr, u16* __restrict__ grT,
                                               double* __restrict__ lpart) {
    __shared__ double red[NTPB];
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (q < g.tpitch) {
        for (int k = 0; k < 8; ++k) {
            u16 h = 0;
            if (q < g.total && k < 2) {
                const double d = (double)pred[q * 2 + k] - (double)tgt[q * 2 + k];
                l += d * d;
                h = f2bf((float)(2.0 * d * inv));
            }
            if (q < g.total) gr[q * 8 + k] = h;
            grT[(size_t)k * g.tpitch + q] = h;
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = NTPB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = red[0];
}

/* Adam on the float64 master parameters */
__global__ void k_adam(double* __restrict__ th, double* __restrict__ m, double* __restrict__ v,
                       const double* __restrict__ gr, int n, double lr, double c1, double c2) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double gg = gr[e];
    const double mm = 0.9 * m[e] + 0.1 * gg;
    const double vv = 0.999 * v[e] + 0.001 * gg * gg;
    m[e] = mm;
    v[e] = vv;
    th[e] -= lr * (mm / c1) / (sqrt(vv / c2) + 1e-8);
}

/* bfloat16 forward weights Wf[16*NB][9][CINP] (zero padded) and the
 * transposed, tap-flipped weights of the input-gradient pass
 * Wt[ci (32)][9][COP] with Wt[ci][8-tap][co] = W[co][tap][ci] */
__global__ void k_pack_weights(const double* __restrict__ th, int co, int ci, int coPad, int ciPad, int coPadT,
                               u16* __restrict__ wf, float* __restrict__ bias, u16* __restrict__ wt) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int nF = coPad * 9 * ciPad;
    if (e < nF) {
        const int o = e / (9 * ciPad), t = (e / ciPad) % 9, i = e % ciPad;
        wf[e] = (o < co && i < ci) ? f2bf((float)th[((size_t)o * 9 + t) * ci + i]) : (u16)0;
    }
    if (wt) {
        const int nT = 32 * 9 * coPadT;
        if (e < nT) {
            const int i = e / (9 * coPadT), t = (e / coPadT) % 9, o = e % coPadT;
            wt[e] = (o < co && i < ci) ? f2bf((float)th[((size_t)o * 9 + (8 - t)) * ci + i]) : (u16)0;
        }
    }
    if (e < co) bias[e] = (float)th[(size_t)co * 9 * ci + e];
}

/* one 16x16x32 product with the operands as given, for the layout check */
__global__ void k_mfma_probe(const u16* __restrict__ a, const u16* __restrict__ b, float* __restrict__ d) {
    const int l = threadIdx.x;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld8(a + 8 * l), ld8(b + 8 * l), acc, 0, 0, 0);
    for (int r = 0; r < 4; ++r) d[4 * l + r] = acc[r];
}

inline unsigned nb(long long n, int t = NTPB) { return (unsigned)((n + t - 1) / t); }

}  // namespace

/* ===================================================================== */

struct mpcd_gpu_net {
    Grid g{};
    int maxB = 0;
    hipStream_t stream = nullptr;
    int nParts = 0;
    /* parameters */
    double *theta = nullptr, *mom = nullptr, *vel = nullptr, *grad = nullptr;
    int step = 0;
    u16* wf[NLAYERS] = {};
    u16* wt[NLAYERS] = {};
    float* bias[NLAYERS] = {};
    /* activations, [cell][pitch] and transposed */
    u16 *x = nullptr, *xT = nullptr;
    u16 *a[3] = {}, *aT[3] = {};
    u16 *d[3] = {}, *dT[3] = {};
    u16 *g4 = nullptr, *g4T = nullptr;
    float *pred = nullptr, *feat = nullptr, *tgt = nullptr;
    float *part = nullptr, *bpart = nullptr;
    double* lpart = nullptr;
    /* training set (device) */
    float *setF = nullptr, *setT = nullptr;
    long long setN = 0;

    static int n_params() {
        int n = 0;
        for (int l = 0; l < NLAYERS; ++l) n += CH[l + 1] * 9 * CH[l] + CH[l + 1];
        return n;
    }
    static int offset(int l) {
        int o = 0;
        for (int k = 0; k < l; ++k) o += CH[k + 1] * 9 * CH[k] + CH[k + 1];
        return o;
    }
    static int cinPad(int l) { return l == 0 ? 8 : 32; }
    static int coPad(int l) { return l == 3 ? 16 : 32; }

    mpcd_gpu_net(int Nx, int Ny, int perx, int pery, int maxBatch) {
        if (Nx <= 0 || Ny <= 0 || maxBatch <= 0) throw ApiError(MPCD_E_INVALID, "grid and batch must be positive");
        g.Nx = Nx;
        g.Ny = Ny;
        g.Nc = Nx * Ny;
        g.perx = perx ? 1 : 0;
        g.pery = pery ? 1 : 0;
        maxB = maxBatch;
        HIP_OK(hipStreamCreate(&stream));
        const long long cells = (long long)g.Nc * maxB;
        const long long tp = (cells + 15) / 16 * 16;
        const int np = n_params();
        theta = dalloc<double>(np);
        mom = dalloc<double>(np);
        vel = dalloc<double>(np);
        grad = dalloc<double>(np);
        for (int l = 0; l < NLAYERS; ++l) {
            wf[l] = dalloc<u16>((size_t)coPad(l) * 9 * cinPad(l));
            wt[l] = dalloc<u16>((size_t)32 * 9 * (l == 3 ? 8 : 32));
            bias[l] = dalloc<float>(32);
        }
        x = dalloc<u16>((size_t)cells * 8);
        xT = dalloc<u16>((size_t)8 * tp);
        for (int k = 0; k < 3; ++k) {
            a[k] = dalloc<u16>((size_t)cells * 32);
            aT[k] = dalloc<u16>((size_t)32 * tp);
            d[k] = dalloc<u16>((size_t)cells * 32);
            dT[k] = dalloc<u16>((size_t)32 * tp);
        }
        g4 = dalloc<u16>((size_t)cells * 8);
        g4T = dalloc<u16>((size_t)8 * tp);
        pred = dalloc<float>((size_t)cells * 2);
        feat = dalloc<float>((size_t)cells * 4);
        tgt = dalloc<float>((size_t)cells * 2);
        nParts = 512;
        part = dalloc<float>((size_t)nParts * 32 * 9 * 32);
        bpart = dalloc<float>((size_t)nParts * 32);
        lpart = dalloc<double>(nb(tp) + 1);
    }
    ~mpcd_gpu_net() {
        void* ptrs[] = {theta, mom, vel, grad, x, xT, g4, g4T, pred, feat, tgt, part, bpart, lpart, setF, setT};
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
        for (int l = 0; l < NLAYERS; ++l) {
            (void)hipFree(wf[l]);
            (void)hipFree(wt[l]);
            (void)hipFree(bias[l]);
        }
        for (int k = 0; k < 3; ++k) {
            (void)hipFree(a[k]);
            (void)hipFree(aT[k]);
            (void)hipFree(d[k]);
            (void)hipFree(dT[k]);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }

    void set_batch(int B) {
        if (B <= 0 || B > maxB) throw ApiError(MPCD_E_INVALID, "batch outside [1, maxBatch]");
        g.B = B;
        g.total = (long long)g.Nc * B;
        g.tpitch = (g.total + 15) / 16 * 16;
    }

    void pack_weights() {
        for (int l = 0; l < NLAYERS; ++l) {
            const int nT = 32 * 9 * (l == 3 ? 8 : 32);
            const int nF = coPad(l) * 9 * cinPad(l);
            hipLaunchKernelGGL(k_pack_weights, dim3(nb(std::max(nT, nF))), dim3(NTPB), 0, stream, theta + offset(l),
                               CH[l + 1], CH[l], coPad(l), cinPad(l), l == 3 ? 8 : 32, wf[l], bias[l],
                               l == 0 ? (u16*)nullptr : wt[l]);
        }
        HIP_OK(hipGetLastError());
    }

    void set_params(const double* h) {
        const int np = n_params();
        HIP_OK(hipMemcpyAsync(theta, h, np * sizeof(double), hipMemcpyHostToDevice, stream));
        HIP_OK(hipMemsetAsync(mom, 0, np * sizeof(double), stream));
        HIP_OK(hipMemsetAsync(vel, 0, np * sizeof(double), stream));
        step = 0;
        pack_weights();
        HIP_OK(hipStreamSynchronize(stream));
    }

    unsigned conv_grid() const { return (unsigned)std::min<long long>((g.total + 63) / 64, 4096); }

    /* features `f` (device, g.B samples; default the net's own buffer) */
    void forward(bool keepT, const float* f = nullptr) {
        hipLaunchKernelGGL(k_pack_features, dim3(nb(g.tpitch)), dim3(NTPB), 0, stream, g, f ? f : feat, x, xT);
        const u16* in = x;
        for (int l = 0; l < 3; ++l) {
            ConvArgs c{};
            c.g = g;
            c.in = in;
            c.w = wf[l];
            c.bias = bias[l];
            c.coReal = 32;
            c.out = a[l];
            c.outT = keepT ? aT[l] : nullptr;
            if (l == 0) hipLaunchKernelGGL((k_conv<8, 2, EPI_RELU>), dim3(conv_grid()), dim3(NTPB), 0, stream, c);
            else hipLaunchKernelGGL((k_conv<32, 2, EPI_RELU>), dim3(conv_grid()), dim3(NTPB), 0, stream, c);
            in = a[l];
        }
        ConvArgs c{};
        c.g = g;
        c.in = a[2];
        c.w = wf[3];
        c.bias = bias[3];
        c.coReal = 2;
        c.outf = pred;
        hipLaunchKernelGGL((k_conv<32, 1, EPI_LINEAR>), dim3(conv_grid()), dim3(NTPB), 0, stream, c);
        HIP_OK(hipGetLastError());
    }

    /* loss over `tgt` and the flat gradient in `grad`; returns the loss */
    void backward_enqueue() {
        const double inv = 1.0 / (double)(g.total * 2);
        const unsigned nl = nb(g.tpitch);
        hipLaunchKernelGGL(k_loss, dim3(nl), dim3(NTPB), 0, stream, g, pred, tgt, inv, g4, g4T, lpart);
        const unsigned np = (unsigned)std::min<long long>(nParts, (g.total + WG_PITCH - 1) / WG_PITCH);
        /* layer 4: gradient g4 (2 channels, pitch 8) */
        const u16* gT = g4T;
        const u16* gRow = g4;
        for (int l = 3; l >= 0; --l) {
            const u16* inT = l == 0 ? xT : aT[l - 1];
            /* rows of the transposed output gradient: 8 for the 2-channel output */
            const int gRows = l == 3 ? 8 : 32;
            if (l == 0)
                hipLaunchKernelGGL((k_conv_wgrad<8, 2>), dim3(np), dim3(NTPB), 0, stream, g, inT, gT, gRows, part, bpart);
            else if (l == 3)
                hipLaunchKernelGGL((k_conv_wgrad<32, 1>), dim3(np), dim3(NTPB), 0, stream, g, inT, gT, gRows, part,
                                   bpart);
            else
                hipLaunchKernelGGL((k_conv_wgrad<32, 2>), dim3(np), dim3(NTPB), 0, stream, g, inT, gT, gRows, part,
                                   bpart);
            const int nW = CH[l + 1] * 9 * CH[l] + CH[l + 1];
            hipLaunchKernelGGL(k_wgrad_reduce, dim3(nb(nW)), dim3(NTPB), 0, stream, part, bpart, (int)np, CH[l + 1],
                               CH[l], grad + offset(l));
            if (l == 0) break;
            /* input gradient of layer l, masked by the forward activation */
            ConvArgs c{};
            c.g = g;
            c.in = gRow;
            c.w = wt[l];
            c.coReal = 32;
            c.out = d[l - 1];
            c.outT = dT[l - 1];
            c.mask = a[l - 1];
            if (l == 3) hipLaunchKernelGGL((k_conv<8, 2, EPI_MASK>), dim3(conv_grid()), dim3(NTPB), 0, stream, c);
            else hipLaunchKernelGGL((k_conv<32, 2, EPI_MASK>), dim3(conv_grid()), dim3(NTPB), 0, stream, c);
            gRow = d[l - 1];
            gT = dT[l - 1];
        }
        HIP_OK(hipGetLastError());
    }

    double read_loss() {
        const unsigned nl = nb(g.tpitch);
        std::vector<double> h(nl);
        HIP_OK(hipMemcpyAsync(h.data(), lpart, nl * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        double s = 0.0;
        for (double v : h) s += v;
        return s / (double)(g.total * 2);
    }

    void adam(double lr) {
        ++step;
        const int np = n_params();
        hipLaunchKernelGGL(k_adam, dim3(nb(np)), dim3(NTPB), 0, stream, theta, mom, vel, grad, np, lr,
                           1.0 - std::pow(0.9, step), 1.0 - std::pow(0.999, step));
        pack_weights();
    }
};

namespace mpcdk {

const float* net_forward_device(mpcd_gpu_net* net, const float* featDev, hipStream_t st) {
    const hipStream_t own = net->stream;
    net->stream = st;
    try {
        net->set_batch(1);
        net->forward(false, featDev);
    } catch (...) {
        net->stream = own;
        throw;
    }
    net->stream = own;
    return net->pred;
}

void net_grid(const mpcd_gpu_net* net, int* Nx, int* Ny, int* perx, int* pery) {
    *Nx = net->g.Nx;
    *Ny = net->g.Ny;
    *perx = net->g.perx;
    *pery = net->g.pery;
}

}  // namespace mpcdk

extern "C" {

int mpcd_gpu_net_num_params(void) { return mpcd_gpu_net::n_params(); }

int mpcd_gpu_net_create(int Nx, int Ny, int perx, int pery, int maxBatch, mpcd_gpu_net** out) {
    return guarded([&] {
        if (!out) throw ApiError(MPCD_E_INVALID, "null argument");
        *out = new mpcd_gpu_net(Nx, Ny, perx, pery, maxBatch);
    });
}

int mpcd_gpu_net_destroy(mpcd_gpu_net* net) {
    return guarded([&] { delete net; });
}

int mpcd_gpu_net_set_params(mpcd_gpu_net* net, const double* theta) {
    return guarded([&] {
        if (!net || !theta) throw ApiError(MPCD_E_INVALID, "null argument");
        net->set_params(theta);
    });
}

int mpcd_gpu_net_get_params(mpcd_gpu_net* net, double* theta) {
    return guarded([&] {
        if (!net || !theta) throw ApiError(MPCD_E_INVALID, "null argument");
        HIP_OK(hipMemcpyAsync(theta, net->theta, mpcd_gpu_net::n_params() * sizeof(double), hipMemcpyDeviceToHost,
                              net->stream));
        HIP_OK(hipStreamSynchronize(net->stream));
    });
}

int mpcd_gpu_net_forward(mpcd_gpu_net* net, const float* feat, int B, float* out) {
    return guarded([&] {
        if (!net || !feat || !out) throw ApiError(MPCD_E_INVALID, "null argument");
        net->set_batch(B);
        HIP_OK(hipMemcpyAsync(net->feat, feat, (size_t)net->g.total * 4 * sizeof(float), hipMemcpyHostToDevice,
                              net->stream));
        net->forward(false);
        HIP_OK(hipMemcpyAsync(out, net->pred, (size_t)net->g.total * 2 * sizeof(float), hipMemcpyDeviceToHost,
                              net->stream));
        HIP_OK(hipStreamSynchronize(net->stream));
    });
}

int mpcd_gpu_net_loss_grad(mpcd_gpu_net* net, const float* feat, const float* target, int B, double* loss,
                           double* grad) {
    return guarded([&] {
        if (!net || !feat || !target) throw ApiError(MPCD_E_INVALID, "null argument");
        net->set_batch(B);
        HIP_OK(hipMemcpyAsync(net->feat, feat, (size_t)net->g.total * 4 * sizeof(float), hipMemcpyHostToDevice,
                              net->stream));
        HIP_OK(hipMemcpyAsync(net->tgt, target, (size_t)net->g.total * 2 * sizeof(float), hipMemcpyHostToDevice,
                              net->stream));
        net->forward(true);
        net->backward_enqueue();
        const double l = net->read_loss();
        if (loss) *loss = l;
        if (grad) {
            HIP_OK(hipMemcpyAsync(grad, net->grad, mpcd_gpu_net::n_params() * sizeof(double), hipMemcpyDeviceToHost,
                                  net->stream));
            HIP_OK(hipStreamSynchronize(net->stream));
        }
    });
}

int mpcd_gpu_net_adam_step(mpcd_gpu_net* net, const double* grad, double lr) {
    return guarded([&] {
        if (!net) throw ApiError(MPCD_E_INVALID, "null argument");
        if (grad)
            HIP_OK(hipMemcpyAsync(net->grad, grad, mpcd_gpu_net::n_params() * sizeof(double), hipMemcpyHostToDevice,
                                  net->stream));
        net->adam(lr);
        HIP_OK(hipStreamSynchronize(net->stream));
    });
}

int mpcd_gpu_net_train(mpcd_gpu_net* net, const float* feat, const float* target, int nSamples, int batch,
                       int iters, double lr, uint64_t seed, double* losses) {
    return guarded([&] {
        if (!net || !feat || !target || nSamples <= 0 || iters < 0) throw ApiError(MPCD_E_INVALID, "bad argument");
        net->set_batch(batch);
        const size_t nc = (size_t)net->g.Nc;
        if (net->setN != nSamples) {
            if (net->setF) HIP_OK(hipFree(net->setF));
            if (net->setT) HIP_OK(hipFree(net->setT));
            net->setF = dalloc<float>((size_t)nSamples * nc * 4);
            net->setT = dalloc<float>((size_t)nSamples * nc * 2);
            net->setN = nSamples;
        }
        HIP_OK(hipMemcpyAsync(net->setF, feat, (size_t)nSamples * nc * 4 * sizeof(float), hipMemcpyHostToDevice,
                              net->stream));
        HIP_OK(hipMemcpyAsync(net->setT, target, (size_t)nSamples * nc * 2 * sizeof(float), hipMemcpyHostToDevice,
                              net->stream));
        for (int it = 0; it < iters; ++it) {
            /* mini-batch of samples drawn with the counter generator */
            for (int b = 0; b < batch; ++b) {
                const mpcd_u32x4 r = mpcd_philox(seed, 0x20000000u, (uint32_t)it, (uint32_t)b, 0u);
                const size_t sidx = mpcd_uint_below(r.w[0], (uint32_t)nSamples);
                HIP_OK(hipMemcpyAsync(net->feat + b * nc * 4, net->setF + sidx * nc * 4, nc * 4 * sizeof(float),
                                      hipMemcpyDeviceToDevice, net->stream));
                HIP_OK(hipMemcpyAsync(net->tgt + b * nc * 2, net->setT + sidx * nc * 2, nc * 2 * sizeof(float),
                                      hipMemcpyDeviceToDevice, net->stream));
            }
            net->forward(true);
            net->backward_enqueue();
            if (losses) losses[it] = net->read_loss();
            net->adam(lr);
        }
        HIP_OK(hipStreamSynchronize(net->stream));
    });
}

int mpcd_gpu_net_mfma_probe(const uint16_t* a, const uint16_t* b, float* d) {
    return guarded([&] {
        u16 *da = dalloc<u16>(512), *db = dalloc<u16>(512);
        float* dd = dalloc<float>(256);
        HIP_OK(hipMemcpy(da, a, 512 * sizeof(u16), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(db, b, 512 * sizeof(u16), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, 0, da, db, dd);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpy(d, dd, 256 * sizeof(float), hipMemcpyDeviceToHost));
        (void)hipFree(da);
        (void)hipFree(db);
        (void)hipFree(dd);
    });
}

}  // extern "C"
