/*
 * mpcd_rng.h -- counter-based random draws for the SRD/MPCD step.
 *
 * The reference draws every random number from std::mt19937_64 streams fed
 * through libstdc++ distributions (boundary_conditions.cpp:24-25,
 * srd_collision.cpp:13-15 and :56-57, zone_pass.cpp:701).  Those streams are
 * sequential, so a GPU cannot replay them.  This header defines the
 * replacement: every draw is a pure function of
 *
 *     key     = the 64-bit seed the reference passes to the stage
 *               (e.g. 104729 + 10007*step for the wall BC), and
 *     counter = (site tag | aux, a, b, c), the semantic coordinates of the
 *               draw (particle id, cell, pass, ...),
 *
 * evaluated with Philox4x32-10 (Salmon, Moraes, Dror & Shaw, "Parallel random
 * numbers: as easy as 1, 2, 3", SC'11).  The same header is compiled for the
 * host (the CPU oracle) and for gfx950 (the HIP kernels), so both sides
 * produce bit-identical draws.  All arithmetic below is IEEE add/sub/mul/div,
 * sqrt and floor, which are correctly rounded on both sides provided FMA
 * contraction is disabled (-ffp-contract=off on every compile line).
 *
 * Draw sites (tag, counter words) -- the single source of truth for the
 * counter mode; see the MPCD_TAG_* constants:
 *   BC       (tag|sweep<<2|wall, pid, 0, 0)    -> normal pair (first, second)
 *   SHIFT    (tag, 0, 0, 0)                   -> sx, sy uniform
 *   ROT      (tag, cell, 0, 0)                -> u01, sign = u<0.5 ? -1 : +1
 *   DSHUF    (tag|pass, i, 0, 0)              -> Fisher-Yates index j in [0,i]
 *   RBIN     (tag|pass<<1|phase, src, 0, 0)   -> start bin in [0,8)
 *   RTIE     (tag|pass<<1|phase, src, nb, 0)  -> u01 tie-break
 *   DPOS     (tag|pass, pid, 0, 0)            -> (ux, uy) position resample
 *   LSHUF    (tag, i, 0, 0)                   -> Fisher-Yates of the sparse list
 *   DONSHUF  (tag, cell, i, 0)                -> Fisher-Yates of the donor list
 *   DONNOISE (tag, cell, donor, 0)            -> N(0,1) score noise
 *   PSHUF    (tag, cell, donor, i)            -> Fisher-Yates of donor particles
 *   SPOS     (tag, pid, 0, 0)                 -> (ux, uy) position resample
 */
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define MPCD_HD __host__ __device__ inline
#else
#define MPCD_HD inline
#endif

#if defined(__HIP_DEVICE_COMPILE__)
#define MPCD_SQRT(x) __builtin_sqrt(x)
#define MPCD_FLOOR(x) __builtin_floor(x)
#else
#include <cmath>
#define MPCD_SQRT(x) std::sqrt(x)
#define MPCD_FLOOR(x) std::floor(x)
#endif

enum {
    MPCD_TAG_BC = 0x01,
    MPCD_TAG_SHIFT = 0x02,
    MPCD_TAG_ROT = 0x03,
    MPCD_TAG_DSHUF = 0x10,
    MPCD_TAG_RBIN = 0x11,
    MPCD_TAG_RTIE = 0x12,
    MPCD_TAG_DPOS = 0x13,
    MPCD_TAG_LSHUF = 0x14,
    MPCD_TAG_DONSHUF = 0x15,
    MPCD_TAG_DONNOISE = 0x16,
    MPCD_TAG_PSHUF = 0x17,
    MPCD_TAG_SPOS = 0x18,
};

struct mpcd_u32x4 {
    uint32_t w[4];
};

/* Philox4x32-10 on counter (c0..c3) with key (low, high 32 bits of seed). */
MPCD_HD mpcd_u32x4 mpcd_philox(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3) {
    uint32_t k0 = (uint32_t)(seed & 0xffffffffu);
    uint32_t k1 = (uint32_t)(seed >> 32);
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)M0 * (uint64_t)c0;
        const uint64_t p1 = (uint64_t)M1 * (uint64_t)c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += W0;
        k1 += W1;
    }
    mpcd_u32x4 out;
    out.w[0] = c0;
    out.w[1] = c1;
    out.w[2] = c2;
    out.w[3] = c3;
    return out;
}

/* 53-bit uniform in [0,1) from two 32-bit words. */
MPCD_HD double mpcd_u01(uint32_t hi, uint32_t lo) {
    const uint64_t b = (((uint64_t)hi << 32) | (uint64_t)lo) >> 11;
    return (double)b * (1.0 / 9007199254740992.0);
}

/* Uniform integer in [0, m) by multiply-shift (m >= 1). */
MPCD_HD uint32_t mpcd_uint_below(uint32_t w, uint32_t m) {
    return (uint32_t)(((uint64_t)w * (uint64_t)m) >> 32);
}

MPCD_HD uint32_t mpcd_ctr0(uint32_t tag, uint32_t aux) {
    return (tag << 24) | (aux & 0x00ffffffu);
}

/* ---- portable elementary functions (same bits on host and gfx950) ---- */

/* Natural log for finite x > 0 (normal range).  x = m*2^e, m in
 * [sqrt(1/2), sqrt(2)), log m = 2*atanh(s) with s = (m-1)/(m+1). */
MPCD_HD double mpcd_log(double x) {
    union { double d; uint64_t u; } bits;
    bits.d = x;
    int e = (int)((bits.u >> 52) & 0x7ff) - 1023;
    bits.u = (bits.u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
    double m = bits.d; /* [1,2) */
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    /* sum_{k=0}^{12} z^k/(2k+1), Horner */
    double p = 1.0 / 25.0;
    p = p * z + 1.0 / 23.0;
    p = p * z + 1.0 / 21.0;
    p = p * z + 1.0 / 19.0;
    p = p * z + 1.0 / 17.0;
    p = p * z + 1.0 / 15.0;
    p = p * z + 1.0 / 13.0;
    p = p * z + 1.0 / 11.0;
    p = p * z + 1.0 / 9.0;
    p = p * z + 1.0 / 7.0;
    p = p * z + 1.0 / 5.0;
    p = p * z + 1.0 / 3.0;
    const double r = 2.0 * s + 2.0 * s * (z * p);
    const double ln2_hi = 6.93147180369123816490e-01; /* 0x3fe62e42fee00000 */
    const double ln2_lo = 1.90821492927058770002e-10; /* 0x3dea39ef35793c76 */
    const double de = (double)e;
    return de * ln2_hi + (de * ln2_lo + r);
}

/* sin and cos of 2*pi*u for u in [0,1). */
MPCD_HD void mpcd_sincos_2pi(double u, double* s_out, double* c_out) {
    const double q = MPCD_FLOOR(u * 4.0 + 0.5);   /* nearest quarter turn, 0..4 */
    const double r = u - q * 0.25;                /* exact, |r| <= 1/8 */
    const double a = r * 6.283185307179586;       /* |a| <= pi/4 */
    const double a2 = a * a;
    /* Taylor series to a^17 / a^18 (truncation < 1e-17 on |a| <= pi/4) */
    double sp = 1.0 / 355687428096000.0;                  /* 1/17! */
    sp = sp * a2 - 1.0 / 1307674368000.0;                 /* 1/15! */
    sp = sp * a2 + 1.0 / 6227020800.0;                    /* 1/13! */
    sp = sp * a2 - 1.0 / 39916800.0;                      /* 1/11! */
    sp = sp * a2 + 1.0 / 362880.0;                        /* 1/9!  */
    sp = sp * a2 - 1.0 / 5040.0;                          /* 1/7!  */
    sp = sp * a2 + 1.0 / 120.0;                           /* 1/5!  */
    sp = sp * a2 - 1.0 / 6.0;                             /* 1/3!  */
    const double sa = a + a * (a2 * sp);
    double cp = 1.0 / 6402373705728000.0;                 /* 1/18! */
    cp = cp * a2 - 1.0 / 20922789888000.0;                /* 1/16! */
    cp = cp * a2 + 1.0 / 87178291200.0;                   /* 1/14! */
    cp = cp * a2 - 1.0 / 479001600.0;                     /* 1/12! */
    cp = cp * a2 + 1.0 / 3628800.0;                       /* 1/10! */
    cp = cp * a2 - 1.0 / 40320.0;                         /* 1/8!  */
    cp = cp * a2 + 1.0 / 720.0;                           /* 1/6!  */
    cp = cp * a2 - 1.0 / 24.0;                            /* 1/4!  */
    cp = cp * a2 + 0.5;                                   /* 1/2!  */
    const double ca = 1.0 - a2 * cp;
    const int k = ((int)q) & 3;
    double s, c;
    if (k == 0) { s = sa; c = ca; }
    else if (k == 1) { s = ca; c = -sa; }
    else if (k == 2) { s = -sa; c = -ca; }
    else { s = -ca; c = sa; }
    *s_out = s;
    *c_out = c;
}

/* Box-Muller pair from one Philox block: first = r*cos, second = r*sin. */
MPCD_HD void mpcd_normal_pair(mpcd_u32x4 b, double* first, double* second) {
    const double u1 = mpcd_u01(b.w[0], b.w[1]);
    const double u2 = mpcd_u01(b.w[2], b.w[3]);
    const double rad = MPCD_SQRT(-2.0 * mpcd_log(1.0 - u1));
    double s, c;
    mpcd_sincos_2pi(u2, &s, &c);
    *first = rad * c;
    *second = rad * s;
}

/* sqrt(x*x + y*y): the hypot used by the counter mode on both sides. */
MPCD_HD double mpcd_hypot(double x, double y) {
    return MPCD_SQRT(x * x + y * y);
}

/* ---- the draw sites ---- */

/* Wall BC: the two normals a thermalising wall hit consumes. */
MPCD_HD void mpcd_draw_bc(uint64_t seed, uint32_t pid, int sweep, int wall, double* n1, double* n2) {
    const mpcd_u32x4 b = mpcd_philox(seed, mpcd_ctr0(MPCD_TAG_BC, (uint32_t)((sweep << 2) | wall)), pid, 0u, 0u);
    mpcd_normal_pair(b, n1, n2);
}

/* Random grid shift: (hi-lo)*u + lo per axis, lo = -a0/2, hi = a0/2. */
MPCD_HD void mpcd_draw_shift(uint64_t seed, double a0, double* sx, double* sy) {
    const mpcd_u32x4 b = mpcd_philox(seed, mpcd_ctr0(MPCD_TAG_SHIFT, 0u), 0u, 0u, 0u);
    const double lo = -0.5 * a0, hi = 0.5 * a0;
    *sx = mpcd_u01(b.w[0], b.w[1]) * (hi - lo) + lo;
    *sy = mpcd_u01(b.w[2], b.w[3]) * (hi - lo) + lo;
}

/* Rotation sign of one collision cell: -1 if u < 0.5 else +1. */
MPCD_HD double mpcd_draw_rot_sign(uint64_t seed, uint32_t cell) {
    const mpcd_u32x4 b = mpcd_philox(seed, mpcd_ctr0(MPCD_TAG_ROT, 0u), cell, 0u, 0u);
    return (mpcd_u01(b.w[0], b.w[1]) < 0.5) ? -1.0 : 1.0;
}

/* Fisher-Yates partner indices j in [0, i] for position i of each shuffle
 * (the shuffle walks i = m-1 down to 1 and swaps a[i] with a[j]). */
MPCD_HD uint32_t mpcd_draw_dense_shuffle(uint64_t seed, int pass, uint32_t i) {
    const mpcd_u32x4 r = mpcd_philox(seed, mpcd_ctr0(MPCD_TAG_DSHUF, (uint32_t)pass), i, 0u, 0u);
    return mpcd_uint_below(r.w[0], i + 1u);
}

MPCD_HD uint32_t mpcd_draw_low_shuffle(uint64_t seed, uint32_t i) {
    const mpcd_u32x4 r = mpcd_philox(seed, mpcd_ctr0(MPCD_TAG_LSHUF, 0u), i, 0u, 0u);
    return mpcd_uint_below(r.w[0], i + 1u);
}

MPCD_HD uint32_t mpcd_draw_donor_shuffle(uint64_t seed, uint32_t cell, uint32_t i) {
    const mpcd_u32x4 r = mpcd_philox(seed, mpcd_ctr0(MPCD_TAG_DONSHUF, 0u), cell, i, 0u);
    return mpcd_uint_below(r.w[0], i + 1u);
}

MPCD_HD uint32_t mpcd_draw_particle_shuffle(uint64_t seed, uint32_t cell, uint32_t donor, uint32_t i) {
    const mpcd_u32x4 r = mpcd_philox(seed, mpcd_ctr0(MPCD_TAG_PSHUF, 0u), cell, donor, i);
    return mpcd_uint_below(r.w[0], i + 1u);
}

MPCD_HD int mpcd_draw_start_bin(uint64_t seed, int pass, int phase, uint32_t src) {
    const mpcd_u32x4 r = mpcd_philox(seed, mpcd_ctr0(MPCD_TAG_RBIN, (uint32_t)((pass << 1) | phase)), src, 0u, 0u);
    return (int)mpcd_uint_below(r.w[0], 8u);
}

MPCD_HD double mpcd_draw_tie(uint64_t seed, int pass, int phase, uint32_t src, uint32_t nb) {
    const mpcd_u32x4 r = mpcd_philox(seed, mpcd_ctr0(MPCD_TAG_RTIE, (uint32_t)((pass << 1) | phase)), src, nb, 0u);
    return mpcd_u01(r.w[0], r.w[1]);
}

MPCD_HD void mpcd_draw_dense_pos(uint64_t seed, int pass, uint32_t pid, double* ux, double* uy) {
    const mpcd_u32x4 r = mpcd_philox(seed, mpcd_ctr0(MPCD_TAG_DPOS, (uint32_t)pass), pid, 0u, 0u);
    *ux = mpcd_u01(r.w[0], r.w[1]);
    *uy = mpcd_u01(r.w[2], r.w[3]);
}

MPCD_HD void mpcd_draw_sparse_pos(uint64_t seed, uint32_t pid, double* ux, double* uy) {
    const mpcd_u32x4 r = mpcd_philox(seed, mpcd_ctr0(MPCD_TAG_SPOS, 0u), pid, 0u, 0u);
    *ux = mpcd_u01(r.w[0], r.w[1]);
    *uy = mpcd_u01(r.w[2], r.w[3]);
}

MPCD_HD double mpcd_draw_donor_noise(uint64_t seed, uint32_t cell, uint32_t donor) {
    const mpcd_u32x4 r = mpcd_philox(seed, mpcd_ctr0(MPCD_TAG_DONNOISE, 0u), cell, donor, 0u);
    double n1, n2;
    mpcd_normal_pair(r, &n1, &n2);
    return n1;
}
