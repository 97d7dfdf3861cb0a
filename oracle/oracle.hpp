/*
 * oracle.hpp -- CPU restatement of the reference SRD/MPCD step.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker the GPU library is
 * compared against; nothing in the product (src_incompressible_openmp_amd/,
 * include/) links or calls it.  It follows the reference files line by line
 * (citations below are relative to the reference repository) and runs every
 * random draw through one of two policies:
 *
 *   StdRng*  -- libstdc++ std::mt19937_64 + <random> distributions in the
 *               reference's exact draw order; used to pin this restatement
 *               bitwise against the compiled original (oracle/_ref).
 *   CtrRng*  -- the counter-based draws of include/mpcd_rng.h, keyed by the
 *               same seeds; the mode the GPU library implements.
 */
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <random>
#include <string>
#include <vector>

#include "mpcd_gpu.h"
#include "mpcd_rng.h"

namespace orc {

using P = mpcd_params;

inline bool per_x(const P& p) { return p.bc_left == MPCD_BC_PERIODIC && p.bc_right == MPCD_BC_PERIODIC; }
inline bool per_y(const P& p) { return p.bc_bottom == MPCD_BC_PERIODIC && p.bc_top == MPCD_BC_PERIODIC; }

/* ---------------- RNG policies ---------------- */

/* Wall BC draws: boundary_conditions.cpp:24-25 (one engine and one
 * normal_distribution for the whole call, so the cached polar pair carries
 * over between particles). */
struct StdRngBC {
    std::mt19937_64 eng;
    std::normal_distribution<double> nd{0.0, 1.0};
    explicit StdRngBC(uint64_t seed) : eng(seed) {}
    double first(int, int, int) { return nd(eng); }
    double second(int, int, int) { return nd(eng); }
};
struct CtrRngBC {
    uint64_t seed;
    double n1 = 0, n2 = 0;
    explicit CtrRngBC(uint64_t s) : seed(s) {}
    double first(int pid, int sweep, int wall) {
        mpcd_draw_bc(seed, (uint32_t)pid, sweep, wall, &n1, &n2);
        return n1;
    }
    double second(int, int, int) { return n2; }
};

/* Grid shift: srd_collision.cpp:13-15. */
inline void std_shift(uint64_t seed, double a0, double& sx, double& sy) {
    std::mt19937_64 eng(seed);
    std::uniform_real_distribution<double> u(-0.5 * a0, 0.5 * a0);
    sx = u(eng);
    sy = u(eng);
}
inline void ctr_shift(uint64_t seed, double a0, double& sx, double& sy) { mpcd_draw_shift(seed, a0, &sx, &sy); }

/* Rotation signs: srd_collision.cpp:56-57, :77-83 (one draw per cell with
 * cnt >= 2, in cell order). */
struct StdRngRot {
    std::mt19937_64 eng;
    std::uniform_real_distribution<double> u{0.0, 1.0};
    explicit StdRngRot(uint64_t s) : eng(s) {}
    double sign(int) { return (u(eng) < 0.5) ? -1.0 : 1.0; }
};
struct CtrRngRot {
    uint64_t seed;
    explicit CtrRngRot(uint64_t s) : seed(s) {}
    double sign(int cell) { return mpcd_draw_rot_sign(seed, (uint32_t)cell); }
};

/* Zone kernel draws: zone_pass.cpp:701 (one engine per zone kernel call). */
struct StdRngZone {
    std::mt19937_64 eng;
    std::normal_distribution<double> nd{0.0, 1.0};
    explicit StdRngZone(uint64_t s) : eng(s) {}
    void shuffle_dense(std::vector<int>& a, int) { std::shuffle(a.begin(), a.end(), eng); }
    int start_bin(int, int, int) {
        std::uniform_int_distribution<int> ub(1, 8);
        return ub(eng) - 1;
    }
    double tie(int, int, int, int) {
        std::uniform_real_distribution<double> u(0.0, 1.0);
        return u(eng);
    }
    void pos_dense(int, int, double& ux, double& uy) {
        std::uniform_real_distribution<double> u(0.0, 1.0);
        ux = u(eng);
        uy = u(eng);
    }
    void shuffle_low(std::vector<int>& a) { std::shuffle(a.begin(), a.end(), eng); }
    void shuffle_donors(std::vector<int>& a, int) { std::shuffle(a.begin(), a.end(), eng); }
    void begin_noise() { nd.reset(); }
    double noise(int, int) { return nd(eng); }
    void shuffle_ids(std::vector<int>& a, int, int) { std::shuffle(a.begin(), a.end(), eng); }
    void pos_sparse(int, double& ux, double& uy) {
        std::uniform_real_distribution<double> u(0.0, 1.0);
        ux = u(eng);
        uy = u(eng);
    }
    static double hyp(double a, double b) { return std::hypot(a, b); }
    static constexpr bool stable_sorts = false;
};

struct CtrRngZone {
    uint64_t seed;
    explicit CtrRngZone(uint64_t s) : seed(s) {}
    template <class F>
    static void fisher_yates(std::vector<int>& a, F draw) {
        for (int i = (int)a.size() - 1; i >= 1; --i) {
            const int j = (int)draw((uint32_t)i);
            std::swap(a[(size_t)i], a[(size_t)j]);
        }
    }
    void shuffle_dense(std::vector<int>& a, int pass) {
        fisher_yates(a, [&](uint32_t i) { return mpcd_draw_dense_shuffle(seed, pass, i); });
    }
    int start_bin(int pass, int phase, int src) { return mpcd_draw_start_bin(seed, pass, phase, (uint32_t)src); }
    double tie(int pass, int phase, int src, int nb) { return mpcd_draw_tie(seed, pass, phase, (uint32_t)src, (uint32_t)nb); }
    void pos_dense(int pass, int pid, double& ux, double& uy) { mpcd_draw_dense_pos(seed, pass, (uint32_t)pid, &ux, &uy); }
    void shuffle_low(std::vector<int>& a) {
        fisher_yates(a, [&](uint32_t i) { return mpcd_draw_low_shuffle(seed, i); });
    }
    void shuffle_donors(std::vector<int>& a, int cell) {
        fisher_yates(a, [&](uint32_t i) { return mpcd_draw_donor_shuffle(seed, (uint32_t)cell, i); });
    }
    void begin_noise() {}
    double noise(int cell, int donor) { return mpcd_draw_donor_noise(seed, (uint32_t)cell, (uint32_t)donor); }
    void shuffle_ids(std::vector<int>& a, int cell, int donor) {
        fisher_yates(a, [&](uint32_t i) { return mpcd_draw_particle_shuffle(seed, (uint32_t)cell, (uint32_t)donor, i); });
    }
    void pos_sparse(int pid, double& ux, double& uy) { mpcd_draw_sparse_pos(seed, (uint32_t)pid, &ux, &uy); }
    static double hyp(double a, double b) { return mpcd_hypot(a, b); }
    static constexpr bool stable_sorts = true;
};

/* ---------------- data ---------------- */

struct Fields {
    int Nx = 0, Ny = 0;
    std::vector<int32_t> N;
    std::vector<double> Ux, Uy, Px, Py, T, rho, Pkin, Pvir, P, rhoTarget;
};

/* a zone: the cells [x0,x1) x [y0,y1) (zone_pass.cpp:191-219), listed in
 * ascending cell id; or, with a mask (one entry per grid cell, as
 * Zone.maskVec, zone_pass.cpp:20-24), the cells whose entry is nonzero
 * within their bounding rectangle [x0,x1) x [y0,y1) */
struct Zone {
    int id = 0;
    int x0 = 0, x1 = 0, y0 = 0, y1 = 0;
    std::vector<int> cells;
    std::vector<char> mask;
    bool in(int c, int Ny) const {
        if (!mask.empty()) return mask[(size_t)c] != 0;
        const int ix = c / Ny, iy = c - ix * Ny;
        return ix >= x0 && ix < x1 && iy >= y0 && iy < y1;
    }
};

struct Layout {
    std::vector<Zone> zones;
    std::vector<int32_t> owner;
};

struct ZoneStats {
    std::array<double, 6> stats{};
    std::array<double, 10> diag{};
    int movedDense = 0;
    int movedSparse = 0;
};

struct PassMetrics {
    int nZones = 0;
    int movedDense = 0;
    int movedSparse = 0;
    double corrRmsDU = 0.0;
    double corrMaxDU = 0.0;
};

struct ClosureOut {
    Fields ref, red, repaired, out;
    std::vector<double> pdriveRaw, pdrive, duRepairX, duRepairY, duEOSX, duEOSY;
    std::vector<int32_t> mask;
    mpcd_closure_metrics m{};
    std::vector<double> vRepaired; /* v of stateRepaired */
};

/* ---------------- stages ---------------- */

template <class R>
void apply_bc(const P& p, std::vector<double>& x, std::vector<double>& v, mpcd_wall_info& info, R& rng);
void shift_cell_ids(const P& p, const std::vector<double>& x, double sx, double sy, std::vector<int>& cid);
template <class R>
void collide(const P& p, std::vector<double>& v, const std::vector<int>& cid, R& rng);
Fields cell_fields(const P& p, const std::vector<double>& x, const std::vector<double>& v, double rhoTarget);
int cell_of(const P& p, double x, double y);
Layout build_layout(const P& p, int mode);
template <class R>
ZoneStats zone_kernel(const P& p, std::vector<double>& x, std::vector<double>& v, const Zone& z, R& rng);
template <class R>
PassMetrics zone_pass_snapshot(const P& p, std::vector<double>& x, std::vector<double>& v, int mode, uint64_t seedBase,
                               std::vector<int32_t>* owner);
template <class R>
PassMetrics zone_pass_snapshot_full(const P& p, std::vector<double>& x, std::vector<double>& v, int mode,
                                    uint64_t seedBase, std::vector<int32_t>* owner);
template <class R>
PassMetrics zone_pass_gauss_seidel(const P& p, std::vector<double>& x, std::vector<double>& v, int mode,
                                   uint64_t seedBase);
ClosureOut liquid_closure(const P& p, const std::vector<double>& xRef, const std::vector<double>& vRef,
                          const std::vector<double>& xRed, const std::vector<double>& vRed,
                          std::vector<double>& xOut, std::vector<double>& vOut);
std::vector<int32_t> interzone_mask(const P& p);
double rho_target(const P& p);
mpcd_occupancy_stats occupancy_stats(const P& p, const std::vector<int>& cnt);

}  // namespace orc
