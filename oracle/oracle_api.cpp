/*
 * oracle_api.cpp -- extern "C" entry points of the CPU oracle (TEST
 * INFRASTRUCTURE ONLY; loaded by tests/ and by bench.py's CPU-baseline leg).
 *
 * mode = 0: libstdc++ draws in the reference's order (pinned against the
 *           compiled original);
 * mode = 1: counter-based draws of include/mpcd_rng.h (the GPU's contract).
 * Arrays are interleaved [x1,y1,...] in particle-id order, as in the
 * reference's State (types.h:74-79).
 */
#include <cstring>
#include <stdexcept>
#include <string>

#include "oracle.hpp"

using namespace orc;

static thread_local std::string g_err;

extern "C" const char* oracle_last_error() { return g_err.c_str(); }

#define OR_TRY try {
#define OR_CATCH                         \
    }                                    \
    catch (const std::exception& e) {    \
        g_err = e.what();                \
        return -1;                       \
    }                                    \
    return 0;

static std::vector<double> vec(const double* a, size_t n) { return std::vector<double>(a, a + n); }
static void put(const std::vector<double>& a, double* out) { std::memcpy(out, a.data(), a.size() * sizeof(double)); }

static void put_fields(const Fields& G, mpcd_cell_fields* f) {
    if (!f) return;
    const size_t n = G.N.size();
    if (f->N) std::memcpy(f->N, G.N.data(), n * sizeof(int32_t));
    auto cp = [&](double* dst, const std::vector<double>& s) {
        if (dst) std::memcpy(dst, s.data(), n * sizeof(double));
    };
    cp(f->Ux, G.Ux);
    cp(f->Uy, G.Uy);
    cp(f->Px, G.Px);
    cp(f->Py, G.Py);
    cp(f->T, G.T);
    cp(f->rho, G.rho);
    cp(f->Pkin, G.Pkin);
    cp(f->Pvir, G.Pvir);
    cp(f->P, G.P);
    cp(f->rhoTarget, G.rhoTarget);
}

/* force + stream: the driver loop, e.g.
 * main_benchmark_poiseuille_openmp_profiled_shifted_profiled.cpp:242-247 */
static void force_stream(const P& p, std::vector<double>& x, std::vector<double>& v) {
    for (int i = 0; i < p.n; ++i) {
        v[2 * (size_t)i] += p.bodyForceX * p.dt;
        v[2 * (size_t)i + 1] += p.g * p.dt;
        x[2 * (size_t)i] += p.dt * v[2 * (size_t)i];
        x[2 * (size_t)i + 1] += p.dt * v[2 * (size_t)i + 1];
    }
}

extern "C" int oracle_force_stream(const mpcd_params* p, double* x, double* v) {
    OR_TRY
    auto xv = vec(x, 2 * (size_t)p->n), vv = vec(v, 2 * (size_t)p->n);
    force_stream(*p, xv, vv);
    put(xv, x);
    put(vv, v);
    OR_CATCH
}

extern "C" int oracle_apply_bc(int mode, const mpcd_params* p, double* x, double* v, uint64_t seed,
                               mpcd_wall_info* info) {
    OR_TRY
    auto xv = vec(x, 2 * (size_t)p->n), vv = vec(v, 2 * (size_t)p->n);
    mpcd_wall_info wi{};
    if (mode == 0) {
        StdRngBC r(seed);
        apply_bc(*p, xv, vv, wi, r);
    } else {
        CtrRngBC r(seed);
        apply_bc(*p, xv, vv, wi, r);
    }
    put(xv, x);
    put(vv, v);
    if (info) *info = wi;
    OR_CATCH
}

extern "C" int oracle_shift_cell_ids(int mode, const mpcd_params* p, const double* x, uint64_t seed, int32_t* cid,
                                     double* shift_out) {
    OR_TRY
    double sx, sy;
    if (mode == 0) std_shift(seed, p->a0, sx, sy);
    else ctr_shift(seed, p->a0, sx, sy);
    std::vector<int> c;
    shift_cell_ids(*p, vec(x, 2 * (size_t)p->n), sx, sy, c);
    std::memcpy(cid, c.data(), c.size() * sizeof(int32_t));
    if (shift_out) {
        shift_out[0] = sx;
        shift_out[1] = sy;
    }
    OR_CATCH
}

extern "C" int oracle_collide(int mode, const mpcd_params* p, double* v, const int32_t* cid, uint64_t seed) {
    OR_TRY
    auto vv = vec(v, 2 * (size_t)p->n);
    std::vector<int> c(cid, cid + p->n);
    if (mode == 0) {
        StdRngRot r(seed);
        collide(*p, vv, c, r);
    } else {
        CtrRngRot r(seed);
        collide(*p, vv, c, r);
    }
    put(vv, v);
    OR_CATCH
}

extern "C" int oracle_cell_fields(const mpcd_params* p, const double* x, const double* v, double rhoT,
                                  mpcd_cell_fields* f) {
    OR_TRY
    put_fields(cell_fields(*p, vec(x, 2 * (size_t)p->n), vec(v, 2 * (size_t)p->n), rhoT), f);
    OR_CATCH
}

extern "C" int oracle_occupancy_stats(const mpcd_params* p, const int32_t* counts, mpcd_occupancy_stats* out) {
    OR_TRY
    *out = occupancy_stats(*p, std::vector<int>(counts, counts + p->Nc));
    OR_CATCH
}

extern "C" int oracle_layout(const mpcd_params* p, int mode, int32_t* owner, int32_t* nzones) {
    OR_TRY
    const Layout L = build_layout(*p, mode);
    if (owner) std::memcpy(owner, L.owner.data(), L.owner.size() * sizeof(int32_t));
    if (nzones) *nzones = (int32_t)L.zones.size();
    OR_CATCH
}

extern "C" int oracle_zone_kernel(int mode, const mpcd_params* p, double* x, double* v, int layout, int zoneIndex,
                                  uint64_t seed, double* stats6, double* diag10, int32_t* moved2) {
    OR_TRY
    const Layout L = build_layout(*p, layout);
    if (zoneIndex < 0 || zoneIndex >= (int)L.zones.size()) throw std::runtime_error("zone index out of range");
    auto xv = vec(x, 2 * (size_t)p->n), vv = vec(v, 2 * (size_t)p->n);
    ZoneStats st;
    if (mode == 0) {
        StdRngZone r(seed);
        st = zone_kernel(*p, xv, vv, L.zones[(size_t)zoneIndex], r);
    } else {
        CtrRngZone r(seed);
        st = zone_kernel(*p, xv, vv, L.zones[(size_t)zoneIndex], r);
    }
    put(xv, x);
    put(vv, v);
    if (stats6) std::memcpy(stats6, st.stats.data(), 6 * sizeof(double));
    if (diag10) std::memcpy(diag10, st.diag.data(), 10 * sizeof(double));
    if (moved2) {
        moved2[0] = st.movedDense;
        moved2[1] = st.movedSparse;
    }
    OR_CATCH
}

/* run_zone_kernel_inplace on a zone given by its mask (p->Nc bytes, nonzero
 * in the zone), zone_pass.cpp:1065-1081 with ZoneDescriptor.maskVec */
extern "C" int oracle_zone_kernel_mask(int mode, const mpcd_params* p, double* x, double* v, const uint8_t* mask,
                                       uint64_t seed, double* stats6, double* diag10, int32_t* moved2) {
    OR_TRY
    Zone z;
    z.mask.assign((size_t)p->Nc, 0);
    z.x0 = p->Nx;
    z.y0 = p->Ny;
    for (int c = 0; c < p->Nc; ++c) {
        if (!mask[c]) continue;
        z.mask[(size_t)c] = 1;
        z.cells.push_back(c);
        const int ix = c / p->Ny, iy = c - ix * p->Ny;
        z.x0 = std::min(z.x0, ix);
        z.x1 = std::max(z.x1, ix + 1);
        z.y0 = std::min(z.y0, iy);
        z.y1 = std::max(z.y1, iy + 1);
    }
    if (z.cells.empty()) throw std::runtime_error("empty zone mask");
    auto xv = vec(x, 2 * (size_t)p->n), vv = vec(v, 2 * (size_t)p->n);
    ZoneStats st;
    if (mode == 0) {
        StdRngZone r(seed);
        st = zone_kernel(*p, xv, vv, z, r);
    } else {
        CtrRngZone r(seed);
        st = zone_kernel(*p, xv, vv, z, r);
    }
    put(xv, x);
    put(vv, v);
    if (stats6) std::memcpy(stats6, st.stats.data(), 6 * sizeof(double));
    if (diag10) std::memcpy(diag10, st.diag.data(), 10 * sizeof(double));
    if (moved2) {
        moved2[0] = st.movedDense;
        moved2[1] = st.movedSparse;
    }
    OR_CATCH
}

extern "C" int oracle_zone_pass(int mode, const mpcd_params* p, double* x, double* v, int layout, uint64_t seedBase,
                                int gaussSeidel, mpcd_zone_pass_metrics* m, int32_t* owner) {
    OR_TRY
    auto xv = vec(x, 2 * (size_t)p->n), vv = vec(v, 2 * (size_t)p->n);
    PassMetrics pm;
    std::vector<int32_t> own;
    if (gaussSeidel) {
        pm = (mode == 0) ? zone_pass_gauss_seidel<StdRngZone>(*p, xv, vv, layout, seedBase)
                         : zone_pass_gauss_seidel<CtrRngZone>(*p, xv, vv, layout, seedBase);
        own = build_layout(*p, layout).owner;
    } else {
        pm = (mode == 0) ? zone_pass_snapshot<StdRngZone>(*p, xv, vv, layout, seedBase, &own)
                         : zone_pass_snapshot<CtrRngZone>(*p, xv, vv, layout, seedBase, &own);
    }
    put(xv, x);
    put(vv, v);
    if (m) {
        m->nZones = pm.nZones;
        m->nParticlesMovedDense = pm.movedDense;
        m->nParticlesMovedSparse = pm.movedSparse;
        m->corrRmsDU = pm.corrRmsDU;
        m->corrMaxDU = pm.corrMaxDU;
    }
    if (owner) std::memcpy(owner, own.data(), own.size() * sizeof(int32_t));
    OR_CATCH
}

/* the snapshot pass in the reference's structure (a private copy of the
 * domain per zone, O(n) per zone): the check of the tile-local form */
extern "C" int oracle_zone_pass_full(int mode, const mpcd_params* p, double* x, double* v, int layout,
                                     uint64_t seedBase, mpcd_zone_pass_metrics* m) {
    OR_TRY
    auto xv = vec(x, 2 * (size_t)p->n), vv = vec(v, 2 * (size_t)p->n);
    const PassMetrics pm = (mode == 0) ? zone_pass_snapshot_full<StdRngZone>(*p, xv, vv, layout, seedBase, nullptr)
                                       : zone_pass_snapshot_full<CtrRngZone>(*p, xv, vv, layout, seedBase, nullptr);
    put(xv, x);
    put(vv, v);
    if (m) {
        m->nZones = pm.nZones;
        m->nParticlesMovedDense = pm.movedDense;
        m->nParticlesMovedSparse = pm.movedSparse;
        m->corrRmsDU = pm.corrRmsDU;
        m->corrMaxDU = pm.corrMaxDU;
    }
    OR_CATCH
}

/* grids: 6 arrays of Nc doubles (pdriveRaw, pdrive, duRepairX, duRepairY,
 * duEOSX, duEOSY), any may be NULL; fields: 4 structs (ref, red, repaired,
 * out), any may be NULL. */
extern "C" int oracle_closure(const mpcd_params* p, const double* xRef, const double* vRef, const double* xRed,
                              const double* vRed, double* xOut, double* vOut, mpcd_closure_metrics* m,
                              double** grids6, int32_t* mask, mpcd_cell_fields* fields4, double* vRepaired) {
    OR_TRY
    const size_t n2 = 2 * (size_t)p->n;
    std::vector<double> xo, vo;
    ClosureOut o = liquid_closure(*p, vec(xRef, n2), vec(vRef, n2), vec(xRed, n2), vec(vRed, n2), xo, vo);
    if (xOut) put(xo, xOut);
    if (vOut) put(vo, vOut);
    if (vRepaired) put(o.vRepaired, vRepaired);
    if (m) *m = o.m;
    if (grids6) {
        const std::vector<double>* g[6] = {&o.pdriveRaw, &o.pdrive, &o.duRepairX, &o.duRepairY, &o.duEOSX, &o.duEOSY};
        for (int k = 0; k < 6; ++k)
            if (grids6[k]) put(*g[k], grids6[k]);
    }
    if (mask) std::memcpy(mask, o.mask.data(), o.mask.size() * sizeof(int32_t));
    if (fields4) {
        put_fields(o.ref, &fields4[0]);
        put_fields(o.red, &fields4[1]);
        put_fields(o.repaired, &fields4[2]);
        put_fields(o.out, &fields4[3]);
    }
    OR_CATCH
}

/* One full time step exactly as the benchmark driver composes it
 * (main_benchmark_poiseuille_openmp_profiled_shifted_profiled.cpp:240-279),
 * with the driver's seeds for `step`.  Optional stage snapshots (NULL to
 * skip): ref = after force/stream/BC/collision, base = after the base pass,
 * red = after the shifted pass.  x, v are advanced in place. */
extern "C" int oracle_step(int mode, const mpcd_params* p, double* x, double* v, int step, double* xRef,
                           double* vRef, double* xBase, double* vBase, double* xRed, double* vRed,
                           mpcd_zone_pass_metrics* mBase, mpcd_zone_pass_metrics* mShift, mpcd_closure_metrics* mc,
                           mpcd_wall_info* wi) {
    OR_TRY
    const size_t n2 = 2 * (size_t)p->n;
    const uint64_t s = (uint64_t)step;
    auto xr = vec(x, n2), vr = vec(v, n2);
    force_stream(*p, xr, vr);
    mpcd_wall_info w{};
    std::vector<int> cid;
    double sx, sy;
    if (mode == 0) {
        StdRngBC rb(104729ULL + 10007ULL * s);
        apply_bc(*p, xr, vr, w, rb);
        std_shift(130363ULL + 10007ULL * s, p->a0, sx, sy);
        shift_cell_ids(*p, xr, sx, sy, cid);
        StdRngRot rr(433494437ULL + 10007ULL * s);
        collide(*p, vr, cid, rr);
    } else {
        CtrRngBC rb(104729ULL + 10007ULL * s);
        apply_bc(*p, xr, vr, w, rb);
        ctr_shift(130363ULL + 10007ULL * s, p->a0, sx, sy);
        shift_cell_ids(*p, xr, sx, sy, cid);
        CtrRngRot rr(433494437ULL + 10007ULL * s);
        collide(*p, vr, cid, rr);
    }
    if (wi) *wi = w;
    if (xRef) put(xr, xRef);
    if (vRef) put(vr, vRef);
    auto xb = xr, vb = vr;
    const uint64_t zs = 88172645463325252ULL + 10007ULL * s;
    PassMetrics pb = (mode == 0) ? zone_pass_snapshot<StdRngZone>(*p, xb, vb, MPCD_LAYOUT_BASE, zs, nullptr)
                                 : zone_pass_snapshot<CtrRngZone>(*p, xb, vb, MPCD_LAYOUT_BASE, zs, nullptr);
    if (xBase) put(xb, xBase);
    if (vBase) put(vb, vBase);
    auto xs = xb, vs = vb;
    PassMetrics ps = (mode == 0)
                         ? zone_pass_snapshot<StdRngZone>(*p, xs, vs, MPCD_LAYOUT_SHIFTED, zs + 4099ULL, nullptr)
                         : zone_pass_snapshot<CtrRngZone>(*p, xs, vs, MPCD_LAYOUT_SHIFTED, zs + 4099ULL, nullptr);
    if (xRed) put(xs, xRed);
    if (vRed) put(vs, vRed);
    std::vector<double> xo, vo;
    ClosureOut o = liquid_closure(*p, xr, vr, xs, vs, xo, vo);
    put(xo, x);
    put(vo, v);
    auto fill = [](const PassMetrics& a, mpcd_zone_pass_metrics* b) {
        if (!b) return;
        b->nZones = a.nZones;
        b->nParticlesMovedDense = a.movedDense;
        b->nParticlesMovedSparse = a.movedSparse;
        b->corrRmsDU = a.corrRmsDU;
        b->corrMaxDU = a.corrMaxDU;
    };
    fill(pb, mBase);
    fill(ps, mShift);
    if (mc) *mc = o.m;
    OR_CATCH
}

extern "C" int oracle_interzone_mask(const mpcd_params* p, int32_t* out) {
    OR_TRY
    const auto m = interzone_mask(*p);
    std::memcpy(out, m.data(), m.size() * sizeof(int32_t));
    OR_CATCH
}

/* Draw-level access to the counter generator for the RNG checks. */
extern "C" void oracle_ctr_normal_pair(uint64_t seed, uint32_t pid, int sweep, int wall, double* out2) {
    mpcd_draw_bc(seed, pid, sweep, wall, &out2[0], &out2[1]);
}
extern "C" void oracle_ctr_philox(uint64_t seed, const uint32_t* ctr4, uint32_t* out4) {
    const mpcd_u32x4 r = mpcd_philox(seed, ctr4[0], ctr4[1], ctr4[2], ctr4[3]);
    for (int k = 0; k < 4; ++k) out4[k] = r.w[k];
}
extern "C" double oracle_ctr_log(double x) { return mpcd_log(x); }
extern "C" void oracle_ctr_sincos_2pi(double u, double* sc) { mpcd_sincos_2pi(u, &sc[0], &sc[1]); }
