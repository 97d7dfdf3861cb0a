/*
 * core_host.cpp -- host build of the GPU library's per-particle, per-cell and
 * per-zone logic (src_incompressible_openmp_amd/csrc/mpcd_core.h), driven the
 * way mpcd_engine.hip's kernels drive it: one zone at a time with the lanes of
 * a workgroup run in sequence.  Test infrastructure: the host logic check
 * compares it with the oracle's counter mode, so a logic error in the shared
 * kernel bodies shows on the CPU before any GPU run.
 */
#include <cstring>
#include <numeric>
#include <vector>

#include "mpcd_core.h"
#include "mpcd_host.h"

using namespace mpcdk;

namespace {

struct Bin {
    std::vector<int> cnt, start, list;
};

/* counting sort of slots by cell; runs in ascending pid (slot == pid here) */
Bin bin_by(const std::vector<int>& cid, int Nc) {
    Bin b;
    b.cnt.assign((size_t)Nc + 1, 0);
    for (int c : cid) b.cnt[(size_t)c]++;
    b.start.assign((size_t)Nc + 1, 0);
    for (int c = 0; c < Nc; ++c) b.start[(size_t)c + 1] = b.start[(size_t)c] + b.cnt[(size_t)c];
    b.list.assign(cid.size(), 0);
    std::vector<int> cur(b.start.begin(), b.start.end() - 1);
    for (size_t i = 0; i < cid.size(); ++i) b.list[(size_t)cur[(size_t)cid[i]]++] = (int)i;
    return b;
}

void split(const double* a, int n, std::vector<double>& x, std::vector<double>& y) {
    x.resize((size_t)n);
    y.resize((size_t)n);
    for (int i = 0; i < n; ++i) {
        x[(size_t)i] = a[2 * (size_t)i];
        y[(size_t)i] = a[2 * (size_t)i + 1];
    }
}
void join(const std::vector<double>& x, const std::vector<double>& y, double* a) {
    for (size_t i = 0; i < x.size(); ++i) {
        a[2 * i] = x[i];
        a[2 * i + 1] = y[i];
    }
}

}  // namespace

extern "C" int core_force_stream_bc(const mpcd_params* p, double* xa, double* va, uint64_t seed, int do_stream) {
    const Geom g = make_geom(*p);
    const BCParams b = make_bc_params(*p);
    int rc = 0;
    for (int i = 0; i < p->n; ++i) {
        WallAcc acc;
        wall_acc_zero(acc);
        double x = xa[2 * (size_t)i], y = xa[2 * (size_t)i + 1], vx = va[2 * (size_t)i], vy = va[2 * (size_t)i + 1];
        if (bc_particle(g, b, seed, (uint32_t)i, do_stream, x, y, vx, vy, acc) != 0) rc = MPCD_E_BC_MODE;
        xa[2 * (size_t)i] = x;
        xa[2 * (size_t)i + 1] = y;
        va[2 * (size_t)i] = vx;
        va[2 * (size_t)i + 1] = vy;
    }
    return rc;
}

extern "C" int core_collide(const mpcd_params* p, double* va, const int32_t* cid, uint64_t seed) {
    const CollideParams cp = make_collide_params(*p);
    std::vector<int> c(cid, cid + p->n);
    const Bin b = bin_by(c, p->Nc);
    std::vector<double> vx, vy;
    split(va, p->n, vx, vy);
    for (int cell = 0; cell < p->Nc; ++cell)
        collide_cell(cp, seed, cell, b.list.data() + b.start[(size_t)cell], b.cnt[(size_t)cell], vx.data(), vy.data());
    if (p->keepMeanFlow) { /* the engine reduces with a tree; here in particle order */
        double mx = 0.0, my = 0.0;
        for (int i = 0; i < p->n; ++i) {
            mx += vx[(size_t)i];
            my += vy[(size_t)i];
        }
        mx /= (double)p->n;
        my /= (double)p->n;
        for (int i = 0; i < p->n; ++i) {
            vx[(size_t)i] -= mx;
            vy[(size_t)i] -= my;
        }
    }
    join(vx, vy, va);
    return 0;
}

/* One zone of a pass, as k_zone runs it, with working-set index type IT and
 * `bytes` of working set (links beyond it go to the zone's global region). */
template <class IT>
void run_zone(const ZoneIn& in, ZoneRect r, int zid, uint64_t seed, size_t bytes, std::vector<IT>& nxG,
              std::vector<IT>& tmpG, std::vector<double>& duG, ZoneShared& sh) {
    ZoneCtxT<IT> z{};
    zone_init(z, in, r, zid, seed, &sh);
    std::vector<unsigned char> rec(nxG.size(), 0xEE);
    z.rec = rec.data();
    z.nxG = nxG.data();
    z.tmpG = tmpG.data();
    z.duG = duG.data();
    const size_t fixed = zone_fixed_bytes<IT>(z.C, z.zw, in.zp.useLocal);
    std::vector<unsigned char> buf(std::max(bytes, fixed) + 16);
    unsigned char* base = buf.data() + ((16 - ((uintptr_t)buf.data() & 15)) & 15);
    zone_carve(z, base, std::max(bytes, fixed));
    zone_setup(z, in, 0, 1);
    zone_run(z, 0, 1);
}

/* One snapshot zone pass as mpcd_engine.hip runs it: per-zone scratch,
 * owned particles written, velocities in place except for cells several
 * zones cover (side buffer, copied back after the pass).  wide = 1: int
 * working set, else uint16_t; ldsBytes: the working set's size (small values
 * push list links to the global region). */
extern "C" int core_zone_pass(const mpcd_params* p, double* xa, double* va, int layout, uint64_t seedBase,
                              int32_t* moved2, double* corr2, int wide, long long ldsBytes) {
    const Geom g = make_geom(*p);
    const ZoneParams zp = make_zone_params(*p);
    std::vector<double> x, y, vx, vy;
    split(xa, p->n, x, y);
    split(va, p->n, vx, vy);
    std::vector<double> vsx(vx.size(), -777.0), vsy(vy.size(), -777.0);
    std::vector<int> cid((size_t)p->n), pid((size_t)p->n);
    std::iota(pid.begin(), pid.end(), 0);
    for (int i = 0; i < p->n; ++i) cid[(size_t)i] = cell_of(g, x[(size_t)i], y[(size_t)i]);
    const Bin b = bin_by(cid, p->Nc);
    std::vector<int32_t> owner;
    const auto rects = layout_rects(*p, layout, &owner);
    std::vector<int> cov((size_t)p->Nc, 0);
    for (const auto& r : rects)
        for (int gx = r.xs; gx < r.xe; ++gx)
            for (int gy = r.ys; gy < r.ye; ++gy) cov[(size_t)(gy + p->Ny * gx)]++;
    std::vector<unsigned char> over((size_t)p->Nc, 0);
    bool anyOver = false;
    for (int c = 0; c < p->Nc; ++c)
        if (cov[(size_t)c] > 1) {
            over[(size_t)c] = 1;
            anyOver = true;
        }
    ZoneIn in{};
    in.g = g;
    in.zp = zp;
    in.gcnt = b.cnt.data();
    in.gstart = b.start.data();
    in.glist = b.list.data();
    std::vector<unsigned long long> key(b.list.size());
    for (size_t j = 0; j < key.size(); ++j)
        key[j] = ((unsigned long long)(uint32_t)pid[(size_t)b.list[j]] << 32) | (uint32_t)b.list[j];
    in.gkey = key.data();
    in.owner = owner.data();
    in.gcover = anyOver ? over.data() : nullptr;
    in.x = x.data();
    in.y = y.data();
    in.vx = vx.data();
    in.vy = vy.data();
    in.vsx = vsx.data();
    in.vsy = vsy.data();
    int md = 0, ms = 0;
    double acc = 0.0, mx = 0.0;
    int nc = 0;
    for (size_t zi = 0; zi < rects.size(); ++zi) {
        const ZoneRect r = rects[zi];
        int P = 0;
        for (int gx = r.xs; gx < r.xe; ++gx)
            P += b.start[(size_t)(r.ye - 1 + p->Ny * gx) + 1] - b.start[(size_t)(r.ys + p->Ny * gx)];
        const int C = (r.xe - r.xs) * (r.ye - r.ys);
        std::vector<double> duG(2 * (size_t)C, -555.0);
        ZoneShared sh{};
        const uint64_t seed = seedBase + 104729ULL * (uint64_t)(zi + 1);
        if (wide) {
            std::vector<int> nxG((size_t)P + 1, -999), tmpG((size_t)P + 1, -999);
            run_zone<int>(in, r, (int)zi + 1, seed, (size_t)ldsBytes, nxG, tmpG, duG, sh);
        } else {
            if (P > 65534 || C > 65534) return -3;
            std::vector<uint16_t> nxG((size_t)P + 1, 0xABCD), tmpG((size_t)P + 1, 0xABCD);
            run_zone<uint16_t>(in, r, (int)zi + 1, seed, (size_t)ldsBytes, nxG, tmpG, duG, sh);
        }
        if (sh.P != P) return -1;
        md += sh.movedDense;
        ms += sh.movedSparse;
        if (sh.stats4 > 0.0) {
            acc += sh.stats4 * sh.stats4;
            ++nc;
        }
        mx = std::max(mx, sh.stats5);
    }
    for (int c = 0; c < p->Nc; ++c) {
        if (!over[(size_t)c]) continue;
        for (int j = b.start[(size_t)c]; j < b.start[(size_t)c + 1]; ++j) {
            const int slot = b.list[(size_t)j];
            if (vsx[(size_t)slot] == -777.0) return -2; /* an overlap particle no zone wrote */
            vx[(size_t)slot] = vsx[(size_t)slot];
            vy[(size_t)slot] = vsy[(size_t)slot];
        }
    }
    join(x, y, xa);
    join(vx, vy, va);
    if (moved2) {
        moved2[0] = md;
        moved2[1] = ms;
    }
    if (corr2) {
        corr2[0] = nc > 0 ? std::sqrt(acc / (double)nc) : 0.0;
        corr2[1] = mx;
    }
    return 0;
}
