/*
 * zone_kernel_probe -- runs the compiled original's run_zone_kernel_inplace
 * (zone_pass.h:64-68) on one zone of a layout and writes the resulting
 * state and its ZoneKernelStats.  It is linked against the original's own
 * objects by oracle/Makefile into oracle/_ref/, and only the fixture
 * generator (tests/golden/make_goldens.py) runs it: the original's drivers do
 * not expose ZoneKernelStats, so this is how the oracle's zone-kernel
 * statistics and diagnostics are pinned.
 *
 * Usage: zone_kernel_probe params.kv x.bin v.bin layout zone_index seed out_prefix [mask.bin]
 * With mask.bin (Nc bytes, nonzero in the zone) the zone is that cell set
 * (ZoneDescriptor.cellIds in ascending order and maskVec, zone_pass.h:12-16)
 * instead of the layout's zone.
 * Writes <out>_x.bin, <out>_v.bin (interleaved f64) and <out>_stats.kv
 * (stats0..5, diag0..9, movedDense, movedSparse, nZones; %.17g).
 */
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <vector>

#include "params_io.h"
#include "state_io.h"
#include "zone_pass.h"

int main(int argc, char** argv) {
    if (argc != 8 && argc != 9) {
        std::fprintf(stderr, "usage: %s params.kv x.bin v.bin layout zone_index seed out_prefix [mask.bin]\n", argv[0]);
        return 2;
    }
    try {
        const mpcd::Params p = mpcd::read_params_kv(argv[1]);
        mpcd::State st = mpcd::read_state(argv[2], argv[3], "", "", false, false, p.n);
        const mpcd::ZoneLayout lay = mpcd::build_zone_layout(p, argv[4]);
        const int zi = std::atoi(argv[5]);
        const std::uint64_t seed = std::strtoull(argv[6], nullptr, 10);
        if (zi < 0 || zi >= (int)lay.zones.size()) {
            std::fprintf(stderr, "zone index %d out of range (%zu zones)\n", zi, lay.zones.size());
            return 2;
        }
        mpcd::ZoneDescriptor zone = lay.zones[(size_t)zi];
        if (argc == 9) {
            std::ifstream f(argv[8], std::ios::binary);
            const std::vector<char> m((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
            if ((int)m.size() != p.Nc) throw std::runtime_error("mask.bin must hold Nc bytes");
            zone.cellIds.clear();
            zone.maskVec.assign((size_t)p.Nc, 0);
            for (int c = 0; c < p.Nc; ++c)
                if (m[(size_t)c]) {
                    zone.cellIds.push_back(c);
                    zone.maskVec[(size_t)c] = 1;
                }
        }
        const mpcd::ZoneKernelStats r = mpcd::run_zone_kernel_inplace(st.x, st.v, p, zone, seed);
        const std::string out = argv[7];
        mpcd::write_xy_interleaved(out + "_x.bin", st.x, p.n);
        mpcd::write_xy_interleaved(out + "_v.bin", st.v, p.n);
        FILE* f = std::fopen((out + "_stats.kv").c_str(), "w");
        if (!f) throw std::runtime_error("cannot open " + out + "_stats.kv");
        for (int k = 0; k < 6; ++k) std::fprintf(f, "stats%d=%.17g\n", k, r.stats[(size_t)k]);
        for (int k = 0; k < 10; ++k) std::fprintf(f, "diag%d=%.17g\n", k, r.diag[(size_t)k]);
        std::fprintf(f, "movedDense=%d\nmovedSparse=%d\nnZones=%zu\n", r.movedDense, r.movedSparse, lay.zones.size());
        std::fclose(f);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "Fatal error: %s\n", e.what());
        return 1;
    }
    return 0;
}
