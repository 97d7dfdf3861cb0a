/*
 * main_benchmark_poiseuille_gpu -- the benchmark driver
 * (main_benchmark_poiseuille_openmp_profiled_shifted_profiled.cpp) on the
 * GPU-resident simulation: the same command line, the same input files, the
 * same outputs (<out>_metrics.csv with the same header and columns,
 * <out>_runout.kv with the same keys, <out>_stepNNNN_{x,v,type,r0}.bin and
 * _cellfields.bin dumps in the reference formats), with the state kept in
 * HBM between steps.  The step is the reference driver's loop body
 * (:240-279) with its seeds; the metrics of a row are reduced on the GPU
 * (mpcd_gpu_sim_step_diag) and only a few scalars come back per row.
 *
 * Usage (as the reference driver, :187-191):
 *   main_benchmark_poiseuille_gpu params.kv state_x.bin state_v.bin
 *       state_type.bin state_r0.bin out_prefix has_type has_r0 nthreads
 * params.kv also holds benchmark_nSteps, benchmark_metricsEvery and
 * benchmark_dumpSteps (comma-separated steps; 0 dumps the initial state).
 *
 * Training data for the closure surrogate: an optional params.kv key
 * benchmark_closureDumpSteps (comma-separated steps; the reference driver
 * does not read it) writes each listed step's closure input and output in
 * main_step04_liquid_closure's formats and file suffixes, under the stem
 * <out>_closure_stepNNNN.
 *
 * Differences, all in the timing fields: the phase times are HIP event
 * times of the GPU work (reference build = force, stream, walls, collision
 * and the reference fields; each pass includes its binning update; between
 * metrics rows the kick of one step also performs the next step's stream,
 * which is then timed in the closure phase).  A pass has no separate layout,
 * extract, snapshot, alloc, pack or merge phases on the GPU: its whole time
 * is reported as the kernel time and the others are 0.  nThreadsUsed* is 1.
 */
#include <algorithm>
#include <cctype>
#include <chrono>
#include <cstdint>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "dump_io.h"
#include "gpu_params.h"
#include "liquid_closure.h"
#include "mpcd_gpu.h"
#include "params_io.h"
#include "state_io.h"
#include "types.h"

namespace {

using Clock = std::chrono::steady_clock;

double seconds(const Clock::time_point& a, const Clock::time_point& b) {
    return std::chrono::duration<double>(b - a).count();
}

void ok(int rc) {
    if (rc != MPCD_OK) {
        const char* m = mpcd_gpu_last_error();
        throw std::runtime_error(m && *m ? m : "libmpcd_gpu call failed");
    }
}

std::string trim(const std::string& s) {
    size_t b = 0, e = s.size();
    while (b < e && std::isspace((unsigned char)s[b])) ++b;
    while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
    return s.substr(b, e - b);
}

/* the key=value lines of params.kv (the benchmark keys are read from the
 * same file as the physics, :200-204) */
std::unordered_map<std::string, std::string> kv_lines(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Cannot open params file: " + path);
    std::unordered_map<std::string, std::string> kv;
    std::string line;
    while (std::getline(in, line)) {
        line = trim(line);
        if (line.empty() || line[0] == '#') continue;
        const size_t eq = line.find('=');
        if (eq == std::string::npos) throw std::runtime_error("Invalid key=value line: " + line);
        kv[trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
    }
    return kv;
}

const std::string& key(const std::unordered_map<std::string, std::string>& kv, const std::string& k) {
    const auto it = kv.find(k);
    if (it == kv.end()) throw std::runtime_error("Missing benchmark key: " + k);
    return it->second;
}

std::vector<int> steps_csv(const std::string& s) {
    std::vector<int> out;
    std::stringstream ss(s);
    std::string tok;
    while (std::getline(ss, tok, ',')) {
        tok = trim(tok);
        if (!tok.empty()) out.push_back(std::stoi(tok));
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
}

/* compute_cell_fields of the resident state, as a CellFields */
mpcd::CellFields fields_of(mpcd_gpu_sim* sim, const mpcd::Params& prm, double rhoT) {
    mpcd::CellFields G;
    G.Nx = prm.Nx;
    G.Ny = prm.Ny;
    const size_t C = (size_t)prm.Nc;
    G.N.assign(C, 0);
    for (auto* f : {&G.Ux, &G.Uy, &G.Px, &G.Py, &G.T, &G.rho, &G.Pkin, &G.Pvir, &G.P, &G.rhoTarget}) f->assign(C, 0.0);
    mpcd_cell_fields c{G.N.data(), G.Ux.data(),  G.Uy.data(),   G.Px.data(), G.Py.data(), G.T.data(),
                       G.rho.data(), G.Pkin.data(), G.Pvir.data(), G.P.data(),  G.rhoTarget.data()};
    ok(mpcd_gpu_sim_cell_fields(sim, rhoT, &c));
    return G;
}

struct Timers {
    double loop = 0, ref = 0, base = 0, shifted = 0, closure = 0, diag = 0, dumps = 0;
};

}  // namespace

int main(int argc, char** argv) {
    if (argc != 10) {
        std::cerr << "Usage: " << argv[0]
                  << " params.kv state_x.bin state_v.bin state_type.bin state_r0.bin out_prefix has_type has_r0 nthreads\n";
        return 2;
    }
    const std::string paramsPath = argv[1], xPath = argv[2], vPath = argv[3], typePath = argv[4], r0Path = argv[5],
                      out = argv[6];
    const bool hasType = std::string(argv[7]) == "1", hasR0 = std::string(argv[8]) == "1";
    const int nthreads = std::max(1, std::stoi(argv[9]));
    mpcd_gpu_sim* sim = nullptr;
    try {
        const mpcd::Params prm = mpcd::read_params_kv(paramsPath);
        const auto kv = kv_lines(paramsPath);
        const int nSteps = std::stoi(key(kv, "benchmark_nSteps"));
        const int metricsEvery = std::stoi(key(kv, "benchmark_metricsEvery"));
        const std::vector<int> dumps = steps_csv(key(kv, "benchmark_dumpSteps"));
        auto dumpAt = [&](int s) { return std::binary_search(dumps.begin(), dumps.end(), s); };
        /* optional: steps whose closure input and output are written as
         * training data (absent from the reference driver's keys, which it
         * ignores) */
        const auto cdIt = kv.find("benchmark_closureDumpSteps");
        const std::vector<int> closureDumps = cdIt == kv.end() ? std::vector<int>{} : steps_csv(cdIt->second);
        auto closureAt = [&](int s) { return std::binary_search(closureDumps.begin(), closureDumps.end(), s); };
        auto stopAt = [&](int s) { return dumpAt(s) || closureAt(s); };
        auto rowAt = [&](int s) { return metricsEvery > 0 && (s % metricsEvery == 0 || s == nSteps); };

        mpcd::State st = mpcd::read_state(xPath, vPath, typePath, r0Path, hasType, hasR0, prm.n);
        const double rhoT = prm.gamma / ((prm.Lx / (double)prm.Nx) * (prm.Ly / (double)prm.Ny));
        const mpcd_params p = mpcd::to_gpu_params(prm);
        ok(mpcd_gpu_sim_create(&p, &sim));
        ok(mpcd_gpu_sim_set_state(sim, st.x.data(), st.v.data()));

        std::ofstream metrics(out + "_metrics.csv");
        if (!metrics) throw std::runtime_error("Cannot open metrics CSV for writing");
        metrics << "step,occStd,outBand,meanKinetic,meanUx,meanUy,Qx,baseOccStd,baseOutBand,shiftedOccStd,"
                   "shiftedOutBand,meanPdrive,meanPdriveRaw,meanVelocityX,meanVelocityY,particlesMovedDense,"
                   "particlesMovedSparse,betaRepairOpt,betaRepairApplied,betaRepairNum,betaRepairDen,rmsRepairDU,"
                   "maxRepairDU,meanAbsRepairDU,nMaskedInterzoneCells,momentumDeltaX,momentumDeltaY,"
                   "nThreadsUsedBase,nThreadsUsedShifted,timeStepTotal,timeRefBuild,timeBase,timeShifted,"
                   "timeClosure,timeDiagnostics,timeDump\n";
        std::ofstream man(out + "_runout.kv");
        if (!man) throw std::runtime_error("Cannot open benchmark runout file");
        man << "inputTag=" << xPath << "\n";
        man << "outputTag=" << out << "\n";
        man << "benchmark_nSteps=" << nSteps << "\n";
        man << "benchmark_metricsEvery=" << metricsEvery << "\n";
        man << "benchmark_dumpSteps=" << key(kv, "benchmark_dumpSteps") << "\n";
        man << "nThreadsRequested=" << nthreads << "\n";

        Timers T;
        const auto tAll = Clock::now();
        auto dump = [&](int s) {
            const auto t0 = Clock::now();
            ok(mpcd_gpu_sim_get_state(sim, st.x.data(), st.v.data()));
            std::ostringstream o;
            o << out << "_step" << std::setw(4) << std::setfill('0') << s;
            const std::string pre = o.str();
            mpcd::write_xy_interleaved(pre + "_x.bin", st.x, prm.n);
            mpcd::write_xy_interleaved(pre + "_v.bin", st.v, prm.n);
            if (hasType) mpcd::write_u8(pre + "_type.bin", st.type, prm.n);
            if (hasR0) mpcd::write_xy_interleaved(pre + "_r0.bin", st.r0, prm.n);
            mpcd::write_cell_fields_bin(pre + "_cellfields.bin", fields_of(sim, prm, rhoT));
            T.dumps += seconds(t0, Clock::now());
        };
        if (dumpAt(0)) dump(0);
        /* the closure of step s as main_step04_liquid_closure writes it
         * (:61-71; the repaired and out fields are not kept), under the stem
         * <out>_closure_stepNNNN, with its beta in _closure.kv */
        auto closure_dump = [&](int s) {
            const auto t0 = Clock::now();
            const size_t C = (size_t)prm.Nc;
            mpcd::CellFields F[2];
            mpcd_cell_fields cf[2];
            for (int k = 0; k < 2; ++k) {
                mpcd::CellFields& G = F[k];
                G.Nx = prm.Nx;
                G.Ny = prm.Ny;
                G.N.assign(C, 0);
                for (auto* f : {&G.Ux, &G.Uy, &G.Px, &G.Py, &G.T, &G.rho, &G.Pkin, &G.Pvir, &G.P, &G.rhoTarget})
                    f->assign(C, 0.0);
                cf[k] = mpcd_cell_fields{G.N.data(),   G.Ux.data(),   G.Uy.data(),   G.Px.data(),
                                         G.Py.data(),  G.T.data(),    G.rho.data(),  G.Pkin.data(),
                                         G.Pvir.data(), G.P.data(),   G.rhoTarget.data()};
            }
            std::vector<double> gr[6];
            for (auto& q : gr) q.assign(C, 0.0);
            std::vector<std::int32_t> mask(C, 0);
            mpcd_closure_grids cg{gr[0].data(), gr[1].data(), gr[2].data(), gr[3].data(),
                                  gr[4].data(), gr[5].data(), mask.data()};
            double beta = 0.0;
            int32_t got = -1;
            ok(mpcd_gpu_sim_closure_snapshot(sim, 1, &cf[0], &cf[1], &cg, &beta, &got));
            if (got != s) throw std::runtime_error("closure snapshot of step " + std::to_string(s) + " missing");
            std::ostringstream o;
            o << out << "_closure_step" << std::setw(4) << std::setfill('0') << s;
            const std::string pre = o.str();
            mpcd::write_cell_fields_bin(pre + "_cellfields_ref.bin", F[0]);
            mpcd::write_cell_fields_bin(pre + "_cellfields_red.bin", F[1]);
            const char* names[6] = {"_pdrive_raw.bin", "_pdrive.bin", "_durepairx.bin",
                                    "_durepairy.bin",  "_dueosx.bin", "_dueosy.bin"};
            for (int j = 0; j < 6; ++j) mpcd::write_f64_grid_bin(pre + names[j], prm.Nx, prm.Ny, gr[j]);
            mpcd::write_i32_grid_bin_lc(pre + "_interzone_mask.bin", prm.Nx, prm.Ny, mask);
            std::ofstream kvo(pre + "_closure.kv");
            if (!kvo) throw std::runtime_error("Cannot open closure snapshot file");
            kvo << std::setprecision(17) << "step=" << s << "\nbetaRepairApplied=" << beta << "\n";
            T.dumps += seconds(t0, Clock::now());
        };
        if (!closureDumps.empty()) ok(mpcd_gpu_sim_closure_snapshot(sim, 1, nullptr, nullptr, nullptr, nullptr, nullptr));

        ok(mpcd_gpu_sim_profile(sim, MPCD_PROFILE_PHASES)); /* phase events on every step */
        double phaseSeen[4] = {0, 0, 0, 0};
        int step = 1;
        while (step <= nSteps) {
            const auto t0 = Clock::now();
            const bool last = step == nSteps;
            if (!rowAt(step) && !last) {
                /* steps up to the next row, dump or the last step in one call
                 * (the kick of each performs the next one's stream) */
                int e = step;
                while (e < nSteps && !rowAt(e) && !stopAt(e)) ++e;
                const int k = stopAt(e) && !rowAt(e) && e < nSteps ? e - step + 1 : e - step;
                ok(mpcd_gpu_sim_step(sim, step, k, nullptr));
                T.loop += seconds(t0, Clock::now());
                step += k;
                if (dumpAt(step - 1)) dump(step - 1);
                if (closureAt(step - 1)) closure_dump(step - 1);
                continue;
            }
            mpcd_step_metrics m{};
            mpcd_step_diag d{};
            ok(mpcd_gpu_sim_step_diag(sim, step, &m, &d));
            const auto t1 = Clock::now();
            double ph[4];
            ok(mpcd_gpu_sim_phase_read(sim, ph));
            double dph[4];
            for (int k = 0; k < 4; ++k) {
                dph[k] = (ph[k] - phaseSeen[k]) * 1e-3;
                phaseSeen[k] = ph[k];
            }
            const double tStep = seconds(t0, Clock::now());
            T.diag += seconds(t1, Clock::now());
            if (rowAt(step)) {
                const auto& c = m.closure;
                metrics << step << ',' << d.shifted.occStd << ',' << d.shifted.outBand << ',' << d.meanKinetic << ','
                        << d.meanUx << ',' << d.meanUy << ',' << d.Qx << ',' << d.base.occStd << ','
                        << d.base.outBand << ',' << d.shifted.occStd << ',' << d.shifted.outBand << ','
                        << c.meanPdrive << ',' << c.meanPdriveRaw << ',' << d.meanVelocityX << ','
                        << d.meanVelocityY << ','
                        << (m.base.nParticlesMovedDense + m.shifted.nParticlesMovedDense) << ','
                        << (m.base.nParticlesMovedSparse + m.shifted.nParticlesMovedSparse) << ','
                        << c.betaRepairOpt << ',' << c.betaRepairApplied << ',' << c.betaRepairNum << ','
                        << c.betaRepairDen << ',' << c.rmsRepairDU << ',' << c.maxRepairDU << ','
                        << c.meanAbsRepairDU << ',' << c.nMaskedInterzoneCells << ',' << c.momentumDeltaX << ','
                        << c.momentumDeltaY << ',' << 1 << ',' << 1 << ',' << tStep << ',' << dph[0] << ','
                        << dph[1] << ',' << dph[2] << ',' << dph[3] << ',' << seconds(t1, Clock::now()) << ','
                        << 0.0 << '\n';
            }
            T.loop += tStep;
            if (dumpAt(step)) dump(step);
            if (closureAt(step)) closure_dump(step);
            if (last) {
                const auto& c = m.closure;
                man << "finalOccStd=" << d.shifted.occStd << "\n";
                man << "finalOutBand=" << d.shifted.outBand << "\n";
                man << "finalMeanKinetic=" << d.meanKinetic << "\n";
                man << "finalMeanVelocityX=" << d.meanVelocityX << "\n";
                man << "finalMeanVelocityY=" << d.meanVelocityY << "\n";
                man << "finalQx=" << d.Qx << "\n";
                man << "totalParticlesMovedDense="
                    << (m.base.nParticlesMovedDense + m.shifted.nParticlesMovedDense) << "\n";
                man << "totalParticlesMovedSparse="
                    << (m.base.nParticlesMovedSparse + m.shifted.nParticlesMovedSparse) << "\n";
                man << "finalBetaRepairOpt=" << c.betaRepairOpt << "\n";
                man << "finalBetaRepairApplied=" << c.betaRepairApplied << "\n";
                man << "finalBetaRepairNum=" << c.betaRepairNum << "\n";
                man << "finalBetaRepairDen=" << c.betaRepairDen << "\n";
                man << "finalRmsRepairDU=" << c.rmsRepairDU << "\n";
                man << "finalMaxRepairDU=" << c.maxRepairDU << "\n";
                man << "finalMeanAbsRepairDU=" << c.meanAbsRepairDU << "\n";
                man << "finalMomentumDeltaX=" << c.momentumDeltaX << "\n";
                man << "finalMomentumDeltaY=" << c.momentumDeltaY << "\n";
                man << "nThreadsUsedBase=" << 1 << "\n";
                man << "nThreadsUsedShifted=" << 1 << "\n";
            }
            ++step;
        }
        {
            double ph[4];
            ok(mpcd_gpu_sim_phase_read(sim, ph));
            T.ref = ph[0] * 1e-3;
            T.base = ph[1] * 1e-3;
            T.shifted = ph[2] * 1e-3;
            T.closure = ph[3] * 1e-3;
        }
        const double wall = seconds(tAll, Clock::now());
        auto frac = [](double a, double b) { return b > 0 ? a / b : 0.0; };
        man << "timeTotalWall=" << wall << "\n";
        man << "timeLoopTotal=" << T.loop << "\n";
        man << "timeRefBuild=" << T.ref << "\n";
        man << "timeBase=" << T.base << "\n";
        man << "timeShifted=" << T.shifted << "\n";
        man << "timeClosure=" << T.closure << "\n";
        man << "timeDiagnostics=" << T.diag << "\n";
        man << "timeDumps=" << T.dumps << "\n";
        man << "timePerStepMean=" << (T.loop / std::max(1, nSteps)) << "\n";
        man << "fracRefBuild=" << frac(T.ref, T.loop) << "\n";
        man << "fracBase=" << frac(T.base, T.loop) << "\n";
        man << "fracShifted=" << frac(T.shifted, T.loop) << "\n";
        man << "fracClosure=" << frac(T.closure, T.loop) << "\n";
        man << "fracDiagnostics=" << frac(T.diag, T.loop) << "\n";
        man << "fracDumps=" << frac(T.dumps, T.loop) << "\n";
        for (const char* L : {"Base", "Shifted"}) {
            const double k = std::string(L) == "Base" ? T.base : T.shifted;
            man << "time" << L << "Layout=" << 0.0 << "\n";
            man << "time" << L << "Extract=" << 0.0 << "\n";
            man << "time" << L << "Snapshot=" << 0.0 << "\n";
            man << "time" << L << "Alloc=" << 0.0 << "\n";
            man << "time" << L << "Kernel=" << k << "\n";
            man << "time" << L << "Pack=" << 0.0 << "\n";
            man << "time" << L << "Merge=" << 0.0 << "\n";
        }
        for (const char* L : {"Base", "Shifted"}) {
            const double k = std::string(L) == "Base" ? T.base : T.shifted;
            man << "frac" << L << "Layout=" << 0.0 << "\n";
            man << "frac" << L << "Extract=" << 0.0 << "\n";
            man << "frac" << L << "Snapshot=" << 0.0 << "\n";
            man << "frac" << L << "Alloc=" << 0.0 << "\n";
            man << "frac" << L << "Kernel=" << (k > 0 ? 1.0 : 0.0) << "\n";
            man << "frac" << L << "Pack=" << 0.0 << "\n";
            man << "frac" << L << "Merge=" << 0.0 << "\n";
        }
        ok(mpcd_gpu_sim_destroy(sim));
        return 0;
    } catch (const std::exception& e) {
        if (sim) (void)mpcd_gpu_sim_destroy(sim);
        std::cerr << "Fatal error: " << e.what() << '\n';
        return 1;
    }
}
