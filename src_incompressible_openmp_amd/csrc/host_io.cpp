/*
 * host_io.cpp -- the file formats of the mpcd:: API: params.kv
 * (params_io.cpp:21-135), runout.kv (params_io.cpp:138-151), raw state
 * files (state_io.cpp:9-69), cellfields.bin (dump_io.cpp:19-39) and the
 * i32/f64 grids (zone_pass.cpp:1165-1181, liquid_closure.cpp:446-480).
 * The formats, key names and error messages are the reference's, so files
 * and drivers move between the two unchanged.
 */
#include <cctype>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "dump_io.h"
#include "liquid_closure.h"
#include "params_io.h"
#include "state_io.h"
#include "zone_pass.h"

namespace mpcd {
namespace {

std::string strip(const std::string& s) {
    size_t b = 0, e = s.size();
    while (b < e && std::isspace((unsigned char)s[b])) ++b;
    while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
    return s.substr(b, e - b);
}

/* One Params member: its key and a member pointer of the matching type. */
struct Field {
    const char* key;
    int Params::*i;
    double Params::*d;
    bool Params::*b;
    std::string Params::*s;
};
#define FI(m) {#m, &Params::m, nullptr, nullptr, nullptr}
#define FD(m) {#m, nullptr, &Params::m, nullptr, nullptr}
#define FB(m) {#m, nullptr, nullptr, &Params::m, nullptr}
#define FS(m) {#m, nullptr, nullptr, nullptr, &Params::m}

/* every key of Params, all required */
const Field kFields[] = {
    FI(Nx),
    FI(Ny),
    FI(Nc),
    FI(n),
    FD(Lx),
    FD(Ly),
    FD(dt),
    FD(a0),
    FD(alpha),
    FD(kBT),
    FD(g),
    FD(bodyForceX),
    FB(useThermostat),
    FB(keepMeanFlow),
    FS(boundary_left),
    FS(boundary_right),
    FS(boundary_bottom),
    FS(boundary_top),
    FD(Utop),
    FD(Ubottom),
    FD(wallSigma),
    FB(noSurfaceCase),
    FB(redistributionEnableSurfaceTopology),
    FB(redistributionWallWettingEnabled),
    FB(useInterfaceVelocityReorientation),
    FB(useIncompressibleRedistribution),
    FB(redistribAfterCollision),
    FB(useZoneRedistribution),
    FI(zoneTileNx),
    FI(zoneTileNy),
    FB(zoneUseShiftedSecondPass),
    FB(zonePassthroughMode),
    FB(zoneSingleFullDomainMode),
    FB(enableMomentumCorrectionPostRedistribution),
    FD(gamma),
    FD(coef),
    FS(highMode),
    FS(lowMode),
    FI(maxRedistribPasses),
    FB(useLocalFluidFractionThresholds),
    FD(fluidFracGain),
    FD(lowThrFloor),
    FD(highThrFloor),
    FD(lowThrBulkOverride),
    FB(useLiquidClosure),
    FB(useOptimalBetaRepair),
    FD(betaRepair),
    FD(betaEOS),
    FD(Kvirial),
    FB(smoothPdriveAtInterzone),
    FI(smoothPdriveInterzoneWidthCells),
    FD(smoothPdriveInterzoneBlend),
    FB(smoothPdriveIncludeShiftedLayouts),
};
#undef FI
#undef FD
#undef FB
#undef FS

bool to_bool(const std::string& s) {
    if (s == "true" || s == "1") return true;
    if (s == "false" || s == "0") return false;
    throw std::runtime_error("Invalid boolean value: " + s);
}

template <class T>
void write_raw(const std::string& path, const T* data, size_t count, const char* what) {
    std::ofstream out(path, std::ios::binary);
    if (!out) throw std::runtime_error(std::string("Cannot open ") + what + " for writing: " + path);
    out.write(reinterpret_cast<const char*>(data), (std::streamsize)(count * sizeof(T)));
    if (!out) throw std::runtime_error(std::string("Cannot write payload to: ") + path);
}

template <class T>
std::vector<T> read_raw(const std::string& path, size_t count) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot open binary file: " + path);
    std::vector<T> v(count);
    in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(count * sizeof(T)));
    if (!in) throw std::runtime_error("Cannot read expected payload from: " + path);
    return v;
}

/* int32 Nx, Ny, then Nx*Ny values */
template <class T>
void write_grid(const std::string& path, int Nx, int Ny, const std::vector<T>& data, const char* fn,
                const char* kind) {
    if (data.size() != (size_t)Nx * (size_t)Ny) throw std::runtime_error(std::string(fn) + ": unexpected data size");
    std::ofstream out(path, std::ios::binary);
    if (!out) throw std::runtime_error(std::string("Cannot open ") + kind + " grid file for writing: " + path);
    const int32_t hdr[2] = {(int32_t)Nx, (int32_t)Ny};
    out.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
    out.write(reinterpret_cast<const char*>(data.data()), (std::streamsize)(data.size() * sizeof(T)));
    if (!out) throw std::runtime_error(std::string("Failed while writing ") + kind + " grid payload");
}

}  // namespace

Params read_params_kv(const std::string& filepath) {
    std::ifstream in(filepath);
    if (!in) throw std::runtime_error("Cannot open params file: " + filepath);
    std::unordered_map<std::string, std::string> kv;
    std::string line;
    while (std::getline(in, line)) {
        line = strip(line);
        if (line.empty() || line[0] == '#') continue;
        const size_t eq = line.find('=');
        if (eq == std::string::npos) throw std::runtime_error("Invalid key=value line in params file: " + line);
        kv[strip(line.substr(0, eq))] = strip(line.substr(eq + 1)); /* a repeated key keeps the last value */
    }
    Params p{};
    for (const Field& f : kFields) {
        const auto it = kv.find(f.key);
        if (it == kv.end()) throw std::runtime_error(std::string("Missing required params key: ") + f.key);
        const std::string& s = it->second;
        if (f.i) p.*f.i = std::stoi(s);
        else if (f.d) p.*f.d = std::stod(s);
        else if (f.b) p.*f.b = to_bool(s);
        else p.*f.s = s;
    }
    return p;
}

void write_runout_kv(const std::string& filepath, const RunOutStep& r) {
    std::ofstream out(filepath);
    if (!out) throw std::runtime_error("Cannot open runout file for writing: " + filepath);
    out << "inputTag=" << r.inputTag << '\n'
        << "outputTag=" << r.outputTag << '\n'
        << "pBotInst=" << r.pBotInst << '\n'
        << "pTopInst=" << r.pTopInst << '\n'
        << "pMeanInst=" << r.pMeanInst << '\n'
        << "layoutMode=" << r.layoutMode << '\n'
        << "nZonesExecuted=" << r.nZonesExecuted << '\n';
}

State read_state(const std::string& x_path, const std::string& v_path, const std::string& type_path,
                 const std::string& r0_path, bool has_type, bool has_r0, int n_expected) {
    State s;
    const size_t n = (size_t)n_expected;
    s.x = read_raw<double>(x_path, 2 * n);
    s.v = read_raw<double>(v_path, 2 * n);
    if (has_type) s.type = read_raw<std::uint8_t>(type_path, n);
    if (has_r0) s.r0 = read_raw<double>(r0_path, 2 * n);
    return s;
}

void write_xy_interleaved(const std::string& filepath, const std::vector<double>& data, int n_expected) {
    if (data.size() != 2 * (size_t)n_expected)
        throw std::runtime_error("Interleaved XY buffer has unexpected size for write: " + filepath);
    write_raw(filepath, data.data(), data.size(), "binary file");
}

void write_u8(const std::string& filepath, const std::vector<std::uint8_t>& data, int n_expected) {
    if (data.size() != (size_t)n_expected) throw std::runtime_error("u8 buffer has unexpected size for write: " + filepath);
    write_raw(filepath, data.data(), data.size(), "binary file");
}

void write_cell_fields_bin(const std::string& filepath, const CellFields& G) {
    std::ofstream out(filepath, std::ios::binary);
    if (!out) throw std::runtime_error("Cannot open cell field file for writing: " + filepath);
    const int32_t hdr[2] = {(int32_t)G.Nx, (int32_t)G.Ny};
    out.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
    out.write(reinterpret_cast<const char*>(G.N.data()), (std::streamsize)(G.N.size() * sizeof(int32_t)));
    for (const std::vector<double>* f : {&G.Ux, &G.Uy, &G.Px, &G.Py, &G.T, &G.rho, &G.Pkin, &G.Pvir, &G.P, &G.rhoTarget})
        out.write(reinterpret_cast<const char*>(f->data()), (std::streamsize)(f->size() * sizeof(double)));
    if (!out) throw std::runtime_error("Failed while writing cell field payload");
}

void write_i32_grid_bin(const std::string& filepath, int Nx, int Ny, const std::vector<std::int32_t>& data) {
    write_grid(filepath, Nx, Ny, data, "write_i32_grid_bin", "i32");
}

void write_f64_grid_bin(const std::string& filepath, int Nx, int Ny, const std::vector<double>& data) {
    write_grid(filepath, Nx, Ny, data, "write_f64_grid_bin", "f64");
}

void write_i32_grid_bin_lc(const std::string& filepath, int Nx, int Ny, const std::vector<std::int32_t>& data) {
    write_grid(filepath, Nx, Ny, data, "write_i32_grid_bin_lc", "i32");
}

}  // namespace mpcd
