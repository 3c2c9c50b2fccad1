// Host side: configuration, input files and dual-grid geometry (see tbfm_host.hpp).
#include "tbfm_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace tbfm {

namespace {

std::string trim(const std::string& s, const char* ws = " \t\r\n") {
    size_t a = s.find_first_not_of(ws);
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(ws);
    return s.substr(a, b - a + 1);
}

std::string lower(std::string s) {
    for (auto& ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}

double toDouble(const std::string& key, const std::string& v) {
    char* end = nullptr;
    double d = std::strtod(v.c_str(), &end);
    if (end == v.c_str())
        throw std::runtime_error("Invalid float format for key \"" + key + "\": \"" + v + "\"");
    return d;
}

struct Vec {
    double x, y, z;
};
inline Vec operator+(Vec a, Vec b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec operator-(Vec a, Vec b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec operator*(Vec a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline Vec operator/(Vec a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline Vec operator-(Vec a) { return {-a.x, -a.y, -a.z}; }
inline double norm(Vec a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline Vec cross(Vec a, Vec b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

}  // namespace

// ------------------------------------------------------------------------ Config
Config::Config(const std::string& path) {
    std::ifstream f(path);
    if (!f.is_open()) throw std::runtime_error("Failed to load configuration file: " + path);
    size_t slash = path.find_last_of('/');
    baseDir = (slash == std::string::npos) ? std::string(".") : path.substr(0, slash);
    std::string line;
    while (std::getline(f, line)) {
        line = trim(line);
        if (line.empty() || line[0] == ';' || line[0] == '[') continue;
        size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        data_[trim(line.substr(0, eq), " \t")] = trim(line.substr(eq + 1), " \t");
    }
}

std::string Config::get(const std::string& key) const {
    auto it = data_.find(key);
    return it == data_.end() ? std::string() : it->second;
}
double Config::getFloat(const std::string& key) const {
    if (!has(key)) throw std::runtime_error("Not found value for key \"" + key + "\" in configuration.");
    return toDouble(key, get(key));
}
double Config::getFloat(const std::string& key, double def) const {
    return has(key) ? toDouble(key, get(key)) : def;
}
bool Config::getBool(const std::string& key, bool def) const {
    if (!has(key)) return def;
    std::string v = lower(get(key));
    if (v == "true" || v == "yes" || v == "1") return true;
    if (v == "false" || v == "no" || v == "0") return false;
    throw std::runtime_error("Invalid value for key \"" + key + "\" in configuration.");
}
std::vector<std::string> Config::getList(const std::string& key) const {
    if (!has(key)) throw std::runtime_error("Missing value for key \"" + key + "\" in configuration.");
    std::vector<std::string> out;
    std::stringstream ss(get(key));
    std::string item;
    while (std::getline(ss, item, ',')) {
        item = trim(item);
        if (!item.empty()) out.push_back(item);
    }
    return out;
}
Topology Config::getTopology() const {
    std::string v = get("TOPOLOGY");
    if (v == "2D" || v == "2d") return Topology::TwoD;
    if (v == "3D" || v == "3d") return Topology::ThreeD;
    if (v == "axisymmetric" || v == "Axisymmetric" || v == "AXISYMMETRIC") return Topology::Axisymmetric;
    if (v == "1D" || v == "1d") return Topology::OneD;
    throw std::runtime_error("Invalid value for key \"TOPOLOGY\" in configuration.");
}
Scheme Config::getAdvectionScheme() const {
    std::string v = get("CONVECTION_SCHEME");
    if (v == "JST") return Scheme::JST;
    if (v == "ROE" || v == "roe" || v == "Roe") return Scheme::ROE;
    throw std::runtime_error("Invalid value for key \"CONVECTION_SCHEME\" in configuration.");
}
Limiter Config::getFluxLimiter() const {
    std::string v = has("FLUX_LIMITER") ? get("FLUX_LIMITER") : "None";
    if (v == "None") return Limiter::None;
    if (v == "van albada") return Limiter::VanAlbada;
    if (v == "van leer") return Limiter::VanLeer;
    if (v == "min mod") return Limiter::MinMod;
    throw std::runtime_error("Invalid value for key \"FLUX_LIMITER\" in configuration.");
}
std::vector<double> Config::getTimeIntegrationCoeffs() const {
    std::string key = has("TIME_INTEGRATION") ? "TIME_INTEGRATION" : "TIME_INTEGRATION_TYPE";
    std::string s;
    for (char ch : get(key))
        if (!std::isspace((unsigned char)ch)) s.push_back((char)std::tolower((unsigned char)ch));
    if (!has(key) || s == "0" || s == "explicit_rk4" || s == "rk4" || s == "runge_kutta_4")
        return {1.0 / 4.0, 5.0 / 14.0, 14.0 / 25.0, 1.0};
    if (s == "1" || s == "explicit_rk3" || s == "rk3" || s == "runge_kutta_3") return {8.0 / 17.0, 17.0 / 20.0, 1.0};
    throw std::runtime_error("TIME_INTEGRATION_TYPE \"" + get(key) +
                             "\": only the explicit RK41A (0) and RK31B (1) schemes run on the GPU path.");
}
TimeStepMethod Config::getTimeStepMethod() const {
    std::string v = get("TIME_STEP_METHOD");
    if (v == "local") return TimeStepMethod::Local;
    if (v == "global") return TimeStepMethod::Global;
    if (v == "fixed") return TimeStepMethod::Fixed;
    throw std::runtime_error("Invalid value for key \"TIME_STEP_METHOD\" in configuration.");
}
double Config::getCFL() const {
    if (has("CFL")) return getFloat("CFL");
    if (has("CFL_START")) return getFloat("CFL_START");
    if (has("CFL_MAX")) return getFloat("CFL_MAX");
    return getFloat("CFL");
}
std::vector<double> Config::getInitDirection() const {
    if (!has("INIT_DIRECTION")) return {1.0, 0.0, 0.0};
    if (get("INIT_DIRECTION") == "adaptive")
        throw std::runtime_error("INIT_DIRECTION = adaptive is not supported on the GPU path");
    std::vector<double> d;
    for (auto& s : getList("INIT_DIRECTION")) d.push_back(toDouble("INIT_DIRECTION", s));
    if (d.size() != 3) throw std::runtime_error("INIT_DIRECTION must have 3 components or be 'adaptive'");
    return d;
}
BfmModel Config::getBFMModel() const {
    std::string v = has("BFM_MODEL") ? get("BFM_MODEL") : "None";
    if (v == "Hall") return BfmModel::Hall;
    if (v == "Lift-Drag") return BfmModel::LiftDrag;
    if (v == "None") return BfmModel::None;
    if (v == "Blockage") return BfmModel::Blockage;
    if (v == "Hall-Thollet") return BfmModel::HallThollet;
    if (v == "Chima" || v == "Correlations")
        throw std::runtime_error("BFM_MODEL = " + v + " is outside the GPU residual path (Hall, Hall-Thollet, Lift-Drag)");
    throw std::runtime_error("Invalid value for key \"BFM_MODEL\" in configuration.");
}
bool Config::isInletCylindrical() const {
    std::string v = has("INLET_REFERENCE_FRAME") ? get("INLET_REFERENCE_FRAME") : "None";
    if (v == "Cylindrical" || v == "cylindrical" || v == "CYLINDRICAL") return true;
    if (v == "Cartesian" || v == "cartesian" || v == "CARTESIAN" || v == "None") return false;
    throw std::runtime_error("Invalid value for key \"INLET_REFERENCE_FRAME\" in configuration.");
}
int Config::getTurboDirection() const {
    std::string v = has("TURBO_DIRECTION") ? get("TURBO_DIRECTION") : "i";
    if (v == "-i") return 1;
    if (v == "j" || v == "+j") return 2;
    if (v == "-j") return 3;
    return 0;
}
Precision Config::getGpuPrecision() const {
    std::string v = has("GPU_PRECISION") ? lower(get("GPU_PRECISION")) : "fp64";
    if (v == "fp64" || v == "double") return Precision::FP64;
    if (v == "mixed" || v == "fp32_flux") return Precision::Mixed;
    throw std::runtime_error("Invalid value for key \"GPU_PRECISION\" (fp64 | mixed)");
}

// Config::getFluidModel / getFluidName / getFluidTableFile (src/config.cpp:468-497)
FluidModel Config::getFluidModel() const {
    if (!has("FLUID_MODEL")) return FluidModel::Ideal;
    std::string v = get("FLUID_MODEL");
    if (v == "None" || v == "Ideal" || v == "ideal" || v == "IDEAL") return FluidModel::Ideal;
    if (v == "real" || v == "Real" || v == "REAL") return FluidModel::Real;
    throw std::runtime_error("Invalid value for key \"FLUID_MODEL\" in configuration.");
}
std::string Config::getFluidName() const {
    if (has("FLUID_NAME")) return get("FLUID_NAME");
    if (has("FLUID")) return get("FLUID");
    return "CO2";
}
std::string Config::getFluidTableFile() const {
    if (has("FLUID_TABLE_FILE")) return get("FLUID_TABLE_FILE");
    return "fluid_table_" + getFluidName() + ".lut";
}

FluidTable readFluidTable(const std::string& path) {
    std::ifstream f(path);
    if (!f.is_open()) throw std::runtime_error("FluidReal: Could not open table file at: " + path);
    FluidTable out;
    // section name, its six bound keys (the last one starts the rows), fields, first table
    struct Sec {
        const char* name;
        const char* keys[6];
        int nf, base;
    };
    static const Sec secs[3] = {
        {"RHO_E", {"N_RHO", "N_E", "RHO_MIN", "RHO_MAX", "E_MIN", "E_MAX"}, 8, 0},
        {"P_T", {"N_P", "N_T", "P_MIN", "P_MAX", "T_MIN", "T_MAX"}, 4, 8},
        {"P_S", {"N_P", "N_S", "P_MIN", "P_MAX", "S_MIN", "S_MAX"}, 3, 12},
    };
    const Sec* cur = nullptr;
    std::map<std::string, std::string> meta;
    auto trimKey = [](std::string k) { k.erase(k.find_last_not_of(" \t\r\n") + 1); return k; };
    auto trimVal = [](std::string v) {
        size_t a = v.find_first_not_of(" \t\r\n");
        return a == std::string::npos ? std::string() : v.substr(a);
    };
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '\r') continue;
        bool marker = false;
        for (const Sec& s : secs)
            if (line.find(std::string("# SECTION: ") + s.name) != std::string::npos) {
                cur = &s;
                meta.clear();
                marker = true;
                break;
            }
        if (marker || line[0] == '#') continue;
        const size_t colon = line.find(':');
        if (colon == std::string::npos) continue;
        const std::string key = trimKey(line.substr(0, colon)), val = trimVal(line.substr(colon + 1));
        if (!cur) {
            out.header[key] = val;
            continue;
        }
        meta[key] = val;
        if (key != cur->keys[5]) continue;
        const long nx = std::stol(meta.at(cur->keys[0])), ny = std::stol(meta.at(cur->keys[1]));
        // bilinear interpolation reads the cell (i, j)..(i+1, j+1): two nodes per direction
        if (nx < 2 || ny < 2)
            throw std::runtime_error("FluidReal: the " + std::string(cur->name) + " table in " + path +
                                     " needs at least 2 x 2 nodes");
        for (int q = 0; q < cur->nf; ++q) {
            FluidTable::Tab& T = out.t[cur->base + q];
            T.nx = nx; T.ny = ny;
            T.xmin = std::stod(meta.at(cur->keys[2])); T.xmax = std::stod(meta.at(cur->keys[3]));
            T.ymin = std::stod(meta.at(cur->keys[4])); T.ymax = std::stod(val);
            T.v.assign((size_t)(nx * ny), 0.0);
        }
        // nx * ny rows; empty, '#' and '\r' lines in between are skipped
        for (long r = 0; r < nx * ny; ++r) {
            if (!std::getline(f, line)) break;
            while (line.empty() || line[0] == '#' || line[0] == '\r')
                if (!std::getline(f, line)) break;
            std::istringstream iss(line);
            for (int q = 0; q < cur->nf; ++q) {
                double x = 0.0;
                iss >> x;
                out.t[cur->base + q].v[(size_t)r] = x;
            }
        }
    }
    for (int t = 0; t < FluidTable::kTables; ++t)
        if (out.t[t].v.empty()) throw std::runtime_error("FluidReal: table file " + path + " lacks a section");
    return out;
}

PatchType patchTypeFromString(const std::string& s) {
    std::string k = lower(s);
    if (k == "inlet") return PatchType::Inlet;
    if (k == "outlet") return PatchType::Outlet;
    if (k == "radial_equilibrium") return PatchType::RadialEquilibrium;
    if (k == "inviscid_wall") return PatchType::Wall;
    if (k == "no_slip_wall") return PatchType::NoSlipWall;
    if (k == "periodic") return PatchType::Periodic;
    if (k == "wedge") return PatchType::Wedge;
    if (k == "throttle") return PatchType::Throttle;
    if (k == "mass_flow") return PatchType::MassFlow;
    if (k == "transparent") return PatchType::Transparent;
    throw std::runtime_error("BoundaryType " + s + " is outside the GPU residual path");
}

// ---------------------------------------------------------------------- readers
GridInput readGridCsv(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("Error opening coordinates CSV file '" + path + "'");
    GridInput g;
    std::string line;
    std::vector<std::string> header;
    while (std::getline(f, line)) {
        if (line.rfind("NI=", 0) == 0) g.ni = std::stol(line.substr(3));
        else if (line.rfind("NJ=", 0) == 0) g.nj = std::stol(line.substr(3));
        else if (line.rfind("NK=", 0) == 0) g.nk = std::stol(line.substr(3));
        else if (line.rfind("x,y,z", 0) == 0) {
            std::stringstream hs(trim(line));
            std::string col;
            while (std::getline(hs, col, ',')) header.push_back(trim(col));
            break;
        }
    }
    const long N = g.ni * g.nj * g.nk;
    if (N <= 0 || header.size() < 3) throw std::runtime_error("Malformed grid file " + path);
    std::vector<std::vector<double>*> cols;
    g.x.assign(N, 0.0);
    g.y.assign(N, 0.0);
    g.z.assign(N, 0.0);
    for (auto& h : header) {
        if (h == "x") cols.push_back(&g.x);
        else if (h == "y") cols.push_back(&g.y);
        else if (h == "z") cols.push_back(&g.z);
        else {
            g.fields[h].assign(N, 0.0);
            cols.push_back(&g.fields[h]);
        }
    }
    long n = 0;
    std::vector<double> vals;
    vals.reserve(header.size());
    while (std::getline(f, line) && n < N) {
        vals.clear();
        const char* p = line.c_str();
        char* end = nullptr;
        while (*p) {
            double v = std::strtod(p, &end);
            if (end == p) break;
            vals.push_back(v);
            p = end;
            while (*p == ',' || *p == ' ' || *p == '\r') ++p;
        }
        if (vals.size() != header.size()) continue;
        for (size_t c = 0; c < cols.size(); ++c) (*cols[c])[n] = vals[c];
        ++n;
    }
    if (n != N) throw std::runtime_error("Number of points in the grid file is not equal to ni*nj*nk");
    return g;
}

std::vector<Patch> readBoundaryCsv(const std::string& path, const Config& cfg) {
    std::ifstream f(path);
    if (!f.is_open()) throw std::runtime_error("Cannot open file: " + path);
    std::vector<Patch> out;
    std::string line;
    bool data = false;
    while (std::getline(f, line)) {
        if (line.empty() || line.find('=') != std::string::npos) continue;
        if (line.find("PATCH_NAME") != std::string::npos) {
            data = true;
            continue;
        }
        if (!data) continue;
        std::stringstream ss(line);
        std::string tok;
        std::vector<std::string> t;
        while (std::getline(ss, tok, ',')) {
            tok.erase(std::remove(tok.begin(), tok.end(), ' '), tok.end());
            t.push_back(tok);
        }
        if (t.size() < 7) continue;
        Patch p;
        p.name = t[0];
        p.i0 = std::stol(t[1]); p.i1 = std::stol(t[2]);
        p.j0 = std::stol(t[3]); p.j1 = std::stol(t[4]);
        p.k0 = std::stol(t[5]); p.k1 = std::stol(t[6]);
        auto info = cfg.getList(p.name);
        if (info.empty()) throw std::runtime_error("Boundary info for " + p.name + " is missing in configuration.");
        p.type = patchTypeFromString(info[0]);
        for (size_t q = 1; q < info.size(); ++q) p.values.push_back(toDouble(p.name, info[q]));
        out.push_back(p);
    }
    return out;
}

// --------------------------------------------------------------------- geometry
namespace {

// Face vector and centre of the quadrilateral p1-p2-p3-p4 split in two triangles
// (src/mesh.cpp:165-190).
void quadFace(Vec p1, Vec p2, Vec p3, Vec p4, double* S, double* C) {
    Vec t1 = cross(p2 - p1, p3 - p2) / 2.0;
    Vec t2 = cross(p3 - p1, p4 - p3) / 2.0;
    Vec n = t1 + t2;
    Vec c1 = (p1 + p2 + p3) / 3.0;
    Vec c2 = (p1 + p3 + p4) / 3.0;
    double a1 = norm(t1), a2 = norm(t2);
    Vec c = (c1 * a1 + c2 * a2) / (a1 + a2);
    S[0] = n.x; S[1] = n.y; S[2] = n.z;
    C[0] = c.x; C[1] = c.y; C[2] = c.z;
}

}  // namespace

Geometry buildGeometry(const GridInput& gr, Topology topo, bool withGrad) {
    Geometry g;
    const long NI = gr.ni, NJ = gr.nj, NK = gr.nk, N = NI * NJ * NK;
    const long DI = NI + 1, DJ = NJ + 1, DK = NK + 1;
    g.ni = NI; g.nj = NJ; g.nk = NK;
    auto P = [&](long i, long j, long k) {
        long n = (i * NJ + j) * NK + k;
        return Vec{gr.x[n], gr.y[n], gr.z[n]};
    };
    g.dual.assign(DI * DJ * DK * 3, 0.0);
    auto D = [&](long i, long j, long k, Vec v) {
        double* d = &g.dual[((i * DJ + j) * DK + k) * 3];
        d[0] = v.x; d[1] = v.y; d[2] = v.z;
    };
    if (topo == Topology::ThreeD) {
        // src/mesh.cpp:478-557: volume-interior dual nodes average 8 points, boundary
        // faces 4, edges 2, corners copy the grid corner.
        for (long i = 0; i < DI; ++i) {
            const bool iLo = (i == 0), iHi = (i == DI - 1);
            for (long j = 0; j < DJ; ++j) {
                const bool jLo = (j == 0), jHi = (j == DJ - 1);
                for (long k = 0; k < DK; ++k) {
                    const bool kLo = (k == 0), kHi = (k == DK - 1);
                    const int nb = (int)(iLo || iHi) + (int)(jLo || jHi) + (int)(kLo || kHi);
                    const long ia = iLo ? 0 : (iHi ? NI - 1 : i - 1), ib = iLo ? 0 : (iHi ? NI - 1 : i);
                    const long ja = jLo ? 0 : (jHi ? NJ - 1 : j - 1), jb = jLo ? 0 : (jHi ? NJ - 1 : j);
                    const long ka = kLo ? 0 : (kHi ? NK - 1 : k - 1), kb = kLo ? 0 : (kHi ? NK - 1 : k);
                    Vec v;
                    if (nb == 0) {
                        v = (P(ia, ja, ka) + P(ib, ja, ka) + P(ia, jb, ka) + P(ia, ja, kb) + P(ib, jb, ka) +
                             P(ib, ja, kb) + P(ia, jb, kb) + P(ib, jb, kb)) / 8.0;
                    } else if (nb == 3) {
                        v = P(ia, ja, ka);
                    } else if (nb == 2) {
                        // edge along the one interior direction
                        if (!(iLo || iHi)) v = (P(ia, ja, ka) + P(ib, ja, ka)) / 2.0;
                        else if (!(jLo || jHi)) v = (P(ia, ja, ka) + P(ia, jb, ka)) / 2.0;
                        else v = (P(ia, ja, ka) + P(ia, ja, kb)) / 2.0;
                    } else {
                        // boundary face: average of the 4 points of the face
                        if (iLo || iHi) v = (P(ia, ja, ka) + P(ia, jb, ka) + P(ia, ja, kb) + P(ia, jb, kb)) / 4.0;
                        else if (jLo || jHi) v = (P(ia, ja, ka) + P(ib, ja, ka) + P(ia, ja, kb) + P(ib, ja, kb)) / 4.0;
                        else v = (P(ia, ja, ka) + P(ib, ja, ka) + P(ia, jb, ka) + P(ib, jb, ka)) / 4.0;
                    }
                    D(i, j, k, v);
                }
            }
        }
    } else {
        // src/mesh.cpp:373-475: planar dual nodes, then extruded into 2 layers
        std::vector<Vec> nodes(DI * DJ);
        auto ND = [&](long i, long j) -> Vec& { return nodes[i * DJ + j]; };
        if (topo == Topology::OneD) {
            for (long i = 1; i < DI - 1; ++i) ND(i, 0) = (P(i - 1, 0, 0) + P(i, 0, 0)) / 2.0;
            ND(0, 0) = P(0, 0, 0);
            ND(DI - 1, 0) = P(NI - 1, 0, 0);
            for (long i = 0; i < DI; ++i) {
                double xx = ND(i, 0).x;
                D(i, 0, 0, Vec{xx, -0.5, -0.5});
                D(i, 1, 0, Vec{xx, 0.5, -0.5});
                D(i, 0, 1, Vec{xx, -0.5, 0.5});
                D(i, 1, 1, Vec{xx, 0.5, 0.5});
            }
        } else {
            for (long i = 0; i < DI; ++i)
                for (long j = 0; j < DJ; ++j) {
                    const bool ie = (i == 0 || i == DI - 1), je = (j == 0 || j == DJ - 1);
                    const long ia = (i == 0) ? 0 : (i == DI - 1 ? NI - 1 : i - 1), ib = (i == 0) ? 0 : (i == DI - 1 ? NI - 1 : i);
                    const long ja = (j == 0) ? 0 : (j == DJ - 1 ? NJ - 1 : j - 1), jb = (j == 0) ? 0 : (j == DJ - 1 ? NJ - 1 : j);
                    if (!ie && !je) ND(i, j) = (P(ia, ja, 0) + P(ib, ja, 0) + P(ib, jb, 0) + P(ia, jb, 0)) / 4.0;
                    else if (ie && je) ND(i, j) = P(ia, ja, 0);
                    else if (je) ND(i, j) = (P(ia, ja, 0) + P(ib, ja, 0)) / 2.0;
                    else ND(i, j) = (P(ia, ja, 0) + P(ia, jb, 0)) / 2.0;
                }
            if (topo == Topology::Axisymmetric) {
                g.wedgeAngle = 1.0 * M_PI / 180.0;
                const double w = g.wedgeAngle;
                for (long i = 0; i < DI; ++i)
                    for (long j = 0; j < DJ; ++j) {
                        Vec n = ND(i, j);
                        D(i, j, 0, Vec{n.x, n.y * std::cos(-w / 2.0), n.y * std::sin(-w / 2.0)});
                        D(i, j, 1, Vec{n.x, n.y * std::cos(+w / 2.0), n.y * std::sin(+w / 2.0)});
                    }
            } else {
                for (long i = 0; i < DI; ++i)
                    for (long j = 0; j < DJ; ++j) {
                        Vec n = ND(i, j);
                        D(i, j, 0, Vec{n.x, n.y, -1.0 / 2.0});
                        D(i, j, 1, Vec{n.x, n.y, +1.0 / 2.0});
                    }
            }
        }
    }
    auto DN = [&](long i, long j, long k) {
        const double* d = &g.dual[((i * DJ + j) * DK + k) * 3];
        return Vec{d[0], d[1], d[2]};
    };
    g.surfI.assign(DI * NJ * NK * 3, 0.0); g.centI = g.surfI;
    g.surfJ.assign(NI * DJ * NK * 3, 0.0); g.centJ = g.surfJ;
    g.surfK.assign(NI * NJ * DK * 3, 0.0); g.centK = g.surfK;
    for (long i = 0; i < DI; ++i)
        for (long j = 0; j < NJ; ++j)
            for (long k = 0; k < NK; ++k) {
                long f = ((i * NJ + j) * NK + k) * 3;
                quadFace(DN(i, j, k), DN(i, j + 1, k), DN(i, j + 1, k + 1), DN(i, j, k + 1), &g.surfI[f], &g.centI[f]);
            }
    for (long i = 0; i < NI; ++i)
        for (long j = 0; j < DJ; ++j)
            for (long k = 0; k < NK; ++k) {
                long f = ((i * DJ + j) * NK + k) * 3;
                quadFace(DN(i, j, k), DN(i, j, k + 1), DN(i + 1, j, k + 1), DN(i + 1, j, k), &g.surfJ[f], &g.centJ[f]);
            }
    for (long i = 0; i < NI; ++i)
        for (long j = 0; j < NJ; ++j)
            for (long k = 0; k < DK; ++k) {
                long f = ((i * NJ + j) * DK + k) * 3;
                quadFace(DN(i, j, k), DN(i + 1, j, k), DN(i + 1, j + 1, k), DN(i, j + 1, k), &g.surfK[f], &g.centK[f]);
            }
    auto fI = [&](long i, long j, long k) { return ((i * NJ + j) * NK + k) * 3; };
    auto fJ = [&](long i, long j, long k) { return ((i * DJ + j) * NK + k) * 3; };
    auto fK = [&](long i, long j, long k) { return ((i * NJ + j) * DK + k) * 3; };
    // volumes: divergence theorem with the x components only (src/mesh.cpp:192-230)
    g.volume.assign(N, 0.0);
    for (long i = 0; i < NI; ++i)
        for (long j = 0; j < NJ; ++j)
            for (long k = 0; k < NK; ++k) {
                const long faces[6] = {fI(i, j, k), fJ(i, j, k), fK(i, j, k), fI(i + 1, j, k), fJ(i, j + 1, k), fK(i, j, k + 1)};
                const double* Sx[6] = {&g.surfI[faces[0]], &g.surfJ[faces[1]], &g.surfK[faces[2]],
                                       &g.surfI[faces[3]], &g.surfJ[faces[4]], &g.surfK[faces[5]]};
                const double* Cx[6] = {&g.centI[faces[0]], &g.centJ[faces[1]], &g.centK[faces[2]],
                                       &g.centI[faces[3]], &g.centJ[faces[4]], &g.centK[faces[5]]};
                double v = 0.0;
                for (int q = 0; q < 6; ++q) v += (q < 3 ? -Sx[q][0] : Sx[q][0]) * Cx[q][0];
                g.volume[(i * NJ + j) * NK + k] = v;
            }
    g.radius.assign(N, 0.0);
    g.theta.assign(N, 0.0);
    for (long n = 0; n < N; ++n) {
        g.radius[n] = std::sqrt(gr.y[n] * gr.y[n] + gr.z[n] * gr.z[n]);
        g.theta[n] = std::atan2(gr.z[n], gr.y[n]);
    }
    // blockage gradient, Green-Gauss with distance-weighted face values
    // (src/mesh.cpp:571-581, src/math_utils.cpp:173-234 and 395-462)
    const std::vector<double>* b = gr.field("blockage");
    if (withGrad && b) {
        g.gradBlockage.assign(3 * N, 0.0);
        for (long i = 0; i < NI; ++i)
            for (long j = 0; j < NJ; ++j)
                for (long k = 0; k < NK; ++k) {
                    const long n0 = (i * NJ + j) * NK + k;
                    const Vec p0 = P(i, j, k);
                    const double f0 = (*b)[n0];
                    struct Side { long face; const std::vector<double>* S; const std::vector<double>* C; double sign; long ni, nj, nk; bool edge; };
                    const Side sides[6] = {
                        {fI(i, j, k), &g.surfI, &g.centI, -1.0, i - 1, j, k, i == 0},
                        {fI(i + 1, j, k), &g.surfI, &g.centI, 1.0, i + 1, j, k, i == NI - 1},
                        {fJ(i, j + 1, k), &g.surfJ, &g.centJ, 1.0, i, j + 1, k, j == NJ - 1},
                        {fJ(i, j, k), &g.surfJ, &g.centJ, -1.0, i, j - 1, k, j == 0},
                        {fK(i, j, k), &g.surfK, &g.centK, -1.0, i, j, k - 1, k == 0},
                        {fK(i, j, k + 1), &g.surfK, &g.centK, 1.0, i, j, k + 1, k == NK - 1}};
                    Vec acc{0.0, 0.0, 0.0};
                    for (const Side& s : sides) {
                        const double* sv = &(*s.S)[s.face];
                        const double* cv = &(*s.C)[s.face];
                        Vec Sv = (s.sign < 0) ? -Vec{sv[0], sv[1], sv[2]} : Vec{sv[0], sv[1], sv[2]};
                        double val = f0;
                        if (!s.edge) {
                            Vec cc{cv[0], cv[1], cv[2]};
                            double d0 = norm(cc - p0), d1 = norm(cc - P(s.ni, s.nj, s.nk));
                            double f1 = (*b)[(s.ni * NJ + s.nj) * NK + s.nk];
                            val = f0 + (f1 - f0) * d0 / (d0 + d1);
                        }
                        acc = acc + Sv * val;
                    }
                    Vec gv = acc / g.volume[n0];
                    g.gradBlockage[3 * n0 + 0] = gv.x;
                    g.gradBlockage[3 * n0 + 1] = gv.y;
                    g.gradBlockage[3 * n0 + 2] = gv.z;
                }
    }
    return g;
}

}  // namespace tbfm
