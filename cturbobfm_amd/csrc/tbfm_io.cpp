// On-disk formats around the residual path, host side (see tbfm_io.hpp).
//
//   volume solution file      src/output.cpp:87-131 (field sets), :150-245 (values),
//                             :352-364 (file name), :401-447 (header and rows)
//   fluid relations           src/fluid_ideal.cpp, include/fluid_base.hpp:78-124,
//                             src/fluid_real.cpp:227-341
//   initial state             src/solver.cpp:853-888 (scratch), :890-1295 (restart),
//                             include/kdtree.hpp (nearest neighbour)
#include "tbfm_io.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <functional>
#include <iostream>
#include <limits>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <thread>
#include <unordered_map>

namespace tbfm {

// ------------------------------------------------------------------ fluid relations

void HostFluid::load(const Config& cfg, const std::string& baseDir) {
    real = (cfg.getFluidModel() == FluidModel::Real);
    if (!real) {
        gamma = cfg.getFluidGamma();
        R = cfg.getFluidGasConstant();
    } else {
        // Solver::buildFluidModel (solver.cpp:49-74); no table generation here: the
        // table file has to exist
        std::string path = cfg.getFluidTableFile();
        if (!path.empty() && path[0] != '/' && !baseDir.empty()) path = baseDir + "/" + path;
        const FluidTable ft = readFluidTable(path);
        static_assert(FluidTable::kTables == LUT_COUNT, "table count");
        size_t total = 0;
        for (int t = 0; t < LUT_COUNT; ++t) total += ft.t[t].v.size();
        values.resize(total);
        size_t o = 0;
        for (int t = 0; t < LUT_COUNT; ++t) {
            const FluidTable::Tab& F = ft.t[t];
            LutTab& T = lut.t[t];
            T.nx = (int)F.nx; T.ny = (int)F.ny;
            T.xmin = F.xmin; T.xmax = F.xmax; T.ymin = F.ymin; T.ymax = F.ymax;
            // Table2D::init (fluid_real.hpp:22-32)
            T.dx = (F.nx > 1) ? (F.xmax - F.xmin) / (double)(F.nx - 1) : 1.0;
            T.dy = (F.ny > 1) ? (F.ymax - F.ymin) / (double)(F.ny - 1) : 1.0;
            std::copy(F.v.begin(), F.v.end(), values.begin() + o);
            T.d = values.data() + o;
            o += F.v.size();
        }
        header.assign(ft.header.begin(), ft.header.end());
        auto hdr = [&](const char* k, double def) {
            for (auto& kv : header)
                if (kv.first == k) return std::stod(kv.second);
            return def;
        };
        // gamma and R only feed ideal-gas formulas, which this fluid never reaches
        gamma = cfg.has("FLUID_GAMMA") ? cfg.getFluidGamma() : hdr("NOMINAL_GAMMA", 1.4);
        R = cfg.has("FLUID_R_CONSTANT") ? cfg.getFluidGasConstant() : hdr("GAS_CONSTANT", 287.0);
    }
    // FluidIdeal::FluidIdeal (fluid_ideal.cpp:4-8)
    cp = (gamma * R) / (gamma - 1);
    cv = cp / gamma;
}

double HostFluid::e_u_et(double ux, double uy, double uz, double et) const {
    if (!real) {
        const double m = std::sqrt(ux * ux + uy * uy + uz * uz);
        return et - 0.5 * (m * m);
    }
    return et - 0.5 * (ux * ux + uy * uy + uz * uz);
}
double HostFluid::p_rho_e(double rho, double e) const {
    return real ? lut_at(lut, LUT_P, rho, e) : (gamma - 1) * rho * e;
}
double HostFluid::T_rho_e(double rho, double e) const {
    // ideal: p / (rho R) with p = (gamma-1) rho e (computeTemperature_rho_u_et)
    return real ? lut_at(lut, LUT_T, rho, e) : p_rho_e(rho, e) / (rho * R);
}
double HostFluid::a_rho_e(double rho, double e) const {
    // ideal: sqrt(gamma p / rho) (computeSoundSpeed_rho_u_et, fluid_ideal.cpp:83-87)
    return real ? lut_at(lut, LUT_A, rho, e) : std::sqrt(gamma * p_rho_e(rho, e) / rho);
}
double HostFluid::s_rho_e(double rho, double e) const {
    if (real) return lut_at(lut, LUT_S, rho, e);
    // computeEntropy_p_T (fluid_ideal.cpp:193-196) of p and T
    const double p = p_rho_e(rho, e);
    const double T = T_rho_e(rho, e);
    return cp * std::log(T / 288.15) - R * std::log(p / 101325);
}
double HostFluid::p_rho_T(double rho, double T) const {
    if (!real) return rho * R * T;
    return lut_p_rho_T(lut, rho, T);
}
double HostFluid::e_p_rho(double p, double rho) const {
    return real ? lut_e_p_rho(lut, p, rho) : p / ((gamma - 1) * rho);
}
double HostFluid::total_pressure(double rho, double ux, double uy, double uz, double et) const {
    if (real) return lut_total_pressure(lut, rho, ux * ux + uy * uy + uz * uz, et);
    // fluid_ideal.cpp:94-98
    const double e = e_u_et(ux, uy, uz, et);
    const double p = p_rho_e(rho, e);
    const double M = std::sqrt(ux * ux + uy * uy + uz * uz) / a_rho_e(rho, e);
    return p * std::pow(1.0 + (gamma - 1.0) / 2.0 * M * M, gamma / (gamma - 1));
}
double HostFluid::total_temperature(double rho, double ux, double uy, double uz, double et) const {
    if (real) return lut_total_temperature(lut, rho, ux * ux + uy * uy + uz * uz, et);
    // fluid_ideal.cpp:100-104
    const double e = e_u_et(ux, uy, uz, et);
    const double T = T_rho_e(rho, e);
    const double M = std::sqrt(ux * ux + uy * uy + uz * uz) / a_rho_e(rho, e);
    return T * (1.0 + (gamma - 1.0) / 2.0 * M * M);
}

// ------------------------------------------------------------------ volume file

OutputFields getOutputFields(const Config& cfg) {
    const std::string v = cfg.has("OUTPUT_FIELDS_TYPE") ? cfg.get("OUTPUT_FIELDS_TYPE") : "None";
    if (v == "primary" || v == "Primary" || v == "PRIMARY") return OutputFields::Primary;
    if (v == "None" || v == "secondary" || v == "Secondary" || v == "SECONDARY") return OutputFields::Secondary;
    if (v == "turbo_bfm" || v == "Turbo_bfm" || v == "TURBO_BFM") return OutputFields::TurboBfm;
    std::cout << "OUTPUT_FIELDS_TYPE input not recognized, set to SECONDARY fields only." << std::endl;
    return OutputFields::Secondary;
}

std::string volumeFileName(const Config& cfg, long iteration) {
    if (!cfg.has("SOLUTION_NAME"))
        throw std::runtime_error("Not found value for key \"SOLUTION_NAME\" in configuration.");
    const std::string name = cfg.get("SOLUTION_NAME");
    if (cfg.getBool("SAVE_UNSTEADY", true)) {
        char buf[32];
        std::snprintf(buf, sizeof buf, "_%06ld", iteration);
        return name + buf;
    }
    return name;
}

namespace {

// ostream << double after std::setprecision(16), default float field: printf "%.16g"
inline void putG16(std::string& out, double v) {
    char buf[40];
    const int n = std::snprintf(buf, sizeof buf, "%.16g", v);
    out.append(buf, (size_t)n);
}

}  // namespace

void writeVolumeCsv(const std::string& path, const Config& cfg, const HostFluid& fl, const VolumeData& d) {
    const long N = d.ni * d.nj * d.nk;
    const OutputFields kind = getOutputFields(cfg);
    const bool bfm = cfg.isBFMActive();
    if (cfg.getBool("TURBULENCE_ACTIVE", false) || cfg.getBool("VISCOSITY_ACTIVE", false))
        throw std::runtime_error("viscous and turbulence output fields are outside the Euler path");
    const bool secondary = (kind == OutputFields::Secondary || kind == OutputFields::TurboBfm);
    const bool turbo = bfm && kind == OutputFields::TurboBfm;
    if (turbo && (!d.blockage || !d.rpm)) throw std::runtime_error("blockage and rpm grid fields are needed for turbo_bfm output");

    // one array per output field, keyed by the reference's field names
    std::vector<std::pair<std::string, std::vector<double>>> F;
    F.reserve(32);   // at most 26 fields: references returned by add stay valid
    auto add = [&](const char* name) -> std::vector<double>& {
        if (F.size() == F.capacity()) throw std::logic_error("output field table full");
        F.emplace_back(name, std::vector<double>(N));
        return F.back().second;
    };
    std::vector<double>& rho = add("Density");
    std::vector<double>& vx = add("Velocity X");
    std::vector<double>& vy = add("Velocity Y");
    std::vector<double>& vz = add("Velocity Z");
    std::vector<double>& et = add("Total Energy");
    std::vector<double>& p = add("Pressure");
    std::vector<double>& T = add("Temperature");
    for (long n = 0; n < N; ++n) {
        // FlowSolution::getVelocityX / getTotalEnergy (types.hpp:840-858)
        rho[n] = d.U[n];
        vx[n] = d.U[N + n] / d.U[n];
        vy[n] = d.U[2 * N + n] / d.U[n];
        vz[n] = d.U[3 * N + n] / d.U[n];
        et[n] = d.U[4 * N + n] / d.U[n];
        const double e = fl.e_u_et(vx[n], vy[n], vz[n], et[n]);
        p[n] = fl.p_rho_e(rho[n], e);
        T[n] = fl.T_rho_e(rho[n], e);
    }
    if (secondary) {
        std::vector<double>& mach = add("Mach");
        std::vector<double>& pt = add("Total Pressure");
        std::vector<double>& tt = add("Total Temperature");
        std::vector<double>& s = add("Entropy");
        std::vector<double> *blk = nullptr, *rmach = nullptr, *gx = nullptr, *gy = nullptr, *gz = nullptr,
                            *rx = nullptr, *ry = nullptr, *rz = nullptr, *fvx = nullptr, *fvy = nullptr,
                            *fvz = nullptr, *fix = nullptr, *fiy = nullptr, *fiz = nullptr, *dv = nullptr;
        if (turbo) {
            blk = &add("Blockage"); rmach = &add("Relative Mach");
            gx = &add("Grid Velocity X"); gy = &add("Grid Velocity Y"); gz = &add("Grid Velocity Z");
            rx = &add("Relative Velocity X"); ry = &add("Relative Velocity Y"); rz = &add("Relative Velocity Z");
            fvx = &add("Viscous Body Force X"); fvy = &add("Viscous Body Force Y"); fvz = &add("Viscous Body Force Z");
            fix = &add("Inviscid Body Force X"); fiy = &add("Inviscid Body Force Y"); fiz = &add("Inviscid Body Force Z");
            dv = &add("Deviation Angle");
        }
        const double scaling = cfg.getRotationalSpeedScalingFactor();
        for (long n = 0; n < N; ++n) {
            const double r = rho[n], ux = vx[n], uy = vy[n], uz = vz[n], e_t = et[n];
            const double e = fl.e_u_et(ux, uy, uz, e_t);
            const double a = fl.a_rho_e(r, e);
            const double umag = std::sqrt(ux * ux + uy * uy + uz * uz);
            mach[n] = umag / a;
            pt[n] = fl.total_pressure(r, ux, uy, uz, e_t);
            tt[n] = fl.total_temperature(r, ux, uy, uz, e_t);
            s[n] = fl.s_rho_e(r, e);
            if (turbo) {
                (*blk)[n] = d.blockage[n];
                double om = d.rpm[n] * 2.0 * M_PI / 60.0;
                om *= scaling;
                const double th = d.theta[n];
                // computeCartesianComponentsFromCylindrical({0, 0, omega r}, theta)
                const double gcy = 0.0, gcz = om * d.radius[n];
                const double gvx = 0.0;
                const double gvy = gcy * std::cos(th) - gcz * std::sin(th);
                const double gvz = gcy * std::sin(th) + gcz * std::cos(th);
                (*gx)[n] = gvx; (*gy)[n] = gvy; (*gz)[n] = gvz;
                const double wx = ux - gvx, wy = uy - gvy, wz = uz - gvz;
                (*rx)[n] = wx; (*ry)[n] = wy; (*rz)[n] = wz;
                (*rmach)[n] = std::sqrt(wx * wx + wy * wy + wz * wz) / a;
                (*fvx)[n] = d.fvisc ? d.fvisc[3 * n] : 0.0;
                (*fvy)[n] = d.fvisc ? d.fvisc[3 * n + 1] : 0.0;
                (*fvz)[n] = d.fvisc ? d.fvisc[3 * n + 2] : 0.0;
                (*fix)[n] = d.finv ? d.finv[3 * n] : 0.0;
                (*fiy)[n] = d.finv ? d.finv[3 * n + 1] : 0.0;
                (*fiz)[n] = d.finv ? d.finv[3 * n + 2] : 0.0;
                (*dv)[n] = d.dev ? d.dev[n] : 0.0;
            }
        }
    }
    // std::map<std::string, ...> iteration order
    std::sort(F.begin(), F.end(), [](const auto& a, const auto& b) { return a.first < b.first; });

    std::ofstream file(path, std::ios::binary);
    if (!file.is_open()) throw std::runtime_error("Failed to open file: " + path);
    std::string head = "NI=" + std::to_string(d.ni) + "\nNJ=" + std::to_string(d.nj) + "\nNK=" +
                       std::to_string(d.nk) + "\nx,y,z";
    for (auto& f : F) head += "," + f.first;
    head += "\n";
    file << head;
    // rows formatted in parallel, written in order
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const long chunk = 65536;
    for (long c0 = 0; c0 < N; c0 += chunk * hw) {
        std::vector<std::string> parts(hw);
        std::vector<std::thread> th;
        for (unsigned t = 0; t < hw; ++t) {
            const long a = c0 + (long)t * chunk, b = std::min(N, a + chunk);
            if (a >= b) break;
            th.emplace_back([&, t, a, b]() {
                std::string& s = parts[t];
                s.reserve((size_t)(b - a) * (F.size() + 3) * 24);
                for (long n = a; n < b; ++n) {
                    putG16(s, d.x[n]); s += ',';
                    putG16(s, d.y[n]); s += ',';
                    putG16(s, d.z[n]);
                    for (auto& f : F) {
                        s += ',';
                        putG16(s, f.second[n]);
                    }
                    s += '\n';
                }
            });
        }
        for (auto& x : th) x.join();
        for (auto& s : parts) file.write(s.data(), (std::streamsize)s.size());
    }
    if (!file.good()) throw std::runtime_error("Failed to write file: " + path);
}

// ------------------------------------------------------------------ initial state

namespace {

double atan2To2pi(double y, double x) {
    // atan2FromZeroTo2pi (math_utils.cpp:465-470)
    double a = std::atan2(y, x);
    if (a < 0) a += 2 * M_PI;
    return a;
}

// rotateVectorAlongXAxis (math_utils.cpp:109-115)
void rotateX(double& y, double& z, double th) {
    const double ny = y * std::cos(th) - z * std::sin(th);
    const double nz = y * std::sin(th) + z * std::cos(th);
    y = ny;
    z = nz;
}

// Nearest-neighbour search over a point cloud.  The construction (median split with
// std::nth_element on the axis of the depth, the median node first, left half
// [start, mid), right half (mid, end)) and the search (nearer child first, a strict
// improvement replaces the best, the far child visited while its slab is strictly
// closer) are those of the reference's KD tree (include/kdtree.hpp), so that points at
// equal distance resolve to the same index.
template <int D>
class NearestTree {
public:
    using Pt = std::array<double, D>;
    explicit NearestTree(const std::vector<Pt>& pts) {
        if (pts.empty()) return;
        std::vector<std::pair<Pt, size_t>> items(pts.size());
        for (size_t q = 0; q < pts.size(); ++q) items[q] = {pts[q], q};
        nodes_.reserve(pts.size());
        root_ = make(items, 0, items.size(), 0);
    }
    size_t nearest(const Pt& q) const {
        if (root_ < 0) throw std::runtime_error("nearest-neighbour search on an empty point set");
        size_t best = nodes_[root_].id;
        double bd = std::numeric_limits<double>::max();
        visit(root_, q, best, bd);
        return best;
    }

private:
    struct Node {
        Pt p;
        size_t id;
        int axis, lo, hi;
    };
    std::vector<Node> nodes_;
    int root_ = -1;

    int make(std::vector<std::pair<Pt, size_t>>& it, size_t a, size_t b, size_t depth) {
        if (a >= b) return -1;
        const int ax = (int)(depth % D);
        const size_t m = a + (b - a) / 2;
        std::nth_element(it.begin() + a, it.begin() + m, it.begin() + b,
                         [ax](const std::pair<Pt, size_t>& u, const std::pair<Pt, size_t>& v) {
                             return u.first[ax] < v.first[ax];
                         });
        const int me = (int)nodes_.size();
        nodes_.push_back({it[m].first, it[m].second, ax, -1, -1});
        const int l = make(it, a, m, depth + 1);
        const int h = make(it, m + 1, b, depth + 1);
        nodes_[me].lo = l;
        nodes_[me].hi = h;
        return me;
    }
    void visit(int ni, const Pt& q, size_t& best, double& bd) const {
        if (ni < 0) return;
        const Node& nd = nodes_[ni];
        double d2 = 0.0;
        for (int k = 0; k < D; ++k) {
            const double df = q[k] - nd.p[k];
            d2 += df * df;
        }
        if (d2 < bd) {
            bd = d2;
            best = nd.id;
        }
        const double df = q[nd.axis] - nd.p[nd.axis];
        const int nearC = (df <= 0.0) ? nd.lo : nd.hi;
        const int farC = (df <= 0.0) ? nd.hi : nd.lo;
        visit(nearC, q, best, bd);
        if (df * df < bd) visit(farC, q, best, bd);
    }
};

struct RestartFile {
    long ni = 0, nj = 0, nk = 0;
    bool coords = false;
    std::vector<double> x, y, z, rho, u, v, w, T, fi, fv;   // fi, fv: xyz triples
};

// Solver::readRestartFile (solver.cpp:1155-1295)
RestartFile readRestart(const std::string& name, const HostFluid& fl, bool bfm) {
    std::ifstream file(name);
    if (!file.is_open()) throw std::runtime_error("Failed to open restart file: " + name);
    RestartFile r;
    std::string line;
    auto dim = [&]() {
        if (!std::getline(file, line)) throw std::runtime_error("Corrupted restart file '" + name + "': missing NI/NJ/NK");
        return (long)std::stoi(line.substr(line.find('=') + 1));
    };
    r.ni = dim(); r.nj = dim(); r.nk = dim();
    if (r.ni <= 0 || r.nj <= 0 || r.nk <= 0) throw std::runtime_error("Cannot perform restart: restart data is empty.");
    const long N = r.ni * r.nj * r.nk;
    std::getline(file, line);
    std::unordered_map<std::string, size_t> col;
    {
        std::istringstream hs(line);
        std::string c;
        size_t q = 0;
        while (std::getline(hs, c, ',')) {
            const size_t a = c.find_first_not_of(" \t\r\n"), b = c.find_last_not_of(" \t\r\n");
            if (a != std::string::npos) c = c.substr(a, b - a + 1);
            col[c] = q++;
        }
    }
    constexpr size_t UNSET = std::numeric_limits<size_t>::max();
    auto find = [&](const char* a, const char* b) -> size_t {
        if (col.count(a)) return col[a];
        if (b && col.count(b)) return col[b];
        return UNSET;
    };
    const size_t ix = find("x", "X"), iy = find("y", "Y"), iz = find("z", "Z");
    r.coords = (ix != UNSET && iy != UNSET && iz != UNSET);
    auto need = [&](const char* k) {
        auto it = col.find(k);
        if (it == col.end()) throw std::runtime_error(std::string("restart file lacks the column ") + k);
        return it->second;
    };
    const size_t iD = need("Density"), iU = need("Velocity X"), iV = need("Velocity Y"), iW = need("Velocity Z"),
                 iE = need("Total Energy");
    size_t f[6] = {UNSET, UNSET, UNSET, UNSET, UNSET, UNSET};
    if (bfm) {
        const char* names[6] = {"Inviscid Body Force X", "Inviscid Body Force Y", "Inviscid Body Force Z",
                                "Viscous Body Force X", "Viscous Body Force Y", "Viscous Body Force Z"};
        for (int q = 0; q < 6; ++q) f[q] = find(names[q], nullptr);
    }
    r.x.assign(N, 0.0); r.y.assign(N, 0.0); r.z.assign(N, 0.0);
    r.rho.resize(N); r.u.resize(N); r.v.resize(N); r.w.resize(N); r.T.resize(N);
    if (bfm) { r.fi.assign(3 * N, 0.0); r.fv.assign(3 * N, 0.0); }
    long n = 0, i = 0, j = 0, k = 0;
    std::vector<double> row;
    while (std::getline(file, line)) {
        if (n >= N) throw std::runtime_error("Corrupted restart file '" + name + "': more rows than NI*NJ*NK");
        row.clear();
        std::istringstream ss(line);
        std::string val;
        while (std::getline(ss, val, ',')) {
            const double v = std::stod(val);
            if (std::isnan(v) || std::isinf(v))
                throw std::runtime_error("Corrupted restart file '" + name + "': NaN or Inf encountered at row index (" +
                                         std::to_string(i) + "," + std::to_string(j) + "," + std::to_string(k) + ").");
            row.push_back(v);
        }
        const size_t need_cols = std::max({iD, iU, iV, iW, iE}) + 1;
        if (row.size() < need_cols)
            throw std::runtime_error("Corrupted restart file '" + name + "': short row " + std::to_string(n));
        if (r.coords && ix < row.size() && iy < row.size() && iz < row.size()) {
            r.x[n] = row[ix]; r.y[n] = row[iy]; r.z[n] = row[iz];
        }
        r.rho[n] = row[iD]; r.u[n] = row[iU]; r.v[n] = row[iV]; r.w[n] = row[iW];
        const double ete = row[iE];
        // computeTemperature_rho_u_et of the file's state
        r.T[n] = fl.T_rho_e(r.rho[n], fl.e_u_et(r.u[n], r.v[n], r.w[n], ete));
        if (bfm)
            for (int q = 0; q < 3; ++q) {
                r.fi[3 * n + q] = (f[q] != UNSET && f[q] < row.size()) ? row[f[q]] : 0.0;
                r.fv[3 * n + q] = (f[3 + q] != UNSET && f[3 + q] < row.size()) ? row[f[3 + q]] : 0.0;
            }
        ++n;
        if (++k == r.nk) {
            k = 0;
            if (++j == r.nj) { j = 0; ++i; }
        }
    }
    if (n != N) throw std::runtime_error("Corrupted restart file '" + name + "': fewer rows than NI*NJ*NK");
    std::cout << "Data read successfully.\n";
    return r;
}

}  // namespace

void initialState(const Config& cfg, const std::string& baseDir, const HostFluid& fl, const GridInput& grid,
                  Topology topo, bool periodic, std::vector<double>& U, std::vector<double>& finv,
                  std::vector<double>& fvisc, bool& restartForces) {
    const long NI = grid.ni, NJ = grid.nj, NK = grid.nk, N = NI * NJ * NK;
    U.assign(5 * N, 0.0);
    finv.assign(3 * N, 0.0);
    fvisc.assign(3 * N, 0.0);
    restartForces = false;
    auto at = [&](long i, long j, long k) { return (i * NJ + j) * NK + k; };
    // conserved state of (rho, velocity, T) as every restart variant stores it
    // (pressure from rho and T, static energy from p and rho)
    auto store = [&](long n, double rho, double u, double v, double w, double T, double q) {
        U[n] = rho;
        U[N + n] = rho * u;
        U[2 * N + n] = rho * v;
        U[3 * N + n] = rho * w;
        const double p = fl.p_rho_T(rho, T);
        const double e = fl.e_p_rho(p, rho);
        U[4 * N + n] = rho * (e + 0.5 * q);
    };

    if (!cfg.getBool("RESTART_SOLUTION", false)) {
        // Solver::initializeSolutionFromScratch (solver.cpp:853-888): one uniform
        // direction, or (INIT_DIRECTION = 0, 0, 0) the grid's i-direction per node
        // (Mesh::computeAdaptiveFlowDirection, mesh.cpp:635-657)
        std::vector<double> dir = cfg.getInitDirection();
        const bool adaptive = (dir[0] == 0.0 && dir[1] == 0.0 && dir[2] == 0.0);
        const double T = cfg.getInitTemperature(), p = cfg.getInitPressure(), M = cfg.getInitMachNumber();
        auto state = [&](double dx0, double dy0, double dz0, double out[5]) {
            double rho, vx, vy, vz, et;
            const double dm = std::sqrt(dx0 * dx0 + dy0 * dy0 + dz0 * dz0);
            if (fl.real) {
                // FluidReal::computeInitFields (src/fluid_real.cpp:355-373)
                rho = lut_at(fl.lut, LUT_RHO_PT, p, T);
                const double e = lut_at(fl.lut, LUT_E_PT, p, T);
                const double a = lut_at(fl.lut, LUT_A, rho, e);
                double dx = 1.0, dy = 0.0, dz = 0.0;
                if (dm > 1e-12) { dx = dx0 / dm; dy = dy0 / dm; dz = dz0 / dm; }
                vx = dx * a * M; vy = dy * a * M; vz = dz * a * M;
                et = e + 0.5 * (vx * vx + vy * vy + vz * vz);
            } else {
                // FluidIdeal::computeInitFields (fluid_ideal.cpp:157-171)
                const double a = std::sqrt(fl.gamma * fl.R * T);
                vx = dx0 / dm * a * M; vy = dy0 / dm * a * M; vz = dz0 / dm * a * M;
                rho = p / (fl.R * T);
                const double e = p / ((fl.gamma - 1) * rho);
                const double vm = std::sqrt(vx * vx + vy * vy + vz * vz);
                et = e + 0.5 * (vm * vm);
            }
            out[0] = rho; out[1] = rho * vx; out[2] = rho * vy; out[3] = rho * vz; out[4] = rho * et;
        };
        if (!adaptive) {
            double s[5];
            state(dir[0], dir[1], dir[2], s);
            for (long n = 0; n < N; ++n)
                for (int v = 0; v < 5; ++v) U[v * N + n] = s[v];
        } else {
            if (NI < 2) throw std::runtime_error("adaptive flow direction needs NI >= 2");
            for (long i = 0; i < NI; ++i)
                for (long j = 0; j < NJ; ++j)
                    for (long k = 0; k < NK; ++k) {
                        const long a = (i < NI - 1) ? at(i, j, k) : at(NI - 2, j, k);
                        const long b = (i < NI - 1) ? at(i + 1, j, k) : at(NI - 1, j, k);
                        double dx = grid.x[b] - grid.x[a], dy = grid.y[b] - grid.y[a], dz = grid.z[b] - grid.z[a];
                        // Vector3D::normalized (types.hpp:104-107)
                        const double m = std::sqrt(dx * dx + dy * dy + dz * dz);
                        if (m == 0) { dx = dy = dz = 0.0; } else { dx = dx / m; dy = dy / m; dz = dz / m; }
                        double s[5];
                        state(dx, dy, dz, s);
                        const long n = at(i, j, k);
                        for (int v = 0; v < 5; ++v) U[v * N + n] = s[v];
                    }
        }
        return;
    }

    // ---- restart (Solver::initializeSolutionFromRestart, solver.cpp:890-957)
    std::string name = cfg.get("RESTART_SOLUTION_FILEPATH");
    if (!name.empty() && name[0] != '/' && !baseDir.empty()) name = baseDir + "/" + name;
    const bool bfm = cfg.isBFMActive();
    const RestartFile r = readRestart(name, fl, bfm);
    const std::string type = cfg.has("RESTART_TYPE") ? cfg.get("RESTART_TYPE") : "None";
    const long RN = r.ni * r.nj * r.nk;
    auto rat = [&](long i, long j, long k) { return (i * r.nj + j) * r.nk + k; };

    auto nearest = [&]() {
        // Solver::nearestNeighborRestart (solver.cpp:1036-1152)
        const bool axi = (type == "axisymmetric" || type == "nearest_neighbor_axisymmetric" ||
                          topo == Topology::Axisymmetric || (r.nk == 1 && NK > 1 && periodic));
        if (axi) {
            std::vector<std::array<double, 2>> pts(RN);
            for (long q = 0; q < RN; ++q) pts[q] = {r.x[q], std::sqrt(r.y[q] * r.y[q] + r.z[q] * r.z[q])};
            NearestTree<2> tree(pts);
            for (long n = 0; n < N; ++n) {
                const double px = grid.x[n], py = grid.y[n], pz = grid.z[n];
                const double rt = std::sqrt(py * py + pz * pz);
                const double tht = atan2To2pi(pz, py);
                const size_t b = tree.nearest({px, rt});
                const double thc = atan2To2pi(r.z[b], r.y[b]);
                const double rot = tht - thc;
                double u = r.u[b], v = r.v[b], w = r.w[b];
                rotateX(v, w, rot);
                store(n, r.rho[b], u, v, w, r.T[b], u * u + v * v + w * w);
                if (bfm) {
                    double a0 = r.fi[3 * b], a1 = r.fi[3 * b + 1], a2 = r.fi[3 * b + 2];
                    rotateX(a1, a2, rot);
                    finv[3 * n] = a0; finv[3 * n + 1] = a1; finv[3 * n + 2] = a2;
                    double c0 = r.fv[3 * b], c1 = r.fv[3 * b + 1], c2 = r.fv[3 * b + 2];
                    rotateX(c1, c2, rot);
                    fvisc[3 * n] = c0; fvisc[3 * n + 1] = c1; fvisc[3 * n + 2] = c2;
                }
            }
            std::cout << "Axisymmetric nearest-neighbor initialization done.\n";
        } else {
            std::vector<std::array<double, 3>> pts(RN);
            for (long q = 0; q < RN; ++q) pts[q] = {r.x[q], r.y[q], r.z[q]};
            NearestTree<3> tree(pts);
            for (long n = 0; n < N; ++n) {
                const size_t b = tree.nearest({grid.x[n], grid.y[n], grid.z[n]});
                const double u = r.u[b], v = r.v[b], w = r.w[b];
                store(n, r.rho[b], u, v, w, r.T[b], u * u + v * v + w * w);
                if (bfm)
                    for (int q = 0; q < 3; ++q) {
                        finv[3 * n + q] = r.fi[3 * b + q];
                        fvisc[3 * n + q] = r.fv[3 * b + q];
                    }
            }
            std::cout << "Cartesian nearest-neighbor initialization done.\n";
        }
        restartForces = bfm;
    };

    if (type == "nearest_neighbor" || type == "nearest_neighbor_axisymmetric") {
        nearest();
    } else if (r.ni != NI || r.nj != NJ || r.nk != NK) {
        if (r.ni == NI && r.nj == NJ && type == "axisymmetric") {
            // Solver::axisymmetricRestart (solver.cpp:1002-1034): the k = 0 plane of the
            // file rotated to every node's angle
            for (long i = 0; i < NI; ++i)
                for (long j = 0; j < NJ; ++j)
                    for (long k = 0; k < NK; ++k) {
                        const long n0 = at(i, j, 0), n = at(i, j, k);
                        const double th0 = atan2To2pi(grid.z[n0], grid.y[n0]);
                        const double thp = atan2To2pi(grid.z[n], grid.y[n]);
                        const long q = rat(i, j, 0);
                        double u = r.u[q], v = r.v[q], w = r.w[q];
                        rotateX(v, w, thp - th0);
                        store(n, r.rho[q], u, v, w, r.T[q], u * u + v * v + w * w);
                    }
            std::cout << "Axisymmetric initialization done.\n";
        } else if (type == "standard") {
            throw std::runtime_error("Restart file dimensions (" + std::to_string(r.ni) + "," + std::to_string(r.nj) +
                                     "," + std::to_string(r.nk) + ") do not match solver dimensions (" +
                                     std::to_string(NI) + "," + std::to_string(NJ) + "," + std::to_string(NK) +
                                     ") and RESTART_TYPE=standard was requested.");
        } else {
            std::cout << "Restart file dimensions (" << r.ni << "," << r.nj << "," << r.nk
                      << ") do not match solver dimensions (" << NI << "," << NJ << "," << NK
                      << "). Using nearest-neighbor interpolation.\n";
            nearest();
        }
    } else {
        // Solver::standardRestart (solver.cpp:960-1000)
        for (long n = 0; n < N; ++n) {
            const double u = r.u[n], v = r.v[n], w = r.w[n];
            store(n, r.rho[n], u, v, w, r.T[n], u * u + v * v + w * w);
        }
        if (bfm) {
            finv = r.fi;
            fvisc = r.fv;
            restartForces = true;
        }
        std::cout << "Standard initialization done.\n";
    }
}

}  // namespace tbfm
