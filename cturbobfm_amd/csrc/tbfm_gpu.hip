// C-ABI of the MI355X explicit residual path: case set-up on the host, device buffers
// in the i-fastest layout, and the per-iteration launch sequence (include/tbfm_gpu.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tbfm_gpu.h"
#include "tbfm_fused.hpp"
#include "tbfm_host.hpp"
#include "tbfm_io.hpp"

using namespace tbfm;
using tbfm::dev::CaseDev;

namespace {

thread_local std::string g_create_error;

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(hipError_t c, const char* what) : std::runtime_error(std::string(what) + ": " + hipGetErrorString(c)), code(c) {}
};

#define HIPCHK(x)                                  \
    do {                                           \
        hipError_t _e = (x);                       \
        if (_e != hipSuccess) throw HipError(_e, #x); \
    } while (0)

thread_local size_t* g_bytes = nullptr;   // device-byte counter of the case being set up

// Serialises the structural device work of all host threads: allocation, release and
// initialisation of device memory, work-table builds and stream capture.  Cases and
// batches may be driven from several host threads (the serving loop sets requests up
// on a worker while batches advance); replays of captured graphs and stream waits run
// outside the lock, so the GPU work itself still overlaps.
std::recursive_mutex g_dev_mu;
using DevLock = std::lock_guard<std::recursive_mutex>;

// Device allocation, zero-filled before it returns.  The fill runs on the null stream,
// which does not order itself with the non-blocking streams that later kernels use
// (a table kernel launched right after the allocation could otherwise run before the
// fill and have its results zeroed), so the null stream is drained here.
template <typename T>
T* dalloc(size_t n, std::vector<void*>& owned) {
    void* p = nullptr;
    if (n == 0) n = 1;
    HIPCHK(hipMalloc(&p, n * sizeof(T)));
    HIPCHK(hipMemsetAsync(p, 0, n * sizeof(T), nullptr));
    HIPCHK(hipStreamSynchronize(nullptr));
    owned.push_back(p);
    if (g_bytes) *g_bytes += n * sizeof(T);
    return (T*)p;
}

template <typename T>
T* upload(const std::vector<T>& h, std::vector<void*>& owned) {
    T* d = dalloc<T>(h.size(), owned);
    if (!h.empty()) HIPCHK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

}  // namespace

namespace {

// Device work table of one kernel of a launch plan.
struct Table {
    dev::Work* d = nullptr;
    size_t cap = 0;
    int n = 0;
};

// Stage-kernel instantiation shared by the cases of one group.
struct StageGroup {
    int tile = 0, sch = 0;
    bool three = false, mixed = false, real = false;
    Table t;
};

// Iterations captured in the plan's long graph: an advance of n iterations replays it
// n / kGraphIters times and the one-iteration graph for the rest.
constexpr int kGraphIters = 8;

// Launch plan of a set of cases that advance together (one case, or the members of a
// batch): the device case table, the work tables of every kernel and the captured
// graphs of one and of kGraphIters iterations.
struct Plan {
    hipStream_t stream = nullptr;
    std::vector<tbfm_case*> members;
    CaseDev* dcases = nullptr;        // device case table
    size_t dcap = 0;
    bool ownsCases = false;
    Table dt, mn, prof, bnd, per, norm, start, pre, turbo, visc, wpre;
    std::vector<StageGroup> groups;
    int nstages = 0;
    int fluids = 0;                   // bit 0: an ideal-gas member, bit 1: a table-fluid member
    hipGraphExec_t graph = nullptr;   // one iteration
    hipGraphExec_t graphN = nullptr;  // kGraphIters iterations
    bool graphValid = false;
    bool tablesValid = false;
};

}  // namespace

struct tbfm_case {
    long ni = 0, nj = 0, nk = 0, N = 0;
    Config cfg;
    Topology topo = Topology::TwoD;
    Scheme scheme = Scheme::JST;
    Precision precision = Precision::FP64;
    FluidModel fluid = FluidModel::Ideal;
    HostFluid hf;                     // the fluid on the host (tables of FLUID_MODEL = real)
    LutSet lut{};                     // host copy of the fluid tables (points into hf.values)
    GridInput grid;                   // grid points and the blockage / rpm fields, for output files
    double profInit[dev::MAX_PROF] = {0.0, 0.0, 0.0, 0.0};   // > 0: fixed initial outlet profile value
    std::vector<double> rk;
    Geometry geo;
    std::vector<Patch> patches;
    std::vector<int> nswHost;         // no-slip wall nodes (device layout), empty when none
    std::vector<double> wallvHost;    // [3][N] their wall velocities
    std::vector<void*> owned;
    size_t dev_bytes = 0;             // device memory held by the case
    CaseDev c{};
    CaseDev* cdev = nullptr;          // device copy: the case table of the case's own plan
    long hist_cap = 0;
    long iter_host = 1;               // next iteration number
    int tile = 0;                     // 0: 32x8, 1: 16x16, 2: 64x4, 3: 28x9 (i x j columns)
    int forced_tile = -1;             // GPU_STAGE_TILING: tile shape and k pieces, no autotuning
    int forced_pieces = 0;
    bool tuned = false;
    std::string tuning;               // autotuner report
    Plan solo;                        // the case advancing alone
    tbfm_batch* batch = nullptr;      // batch the case is admitted to
    // profiling of the stage-kernel launches (eager launches only)
    bool prof = false;
    // event triples per stage: before the pre-stage kernel, between it and the stage
    // kernel, after the stage kernel
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;             // triples recorded since the last read-out
    double prof_ms = 0.0;             // stage kernel alone
    double prof_pre_ms = 0.0;         // pre-stage (and viscous) kernels before it
    long prof_launches = 0;
    double bytes_per_node = 0.0;
    long body_force_nodes = 0;        // nodes whose stage also runs the blade sources
    std::string err;
};

struct tbfm_batch {
    Plan plan;
    std::string err;
};

namespace {
void hostTurbo(tbfm_case* c, const std::vector<double>& U, double* out);
}

namespace {

void drainProfile(tbfm_case* c);

long nodeDev(const tbfm_case* c, long i, long j, long k) { return (k * c->nj + j) * c->ni + i; }

// reference order (k fastest) <-> device order (i fastest)
void toDevice(const tbfm_case* c, const double* src, std::vector<double>& dst) {
    dst.resize(c->N);
    for (long i = 0; i < c->ni; ++i)
        for (long j = 0; j < c->nj; ++j)
            for (long k = 0; k < c->nk; ++k) dst[nodeDev(c, i, j, k)] = src[(i * c->nj + j) * c->nk + k];
}
void toHost(const tbfm_case* c, const double* src, double* dst) {
    for (long i = 0; i < c->ni; ++i)
        for (long j = 0; j < c->nj; ++j)
            for (long k = 0; k < c->nk; ++k) dst[(i * c->nj + j) * c->nk + k] = src[nodeDev(c, i, j, k)];
}

// face arrays (reference xyz triples) -> three SoA device arrays, i fastest
void faceSoA(const std::vector<double>& src, long ai, long aj, long ak, std::vector<double>& x,
             std::vector<double>& y, std::vector<double>& z) {
    const long n = ai * aj * ak;
    x.resize(n); y.resize(n); z.resize(n);
    for (long i = 0; i < ai; ++i)
        for (long j = 0; j < aj; ++j)
            for (long k = 0; k < ak; ++k) {
                const long s = ((i * aj + j) * ak + k) * 3;
                const long d = (k * aj + j) * ai + i;
                x[d] = src[s]; y[d] = src[s + 1]; z[d] = src[s + 2];
            }
}

// Stage-kernel tile shapes (i x j columns): 0: 32x8, 1: 16x16, 2: 64x4, 3: 28x9 (252
// columns: fewer idle columns on grids such as 193 x 49)
constexpr int kTiles = 4;
int tileI(int t) { static const int v[kTiles] = {32, 16, 64, 28}; return v[t]; }
int tileJ(int t) { static const int v[kTiles] = {8, 16, 4, 9}; return v[t]; }

// Tile shape and the number of k pieces: block z of a column of tiles marches planes
// [z*NK/pieces, (z+1)*NK/pieces).
void setTiling(tbfm_case* cs, int tile, int pieces) {
    const int ti = tileI(tile), tj = tileJ(tile);
    CaseDev& c = cs->c;
    cs->tile = tile;
    c.sgz = std::max(1, std::min(pieces, (int)cs->nk));
    c.kc = (int)((cs->nk + c.sgz - 1) / c.sgz);
    c.sgx = (int)((cs->ni + ti - 1) / ti);
    c.sgy = (int)((cs->nj + tj - 1) / tj);
    c.nparts_stage = c.sgx * c.sgy * c.sgz;
}

double atan2Positive(double y, double x) {
    double a = std::atan2(y, x);
    if (a < 0) a += 2 * M_PI;
    return a;
}

void setup(tbfm_case* cs, const GridInput& grid) {
    Config& cfg = cs->cfg;
    cs->ni = grid.ni; cs->nj = grid.nj; cs->nk = grid.nk;
    cs->N = grid.ni * grid.nj * grid.nk;
    const long NI = cs->ni, NJ = cs->nj, NK = cs->nk, N = cs->N;
    cs->topo = cfg.getTopology();
    cs->scheme = cfg.getAdvectionScheme();
    cs->precision = cfg.getGpuPrecision();
    cs->rk = cfg.getTimeIntegrationCoeffs();
    // kernels index the state (5N values) and the face-normal array (12 values per node)
    // with 32-bit offsets
    if (5 * N >= 2147483647L || 4 * ((NI + 1) * NJ * NK + NI * (NJ + 1) * NK + NI * NJ * (NK + 1)) >= 2147483647L)
        throw std::runtime_error("grid too large for one case (32-bit offsets: about 170 million nodes at most)");
    if (cs->topo == Topology::ThreeD && NK < 4) throw std::runtime_error("3D topology needs NK >= 4");
    if (cs->topo != Topology::ThreeD && NK != 1) throw std::runtime_error("2D/axisymmetric topology needs NK = 1");
    if (NI < 3 || (cs->topo != Topology::OneD && NJ < 3))
        throw std::runtime_error("the residual stencil needs at least 3 points per active direction");
    const bool bfm = cfg.isBFMActive();
    cs->geo = buildGeometry(grid, cs->topo, bfm);
    const Geometry& g = cs->geo;

    HIPCHK(hipStreamCreateWithFlags(&cs->solo.stream, hipStreamNonBlocking));
    g_bytes = &cs->dev_bytes;
    auto& own = cs->owned;
    CaseDev& c = cs->c;
    c.ni = (int)NI; c.nj = (int)NJ; c.nk = (int)NK; c.N = N;
    c.topo = (int)cs->topo;
    c.scheme = (int)cs->scheme;
    c.muscl = cfg.getMUSCLreconstruction() ? 1 : 0;
    c.limiter = (int)cfg.getFluxLimiter();
    cs->fluid = cfg.getFluidModel();
    c.real = 0;
    c.lut = nullptr;
    if (cs->fluid == FluidModel::Real && cs->precision == Precision::Mixed)
        throw std::runtime_error("GPU_PRECISION = mixed needs FLUID_MODEL = ideal: the look-up-table fluid runs in FP64");
    // relative table path from the input.ini folder when read from one, else from the
    // working folder
    cs->hf.load(cfg, cfg.baseDir);
    if (cs->hf.real) {
        cs->lut = cs->hf.lut;
        double* dv = upload(cs->hf.values, cs->owned);
        LutSet dset = cs->lut;
        for (int t = 0; t < LUT_COUNT; ++t) dset.t[t].d = dv + (cs->lut.t[t].d - cs->hf.values.data());
        c.lut = upload(std::vector<LutSet>{dset}, cs->owned);
        c.real = 1;
    }
    c.gamma = cs->hf.gamma;
    c.rgas = cs->hf.R;
    c.gm1 = c.gamma - 1;
    c.cp = c.gamma * c.rgas / (c.gamma - 1.0);
    c.cfl = cfg.getCFL();
    c.dt_mode = (int)cfg.getTimeStepMethod();
    c.fixed_dt = cfg.getFixedTimeStep();
    c.ramp_iter = cfg.getOutletPressureRampMaxIterations();
    c.init_p = cfg.has("INIT_PRESSURE") ? cfg.getInitPressure() : 101325.0;
    c.bfm_active = bfm ? 1 : 0;
    c.bfm_model = (int)cfg.getBFMModel();
    c.blockage_active = cfg.isBlockageActive() ? 1 : 0;
    c.dtheta = (NK > 1) ? (2 * M_PI / (double)(NK - 1)) : 0.0;

    // ---- patches, periodicity, boundary-face maps (solver.cpp:296-367, 491-588)
    double perAngleDeg = 0.0, perTranslation = 0.0;
    bool periodic = false;
    for (auto& p : cs->patches) {
        if (p.type == PatchType::Periodic) {
            if (p.values.size() < 3) throw std::runtime_error("periodic patch " + p.name + " needs 3 values");
            if (p.values[0] == 0) {
                periodic = true;
                perTranslation = p.values[1];
                perAngleDeg = p.values[2];
                if (perAngleDeg == 360.0) perAngleDeg = 0.0;
            }
            if (p.k0 != p.k1) throw std::runtime_error("Periodic boundary only supported on i-j patches");
            if (p.i0 != 0 || p.j0 != 0 || p.i1 != NI || p.j1 != NJ)
                throw std::runtime_error("Periodic boundary only supported on full i-j patches");
        }
    }
    const double perAngleRad = perAngleDeg * M_PI / 180.0;
    if (periodic) {
        // Mesh::checkPeriodicity (src/mesh.cpp:584-619)
        for (long i = 0; i < NI; ++i)
            for (long j = 0; j < NJ; ++j) {
                long a = (i * NJ + j) * NK, b = (i * NJ + j) * NK + NK - 1;
                double ta = atan2Positive(grid.z[a], grid.y[a]);
                double tb = atan2Positive(grid.z[b] - perTranslation, grid.y[b]);
                double d = std::fmod(tb - ta, 2.0 * M_PI);
                if (std::abs(d - perAngleRad) > 1E-4)
                    throw std::invalid_argument("Periodicity check failed at (i,j)=(" + std::to_string(i) + "," +
                                                std::to_string(j) + ")");
            }
    }
    c.periodic = periodic ? 1 : 0;
    c.per_c = std::cos(perAngleRad); c.per_s = std::sin(perAngleRad);
    c.per_cm = std::cos(-perAngleRad); c.per_sm = std::sin(-perAngleRad);
    if ((int)cs->patches.size() > dev::MAX_PATCH) throw std::runtime_error("too many boundary patches");

    std::vector<int> pI(2 * NJ * NK, -1), pJ(2 * NI * NK, -1), pK(2 * NI * NJ, -1);
    std::vector<int> profOfPatch(cs->patches.size(), -1);
    int nprof = 0;
    for (size_t q = 0; q < cs->patches.size(); ++q) {
        const Patch& p = cs->patches[q];
        dev::PatchDev pd{};
        switch (p.type) {
            case PatchType::Inlet:
                if (p.values.size() < 5) throw std::runtime_error("inlet patch " + p.name + " needs p0, T0, direction");
                pd.type = dev::PT_INLET; pd.v0 = p.values[0]; pd.v1 = p.values[1];
                break;
            case PatchType::Outlet:
                if (p.values.empty()) throw std::runtime_error("outlet patch " + p.name + " needs a pressure");
                pd.type = dev::PT_OUTLET; pd.v0 = p.values[0];
                break;
            case PatchType::RadialEquilibrium:
            case PatchType::Throttle:
            case PatchType::MassFlow: {
                // outlet pressure profiles (Solver::buildBoundaryDataStructures,
                // solver.cpp:296-340): radial equilibrium unless a mass-flow outlet opts out
                bool radial = true;
                if (p.type == PatchType::MassFlow) {
                    const bool can = (p.i0 == p.i1 && p.i0 != 0 && p.j1 > p.j0);
                    if (p.values.size() >= 4) radial = (p.values[3] > 0.5) && can;
                    else radial = can && (cs->topo == Topology::Axisymmetric || bfm);
                    if (p.values.empty() || p.values[0] <= 0.0)
                        throw std::runtime_error("Target mass flow for boundary patch " + p.name +
                                                 " must be specified and positive.");
                }
                if (p.type == PatchType::Throttle && p.values.empty())
                    throw std::runtime_error("throttle patch " + p.name + " needs a throttle coefficient");
                if (radial) {
                    if (p.i0 != p.i1) throw std::runtime_error("Radial equilibrium only supported on j-k patches");
                    if (p.i0 == 0) throw std::runtime_error("Radial equilibrium only supported to the last i-position");
                    if (p.j1 <= p.j0) throw std::runtime_error("Invalid boundary indices for radial equilibrium patch: " + p.name);
                }
                if (nprof >= dev::MAX_PROF) throw std::runtime_error("too many outlet pressure profiles");
                pd.type = dev::PT_RADIAL; pd.profile = nprof;
                c.prof_kind[nprof] = (p.type == PatchType::Throttle) ? 1 : (p.type == PatchType::MassFlow ? 2 : 0);
                c.prof_radial[nprof] = radial ? 1 : 0;
                profOfPatch[q] = nprof++;
                break;
            }
            case PatchType::Wall:
            case PatchType::NoSlipWall: pd.type = dev::PT_WALL; break;   // solver.cpp:391-399
            case PatchType::Transparent: pd.type = dev::PT_TRANSPARENT; break;
            default: pd.type = dev::PT_FAKE; break;
        }
        c.patch[q] = pd;
        if (p.i0 == p.i1) {
            long t = (p.i0 == 0) ? 0 : 1;
            for (long j = p.j0; j < p.j1; ++j)
                for (long k = p.k0; k < p.k1; ++k) pI[t * NJ * NK + k * NJ + j] = (int)q;
        } else if (p.j0 == p.j1) {
            long t = (p.j0 == 0) ? 0 : 1;
            for (long i = p.i0; i < p.i1; ++i)
                for (long k = p.k0; k < p.k1; ++k) pJ[t * NI * NK + k * NI + i] = (int)q;
        } else if (p.k0 == p.k1) {
            long t = (p.k0 == 0) ? 0 : 1;
            for (long i = p.i0; i < p.i1; ++i)
                for (long j = p.j0; j < p.j1; ++j) pK[t * NI * NJ + j * NI + i] = (int)q;
        } else {
            throw std::runtime_error("Invalid boundary definition: " + p.name);
        }
    }
    auto missing = [&](const std::vector<int>& m, const char* d) {
        for (size_t q = 0; q < m.size(); ++q)
            if (m[q] < 0) throw std::runtime_error(std::string("Missing boundary condition on a ") + d + " boundary face");
    };
    missing(pI, "I");
    if (cs->topo != Topology::OneD) missing(pJ, "J");
    if (cs->topo == Topology::ThreeD) missing(pK, "K");
    // inlet directions per boundary face (boundary_inlet.cpp:14-35), static geometry
    const bool cyl = cfg.isInletCylindrical();
    auto inletDirs = [&](int dir, const std::vector<int>& pm, long n1, long n2) {
        std::vector<double> d(2 * 3 * n1 * n2, 0.0);
        const long nsz = n1 * n2;
        for (int side = 0; side < 2; ++side)
            for (long q = 0; q < nsz; ++q) {
                int pidx = pm[side * nsz + q];
                if (pidx < 0 || cs->patches[pidx].type != PatchType::Inlet) continue;
                const auto& vals = cs->patches[pidx].values;
                long a = q % n1, b = q / n1;   // (fast, slow) index of the face
                long fi, fj, fk;
                double S[3];
                if (dir == 0) {
                    fi = side ? NI : 0; fj = a; fk = b;
                    const double* s = &g.surfI[((fi * NJ + fj) * NK + fk) * 3];
                    S[0] = s[0]; S[1] = s[1]; S[2] = s[2];
                } else if (dir == 1) {
                    fi = a; fj = side ? NJ : 0; fk = b;
                    const double* s = &g.surfJ[((fi * (NJ + 1) + fj) * NK + fk) * 3];
                    S[0] = s[0]; S[1] = s[1]; S[2] = s[2];
                } else {
                    fi = a; fj = b; fk = side ? NK : 0;
                    const double* s = &g.surfK[((fi * NJ + fj) * (NK + 1) + fk) * 3];
                    S[0] = s[0]; S[1] = s[1]; S[2] = s[2];
                }
                if (!side) { S[0] = -S[0]; S[1] = -S[1]; S[2] = -S[2]; }
                double dx = vals[2], dy = vals[3], dz = vals[4];
                if (cyl) {
                    long vi = std::min(fi, NI - 1), vj = std::min(fj, NJ - 1), vk = std::min(fk, NK - 1);
                    double th = g.theta[(vi * NJ + vj) * NK + vk];
                    double cth = std::cos(th), sth = std::sin(th);
                    double ny = dy * cth - dz * sth, nz = dy * sth + dz * cth;
                    dy = ny; dz = nz;
                }
                if (dx == 1.0 && dy == 1.0 && dz == 1.0) { dx = -S[0]; dy = -S[1]; dz = -S[2]; }
                double m = std::sqrt(dx * dx + dy * dy + dz * dz);
                dx /= m; dy /= m; dz /= m;
                d[(side * 3 + 0) * nsz + q] = dx;
                d[(side * 3 + 1) * nsz + q] = dy;
                d[(side * 3 + 2) * nsz + q] = dz;
            }
        return d;
    };
    c.pI = upload(pI, own); c.pJ = upload(pJ, own); c.pK = upload(pK, own);
    c.dirI = upload(inletDirs(0, pI, NJ, NK), own);
    c.dirJ = upload(inletDirs(1, pJ, NI, NK), own);
    c.dirK = upload(inletDirs(2, pK, NI, NJ), own);
    // boundary-face flux records (6 doubles each): allocated behind the face records below
    c.boffJ = (int)(12 * NJ * NK);
    c.boffK = (int)(12 * NJ * NK + 12 * NI * NK);
    const long nbrec = 2 * (NJ * NK + NI * NK + NI * NJ);

    // ---- radial equilibrium profiles (solver.cpp:296-331)
    c.nprof = nprof;
    std::vector<double> prof_r(dev::MAX_PROF * NJ, 0.0), prof_cs(dev::MAX_PROF * NJ, 1.0), prof_sn(dev::MAX_PROF * NJ, 0.0);
    for (size_t q = 0; q < cs->patches.size(); ++q) {
        int pr = profOfPatch[q];
        if (pr < 0) continue;
        const Patch& p = cs->patches[q];
        c.prof_i[pr] = (int)(p.i0 > 0 ? p.i0 - 1 : 0);
        c.prof_j0[pr] = (int)p.j0;
        c.prof_n[pr] = (int)((p.j1 > p.j0) ? p.j1 - p.j0 : 1);
        if (c.prof_kind[pr] == 1) {
            // hub = p0 of the first inlet + k_t m^2 (solver.cpp:2722-2732)
            double pt_in = 0.0;
            for (auto& b : cs->patches)
                if (b.type == PatchType::Inlet) { pt_in = b.values[0]; break; }
            c.prof_par[pr][0] = p.values[0];
            c.prof_par[pr][1] = pt_in;
            c.prof_need_mf = 1;
        } else if (c.prof_kind[pr] == 2) {
            // solver.cpp:2734-2770
            c.prof_par[pr][0] = p.values[0];
            c.prof_par[pr][1] = (p.values.size() >= 3 && p.values[2] > 0.0) ? p.values[2] : cfg.getFloat("MASS_FLOW_RELAXATION", 0.05);
            c.prof_par[pr][2] = cfg.getFloat("MASS_FLOW_MAX_PRESSURE_CHANGE", 0.02);
            const double ip = cfg.has("INIT_PRESSURE") ? cfg.getInitPressure() : 0.0;
            c.prof_par[pr][3] = ip > 0.0 ? ip : 101325.0;
            if (p.i0 != p.i1) throw std::runtime_error("mass-flow outlet " + p.name + " must lie on an i-boundary");
            c.prof_mfb[pr] = (p.i0 == 0) ? 0 : 1;   // turbo boundaries I_START / I_END, remapped below
            c.prof_need_mf = 1;
            // initial profile (solver.cpp:334-337, 818-821): the patch's pressure when given;
            // an empty j range keeps the set-up value
            if (p.values.size() >= 2 && p.values[1] > 0.0) cs->profInit[pr] = p.values[1];
            else if (p.j1 <= p.j0) cs->profInit[pr] = c.prof_par[pr][3];
        }
        if (!c.prof_radial[pr]) continue;
        for (long j = p.j0; j < p.j1; ++j) {
            prof_r[pr * NJ + (j - p.j0)] = g.radius[((p.i1 - 1) * NJ + j) * NK + 0];
            double th = g.theta[(c.prof_i[pr] * NJ + j) * NK + 0];
            prof_cs[pr * NJ + (j - p.j0)] = std::cos(th);
            prof_sn[pr * NJ + (j - p.j0)] = std::sin(th);
        }
    }
    // Greitzer plenum (solver.cpp:730-747, config.cpp:684-693, config.hpp:347-353)
    c.greitzer = cfg.getBool("ENABLE_GREITZER_MODELING", false) ? 1 : 0;
    if (c.greitzer) {
        if (c.dt_mode != 2) throw std::logic_error("Greitzer modeling is only compatible with fixed time step method");
        bool thr = false;
        for (auto& p : cs->patches)
            if (p.type == PatchType::Throttle) { thr = true; c.gz_kt = p.values[0]; }
        if (!thr) throw std::runtime_error("Greitzer modeling enabled but no throttle boundary condition found!");
        c.gz_vol = cfg.has("GREITZER_PLENUM_VOLUME") ? cfg.getFloat("GREITZER_PLENUM_VOLUME")
                                                      : cfg.getFloat("GREITZER_PENUM_VOLUME");
        c.gz_amb = cfg.getFloat("GREITZER_AMBIENT_PRESSURE", 101325.0);
        c.gz_dt = cfg.getFixedTimeStep();
        c.gz = dalloc<double>(3, own);
        c.prof_need_mf = 1;
    }
    c.prof_p = dalloc<double>(dev::MAX_PROF * NJ, own);
    c.prof_r = upload(prof_r, own);
    c.prof_cs = upload(prof_cs, own);
    c.prof_sn = upload(prof_sn, own);

    // ---- geometry in device layout
    std::vector<double> x, y, z;
    faceSoA(g.surfI, NI + 1, NJ, NK, x, y, z);
    c.six = upload(x, own); c.siy = upload(y, own); c.siz = upload(z, own);
    faceSoA(g.surfJ, NI, NJ + 1, NK, x, y, z);
    c.sjx = upload(x, own); c.sjy = upload(y, own); c.sjz = upload(z, own);
    faceSoA(g.surfK, NI, NJ, NK + 1, x, y, z);
    c.skx = upload(x, own); c.sky = upload(y, own); c.skz = upload(z, own);
    // unit normals and areas, one 32-byte record (nx ny nz |S|) per face: I faces, then
    // J faces from record noffJ, then K faces from record noffK
    c.nfI = (int)((NI + 1) * NJ * NK);
    c.nfJ = (int)(NI * (NJ + 1) * NK);
    c.nfK = (int)(NI * NJ * (NK + 1));
    c.noffJ = c.nfI;
    c.noffK = c.nfI + c.nfJ;
    {
        std::vector<double> nrm(4L * (c.nfI + c.nfJ + c.nfK));
        auto normals = [&](const std::vector<double>& S, long ai, long aj, long ak, long off) {
            std::vector<double> bx, by, bz;
            faceSoA(S, ai, aj, ak, bx, by, bz);
            const long nf = (long)bx.size();
            for (long q = 0; q < nf; ++q) {
                // computeAdvectionFluxFromPrimitive: normal = surface / surface.magnitude()
                const double m = std::sqrt(bx[q] * bx[q] + by[q] * by[q] + bz[q] * bz[q]);
                double* r = &nrm[4 * (off + q)];
                r[0] = bx[q] / m;
                r[1] = by[q] / m;
                r[2] = bz[q] / m;
                r[3] = m;
            }
        };
        normals(g.surfI, NI + 1, NJ, NK, 0);
        normals(g.surfJ, NI, NJ + 1, NK, c.noffJ);
        normals(g.surfK, NI, NJ, NK + 1, c.noffK);
        c.fp32 = (cs->precision == Precision::Mixed) ? 1 : 0;
        // The boundary-face flux records (48 bytes: 5 doubles and a pad) follow the face
        // records in the same allocation, so that the stage kernel reads a face's record
        // or its boundary flux from one base address.
        const size_t nrec = nrm.size();
        if (cs->precision == Precision::Mixed) {
            // the FP32-flux kernels read 16-byte FP32 records (the values they compute
            // with: the FP64 values rounded to float)
            std::vector<float> n32(nrec + 12 * (size_t)nbrec, 0.0f);
            for (size_t q = 0; q < nrec; ++q) n32[q] = (float)nrm[q];
            c.nrm32 = upload(n32, own);
            c.fbase = (const char*)c.nrm32;
            c.bnd_off = (unsigned)(nrec * sizeof(float));
        } else {
            nrm.resize(nrec + 6 * (size_t)nbrec, 0.0);
            c.nrm = upload(nrm, own);
            c.fbase = (const char*)c.nrm;
            c.bnd_off = (unsigned)(nrec * sizeof(double));
        }
        if ((size_t)c.bnd_off + 48 * (size_t)nbrec >= (size_t)1 << 32)
            throw std::runtime_error("grid too large: face records exceed 4 GiB");
        c.bnd = (double*)(c.fbase + c.bnd_off);
    }
    faceSoA(g.dual, NI + 1, NJ + 1, NK + 1, x, y, z);
    c.dux = upload(x, own); c.duy = upload(y, own); c.duz = upload(z, own);
    std::vector<double> tmp;
    toDevice(cs, g.volume.data(), tmp);
    c.vol = upload(tmp, own);

    // ---- body-force slab and source data (solver.cpp:2790-2897, source_bfm_base.cpp)
    auto field = [&](const char* name, bool required) -> const std::vector<double>* {
        const std::vector<double>* f = grid.field(name);
        if (!f && required) throw std::runtime_error(std::string(name) + " not found in the grid file");
        return f;
    };
    std::vector<unsigned char> blade2d(NI * NJ, 0);
    long bi0 = NI, bi1 = 0;
    const bool axi = (cs->topo == Topology::Axisymmetric);
    const bool hall = bfm && (c.bfm_model == 1 || c.bfm_model == 2 || c.bfm_model == 4);
    const bool thollet = bfm && c.bfm_model == 4;
    const std::vector<double>* present = bfm ? field("bladePresent", cs->topo == Topology::ThreeD) : nullptr;
    const std::vector<double>* nblades = bfm ? field("numberBlades", true) : nullptr;
    const std::vector<double>* blockage = bfm ? field("blockage", true) : nullptr;
    if (bfm && cs->topo == Topology::ThreeD)
        for (long i = 0; i < NI; ++i)
            for (long j = 0; j < NJ; ++j) blade2d[j * NI + i] = ((*present)[(i * NJ + j) * NK] > 0.0) ? 1 : 0;
    auto gateOf = [&](long n) {
        const double* gb = &g.gradBlockage[3 * n];
        return std::sqrt(gb[0] * gb[0] + gb[1] * gb[1] + gb[2] * gb[2]) > 1E-10;
    };
    if (axi) { bi0 = 0; bi1 = NI; }
    if (bfm)
        for (long i = 0; i < NI; ++i)
            for (long j = 0; j < NJ; ++j)
                for (long k = 0; k < NK; ++k) {
                    long n = (i * NJ + j) * NK + k;
                    bool pres = present && (*present)[n] > 0.0 && cs->topo == Topology::ThreeD;
                    if (pres || gateOf(n)) { bi0 = std::min(bi0, i); bi1 = std::max(bi1, i + 1); }
                }
    if (bi1 <= bi0) { bi0 = 0; bi1 = 0; }
    c.bi0 = (int)bi0; c.bi1 = (int)bi1;
    // the stage kernel takes the source terms of the slab from srcbuf (geometric sources
    // exist for axisymmetric cases and 3-D blade rows, body forces with BFM_ACTIVE)
    const bool srcOn = (cs->topo == Topology::Axisymmetric) || bfm;
    c.ns = srcOn ? (int)((bi1 - bi0) * NJ * NK) : 0;
    c.srcbuf = c.ns > 0 ? dalloc<double>(12L * c.ns, own) : nullptr;
    const long SW = bi1 - bi0, NS = std::max<long>(SW * NJ * NK, 1);
    std::vector<unsigned char> sflag(NS, 0);
    std::vector<double> s_cs(NS, 1.0), s_sn(NS, 0.0), s_r(NS, 0.0), s_om(NS, 0.0), s_b(NS, 1.0), s_gbx(NS, 0.0),
        s_gby(NS, 0.0), s_gbz(NS, 0.0), s_nax(NS, 0.0), s_nrad(NS, 0.0), s_ntan(NS, 1.0), s_pitch(NS, 1.0),
        s_spitch(NS, 1.0), s_sol(NS, 0.0), s_beta0(NS, 0.0), s_kp(NS, 0.0), s_stwl(NS, 1.0), s_piv(NS, 0.0),
        s_stall(NS, 0.0);
    const std::vector<double>* rpm = (bfm && (hall || cs->topo == Topology::ThreeD)) ? field("rpm", hall) : nullptr;
    const std::vector<double>* nax = hall ? field("normalAxial", true) : nullptr;
    const std::vector<double>* nrad = hall ? field("normalRadial", true) : nullptr;
    const std::vector<double>* ntan = hall ? field("normalTangential", true) : nullptr;
    const std::vector<double>* metal = hall ? field("bladeMetalAngle", false) : nullptr;
    const std::vector<double>* slen = (c.bfm_model == 2 || thollet) ? field("streamwiseLength", true) : nullptr;
    // Hall-Thollet constants (source_bfm_thollet.hpp:48-51); the reference reads each
    // key without a default, so a missing one is an error
    c.stall = (hall && cfg.isStalledBfmActive()) ? 1 : 0;
    if (thollet) {
        c.th_kn = cfg.getHallTholletCoefficient_KN();
        c.th_kf = cfg.getHallTholletCoefficient_KF();
        c.th_kd = cfg.getHallTholletCoefficient_KD();
        c.th_off_design = cfg.getHallTholletOffDesignActive() ? 1 : 0;
        // source_bfm_thollet.cpp:85-95: without FLUID_KINEMATIC_VISCOSITY the reference
        // takes mu(rho, e)/rho from the fluid's viscosity model
        if (!cfg.has("FLUID_KINEMATIC_VISCOSITY"))
            throw std::runtime_error("BFM_MODEL = Hall-Thollet needs FLUID_KINEMATIC_VISCOSITY on the GPU path "
                                     "(the molecular-viscosity model is outside it)");
        c.th_nu = cfg.getFloat("FLUID_KINEMATIC_VISCOSITY");
    }
    const std::vector<double>* pivot = (thollet && c.th_off_design) ? field("deviationAnglePivot", true) : nullptr;
    const std::vector<double>* stallf = c.stall ? field("deviationAngleStall", true) : nullptr;
    const std::vector<double>* beta0 = (c.bfm_model == 2) ? field("liftDragBeta0", true) : nullptr;
    const std::vector<double>* kp = (c.bfm_model == 2) ? field("liftDragKpEtamax", true) : nullptr;
    const long le = cfg.getLeadingEdgeIndex(), te = cfg.getTrailingEdgeIndex();
    const double scaling = cfg.getRotationalSpeedScalingFactor();
    // time-dependent rotational speed (config.hpp:355-367, config.cpp:695-718): the slab
    // keeps rpm 2 pi / 60 and the device scales it by the factor of the iteration's time
    c.accel = cfg.getBool("ACCELERATION_ACTIVE", false) ? 1 : 0;
    c.om_scaling = scaling;
    if (c.accel) {
        if (c.dt_mode != 2)
            throw std::runtime_error("Rotational speed scaling factor is only active for fixed time step method");
        c.acc_t0 = cfg.getFloat("ACCELERATION_INITIAL_TIME");
        c.acc_t1 = cfg.getFloat("ACCELERATION_FINAL_TIME");
        c.acc_s0 = cfg.getFloat("ROTATIONAL_SPEED_SCALING_FACTOR_INITIAL");
        c.acc_s1 = cfg.getFloat("ROTATIONAL_SPEED_SCALING_FACTOR_FINAL");
    }
    // Gaussian body-force perturbation (source_bfm_base.cpp:8-19, 215-243)
    c.pert = (bfm && cfg.getBool("PERTURBATION_BODY_FORCE", false)) ? 1 : 0;
    double pcen[3] = {0.0, 0.0, 0.0}, prad = 1.0;
    if (c.pert) {
        auto cen = cfg.getList("PERTURBATION_CENTER");
        if (cen.size() < 3) throw std::runtime_error("PERTURBATION_CENTER needs three coordinates");
        for (int d = 0; d < 3; ++d) pcen[d] = std::stod(cen[d]);
        prad = cfg.getFloat("PERTURBATION_RADIAL_EXTENSION");
        c.pert_scale = cfg.getFloat("PERTURBATION_SCALING_FACTOR");
        c.pert_start = cfg.getFloat("PERTURBATION_TIME_START");
        c.pert_dur = cfg.getFloat("PERTURBATION_TIME_DURATION");
    }
    std::vector<double> s_psp(c.pert ? NS : 1, -1.0);
    for (long i = bi0; i < bi1; ++i)
        for (long j = 0; j < NJ; ++j)
            for (long k = 0; k < NK; ++k) {
                const long n = (i * NJ + j) * NK + k;
                const long s = (k * NJ + j) * SW + (i - bi0);
                unsigned char fl = 0;
                if (present && (*present)[n] > 0.0) fl |= 1;
                if (bfm && gateOf(n)) fl |= 2;
                if (nblades && (*nblades)[n] != 0) fl |= 4;
                sflag[s] = fl;
                if ((fl & 3) && bfm) cs->body_force_nodes += 1;
                s_cs[s] = std::cos(g.theta[n]);
                s_sn[s] = std::sin(g.theta[n]);
                s_r[s] = g.radius[n];
                if (rpm) {
                    double om = (*rpm)[n] * 2 * M_PI / 60;
                    if (!c.accel) om *= scaling;
                    s_om[s] = om;
                }
                if (c.pert) {
                    const double dx = grid.x[n] - pcen[0], dy = grid.y[n] - pcen[1], dz = grid.z[n] - pcen[2];
                    const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
                    if (!(dist > 3.0 * prad)) s_psp[s] = std::exp(-(dist * dist) / (prad * prad));
                }
                if (bfm) {
                    s_b[s] = (*blockage)[n];
                    s_gbx[s] = g.gradBlockage[3 * n];
                    s_gby[s] = g.gradBlockage[3 * n + 1];
                    s_gbz[s] = g.gradBlockage[3 * n + 2];
                }
                if (hall && (fl & 4)) {
                    s_nax[s] = (*nax)[n]; s_nrad[s] = (*nrad)[n]; s_ntan[s] = (*ntan)[n];
                    double pitch = 2.0 * M_PI * g.radius[n] / (*nblades)[n];
                    s_pitch[s] = pitch;
                    s_spitch[s] = pitch * std::cos(metal ? (*metal)[n] : 0.0);
                    if (c.bfm_model == 2) {
                        if (le < 0 || le >= NI || te < 0 || te >= NI)
                            throw std::runtime_error("LEADING/TRAILING_EDGE_INDEX outside the grid");
                        double chord = (*slen)[(te * NJ + j) * NK + k] - (*slen)[(le * NJ + j) * NK + k];
                        s_sol[s] = chord / pitch;
                        s_beta0[s] = (*beta0)[n];
                        s_kp[s] = (*kp)[n];
                    }
                    if (thollet) {
                        // streamwise length with the leading-edge fallback of
                        // source_bfm_thollet.cpp:97-106
                        double stwl = (*slen)[n];
                        if (stwl < 1E-06) {
                            if (i + 1 < NI) stwl = ((*slen)[((i + 1) * NJ + j) * NK + k] + (*slen)[n]) / 2.0;
                            if (stwl < 1E-06) stwl = 1E-06;
                        }
                        s_stwl[s] = stwl;
                        if (pivot) s_piv[s] = (*pivot)[n];
                    }
                    if (stallf) s_stall[s] = (*stallf)[n];
                }
            }
    std::vector<double> axi_gf(NI * NJ, 0.0);
    if (axi)
        for (long i = 0; i < NI; ++i)
            for (long j = 0; j < NJ; ++j) {
                const double* s = &g.surfK[((i * NJ + j) * (NK + 1) + 0) * 3];
                axi_gf[j * NI + i] = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]) * g.wedgeAngle;
            }
    c.blade2d = upload(blade2d, own);
    c.sflag = upload(sflag, own);
    c.s_cs = upload(s_cs, own); c.s_sn = upload(s_sn, own); c.s_r = upload(s_r, own); c.s_om = upload(s_om, own);
    c.s_b = upload(s_b, own); c.s_gbx = upload(s_gbx, own); c.s_gby = upload(s_gby, own); c.s_gbz = upload(s_gbz, own);
    c.s_nax = upload(s_nax, own); c.s_nrad = upload(s_nrad, own); c.s_ntan = upload(s_ntan, own);
    c.s_pitch = upload(s_pitch, own); c.s_spitch = upload(s_spitch, own); c.s_sol = upload(s_sol, own);
    c.s_beta0 = upload(s_beta0, own); c.s_kp = upload(s_kp, own);
    c.s_stwl = upload(s_stwl, own); c.s_piv = upload(s_piv, own); c.s_stall = upload(s_stall, own);
    c.s_psp = upload(s_psp, own);
    c.axi_gf = upload(axi_gf, own);

    // ---- no-slip walls and laminar viscous terms
    // wall nodes and velocities of the no_slip_wall patches in file order, the last patch
    // holding a node setting its velocity (solver.cpp:369-389, boundary_base.cpp:227-268);
    // fewer than three values leave the wall at rest
    bool moving = false;
    for (const Patch& p : cs->patches) {
        if (p.type != PatchType::NoSlipWall) continue;
        if (cs->nswHost.empty()) { cs->nswHost.assign(N, 0); cs->wallvHost.assign(3 * N, 0.0); }
        long r0[3] = {p.i0, p.j0, p.k0}, r1[3] = {p.i1, p.j1, p.k1};
        const long nn[3] = {NI, NJ, NK};
        for (int d = 0; d < 3; ++d)
            if (r0[d] == r1[d]) {
                if (r0[d] == 0) r1[d] = 1;
                else { r0[d] = nn[d] - 1; r1[d] = nn[d]; }
                break;
            }
        double u[3] = {0.0, 0.0, 0.0};
        if (p.values.size() >= 3) { u[0] = p.values[0]; u[1] = p.values[1]; u[2] = p.values[2]; }
        for (long i = r0[0]; i < r1[0]; ++i)
            for (long j = r0[1]; j < r1[1]; ++j)
                for (long k = r0[2]; k < r1[2]; ++k) {
                    const long n = (k * NJ + j) * NI + i;
                    cs->nswHost[n] = 1;
                    for (int e = 0; e < 3; ++e) cs->wallvHost[e * N + n] = u[e];
                }
        moving = moving || u[0] != 0.0 || u[1] != 0.0 || u[2] != 0.0;
    }
    if (!cs->nswHost.empty()) {
        if (cs->fluid == FluidModel::Real)
            throw std::runtime_error("no_slip_wall patches need FLUID_MODEL = ideal on the GPU path");
        c.nsw = upload(cs->nswHost, own);
        c.wallv = upload(cs->wallvHost, own);
    }
    c.visc = cfg.getBool("VISCOSITY_ACTIVE", false) ? 1 : 0;
    if (cfg.getBool("TURBULENCE_ACTIVE", false))
        throw std::runtime_error("TURBULENCE_ACTIVE is outside the GPU residual path");
    if (c.visc) {
        if (cs->fluid == FluidModel::Real)
            throw std::runtime_error("VISCOSITY_ACTIVE needs FLUID_MODEL = ideal on the GPU path");
        if (cs->topo == Topology::OneD)
            throw std::runtime_error("VISCOSITY_ACTIVE is on the GPU path for 2D, axisymmetric and 3D topologies");
        // Config::getViscosityModel / getFluidMuConstant (config.cpp:575-597),
        // FluidIdeal::setTransportProperties (fluid_ideal.cpp:209-222)
        std::string vm = cfg.has("FLUID_VISCOSITY_MODEL") ? cfg.get("FLUID_VISCOSITY_MODEL") : std::string("None");
        if (vm == "None") {
            c.mu_const = (cfg.has("FLUID_MU_CONSTANT") || cfg.has("FLUID_CONSTANT_MU")) ? 1 : 0;
        } else {
            std::transform(vm.begin(), vm.end(), vm.begin(), [](unsigned char ch) { return (char)std::tolower(ch); });
            if (vm != "sutherland" && vm != "constant") throw std::runtime_error("Unknown viscosity model: " + vm);
            c.mu_const = (vm == "constant") ? 1 : 0;
        }
        if (c.mu_const) {
            c.mu_c = cfg.has("FLUID_MU_CONSTANT") ? cfg.getFloat("FLUID_MU_CONSTANT") : cfg.getFloat("FLUID_CONSTANT_MU");
        } else {
            c.mu_ref = cfg.getFloat("FLUID_SUTHERLAND_MU_REF");
            c.mu_tref = cfg.getFloat("FLUID_SUTHERLAND_T_REF");
            c.mu_s = cfg.getFloat("FLUID_SUTHERLAND_S");
        }
        c.cp_v = cfg.getFloat("FLUID_CP");
        c.pr = cfg.getFloat("FLUID_PRANDTL_NUMBER");
        if (cs->nswHost.empty()) {   // the residual enforcement reads the wall flags
            c.nsw = upload(std::vector<int>(N, 0), own);
            c.wallv = upload(std::vector<double>(3 * N, 0.0), own);
        }
        c.wall_moving = moving ? 1 : 0;
        // face interpolation weights d0 / (d0 + d1), d the distance of a node to the face
        // centre (math_utils.cpp:270-360), for faces W E N S B T of every node
        std::vector<double> gw(6 * N, 0.0);
        const std::vector<double>&X = grid.x, &Y = grid.y, &Z = grid.z;
        auto rid = [&](long i, long j, long k) { return (i * NJ + j) * NK + k; };
        for (long i = 0; i < NI; ++i)
            for (long j = 0; j < NJ; ++j)
                for (long k = 0; k < NK; ++k) {
                    const long r = rid(i, j, k), n = (k * NJ + j) * NI + i;
                    auto weight = [&](const double* mid, long rn) {
                        const double ax = mid[0] - X[r], ay = mid[1] - Y[r], az = mid[2] - Z[r];
                        const double bx = mid[0] - X[rn], by = mid[1] - Y[rn], bz = mid[2] - Z[rn];
                        const double d0 = std::sqrt(ax * ax + ay * ay + az * az);
                        const double d1 = std::sqrt(bx * bx + by * by + bz * bz);
                        return d0 / (d0 + d1);
                    };
                    if (i > 0) gw[0 * N + n] = weight(&g.centI[((i * NJ + j) * NK + k) * 3], rid(i - 1, j, k));
                    if (i < NI - 1) gw[1 * N + n] = weight(&g.centI[(((i + 1) * NJ + j) * NK + k) * 3], rid(i + 1, j, k));
                    if (j < NJ - 1) gw[2 * N + n] = weight(&g.centJ[((i * (NJ + 1) + j + 1) * NK + k) * 3], rid(i, j + 1, k));
                    if (j > 0) gw[3 * N + n] = weight(&g.centJ[((i * (NJ + 1) + j) * NK + k) * 3], rid(i, j - 1, k));
                    if (k > 0) gw[4 * N + n] = weight(&g.centK[((i * NJ + j) * (NK + 1) + k) * 3], rid(i, j, k - 1));
                    if (k < NK - 1) gw[5 * N + n] = weight(&g.centK[((i * NJ + j) * (NK + 1) + k + 1) * 3], rid(i, j, k + 1));
                }
        c.gw = upload(gw, own);
        c.grad = dalloc<double>(12 * N, own);
        c.vf = dalloc<double>(12 * N, own);
    }
    c.finv = dalloc<double>(3 * N, own);
    c.fvisc = dalloc<double>(3 * N, own);
    c.dev = dalloc<double>(N, own);
    c.dt = dalloc<double>(N, own);
    c.dtmin = dalloc<double>(1, own);
    c.stop = dalloc<int>(1, own);
    c.iter = dalloc<long>(1, own);

    // ---- state buffers (roles A, B, C), reductions, convergence
    for (int b = 0; b < 3; ++b) c.U[b] = dalloc<double>(5 * N, own);
    c.nstages = (int)cs->rk.size();
    for (int q = 0; q < 4; ++q) c.alpha[q] = (q < c.nstages) ? cs->rk[q] : 0.0;
    c.rper = dalloc<double>(10 * NI * NJ, own);
    // GPU_STAGE_TILING = <TI>x<TJ>[/z<pieces>] fixes the stage kernel's tiling (one of the
    // tile shapes, default 8 k pieces in 3-D) instead of the autotuner's choice
    if (cfg.has("GPU_STAGE_TILING")) {
        const std::string v = cfg.get("GPU_STAGE_TILING");
        int ti = 0, tj = 0, z = (cs->topo == Topology::ThreeD) ? 8 : 1;
        const int got = std::sscanf(v.c_str(), "%dx%d/z%d", &ti, &tj, &z);
        for (int t = 0; t < kTiles && got >= 2; ++t)
            if (tileI(t) == ti && tileJ(t) == tj) cs->forced_tile = t;
        if (cs->forced_tile < 0 || z < 1)
            throw std::runtime_error("Invalid value for key \"GPU_STAGE_TILING\" (32x8 | 16x16 | 64x4 | 28x9, "
                                     "optionally /z<k pieces>)");
        cs->forced_pieces = z;
    }
    if (cs->forced_tile >= 0) setTiling(cs, cs->forced_tile, cs->forced_pieces);
    else setTiling(cs, 0, (cs->topo == Topology::ThreeD) ? 8 : 1);
    c.nparts_per = (periodic && cs->topo == Topology::ThreeD) ? (int)((NI * NJ + 255) / 256) : 0;
    // norm partials sized for the finest tiling the autotuner may pick
    long maxParts = 0;
    for (int t = 0; t < kTiles; ++t) {
        const int ti = tileI(t), tj = tileJ(t);
        maxParts = std::max(maxParts, ((NI + ti - 1) / ti) * ((NJ + tj - 1) / tj) * NK);
    }
    c.partial = dalloc<double>(5 * (maxParts + c.nparts_per), own);
    c.nblk_dt = (int)((N + 255) / 256);
    c.partial_min = dalloc<double>(c.nblk_dt, own);
    cs->hist_cap = std::max<long>(cfg.has("N_ITERATIONS") ? cfg.getMaxIterations() : 0, 1);
    c.conv.hist = dalloc<double>(5 * cs->hist_cap, own);
    // turbomachinery history of every iteration (TURBO_OUTPUT, solver.cpp:1340-1341)
    // also needed by throttle and mass-flow outlets (solver.cpp:1312-1317)
    // ... and by the time-dependent body force, which needs the physical time (the
    // minimum time step closes each row)
    c.turbo_on = (cfg.getBool("TURBO_OUTPUT", false) || c.prof_need_mf || c.accel || c.pert) ? 1 : 0;
    c.t_now = dalloc<double>(1, own);
    if (c.turbo_on) {
        const Geometry& gg = cs->geo;
        const int td = cfg.getTurboDirection();
        int bi[2] = {0, 1};
        if (td == 1) { bi[0] = 1; bi[1] = 0; } else if (td == 2) { bi[0] = 2; bi[1] = 3; } else if (td == 3) { bi[0] = 3; bi[1] = 2; }
        const bool jdir = (bi[0] >= 2);
        if (jdir && cs->topo == Topology::OneD) throw std::runtime_error("TURBO_DIRECTION j needs a 2-D grid");
        if (jdir && c.prof_need_mf)
            throw std::runtime_error("throttle and mass-flow outlets need TURBO_DIRECTION i or -i on the GPU path");
        // mass-flow outlets: which of the two turbo boundaries is theirs
        for (int q = 0; q < c.nprof; ++q)
            if (c.prof_kind[q] == 2) c.prof_mfb[q] = (bi[0] == c.prof_mfb[q]) ? 0 : 1;
        const long n1 = jdir ? NI : NJ, n2 = NK;
        c.turbo_nf = (int)(n1 * n2);
        c.turbo_nblk = (int)((n1 * n2 + 255) / 256);
        if (c.turbo_nblk > 256) throw std::runtime_error("turbo boundary too large (more than 65536 faces)");
        c.turbo_part = dalloc<double>(2 * (size_t)c.turbo_nblk * dev::TURBO_SUMS, own);
        std::vector<int> tn(2 * n1 * n2);
        std::vector<double> tS(6 * n1 * n2);
        for (int s = 0; s < 2; ++s)
            for (long p = 0; p < n1; ++p)
                for (long q = 0; q < n2; ++q) {
                    const long f = (long)s * n1 * n2 + p * n2 + q;
                    const double* S;
                    long i, j;
                    // boundary surfaces as the mesh stores them (Mesh::getMeshBoundary)
                    switch (bi[s]) {
                        case 0: S = &gg.surfI[((0 * NJ + p) * NK + q) * 3]; i = 0; j = p; break;
                        case 1: S = &gg.surfI[((NI * NJ + p) * NK + q) * 3]; i = NI - 1; j = p; break;
                        case 2: S = &gg.surfJ[((p * (NJ + 1) + 0) * NK + q) * 3]; i = p; j = 0; break;
                        default: S = &gg.surfJ[((p * (NJ + 1) + NJ) * NK + q) * 3]; i = p; j = NJ - 1; break;
                    }
                    tn[f] = (int)nodeDev(cs, i, j, q);
                    for (int d = 0; d < 3; ++d) tS[3 * f + d] = S[d];
                }
        c.turbo_node = upload(tn, own);
        c.turbo_S = upload(tS, own);
        double perDeg = 0.0;
        for (auto& pt : cs->patches)
            if (pt.type == PatchType::Periodic && !pt.values.empty() && pt.values[0] == 0)
                perDeg = (pt.values[2] == 360.0) ? 0.0 : pt.values[2];
        c.turbo_mscale = 1.0;
        if (cs->topo == Topology::Axisymmetric) c.turbo_mscale = 2.0 * M_PI / gg.wedgeAngle;
        else if (perDeg != 0.0) c.turbo_mscale = 360.0 / perDeg;
        c.turbo_hist = dalloc<double>(5 * cs->hist_cap, own);
        c.turbo_time = dalloc<double>(1, own);
    }
    c.conv.init = dalloc<double>(5, own);
    c.conv.conv_at = dalloc<long>(1, own);
    c.conv.diverged = dalloc<int>(1, own);
    c.conv.hist_len = cs->hist_cap;
    c.conv.drop = (double)cfg.getResidualsDropConvergence();
    c.conv.steady = cfg.isSimulationSteady() ? 1 : 0;
    c.conv.density_only = 0;
    long one = 1;
    HIPCHK(hipMemcpy(c.iter, &one, sizeof(long), hipMemcpyHostToDevice));
    // bisection tables of the ideal-gas inlet patches (exact: the same device operations)
    if (!c.real)
        for (size_t q = 0; q < cs->patches.size(); ++q)
            if (c.patch[q].type == dev::PT_INLET) {
                double* tab = dalloc<double>(2 * (size_t)dev::INLET_NODES, own);
                hipLaunchKernelGGL(dev::k_inlet_table, dim3((dev::INLET_NODES + 255) / 256), dim3(256), 0,
                                   cs->solo.stream, tab, c.patch[q].v0, c.patch[q].v1, c.gamma, c.rgas);
                HIPCHK(hipGetLastError());
                c.patch[q].tab = tab;
            }
    HIPCHK(hipStreamSynchronize(cs->solo.stream));
    cs->cdev = dalloc<CaseDev>(1, own);
    HIPCHK(hipMemcpy(cs->cdev, &c, sizeof(CaseDev), hipMemcpyHostToDevice));
    g_bytes = nullptr;
    cs->solo.members = {cs};
    cs->solo.dcases = cs->cdev;
    cs->solo.dcap = 1;
    cs->solo.ownsCases = false;
    // algorithmic bytes per node and stage: U^(k-1), U^(0), U^(k) (5 doubles each), dt, V
    // and one face vector of 3 components per active direction (FP32 for the FP32-flux
    // kernels, which read FP32 face records)
    const int dirs = (cs->topo == Topology::ThreeD) ? 3 : (cs->topo == Topology::OneD ? 1 : 2);
    const double rec = (cs->precision == Precision::Mixed) ? 4.0 : 8.0;
    cs->bytes_per_node = 8.0 * (15 + 2) + rec * 3 * dirs;

    // ---- initial solution: uniform (solver.cpp:853-888) or from a restart file
    // (solver.cpp:890-1295), with the restart's body forces as the side outputs' start
    std::vector<double> U0, fi0, fv0;
    bool restartForces = false;
    initialState(cfg, cfg.baseDir, cs->hf, grid, cs->topo, periodic, U0, fi0, fv0, restartForces);
    if (restartForces) {
        std::vector<double> comp(N), dev(N), h3(3 * N);
        for (const auto& pr : {std::make_pair(&fi0, c.finv), std::make_pair(&fv0, c.fvisc)}) {
            for (int q = 0; q < 3; ++q) {
                for (long n = 0; n < N; ++n) comp[n] = (*pr.first)[3 * n + q];
                toDevice(cs, comp.data(), dev);
                for (long n = 0; n < N; ++n) h3[3 * n + q] = dev[n];
            }
            HIPCHK(hipMemcpy(pr.second, h3.data(), h3.size() * 8, hipMemcpyHostToDevice));
        }
    }
    // what the volume file needs later
    cs->grid.ni = grid.ni; cs->grid.nj = grid.nj; cs->grid.nk = grid.nk;
    cs->grid.x = grid.x; cs->grid.y = grid.y; cs->grid.z = grid.z;
    for (const char* f : {"blockage", "rpm"})
        if (const std::vector<double>* v = grid.field(f)) cs->grid.fields[f] = *v;
    int rc = tbfm_gpu_set_state(cs, U0.data());
    if (rc != TBFM_OK) throw std::runtime_error(cs->err);
}

int fail(tbfm_case* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

int failb(tbfm_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    else g_create_error = msg;
    return code;
}

template <typename F>
int guarded(std::string* err, F&& f) {
    auto set = [&](const std::string& m) {
        if (err) *err = m;
        else g_create_error = m;
    };
    try {
        return f();
    } catch (const HipError& e) {
        set(e.what());
        return e.code == hipErrorOutOfMemory ? TBFM_ERR_OUT_OF_MEMORY : TBFM_ERR_HIP;
    } catch (const std::bad_alloc&) {
        set("host out of memory");
        return TBFM_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set(e.what());
        return TBFM_ERR_ARGUMENT;
    }
}
template <typename F>
int guarded(tbfm_case* c, F&& f) { return guarded(c ? &c->err : nullptr, f); }

// ------------------------------------------------------------------ plans
void freeTable(Table& t) {
    if (t.d) hipFree(t.d);
    t = Table{};
}

void uploadTable(Table& t, const std::vector<dev::Work>& w) {
    if (w.size() > t.cap) {
        freeTable(t);
        t.cap = std::max<size_t>(w.size(), 16);
        HIPCHK(hipMalloc(&t.d, t.cap * sizeof(dev::Work)));
    }
    t.n = (int)w.size();
    if (!w.empty()) HIPCHK(hipMemcpy(t.d, w.data(), w.size() * sizeof(dev::Work), hipMemcpyHostToDevice));
}

void freePlan(Plan& P) {
    if (P.stream) hipStreamSynchronize(P.stream);
    if (P.graph) hipGraphExecDestroy(P.graph);
    if (P.graphN) hipGraphExecDestroy(P.graphN);
    P.graph = P.graphN = nullptr;
    for (Table* t : {&P.dt, &P.mn, &P.prof, &P.bnd, &P.per, &P.norm, &P.start, &P.pre, &P.turbo, &P.visc, &P.wpre})
        freeTable(*t);
    for (auto& g : P.groups) freeTable(g.t);
    P.groups.clear();
    if (P.ownsCases && P.dcases) hipFree(P.dcases);
    P.dcases = nullptr;
}

// Copies the host case descriptions of the plan's members into its device case table.
void uploadCases(Plan& P) {
    if (P.ownsCases) {
        if (P.members.size() > P.dcap) {
            if (P.dcases) hipFree(P.dcases);
            P.dcap = std::max<size_t>(P.members.size(), 8);
            HIPCHK(hipMalloc(&P.dcases, P.dcap * sizeof(CaseDev)));
        }
        std::vector<CaseDev> h;
        for (tbfm_case* m : P.members) h.push_back(m->c);
        if (!h.empty()) HIPCHK(hipMemcpy(P.dcases, h.data(), h.size() * sizeof(CaseDev), hipMemcpyHostToDevice));
    } else {
        HIPCHK(hipMemcpy(P.dcases, &P.members[0]->c, sizeof(CaseDev), hipMemcpyHostToDevice));
    }
}

// Re-publishes a case's host description to every device case table that holds it
// (its own plan's and, once the batch's tables are built, its batch's).
void pushCase(tbfm_case* cs) {
    HIPCHK(hipStreamSynchronize(cs->solo.stream));
    HIPCHK(hipMemcpy(cs->cdev, &cs->c, sizeof(CaseDev), hipMemcpyHostToDevice));
    if (cs->batch && cs->batch->plan.tablesValid) {
        Plan& P = cs->batch->plan;
        HIPCHK(hipStreamSynchronize(P.stream));
        for (size_t m = 0; m < P.members.size(); ++m)
            if (P.members[m] == cs) HIPCHK(hipMemcpy(P.dcases + m, &cs->c, sizeof(CaseDev), hipMemcpyHostToDevice));
    }
}

constexpr int kXcds = 8;   // accelerator complex dies of an MI355X (each with its own L2)

// XCD-aware order of a sequence of blocks whose neighbours share data (planes of the
// state, dual nodes): the GPU hands workgroup b to XCD b % kXcds, so the sequence is cut
// into kXcds contiguous pieces and the table interleaves them; each XCD then walks one
// compact piece and keeps what neighbouring blocks share in its own L2.  Shorter pieces
// are padded with no-op entries (cs = -1).
std::vector<dev::Work> xcdOrder(const std::vector<dev::Work>& seq) {
    const size_t n = seq.size(), len = (n + kXcds - 1) / kXcds;
    std::vector<dev::Work> t(len * kXcds, dev::Work{-1, 0, 0, 0});
    for (int x = 0; x < kXcds; ++x)
        for (size_t q = n * x / kXcds; q < n * (x + 1) / kXcds; ++q) t[(q - n * x / kXcds) * kXcds + x] = seq[q];
    return t;
}

// Builds the work tables of every kernel from the members' current tiling.
void buildTables(Plan& P) {
    std::vector<dev::Work> dt, mn, prof, bnd, per, norm, src, turbo, visc, wpre;
    for (auto& g : P.groups) freeTable(g.t);
    P.groups.clear();
    P.fluids = 0;
    // Stage-kernel blocks, XCD-aware: the GPU hands workgroup b to XCD b % kXcds, and each
    // XCD has its own L2.  Every case's blocks are cut into kXcds contiguous k-major pieces
    // (gl[group][xcd]); the table interleaves them so that each XCD walks one compact
    // piece, and the planes and halos its neighbouring blocks share stay in its L2.
    // Shorter pieces are padded with no-op entries (cs = -1).
    std::vector<std::vector<std::vector<dev::Work>>> gl;
    P.nstages = P.members.empty() ? 0 : P.members[0]->c.nstages;
    for (int m = 0; m < (int)P.members.size(); ++m) {
        tbfm_case* cs = P.members[m];
        const CaseDev& c = cs->c;
        if (c.nstages != P.nstages) throw std::runtime_error("cases of one batch must use the same RK scheme");
        for (int b = 0; b < c.nblk_dt; ++b) dt.push_back({m, b, 0, 0});
        if (c.dt_mode == 1 || c.turbo_on) mn.push_back({m, 0, 0, 0});
        for (int b = 0; c.turbo_on && b < 2 * c.turbo_nblk; ++b) turbo.push_back({m, b, 0, 3});
        if (c.nprof > 0) prof.push_back({m, 0, 0, 1});
        const long nb = 2 * cs->nj * cs->nk + ((cs->topo != Topology::OneD) ? 2 * cs->ni * cs->nk : 0) +
                        ((cs->topo == Topology::ThreeD) ? 2 * cs->ni * cs->nj : 0);
        for (long b = 0; b < (nb + 255) / 256; ++b) bnd.push_back({m, (int)b, 0, 0});
        for (int b = 0; b < c.nparts_per; ++b) per.push_back({m, b, 0, 0});
        for (int b = 0; b < (c.ns + 255) / 256; ++b) src.push_back({m, b, 0, (cs->topo == Topology::ThreeD) ? 1 : 2});
        norm.push_back({m, 0, 0, 0});
        for (long b = 0; c.visc && b < (cs->N + 255) / 256; ++b) visc.push_back({m, (int)b, 0, 0});
        for (long b = 0; c.wall_moving && b < (cs->N + 255) / 256; ++b) wpre.push_back({m, (int)b, 0, 0});
        const int sch = (cs->scheme == Scheme::JST) ? 0 : 1;
        const bool three = (cs->topo == Topology::ThreeD), mixed = (cs->precision == Precision::Mixed);
        const bool real = (c.real != 0);
        P.fluids |= real ? 2 : 1;
        size_t gi = 0;
        for (; gi < P.groups.size(); ++gi) {
            const StageGroup& g = P.groups[gi];
            if (g.tile == cs->tile && g.sch == sch && g.three == three && g.mixed == mixed && g.real == real) break;
        }
        if (gi == P.groups.size()) {
            StageGroup g;
            g.tile = cs->tile; g.sch = sch; g.three = three; g.mixed = mixed; g.real = real;
            P.groups.push_back(g);
            gl.emplace_back(kXcds);
        }
        // the case's blocks in k-major order, cut into one contiguous piece per XCD
        std::vector<dev::Work> seq;
        for (int z = 0; z < c.sgz; ++z)
            for (int y = 0; y < c.sgy; ++y)
                for (int x = 0; x < c.sgx; ++x) seq.push_back({m, x, y, z});
        const size_t n = seq.size();
        for (int x = 0; x < kXcds; ++x)
            for (size_t q = n * x / kXcds; q < n * (x + 1) / kXcds; ++q) gl[gi][x].push_back(seq[q]);
    }
    uploadTable(P.dt, dt);
    uploadTable(P.mn, mn);
    uploadTable(P.prof, prof);
    uploadTable(P.bnd, bnd);
    uploadTable(P.per, per);
    uploadTable(P.norm, norm);
    uploadTable(P.visc, xcdOrder(visc));
    uploadTable(P.wpre, xcdOrder(wpre));
    // combined launches: profiles then time-step blocks; boundary then source blocks.
    // Profile and boundary blocks go first, round robin over the XCDs (the latency-bound
    // inlet bisection spread out); the time-step and source blocks follow in XCD order
    // (shared dual nodes; the Gong source reads the two neighbouring planes).
    auto padded = [](std::vector<dev::Work> v) {
        while (v.size() % kXcds) v.push_back(dev::Work{-1, 0, 0, 0});
        return v;
    };
    std::vector<dev::Work> start = padded(prof), pre = padded(bnd);
    uploadTable(P.turbo, turbo);
    const std::vector<dev::Work> dto = xcdOrder(dt), srco = xcdOrder(src);
    start.insert(start.end(), dto.begin(), dto.end());
    pre.insert(pre.end(), srco.begin(), srco.end());
    uploadTable(P.start, start);
    uploadTable(P.pre, pre);
    for (size_t g = 0; g < P.groups.size(); ++g) {
        size_t len = 0;
        for (auto& l : gl[g]) len = std::max(len, l.size());
        std::vector<dev::Work> t(len * kXcds, dev::Work{-1, 0, 0, 0});
        for (int x = 0; x < kXcds; ++x)
            for (size_t q = 0; q < gl[g][x].size(); ++q) t[q * kXcds + x] = gl[g][x][q];
        uploadTable(P.groups[g].t, t);
    }
    P.tablesValid = true;
    P.graphValid = false;
}

dev::Launch launchArg(const Plan& P, const Table& t, int in, int out, int stage, int last, int rmode) {
    dev::Launch L;
    L.cases = P.dcases;
    L.work = t.d;
    L.in = in; L.out = out; L.stage = stage; L.last = last; L.rmode = rmode;
    L.pre = 0;
    return L;
}

template <class TL, int SCH, bool THREED>
void launchStageT(const Plan& P, const StageGroup& g, const dev::Launch& L) {
    const CaseDev* cs = L.cases;
    const dev::Work* wk = L.work;
    if (g.real)   // the look-up-table fluid: FP64 throughout
        hipLaunchKernelGGL((dev::k_stage<TL, SCH, THREED, double, 1>), dim3((unsigned)g.t.n), dim3(dev::NT), 0, P.stream, cs, wk, L);
    else if (g.mixed)
        hipLaunchKernelGGL((dev::k_stage<TL, SCH, THREED, float, 0>), dim3((unsigned)g.t.n), dim3(dev::NT), 0, P.stream, cs, wk, L);
    else
        hipLaunchKernelGGL((dev::k_stage<TL, SCH, THREED, double, 0>), dim3((unsigned)g.t.n), dim3(dev::NT), 0, P.stream, cs, wk, L);
}

template <int SCH, bool THREED>
void launchStageS(const Plan& P, const StageGroup& g, const dev::Launch& L) {
    switch (g.tile) {
        case 1: launchStageT<dev::Tile<16, 16>, SCH, THREED>(P, g, L); break;
        case 2: launchStageT<dev::Tile<64, 4>, SCH, THREED>(P, g, L); break;
        case 3: launchStageT<dev::Tile<28, 9>, SCH, THREED>(P, g, L); break;
        default: launchStageT<dev::Tile<32, 8>, SCH, THREED>(P, g, L); break;
    }
}

void launchStageGroups(const Plan& P, int in, int out, int stage, int last, int rmode, int pre = 0) {
    for (const StageGroup& g : P.groups) {
        if (g.t.n == 0) continue;
        dev::Launch L = launchArg(P, g.t, in, out, stage, last, rmode);
        L.pre = pre;
        if (g.sch == 0) {
            if (g.three) launchStageS<0, true>(P, g, L); else launchStageS<0, false>(P, g, L);
        } else {
            if (g.three) launchStageS<1, true>(P, g, L); else launchStageS<1, false>(P, g, L);
        }
    }
}

// k_iterstart over a table of time-step and / or profile blocks, then the global-step
// minimum for the cases that use it.
// iteration: the launch of an explicit iteration (also closes the turbomachinery rows)
void launchStart(const Plan& P, const Table& t, int in, bool minimum, bool iteration = false) {
    // boundary partial sums first: the outlet profiles of throttle and mass-flow outlets
    // read them
    if (iteration && P.turbo.n)
        hipLaunchKernelGGL(dev::k_turbo_partial, dim3((unsigned)P.turbo.n), dim3(256), 0, P.stream,
                           launchArg(P, P.turbo, in, 0, 0, 0, 0));
    if (t.n) hipLaunchKernelGGL(dev::k_iterstart, dim3((unsigned)t.n), dim3(256), 0, P.stream, launchArg(P, t, in, 0, 0, 0, 0));
    if (minimum && P.mn.n)
        hipLaunchKernelGGL(dev::k_min_reduce, dim3((unsigned)P.mn.n), dim3(256), 0, P.stream,
                           launchArg(P, P.mn, in, 0, 0, iteration ? 1 : 0, 0));
}
void launchTimestep(const Plan& P, int in) { launchStart(P, P.dt, in, true); }
void launchProfiles(const Plan& P, int in) {
    if (P.turbo.n)   // the mass flows of throttle and mass-flow outlets
        hipLaunchKernelGGL(dev::k_turbo_partial, dim3((unsigned)P.turbo.n), dim3(256), 0, P.stream,
                           launchArg(P, P.turbo, in, 0, 0, 0, 0));
    launchStart(P, P.prof, in, false);
}

// k_prestage over a table of boundary and / or source blocks.
void launchPre(const Plan& P, const Table& t, int in, int out, int stage, int last, int rmode, int pre = 0) {
    if (!t.n) return;
    dev::Launch L = launchArg(P, t, in, out, stage, last, rmode);
    L.pre = pre;
    if (P.fluids & 1) hipLaunchKernelGGL(dev::k_prestage<0>, dim3((unsigned)t.n), dim3(256), 0, P.stream, L);
    if (P.fluids & 2) hipLaunchKernelGGL(dev::k_prestage<1>, dim3((unsigned)t.n), dim3(256), 0, P.stream, L);
}
void launchBoundary(const Plan& P, int in) { launchPre(P, P.bnd, in, 0, 0, 0, 0); }

// One RK stage of every member: boundary fluxes, fused residual + update, periodic
// closure, and after the last stage the norms and convergence test.
void launchStage(Plan& P, int s, int in, int out, bool last, bool rmode, tbfm_case* prof) {
    if (prof) {
        while (prof->pool_used * 3 + 3 > prof->pool.size()) {
            hipEvent_t a;
            HIPCHK(hipEventCreate(&a));
            prof->pool.push_back(a);
        }
        HIPCHK(hipEventRecord(prof->pool[3 * prof->pool_used], P.stream));
    }
    // viscous cases: the first stage's wall momentum, then (after the boundary fluxes and
    // slab source terms) the gradients and viscous face fluxes of the stage input
    dev::Launch V = launchArg(P, P.visc, in, out, s, last ? 1 : 0, rmode ? 1 : 0);
    V.pre = (s == 0 && !rmode) ? 1 : 0;
    if (V.pre && P.wpre.n) {
        dev::Launch Wl = V;
        Wl.work = P.wpre.d;
        hipLaunchKernelGGL(dev::k_wallpre, dim3((unsigned)P.wpre.n), dim3(256), 0, P.stream, Wl);
    }
    // boundary fluxes and slab source terms, then the fused stage kernel (timed together)
    launchPre(P, P.pre, in, out, s, last ? 1 : 0, rmode ? 1 : 0, V.pre);
    if (P.visc.n) {
        hipLaunchKernelGGL(dev::k_vgrad, dim3((unsigned)P.visc.n), dim3(256), 0, P.stream, V);
        hipLaunchKernelGGL(dev::k_vface, dim3((unsigned)P.visc.n), dim3(256), 0, P.stream, V);
    }
    if (prof) HIPCHK(hipEventRecord(prof->pool[3 * prof->pool_used + 1], P.stream));
    launchStageGroups(P, in, out, s, last ? 1 : 0, rmode ? 1 : 0, V.pre);
    if (prof) {
        HIPCHK(hipEventRecord(prof->pool[3 * prof->pool_used + 2], P.stream));
        prof->pool_used += 1;
    }
    if (P.per.n)
        hipLaunchKernelGGL(dev::k_periodic, dim3((unsigned)P.per.n), dim3(256), 0, P.stream,
                           launchArg(P, P.per, in, out, s, last ? 1 : 0, rmode ? 1 : 0));
    if (last) hipLaunchKernelGGL(dev::k_norms, dim3((unsigned)P.norm.n), dim3(256), 0, P.stream, launchArg(P, P.norm, 0, 0, 0, 0, 0));
}

// One explicit iteration of every member (solver.cpp:1337-1381): time step, radial
// profiles, then the RK stages A -> B -> C -> ... -> A with the state back in A.
void launchIteration(Plan& P, tbfm_case* prof) {
    launchStart(P, P.start, 0, true, true);
    for (int s = 0; s < P.nstages; ++s)
        launchStage(P, s, dev::stage_in(s), dev::stage_out(s, P.nstages), s + 1 == P.nstages, false, prof);
    HIPCHK(hipGetLastError());
}

// Captures `iters` iterations of the plan into one graph (kernels only: no memory
// operations inside).  An iteration after the device stopped the case is a no-op, so a
// replay may run past convergence.
hipGraphExec_t captureIterations(Plan& P, int iters) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(P.stream, hipStreamCaptureModeThreadLocal));
    try {
        for (int q = 0; q < iters; ++q) launchIteration(P, nullptr);
    } catch (...) {
        hipStreamEndCapture(P.stream, &g);
        if (g) hipGraphDestroy(g);
        throw;
    }
    HIPCHK(hipStreamEndCapture(P.stream, &g));
    hipGraphExec_t x = nullptr;
    hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    HIPCHK(e);
    return x;
}

void capture(Plan& P) {
    if (P.graphValid) return;
    for (hipGraphExec_t* x : {&P.graph, &P.graphN})
        if (*x) {
            HIPCHK(hipGraphExecDestroy(*x));
            *x = nullptr;
        }
    P.graph = captureIterations(P, 1);
    P.graphN = captureIterations(P, kGraphIters);
    P.graphValid = true;
}

// Replays n iterations of a captured plan on its stream.
void replay(Plan& P, long n) {
    long q = 0;
    for (; q + kGraphIters <= n; q += kGraphIters) HIPCHK(hipGraphLaunch(P.graphN, P.stream));
    for (; q < n; ++q) HIPCHK(hipGraphLaunch(P.graph, P.stream));
}

// Autotuner of the stage kernel's tile shape and k planes per block: times launches of
// the first RK stage of every candidate on the case's current state (reading buffer A,
// writing only buffer B, which holds nothing between iterations), two rounds of five,
// and keeps the fastest minimum.  Results are identical for every tiling, only the time
// differs; the choice is cached per grid size, scheme, topology and precision.
std::map<std::string, std::pair<int, int>> g_tuned;
std::mutex g_tuned_mu;

void autotune(tbfm_case* cs) {
    if (cs->tuned) return;
    cs->tuned = true;
    Plan& P = cs->solo;
    if (cs->forced_tile >= 0) {
        setTiling(cs, cs->forced_tile, cs->forced_pieces);
        pushCase(cs);
        buildTables(P);
        cs->tuning = "fixed -> " + std::to_string(tileI(cs->tile)) + "x" + std::to_string(tileJ(cs->tile)) + "/z" +
                     std::to_string(cs->c.sgz);
        return;
    }
    const std::string key = std::to_string(cs->ni) + "x" + std::to_string(cs->nj) + "x" + std::to_string(cs->nk) + "/" +
                            std::to_string((int)cs->scheme) + std::to_string(cs->c.muscl) + "/" +
                            std::to_string((int)cs->topo) + "/" + std::to_string((int)cs->precision) + "/" +
                            std::to_string(cs->c.real);
    {
        std::lock_guard<std::mutex> lk(g_tuned_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            setTiling(cs, it->second.first, it->second.second);
            cs->tuning = "cached -> " + std::to_string(tileI(cs->tile)) + "x" + std::to_string(tileJ(cs->tile)) +
                         "/z" + std::to_string(cs->c.sgz);
            pushCase(cs);
            return;
        }
    }
    if (cs->nk <= 1 && cs->ni * cs->nj < 4096) {   // tiny 2-D cases: default shape
        pushCase(cs);
        return;
    }
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    float best = 1e30f;
    int bt = cs->tile, bk = cs->c.sgz;
    char buf[96];
    // candidates: per tile shape, k piece counts that fill whole rounds of block slots
    // (m * 256 blocks: one or two resident blocks per CU of 256) and the plain planes-per-
    // piece choices 2, 4, 8, 16
    std::vector<std::pair<int, int>> cand;
    for (int t = 0; t < kTiles; ++t) {
        const long tiles = ((cs->ni + tileI(t) - 1) / tileI(t)) * ((cs->nj + tileJ(t) - 1) / tileJ(t));
        std::vector<int> zs;
        if (cs->topo == Topology::ThreeD) {
            for (int m = 1; m <= 8; ++m) zs.push_back((int)(m * 256 / tiles));
            for (int kc : {2, 4, 8, 16}) zs.push_back((int)((cs->nk + kc - 1) / kc));
        } else {
            zs.push_back(1);
        }
        std::sort(zs.begin(), zs.end());
        zs.erase(std::unique(zs.begin(), zs.end()), zs.end());
        for (int z : zs)
            if (z >= 1 && z <= cs->nk / 2 + (cs->nk == 1)) cand.push_back({t, z});
    }
    std::vector<float> tmin(cand.size(), 1e30f);
    for (int round = 0; round < 2; ++round)
        for (size_t q = 0; q < cand.size(); ++q) {
            setTiling(cs, cand[q].first, cand[q].second);
            pushCase(cs);
            buildTables(P);
            launchTimestep(P, 0);
            launchBoundary(P, 0);
            launchStageGroups(P, 0, 1, 0, 0, 0);   // warm-up
            for (int r = 0; r < 5; ++r) {
                HIPCHK(hipEventRecord(e0, P.stream));
                launchStageGroups(P, 0, 1, 0, 0, 0);
                HIPCHK(hipEventRecord(e1, P.stream));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                tmin[q] = std::min(tmin[q], ms);
            }
        }
    for (size_t q = 0; q < cand.size(); ++q) {
        std::snprintf(buf, sizeof buf, "%dx%d/z%d:%.4fms ", tileI(cand[q].first), tileJ(cand[q].first),
                      cand[q].second, tmin[q]);
        cs->tuning += buf;
        if (tmin[q] < best) { best = tmin[q]; bt = cand[q].first; bk = cand[q].second; }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    setTiling(cs, bt, bk);
    pushCase(cs);
    buildTables(P);
    std::snprintf(buf, sizeof buf, "-> %dx%d/z%d", tileI(bt), tileJ(bt), bk);
    cs->tuning += buf;
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    g_tuned[key] = {bt, bk};
}

// Grows a case's device residual history to hold iteration `last`.
void growHistory(tbfm_case* c, long last) {
    if (last <= c->hist_cap) return;
    DevLock lk(g_dev_mu);
    const long cap = std::max(last, 2 * c->hist_cap);
    size_t* saved = g_bytes;
    g_bytes = &c->dev_bytes;
    double* h = dalloc<double>(5 * cap, c->owned);
    double* t = c->c.turbo_on ? dalloc<double>(5 * cap, c->owned) : nullptr;
    g_bytes = saved;
    HIPCHK(hipStreamSynchronize(c->solo.stream));
    HIPCHK(hipMemcpy(h, c->c.conv.hist, 5 * c->hist_cap * 8, hipMemcpyDeviceToDevice));
    c->c.conv.hist = h;
    if (t) {
        HIPCHK(hipMemcpy(t, c->c.turbo_hist, 5 * c->hist_cap * 8, hipMemcpyDeviceToDevice));
        c->c.turbo_hist = t;
    }
    c->c.conv.hist_len = cap;
    c->hist_cap = cap;
    pushCase(c);
}

// Host bookkeeping after a plan advanced: which iterations each member executed.
void settle(tbfm_case* c) {
    long devIter = 0;
    HIPCHK(hipMemcpy(&devIter, c->c.iter, sizeof(long), hipMemcpyDeviceToHost));
    c->iter_host = devIter;
}

}  // namespace

// =================================================================== C-ABI
extern "C" {

const char* tbfm_gpu_last_error(const tbfm_case* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

void tbfm_gpu_dims(const tbfm_case* c, long* ni, long* nj, long* nk) {
    *ni = c->ni; *nj = c->nj; *nk = c->nk;
}

extern "C++" {
namespace {
double* scratchOfCurrentDevice();
}  // namespace
}

int tbfm_gpu_case_create(const tbfm_case_desc* d, tbfm_case** out) {
    if (!d || !out) return fail(nullptr, TBFM_ERR_ARGUMENT, "null argument");
    DevLock lk(g_dev_mu);
    *out = nullptr;
    auto cs = std::make_unique<tbfm_case>();
    int rc = guarded((std::string*)nullptr, [&]() {
        GridInput g;
        g.ni = d->ni; g.nj = d->nj; g.nk = d->nk;
        const long N = d->ni * d->nj * d->nk;
        if (N <= 0 || !d->x || !d->y || !d->z) throw std::invalid_argument("bad grid dimensions or coordinates");
        g.x.assign(d->x, d->x + N); g.y.assign(d->y, d->y + N); g.z.assign(d->z, d->z + N);
        for (int f = 0; f < d->nfields; ++f) g.fields[d->field_names[f]].assign(d->field_data[f], d->field_data[f] + N);
        for (int q = 0; q < d->nconfig; ++q) cs->cfg.set(d->config_keys[q], d->config_values[q]);
        for (int q = 0; q < d->npatches; ++q) {
            Patch p;
            p.name = d->patch_names[q];
            const long* r = d->patch_ranges + 6 * q;
            p.i0 = r[0]; p.i1 = r[1]; p.j0 = r[2]; p.j1 = r[3]; p.k0 = r[4]; p.k1 = r[5];
            auto info = cs->cfg.getList(p.name);
            if (info.empty()) throw std::runtime_error("Boundary info for " + p.name + " is missing in configuration.");
            p.type = patchTypeFromString(info[0]);
            for (size_t m = 1; m < info.size(); ++m) p.values.push_back(std::stod(info[m]));
            cs->patches.push_back(p);
        }
        setup(cs.get(), g);
        scratchOfCurrentDevice();   // set-up of the allocation-free time step on this GPU
        return TBFM_OK;
    });
    g_bytes = nullptr;
    if (rc != TBFM_OK) {
        tbfm_gpu_case_destroy(cs.release());
        return rc;
    }
    *out = cs.release();
    return TBFM_OK;
}

int tbfm_gpu_case_create_from_ini(const char* ini, tbfm_case** out) {
    if (!ini || !out) return fail(nullptr, TBFM_ERR_ARGUMENT, "null argument");
    DevLock lk(g_dev_mu);
    *out = nullptr;
    auto cs = std::make_unique<tbfm_case>();
    int rc = guarded((std::string*)nullptr, [&]() {
        cs->cfg = Config(ini);
        auto resolve = [&](const std::string& p) { return (!p.empty() && p[0] == '/') ? p : cs->cfg.baseDir + "/" + p; };
        GridInput g = readGridCsv(resolve(cs->cfg.getGridFile()));
        cs->patches = readBoundaryCsv(resolve(cs->cfg.getBoundaryFile()), cs->cfg);
        setup(cs.get(), g);
        scratchOfCurrentDevice();   // set-up of the allocation-free time step on this GPU
        return TBFM_OK;
    });
    g_bytes = nullptr;
    if (rc != TBFM_OK) {
        tbfm_gpu_case_destroy(cs.release());
        return rc;
    }
    *out = cs.release();
    return TBFM_OK;
}

int tbfm_gpu_batch_retire(tbfm_batch* b, tbfm_case* c);

int tbfm_gpu_case_destroy(tbfm_case* c) {
    if (!c) return TBFM_OK;
    DevLock lk(g_dev_mu);
    if (c->batch) tbfm_gpu_batch_retire(c->batch, c);
    freePlan(c->solo);
    for (void* p : c->owned) hipFree(p);
    for (hipEvent_t e : c->pool) hipEventDestroy(e);
    if (c->solo.stream) hipStreamDestroy(c->solo.stream);
    delete c;
    return TBFM_OK;
}

int tbfm_gpu_set_state(tbfm_case* c, const double* U) {
    if (!c || !U) return fail(c, TBFM_ERR_ARGUMENT, "null argument");
    return guarded(c, [&]() {
        DevLock lk(g_dev_mu);
        std::vector<double> dU(5 * c->N), tmp;
        for (int v = 0; v < 5; ++v) {
            toDevice(c, U + v * c->N, tmp);
            std::copy(tmp.begin(), tmp.end(), dU.begin() + v * c->N);
        }
        // a viscous case starts from the state with rho * wall velocity on the no-slip
        // nodes (preprocessSolution before the time loop, solver.cpp:1324)
        if (c->c.visc && !c->nswHost.empty())
            for (long n = 0; n < c->N; ++n)
                if (c->nswHost[n])
                    for (int e = 1; e <= 3; ++e) dU[e * c->N + n] = dU[n] * c->wallvHost[(e - 1) * c->N + n];
        hipStream_t st = c->solo.stream;
        HIPCHK(hipMemcpyAsync(c->c.U[0], dU.data(), dU.size() * 8, hipMemcpyHostToDevice, st));
        // radial-equilibrium profiles from the new state (solver.cpp:818-834)
        std::vector<double> prof(dev::MAX_PROF * c->nj, 0.0);
        const double g = c->c.gamma;
        for (int q = 0; q < c->c.nprof; ++q)
            for (int m = 0; m < c->c.prof_n[q]; ++m) {
                if (c->profInit[q] > 0.0) {
                    prof[q * c->nj + m] = c->profInit[q];
                    continue;
                }
                long n = ((long)c->c.prof_i[q] * c->nj + c->c.prof_j0[q] + m) * c->nk + 0;
                double rho = U[n], W1 = U[c->N + n] / rho, W2 = U[2 * c->N + n] / rho, W3 = U[3 * c->N + n] / rho,
                       W4 = U[4 * c->N + n] / rho;
                if (c->c.real) {
                    prof[q * c->nj + m] = lut_at(c->lut, LUT_P, rho, W4 - 0.5 * (W1 * W1 + W2 * W2 + W3 * W3));
                } else {
                    double vm = std::sqrt(W1 * W1 + W2 * W2 + W3 * W3);
                    double e = W4 - 0.5 * (vm * vm);
                    prof[q * c->nj + m] = (g - 1) * rho * e;
                }
            }
        // a mass-flow outlet clamps the face's j into its profile (boundary_outlet_mass_flow.cpp:13-16)
        for (int q = 0; q < c->c.nprof; ++q)
            if (c->c.prof_kind[q] == 2)
                for (long m = c->c.prof_n[q]; m < c->nj; ++m) prof[q * c->nj + m] = prof[q * c->nj + c->c.prof_n[q] - 1];
        HIPCHK(hipMemcpyAsync(c->c.prof_p, prof.data(), prof.size() * 8, hipMemcpyHostToDevice, st));
        if (c->c.greitzer) {
            // GreitzerModel::initializeState (solver.cpp:802-816): outlet hub node (NI-1, 0,
            // 0) static pressure and temperature, turbomachinery mass flow in and out
            std::vector<double> Uv(U, U + 5 * c->N);
            double t[10];
            hostTurbo(c, Uv, t);
            const long n = ((c->ni - 1) * c->nj + 0) * c->nk + 0;
            const double rho = U[n], W1 = U[c->N + n] / rho, W2 = U[2 * c->N + n] / rho, W3 = U[3 * c->N + n] / rho,
                         W4 = U[4 * c->N + n] / rho;
            double gz[3];
            if (c->c.real) {
                const double e = W4 - 0.5 * (W1 * W1 + W2 * W2 + W3 * W3);
                gz[0] = lut_at(c->lut, LUT_P, rho, e);
                gz[2] = lut_at(c->lut, LUT_T, rho, e);
            } else {
                const double vm = std::sqrt(W1 * W1 + W2 * W2 + W3 * W3);
                const double e = W4 - 0.5 * (vm * vm);
                gz[0] = (g - 1) * rho * e;
                gz[2] = gz[0] / (rho * c->c.rgas);
            }
            gz[1] = t[6];
            HIPCHK(hipMemcpyAsync(c->c.gz, gz, sizeof gz, hipMemcpyHostToDevice, st));
        }
        HIPCHK(hipStreamSynchronize(st));
        return TBFM_OK;
    });
}

int tbfm_gpu_get_state(tbfm_case* c, double* U) {
    if (!c || !U) return fail(c, TBFM_ERR_ARGUMENT, "null argument");
    return guarded(c, [&]() {
        DevLock lk(g_dev_mu);
        std::vector<double> dU(5 * c->N);
        HIPCHK(hipMemcpyAsync(dU.data(), c->c.U[0], dU.size() * 8, hipMemcpyDeviceToHost, c->solo.stream));
        HIPCHK(hipStreamSynchronize(c->solo.stream));
        for (int v = 0; v < 5; ++v) toHost(c, dU.data() + v * c->N, U + v * c->N);
        return TBFM_OK;
    });
}

namespace {
// Residual history rows [first, last] of a case and its convergence / divergence status.
int report(tbfm_case* c, long first, double* logres, long* converged_at) {
    long conv = 0;
    int div = 0;
    HIPCHK(hipMemcpy(&conv, c->c.conv.conv_at, sizeof(long), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&div, c->c.conv.diverged, sizeof(int), hipMemcpyDeviceToHost));
    if (logres) {
        const long rows = std::max<long>(c->iter_host - first, 0);
        if (rows > 0) {
            std::vector<double> h(5 * rows);
            HIPCHK(hipMemcpy(h.data(), c->c.conv.hist + (first - 1) * 5, h.size() * 8, hipMemcpyDeviceToHost));
            std::copy(h.begin(), h.end(), logres);
        }
    }
    if (converged_at) *converged_at = conv;
    if (div) {
        c->err = "Solver diverged: NaN or Inf detected in residual variable " + std::to_string(div - 1);
        return TBFM_ERR_DIVERGED;
    }
    return TBFM_OK;
}
}  // namespace

int tbfm_gpu_advance(tbfm_case* c, long n, double* logres, long* converged_at) {
    if (!c || n < 0) return fail(c, TBFM_ERR_ARGUMENT, "bad argument");
    if (c->batch) return fail(c, TBFM_ERR_ARGUMENT, "case is admitted to a batch: advance the batch");
    return guarded(c, [&]() {
        const long first = c->iter_host;
        Plan& P = c->solo;
        {
            DevLock lk(g_dev_mu);
            growHistory(c, first + n - 1);
            autotune(c);
            if (!P.tablesValid) buildTables(P);
            if (!c->prof) capture(P);
        }
        int stopped = 0;
        {
            DevLock lk(g_dev_mu);
            HIPCHK(hipMemcpy(&stopped, c->c.stop, sizeof(int), hipMemcpyDeviceToHost));
            if (!stopped && c->prof)
                for (long q = 0; q < n; ++q) launchIteration(P, c);
        }
        // graph replays and the wait run outside the lock
        if (!stopped && !c->prof) replay(P, n);
        HIPCHK(hipStreamSynchronize(P.stream));
        DevLock lk(g_dev_mu);
        if (c->prof) drainProfile(c);
        // iterations launched after the device stopped (convergence or divergence) were
        // no-ops; the state in buffer A is the one after the last executed iteration
        settle(c);
        return report(c, first, logres, converged_at);
    });
}

int tbfm_gpu_compute_timestep(tbfm_case* c, const double* U, double* dt) {
    if (!c || !U || !dt) return fail(c, TBFM_ERR_ARGUMENT, "null argument");
    return guarded(c, [&]() {
        DevLock lk(g_dev_mu);
        std::vector<double> dU(5 * c->N), tmp;
        for (int v = 0; v < 5; ++v) {
            toDevice(c, U + v * c->N, tmp);
            std::copy(tmp.begin(), tmp.end(), dU.begin() + v * c->N);
        }
        Plan& P = c->solo;
        if (!P.tablesValid) buildTables(P);
        HIPCHK(hipMemcpy(c->c.U[1], dU.data(), dU.size() * 8, hipMemcpyHostToDevice));
        launchTimestep(P, 1);
        HIPCHK(hipGetLastError());
        std::vector<double> d(c->N);
        if (c->c.dt_mode == 1) {
            double m = 0;
            HIPCHK(hipMemcpyAsync(&m, c->c.dtmin, 8, hipMemcpyDeviceToHost, P.stream));
            HIPCHK(hipStreamSynchronize(P.stream));
            std::fill(d.begin(), d.end(), m);
        } else {
            HIPCHK(hipMemcpyAsync(d.data(), c->c.dt, d.size() * 8, hipMemcpyDeviceToHost, P.stream));
            HIPCHK(hipStreamSynchronize(P.stream));
        }
        toHost(c, d.data(), dt);
        return TBFM_OK;
    });
}

int tbfm_gpu_compute_residual(tbfm_case* c, const double* U, long it, int update_profiles, double* R) {
    if (!c || !U || !R) return fail(c, TBFM_ERR_ARGUMENT, "null argument");
    return guarded(c, [&]() {
        DevLock lk(g_dev_mu);
        std::vector<double> dU(5 * c->N), tmp;
        for (int v = 0; v < 5; ++v) {
            toDevice(c, U + v * c->N, tmp);
            std::copy(tmp.begin(), tmp.end(), dU.begin() + v * c->N);
        }
        Plan& P = c->solo;
        if (!P.tablesValid) buildTables(P);
        HIPCHK(hipMemcpy(c->c.U[1], dU.data(), dU.size() * 8, hipMemcpyHostToDevice));
        long itv = it;
        int stopSaved = 0, zero = 0;
        HIPCHK(hipMemcpy(&stopSaved, c->c.stop, sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(c->c.iter, &itv, sizeof(long), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->c.stop, &zero, sizeof(int), hipMemcpyHostToDevice));
        if (update_profiles) launchProfiles(P, 1);
        launchStage(P, 0, 1, 2, false, true, nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(P.stream));
        std::vector<double> d(5 * c->N);
        HIPCHK(hipMemcpy(d.data(), c->c.U[2], d.size() * 8, hipMemcpyDeviceToHost));
        for (int v = 0; v < 5; ++v) toHost(c, d.data() + v * c->N, R + v * c->N);
        itv = c->iter_host;
        HIPCHK(hipMemcpy(c->c.iter, &itv, sizeof(long), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->c.stop, &stopSaved, sizeof(int), hipMemcpyHostToDevice));
        return TBFM_OK;
    });
}

int tbfm_gpu_get_side_outputs(tbfm_case* c, double* finv, double* fvisc, double* devang, double* dt) {
    if (!c) return fail(c, TBFM_ERR_ARGUMENT, "null argument");
    return guarded(c, [&]() {
        DevLock lk(g_dev_mu);
        HIPCHK(hipStreamSynchronize(c->solo.stream));
        std::vector<double> h3(3 * c->N), h(c->N), comp(c->N), out(c->N);
        auto vec3 = [&](const double* d, double* dst) {
            HIPCHK(hipMemcpy(h3.data(), d, h3.size() * 8, hipMemcpyDeviceToHost));
            for (int q = 0; q < 3; ++q) {
                for (long n = 0; n < c->N; ++n) comp[n] = h3[3 * n + q];
                toHost(c, comp.data(), out.data());
                for (long n = 0; n < c->N; ++n) dst[3 * n + q] = out[n];
            }
        };
        if (finv) vec3(c->c.finv, finv);
        if (fvisc) vec3(c->c.fvisc, fvisc);
        if (devang) {
            HIPCHK(hipMemcpy(h.data(), c->c.dev, h.size() * 8, hipMemcpyDeviceToHost));
            toHost(c, h.data(), devang);
        }
        if (dt) {
            if (c->c.dt_mode == 1) {
                double m = 0;
                HIPCHK(hipMemcpy(&m, c->c.dtmin, 8, hipMemcpyDeviceToHost));
                std::fill(dt, dt + c->N, m);
            } else {
                HIPCHK(hipMemcpy(h.data(), c->c.dt, h.size() * 8, hipMemcpyDeviceToHost));
                toHost(c, h.data(), dt);
            }
        }
        return TBFM_OK;
    });
}

int tbfm_gpu_get_turbo(tbfm_case* c, double* out) {
    if (!c || !out) return fail(c, TBFM_ERR_ARGUMENT, "null argument");
    return guarded(c, [&]() {
        DevLock lk(g_dev_mu);
        // boundary slices are small: evaluate on the host from the device state
        std::vector<double> U(5 * c->N);
        int rc = tbfm_gpu_get_state(c, U.data());
        if (rc != TBFM_OK) return rc;
        hostTurbo(c, U, out);
        return TBFM_OK;
    });
}

namespace {
// Solver::updateMassFlows + updateTurboPerformance (solver.cpp:1633-1781) of a state in
// the reference layout, face by face in the reference's order.
void hostTurbo(tbfm_case* c, const std::vector<double>& U, double* out) {
    {
        const long NI = c->ni, NJ = c->nj, NK = c->nk, N = c->N;
        const Geometry& g = c->geo;
        const double gm = c->c.gamma, Rg = c->c.rgas;
        auto at = [&](long i, long j, long k) { return (i * NJ + j) * NK + k; };
        auto surf = [&](int b, long p, long q) -> const double* {
            if (b == 0) return &g.surfI[((0 * NJ + p) * NK + q) * 3];
            if (b == 1) return &g.surfI[((NI * NJ + p) * NK + q) * 3];
            if (b == 2) return &g.surfJ[((p * (NJ + 1) + 0) * NK + q) * 3];
            if (b == 3) return &g.surfJ[((p * (NJ + 1) + NJ) * NK + q) * 3];
            if (b == 4) return &g.surfK[((p * NJ + q) * (NK + 1) + 0) * 3];
            return &g.surfK[((p * NJ + q) * (NK + 1) + NK) * 3];
        };
        auto node = [&](int b, long p, long q) {
            if (b == 0) return at(0, p, q);
            if (b == 1) return at(NI - 1, p, q);
            if (b == 2) return at(p, 0, q);
            if (b == 3) return at(p, NJ - 1, q);
            if (b == 4) return at(p, q, 0);
            return at(p, q, NK - 1);
        };
        auto extent = [&](int b, long& n1, long& n2) {
            if (b < 2) { n1 = NJ; n2 = NK; } else if (b < 4) { n1 = NI; n2 = NK; } else { n1 = NI; n2 = NJ; }
        };
        double mf[6];
        for (int b = 0; b < 6; ++b) {
            long n1, n2;
            extent(b, n1, n2);
            double s = 0.0;
            for (long p = 0; p < n1; ++p)
                for (long q = 0; q < n2; ++q) {
                    const double* S = surf(b, p, q);
                    long n = node(b, p, q);
                    s += S[0] * U[N + n] + S[1] * U[2 * N + n] + S[2] * U[3 * N + n];
                }
            mf[b] = s;
            out[b] = s;
        }
        int td = c->cfg.getTurboDirection();
        int bi[2] = {0, 1};
        if (td == 1) { bi[0] = 1; bi[1] = 0; } else if (td == 2) { bi[0] = 2; bi[1] = 3; } else if (td == 3) { bi[0] = 3; bi[1] = 2; }
        double m = 0.5 * (std::abs(mf[bi[0]]) + std::abs(mf[bi[1]]));
        double perDeg = 0.0;
        for (auto& p : c->patches)
            if (p.type == PatchType::Periodic && p.values[0] == 0) perDeg = (p.values[2] == 360.0) ? 0.0 : p.values[2];
        if (c->topo == Topology::Axisymmetric) m *= 2.0 * M_PI / g.wedgeAngle;
        else if (perDeg != 0.0) m *= 360.0 / perDeg;
        out[6] = m;
        const bool real = (c->c.real != 0);
        double ptm[2], ttm[2], htm[2] = {0.0, 0.0}, sm[2] = {0.0, 0.0};
        for (int s = 0; s < 2; ++s) {
            int b = bi[s];
            long n1, n2;
            extent(b, n1, n2);
            double sp = 0.0, st = 0.0, sh = 0.0, ss = 0.0;
            for (long p = 0; p < n1; ++p)
                for (long q = 0; q < n2; ++q) {
                    const double* S = surf(b, p, q);
                    long n = node(b, p, q);
                    double rho = U[n], ux = U[N + n] / rho, uy = U[2 * N + n] / rho, uz = U[3 * N + n] / rho,
                           et = U[4 * N + n] / rho;
                    if (real) {
                        // FluidReal total pressure / temperature (fluid_real.cpp:311-341),
                        // h_t = e_t + p / rho and s(rho, e) of the same state
                        const double qq = ux * ux + uy * uy + uz * uz;
                        const double e = et - 0.5 * qq;
                        const double pt = lut_total_pressure(c->lut, rho, qq, et);
                        const double Tt = lut_total_temperature(c->lut, rho, qq, et);
                        const double ht = et + lut_at(c->lut, LUT_P, rho, e) / rho;
                        const double sv = lut_at(c->lut, LUT_S, rho, e);
                        sp += S[0] * (rho * ux * pt) + S[1] * (rho * uy * pt) + S[2] * (rho * uz * pt);
                        st += S[0] * (rho * ux * Tt) + S[1] * (rho * uy * Tt) + S[2] * (rho * uz * Tt);
                        sh += S[0] * (rho * ux * ht) + S[1] * (rho * uy * ht) + S[2] * (rho * uz * ht);
                        ss += S[0] * (rho * ux * sv) + S[1] * (rho * uy * sv) + S[2] * (rho * uz * sv);
                        continue;
                    }
                    double um = std::sqrt(ux * ux + uy * uy + uz * uz);
                    double e = et - 0.5 * (um * um);
                    double pr = (gm - 1) * rho * e;
                    double a = std::sqrt(gm * pr / rho);
                    double M = um / a;
                    double pt = pr * std::pow(1.0 + (gm - 1.0) / 2.0 * M * M, gm / (gm - 1));
                    double T = pr / (rho * Rg);
                    double Tt = T * (1.0 + (gm - 1.0) / 2.0 * M * M);
                    sp += S[0] * (rho * ux * pt) + S[1] * (rho * uy * pt) + S[2] * (rho * uz * pt);
                    st += S[0] * (rho * ux * Tt) + S[1] * (rho * uy * Tt) + S[2] * (rho * uz * Tt);
                }
            double mfl = mf[b];
            if (std::abs(mfl) < 1e-12) mfl = (mfl >= 0.0 ? 1e-12 : -1e-12);
            ptm[s] = sp / mfl;
            ttm[s] = st / mfl;
            htm[s] = sh / mfl;
            sm[s] = ss / mfl;
        }
        double PR = (std::abs(ptm[0]) > 1e-12) ? ptm[1] / ptm[0] : 1.0;
        double TR = (std::abs(ttm[0]) > 1e-12) ? ttm[1] / ttm[0] : 1.0;
        out[7] = PR;
        out[8] = TR;
        double eta;
        if (real) {
            // enthalpy-based total-to-total efficiency (solver.cpp:1766-1776,
            // FluidBase::computeTotalEfficiency_h) with the isentropic outlet enthalpy
            // h(p_t,out, s_in)
            const double ht_in = htm[0], ht_out = htm[1];
            const double ht_out_s = lut_h_p_s(c->lut, ptm[1], sm[0]);
            const double dh = ht_out - ht_in, dhs = ht_out_s - ht_in;
            if (std::abs(dh) < 1e-12) eta = 1.0;
            else if (dh >= 0.0) eta = dhs / dh;
            else eta = (std::abs(dhs) < 1e-12) ? 1.0 : dh / dhs;
        } else {
            double dT = TR - 1.0, dTs = std::pow(PR, (gm - 1.0) / gm) - 1.0;
            if (std::abs(dT) < 1e-12) eta = 1.0;
            else if (dT >= 0.0) eta = dTs / dT;
            else eta = (std::abs(dTs) < 1e-12) ? 1.0 : dT / dTs;
        }
        out[9] = eta;
    }
}
}  // namespace

int tbfm_gpu_get_geometry(tbfm_case* c, int which, double* out) {
    if (!c || !out) return fail(c, TBFM_ERR_ARGUMENT, "null argument");
    const Geometry& g = c->geo;
    const std::vector<double>* src = nullptr;
    switch (which) {
        case 0: src = &g.surfI; break;
        case 1: src = &g.surfJ; break;
        case 2: src = &g.surfK; break;
        case 3: src = &g.volume; break;
        case 4: src = &g.gradBlockage; break;
        default: return fail(c, TBFM_ERR_ARGUMENT, "unknown geometry array");
    }
    if (src->empty()) return fail(c, TBFM_ERR_ARGUMENT, "geometry array not available");
    std::memcpy(out, src->data(), src->size() * 8);
    return TBFM_OK;
}

namespace {
// Folds the recorded event pairs into the running totals (synchronises the stream).
void drainProfile(tbfm_case* c) {
    HIPCHK(hipStreamSynchronize(c->solo.stream));
    for (size_t q = 0; q < c->pool_used; ++q) {
        float pre = 0.f, ms = 0.f;
        HIPCHK(hipEventElapsedTime(&pre, c->pool[3 * q], c->pool[3 * q + 1]));
        HIPCHK(hipEventElapsedTime(&ms, c->pool[3 * q + 1], c->pool[3 * q + 2]));
        c->prof_ms += ms;
        c->prof_pre_ms += pre;
        c->prof_launches += 1;
    }
    c->pool_used = 0;
}
}  // namespace

int tbfm_gpu_profile(tbfm_case* c, int enable, int reset) {
    if (!c) return fail(c, TBFM_ERR_ARGUMENT, "null argument");
    return guarded(c, [&]() {
        DevLock lk(g_dev_mu);
        drainProfile(c);
        c->prof = enable != 0;
        if (reset) { c->prof_ms = 0.0; c->prof_pre_ms = 0.0; c->prof_launches = 0; }
        return TBFM_OK;
    });
}

int tbfm_gpu_profile_get(tbfm_case* c, double* ms, long* launches, double* bytes) {
    if (!c) return fail(c, TBFM_ERR_ARGUMENT, "null argument");
    return guarded(c, [&]() {
        DevLock lk(g_dev_mu);
        drainProfile(c);
        if (ms) *ms = c->prof_ms;
        if (launches) *launches = c->prof_launches;
        // the stage kernel's contract: 208 B per node in 3-D FP64, 172 B with FP32 face
        // records (read U^(k-1), U^(0), dt, V and one face vector per direction, write U^(k))
        if (bytes) *bytes = c->bytes_per_node * (double)c->N;
        return TBFM_OK;
    });
}

int tbfm_gpu_profile_get_prestage(tbfm_case* c, double* ms, double* bytes) {
    if (!c) return fail(c, TBFM_ERR_ARGUMENT, "null argument");
    return guarded(c, [&]() {
        DevLock lk(g_dev_mu);
        drainProfile(c);
        if (ms) *ms = c->prof_pre_ms;
        // 96 B per body-force node: n_ax, n_r, n_theta, N_b, rpm, bladePresent, b,
        // grad b (3), r, theta
        if (bytes) *bytes = 96.0 * (double)c->body_force_nodes;
        return TBFM_OK;
    });
}

namespace {
// Plain streaming copy used to calibrate the HBM byte counters of rocprofv3.
__global__ void __launch_bounds__(256) k_copy_calibration(const double* __restrict__ a, double* __restrict__ b, long n) {
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long)gridDim.x * 256) b[q] = a[q];
}
}  // namespace

int tbfm_gpu_copy_calibration(long nbytes, int repeats, double* ms_per_copy) {
    if (nbytes <= 0 || repeats <= 0) return fail(nullptr, TBFM_ERR_ARGUMENT, "bad argument");
    return guarded((std::string*)nullptr, [&]() {
        DevLock lk(g_dev_mu);
        std::vector<void*> own;
        const long n = nbytes / 8;
        double* a = dalloc<double>(n, own);
        double* b = dalloc<double>(n, own);
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        const unsigned blocks = (unsigned)std::min<long>((n + 255) / 256, 256L * 64);
        hipLaunchKernelGGL(k_copy_calibration, dim3(blocks), dim3(256), 0, 0, (const double*)a, b, n);
        HIPCHK(hipEventRecord(e0, 0));
        for (int r = 0; r < repeats; ++r)
            hipLaunchKernelGGL(k_copy_calibration, dim3(blocks), dim3(256), 0, 0, (const double*)a, b, n);
        HIPCHK(hipEventRecord(e1, 0));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        if (ms_per_copy) *ms_per_copy = ms / repeats;
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        for (void* p : own) hipFree(p);
        return TBFM_OK;
    });
}

const char* tbfm_gpu_tuning(const tbfm_case* c) { return c ? c->tuning.c_str() : ""; }

int tbfm_gpu_sync(tbfm_case* c) {
    if (!c) return fail(c, TBFM_ERR_ARGUMENT, "null argument");
    return guarded(c, [&]() {
        DevLock lk(g_dev_mu);
        HIPCHK(hipStreamSynchronize(c->solo.stream));
        return TBFM_OK;
    });
}

// ------------------------------------------------------- allocation-free time step
// Per-block minima of the global time step for tbfm_gpu_compute_timestep_noalloc, and in
// the last slot the minimum itself.  One copy per GPU, held until the process ends;
// every call on a GPU shares it (calls that overlap on one GPU race).  Each call writes
// every slot it reads before reading it, so nothing has to start at zero.
static __device__ __align__(256) double scratch[TBFM_NOALLOC_MAX_BLOCKS + 1];

extern "C++" {
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};   // device address of `scratch` on each GPU

// Device address of the scratch array on the current GPU, looked up once per device
// (hipGetSymbolAddress allocates nothing).  The array's name is only a host-side
// stand-in: kernels receive this address.
double* scratchOfCurrentDevice() {
    int d = 0;
    HIPCHK(hipGetDevice(&d));
    if (d < 0 || d >= kMaxDevices) throw std::runtime_error("device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[d]) {
        void* ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw std::runtime_error("no device address for the time-step scratch array");
        g_scratch[d] = static_cast<double*>(ptr);
    }
    return g_scratch[d];
}

template <int FL>
__global__ void __launch_bounds__(256) k_timestep_noalloc(const dev::CaseDev c, const double* __restrict__ U,
                                                          double* __restrict__ dt, double* __restrict__ part) {
    dev::timestep_block<FL>(c, U, blockIdx.x, dt, part);
}

// minimum of the nb block minima into part[TBFM_NOALLOC_MAX_BLOCKS] (one block, the same
// tree as k_min_reduce)
__global__ void __launch_bounds__(256) k_min_noalloc(double* __restrict__ part, int nb) {
    __shared__ double smin[256];
    double m = 1.0e300;
    for (int q = threadIdx.x; q < nb; q += 256) m = dev::dmind(m, part[q]);
    smin[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) smin[threadIdx.x] = dev::dmind(smin[threadIdx.x], smin[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[TBFM_NOALLOC_MAX_BLOCKS] = smin[0];
}

// the global minimum at every node
__global__ void __launch_bounds__(256) k_fill_noalloc(const double* __restrict__ part, double* __restrict__ dt,
                                                      long N) {
    const long n = (long)blockIdx.x * 256 + threadIdx.x;
    if (n < N) dt[n] = part[TBFM_NOALLOC_MAX_BLOCKS];
}
}  // namespace
}  // extern "C++"

int tbfm_gpu_compute_timestep_noalloc(tbfm_case* c, const double* dU, double* ddt, void* stream) {
    if (!c || !dU || !ddt) return fail(c, TBFM_ERR_ARGUMENT, "null argument");
    if (c->N > (long)TBFM_NOALLOC_MAX_BLOCKS * 256)
        return fail(c, TBFM_ERR_ARGUMENT,
                    "grid of " + std::to_string(c->N) + " nodes exceeds the allocation-free time step's limit of " +
                        std::to_string((long)TBFM_NOALLOC_MAX_BLOCKS * 256) + " nodes");
    return guarded(c, [&]() {
        double* part = scratchOfCurrentDevice();
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->solo.stream;
        const int nb = (int)((c->N + 255) / 256);
        if (c->c.real)
            hipLaunchKernelGGL(k_timestep_noalloc<1>, dim3((unsigned)nb), dim3(256), 0, st, c->c, dU, ddt, part);
        else
            hipLaunchKernelGGL(k_timestep_noalloc<0>, dim3((unsigned)nb), dim3(256), 0, st, c->c, dU, ddt, part);
        if (c->c.dt_mode == 1) {
            hipLaunchKernelGGL(k_min_noalloc, dim3(1), dim3(256), 0, st, part, nb);
            hipLaunchKernelGGL(k_fill_noalloc, dim3((unsigned)nb), dim3(256), 0, st, (const double*)part, ddt, c->N);
        }
        HIPCHK(hipGetLastError());
        return TBFM_OK;
    });
}


// ------------------------------------------------------------------ batches
int tbfm_gpu_device_memory(long* free_bytes, long* total_bytes) {
    if (!free_bytes || !total_bytes) return fail(nullptr, TBFM_ERR_ARGUMENT, "null argument");
    return guarded((std::string*)nullptr, [&]() {
        DevLock lk(g_dev_mu);
        size_t f = 0, t = 0;
        HIPCHK(hipMemGetInfo(&f, &t));
        *free_bytes = (long)f;
        *total_bytes = (long)t;
        return TBFM_OK;
    });
}

int tbfm_gpu_case_device_bytes(const tbfm_case* c, long* bytes) {
    if (!c || !bytes) return fail(nullptr, TBFM_ERR_ARGUMENT, "null argument");
    *bytes = (long)c->dev_bytes;
    return TBFM_OK;
}

int tbfm_gpu_case_set_convergence(tbfm_case* c, int density_only, double drop) {
    if (!c || !(drop > 0.0)) return fail(c, TBFM_ERR_ARGUMENT, "bad argument");
    return guarded(c, [&]() {
        DevLock lk(g_dev_mu);
        c->c.conv.density_only = density_only ? 1 : 0;
        c->c.conv.drop = drop;
        c->c.conv.steady = 1;
        pushCase(c);
        return TBFM_OK;
    });
}

int tbfm_gpu_case_status(tbfm_case* c, long* iterations, long* converged_at, int* stopped, double* last_logres) {
    if (!c) return fail(c, TBFM_ERR_ARGUMENT, "null argument");
    return guarded(c, [&]() {
        DevLock lk(g_dev_mu);
        long conv = 0;
        int st = 0;
        HIPCHK(hipMemcpy(&conv, c->c.conv.conv_at, sizeof(long), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(&st, c->c.stop, sizeof(int), hipMemcpyDeviceToHost));
        if (iterations) *iterations = c->iter_host - 1;
        if (converged_at) *converged_at = conv;
        if (stopped) *stopped = st;
        if (last_logres) {
            if (c->iter_host > 1)
                HIPCHK(hipMemcpy(last_logres, c->c.conv.hist + (c->iter_host - 2) * 5, 5 * 8, hipMemcpyDeviceToHost));
            else
                for (int v = 0; v < 5; ++v) last_logres[v] = 0.0;
        }
        int div = 0;
        HIPCHK(hipMemcpy(&div, c->c.conv.diverged, sizeof(int), hipMemcpyDeviceToHost));
        if (div) {
            c->err = "Solver diverged: NaN or Inf detected in residual variable " + std::to_string(div - 1);
            return TBFM_ERR_DIVERGED;
        }
        return TBFM_OK;
    });
}

const char* tbfm_gpu_batch_last_error(const tbfm_batch* b) { return b ? b->err.c_str() : g_create_error.c_str(); }

int tbfm_gpu_batch_create(tbfm_batch** out) {
    if (!out) return failb(nullptr, TBFM_ERR_ARGUMENT, "null argument");
    *out = nullptr;
    auto b = std::make_unique<tbfm_batch>();
    int rc = guarded((std::string*)nullptr, [&]() {
        DevLock lk(g_dev_mu);
        HIPCHK(hipStreamCreateWithFlags(&b->plan.stream, hipStreamNonBlocking));
        b->plan.ownsCases = true;
        return TBFM_OK;
    });
    if (rc != TBFM_OK) return rc;
    *out = b.release();
    return TBFM_OK;
}

int tbfm_gpu_batch_admit(tbfm_batch* b, tbfm_case* c) {
    if (!b || !c) return failb(b, TBFM_ERR_ARGUMENT, "null argument");
    if (c->batch) return failb(b, TBFM_ERR_ARGUMENT, "case is already admitted to a batch");
    return guarded(&b->err, [&]() {
        DevLock lk(g_dev_mu);
        Plan& P = b->plan;
        if (!P.members.empty() && P.members[0]->c.nstages != c->c.nstages)
            throw std::runtime_error("cases of one batch must use the same RK scheme");
        autotune(c);
        c->batch = b;
        P.members.push_back(c);
        P.tablesValid = false;
        P.graphValid = false;
        return TBFM_OK;
    });
}

int tbfm_gpu_batch_retire(tbfm_batch* b, tbfm_case* c) {
    if (!b || !c) return failb(b, TBFM_ERR_ARGUMENT, "null argument");
    if (c->batch != b) return failb(b, TBFM_ERR_ARGUMENT, "case is not a member of this batch");
    return guarded(&b->err, [&]() {
        DevLock lk(g_dev_mu);
        Plan& P = b->plan;
        HIPCHK(hipStreamSynchronize(P.stream));
        P.members.erase(std::find(P.members.begin(), P.members.end(), c));
        c->batch = nullptr;
        P.tablesValid = false;
        P.graphValid = false;
        return TBFM_OK;
    });
}

int tbfm_gpu_batch_size(const tbfm_batch* b) { return b ? (int)b->plan.members.size() : 0; }

int tbfm_gpu_batch_advance(tbfm_batch* b, long n) {
    if (!b || n < 0) return failb(b, TBFM_ERR_ARGUMENT, "bad argument");
    return guarded(&b->err, [&]() {
        Plan& P = b->plan;
        if (P.members.empty() || n == 0) return TBFM_OK;
        {
            DevLock lk(g_dev_mu);
            for (tbfm_case* m : P.members) growHistory(m, m->iter_host + n - 1);
            if (!P.tablesValid) {
                uploadCases(P);
                buildTables(P);
            }
            capture(P);
        }
        replay(P, n);
        HIPCHK(hipStreamSynchronize(P.stream));
        DevLock lk(g_dev_mu);
        for (tbfm_case* m : P.members) settle(m);
        return TBFM_OK;
    });
}

int tbfm_gpu_batch_destroy(tbfm_batch* b) {
    if (!b) return TBFM_OK;
    DevLock lk(g_dev_mu);
    if (b->plan.stream) hipStreamSynchronize(b->plan.stream);
    for (tbfm_case* m : b->plan.members) m->batch = nullptr;
    b->plan.members.clear();
    freePlan(b->plan);
    if (b->plan.stream) hipStreamDestroy(b->plan.stream);
    delete b;
    return TBFM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- on-disk formats

namespace {

// Config, grid, boundaries and fluid of an input.ini, host side only
struct HostCase {
    Config cfg;
    GridInput grid;
    std::vector<Patch> patches;
    HostFluid hf;
    explicit HostCase(const char* ini) : cfg(ini) {
        auto resolve = [&](const std::string& p) { return (!p.empty() && p[0] == '/') ? p : cfg.baseDir + "/" + p; };
        grid = readGridCsv(resolve(cfg.getGridFile()));
        patches = readBoundaryCsv(resolve(cfg.getBoundaryFile()), cfg);
        hf.load(cfg, cfg.baseDir);
    }
    bool periodic() const {
        for (auto& p : patches)
            if (p.type == PatchType::Periodic && !p.values.empty() && p.values[0] == 0) return true;
        return false;
    }
};

VolumeData volumeData(const GridInput& g, const Geometry* geo, const double* U, const double* finv,
                      const double* fvisc, const double* dev) {
    VolumeData d;
    d.ni = g.ni; d.nj = g.nj; d.nk = g.nk;
    d.x = g.x.data(); d.y = g.y.data(); d.z = g.z.data();
    d.U = U; d.finv = finv; d.fvisc = fvisc; d.dev = dev;
    if (const std::vector<double>* b = g.field("blockage")) d.blockage = b->data();
    if (const std::vector<double>* r = g.field("rpm")) d.rpm = r->data();
    if (geo) { d.radius = geo->radius.data(); d.theta = geo->theta.data(); }
    return d;
}

}  // namespace

int tbfm_gpu_write_solution(tbfm_case* c, const char* path) {
    if (!c || !path) return fail(c, TBFM_ERR_ARGUMENT, "null argument");
    return guarded(c, [&]() {
        std::vector<double> U(5 * c->N), fi(3 * c->N), fv(3 * c->N), dv(c->N);
        int rc = tbfm_gpu_get_state(c, U.data());
        if (rc == TBFM_OK) rc = tbfm_gpu_get_side_outputs(c, fi.data(), fv.data(), dv.data(), nullptr);
        if (rc != TBFM_OK) return rc;
        writeVolumeCsv(path, c->cfg, c->hf, volumeData(c->grid, &c->geo, U.data(), fi.data(), fv.data(), dv.data()));
        return TBFM_OK;
    });
}

int tbfm_write_solution_from_ini(const char* ini, const double* U, const double* finv, const double* fvisc,
                                 const double* dev, const char* path) {
    if (!ini || !U || !path) return fail(nullptr, TBFM_ERR_ARGUMENT, "null argument");
    return guarded((std::string*)nullptr, [&]() {
        HostCase h(ini);
        const Geometry geo = buildGeometry(h.grid, h.cfg.getTopology(), false);
        writeVolumeCsv(path, h.cfg, h.hf, volumeData(h.grid, &geo, U, finv, fvisc, dev));
        return TBFM_OK;
    });
}

int tbfm_volume_file_name(const char* ini, long iteration, char* name, long size) {
    if (!ini || !name || size <= 0) return fail(nullptr, TBFM_ERR_ARGUMENT, "null argument");
    return guarded((std::string*)nullptr, [&]() {
        const std::string f = volumeFileName(Config(ini), iteration);
        if ((long)f.size() + 1 > size) throw std::invalid_argument("name buffer too small");
        std::memcpy(name, f.c_str(), f.size() + 1);
        return TBFM_OK;
    });
}

int tbfm_initial_state_from_ini(const char* ini, double* U, double* finv, double* fvisc) {
    if (!ini || !U) return fail(nullptr, TBFM_ERR_ARGUMENT, "null argument");
    return guarded((std::string*)nullptr, [&]() {
        HostCase h(ini);
        std::vector<double> u, fi, fv;
        bool forces = false;
        initialState(h.cfg, h.cfg.baseDir, h.hf, h.grid, h.cfg.getTopology(), h.periodic(), u, fi, fv, forces);
        std::copy(u.begin(), u.end(), U);
        if (finv) std::copy(fi.begin(), fi.end(), finv);
        if (fvisc) std::copy(fv.begin(), fv.end(), fvisc);
        return TBFM_OK;
    });
}

int tbfm_gpu_get_turbo_history(tbfm_case* c, long first, long n, double* out) {
    if (!c || !out || first < 1 || n < 0) return fail(c, TBFM_ERR_ARGUMENT, "bad argument");
    return guarded(c, [&]() {
        if (!c->c.turbo_on) throw std::invalid_argument("the case was created without TURBO_OUTPUT");
        if (first + n - 1 >= c->iter_host) throw std::invalid_argument("iterations beyond those executed");
        DevLock lk(g_dev_mu);
        HIPCHK(hipStreamSynchronize(c->solo.stream));
        if (n > 0) HIPCHK(hipMemcpy(out, c->c.turbo_hist + 5 * (first - 1), 5 * n * 8, hipMemcpyDeviceToHost));
        return TBFM_OK;
    });
}

