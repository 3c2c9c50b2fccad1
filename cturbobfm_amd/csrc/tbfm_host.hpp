// Host side of the explicit residual path: input.ini configuration, grid and
// boundary-patch input, and the dual-grid geometry the device kernels consume.
//
// Mirrors the reference surface so that `turbobfm input.ini` reads the same files:
//   Config              src/config.cpp:15-44 (flat KEY = value), getters include/config.hpp
//   GridInput           src/input.cpp:14-82 (NI=/NJ=/NK= header, x,y,z[,field...] rows, k fastest)
//   boundary patches    src/solver.cpp:232-294 (PATCH_NAME header, I/J/K ranges)
//   Geometry            src/mesh.cpp:373-581 (dual grid, face vectors/centres, volumes,
//                       element edges, blockage Green-Gauss gradient)
// All floating-point steps are written in the reference's operation order; the library
// is compiled with -ffp-contract=off so that geometry is bit-identical.
#pragma once

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace tbfm {

enum class Topology { TwoD, ThreeD, Axisymmetric, OneD };
enum class Scheme { JST, ROE };
enum class Limiter { None = 0, VanAlbada = 1, VanLeer = 2, MinMod = 3 };
enum class TimeStepMethod { Local = 0, Global = 1, Fixed = 2 };
enum class BfmModel { None = 0, Hall = 1, LiftDrag = 2, Blockage = 3, HallThollet = 4 };
enum class PatchType { Inlet = 0, Outlet = 1, RadialEquilibrium = 2, Wall = 3, Periodic = 4, Wedge = 5,
                       Throttle = 6, MassFlow = 7, Transparent = 8, NoSlipWall = 9 };
enum class Precision { FP64 = 0, Mixed = 1 };
enum class FluidModel { Ideal = 0, Real = 1 };

class Config {
public:
    Config() = default;
    explicit Config(const std::string& path);
    void set(const std::string& key, const std::string& value) { data_[key] = value; }
    bool has(const std::string& key) const { return data_.count(key) > 0; }
    std::string get(const std::string& key) const;
    double getFloat(const std::string& key) const;
    double getFloat(const std::string& key, double def) const;
    bool getBool(const std::string& key, bool def) const;
    std::vector<std::string> getList(const std::string& key) const;

    // typed getters with the reference's names and defaults (include/config.hpp)
    Topology getTopology() const;
    Scheme getAdvectionScheme() const;
    bool getMUSCLreconstruction() const { return getBool("MUSCL_RECONSTRUCTION", false); }
    Limiter getFluxLimiter() const;
    std::vector<double> getTimeIntegrationCoeffs() const;
    TimeStepMethod getTimeStepMethod() const;
    double getCFL() const;
    double getFixedTimeStep() const { return getFloat("FIXED_TIME_STEP", 0.0); }
    FluidModel getFluidModel() const;
    std::string getFluidName() const;
    std::string getFluidTableFile() const;
    double getFluidGamma() const { return getFloat("FLUID_GAMMA"); }
    double getFluidGasConstant() const { return getFloat("FLUID_R_CONSTANT"); }
    double getInitMachNumber() const { return getFloat("INIT_MACH_NUMBER"); }
    double getInitTemperature() const { return getFloat("INIT_TEMPERATURE"); }
    double getInitPressure() const { return getFloat("INIT_PRESSURE"); }
    std::vector<double> getInitDirection() const;
    double getOutletPressureRampMaxIterations() const { return getFloat("OUTLET_PRESSURE_RAMP_ITERATIONS", 0.0); }
    bool isBFMActive() const { return getBool("BFM_ACTIVE", false); }
    bool isBlockageActive() const { return getBool("BLOCKAGE_ACTIVE", false); }
    BfmModel getBFMModel() const;
    // Hall-Thollet (include/config.hpp:295-313, src/config.cpp:441-450)
    double getHallTholletCoefficient_KN() const { return getFloat("HALL_THOLLET_COEFFICIENT_KN"); }
    double getHallTholletCoefficient_KF() const { return getFloat("HALL_THOLLET_COEFFICIENT_KF"); }
    double getHallTholletCoefficient_KD() const { return getFloat("HALL_THOLLET_COEFFICIENT_KD"); }
    bool getHallTholletOffDesignActive() const {
        if (!has("HALL_THOLLET_OFF_DESIGN_ACTIVE"))
            throw std::runtime_error("Not found value for key \"HALL_THOLLET_OFF_DESIGN_ACTIVE\" in configuration.");
        return getBool("HALL_THOLLET_OFF_DESIGN_ACTIVE", false);
    }
    bool isStalledBfmActive() const { return getBool("STALLED_BFM_ACTIVE", false); }
    long getLeadingEdgeIndex() const { return (long)getFloat("LEADING_EDGE_INDEX", 0.0); }
    long getTrailingEdgeIndex() const { return (long)getFloat("TRAILING_EDGE_INDEX", 0.0); }
    double getRotationalSpeedScalingFactor() const { return getFloat("ROTATIONAL_SPEED_SCALING_FACTOR", 1.0); }
    bool isInletCylindrical() const;
    int getTurboDirection() const;
    long getResidualsDropConvergence() const { return (long)getFloat("RESIDUALS_DROP_CONVERGENCE", 16.0); }
    bool isSimulationSteady() const { return getBool("SIMULATION_IS_STEADY", true); }
    long getMaxIterations() const { return (long)getFloat("N_ITERATIONS"); }
    Precision getGpuPrecision() const;
    std::string getGridFile() const { return get("GRID_FILE"); }
    std::string getBoundaryFile() const { return get("BOUNDARY_CONDITIONS_FILEPATH"); }
    const std::map<std::string, std::string>& data() const { return data_; }
    std::string baseDir;

private:
    std::map<std::string, std::string> data_;
};

struct Patch {
    std::string name;
    PatchType type;
    std::vector<double> values;
    long i0, i1, j0, j1, k0, k1;
};

// Grid points plus optional per-node input fields, in the reference's k-fastest order.
struct GridInput {
    long ni = 0, nj = 0, nk = 0;
    std::vector<double> x, y, z;
    std::map<std::string, std::vector<double>> fields;
    const std::vector<double>* field(const std::string& name) const {
        auto it = fields.find(name);
        return it == fields.end() ? nullptr : &it->second;
    }
};

GridInput readGridCsv(const std::string& path);

// A fluid look-up table file (src/fluid_real.cpp:12-209): header KEY: value lines, then
// the sections RHO_E (p T a s mu kappa dp/drho dp/de over rho, e), P_T (rho e a s over
// p, T) and P_S (rho T e over p, s), each with its node counts and bounds.  The fifteen
// tables are stored in that order, value (i, j) of a table at v[i * ny + j].
struct FluidTable {
    struct Tab {
        long nx = 0, ny = 0;
        double xmin = 0, xmax = 0, ymin = 0, ymax = 0;
        std::vector<double> v;
    };
    static constexpr int kTables = 15;
    std::map<std::string, std::string> header;
    Tab t[kTables];
};
FluidTable readFluidTable(const std::string& path);
std::vector<Patch> readBoundaryCsv(const std::string& path, const Config& cfg);
PatchType patchTypeFromString(const std::string& s);

// Dual-grid geometry, arrays in the reference's (k fastest) index order:
//   surfI: (NI+1, NJ, NK) xyz,  surfJ: (NI, NJ+1, NK),  surfK: (NI, NJ, NK+1)
//   dual:  (NI+1, NJ+1, NK+1) xyz,  volume, radius, theta: (NI, NJ, NK)
struct Geometry {
    long ni, nj, nk;
    std::vector<double> dual, surfI, surfJ, surfK, centI, centJ, centK, volume, radius, theta;
    std::vector<double> gradBlockage;   // 3*N, only when BFM is active
    double wedgeAngle = 0.0;
    long idx(long i, long j, long k) const { return (i * nj + j) * nk + k; }
};

Geometry buildGeometry(const GridInput& grid, Topology topo, bool withBlockageGradient);

}  // namespace tbfm
