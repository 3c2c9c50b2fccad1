// On-disk formats around the residual path, host side: the fluid relations the files use,
// the volume solution file of the reference's Output (src/output.cpp:367-447), the
// residual history file (solver.cpp:2520-2556) and the initial state of a case, from
// scratch or from a restart file (solver.cpp:777-1295, kdtree.hpp).
#pragma once
#include <string>
#include <vector>

#include "tbfm_fluid.hpp"
#include "tbfm_host.hpp"

namespace tbfm {

// The fluid seen from the host: the ideal gas (FluidIdeal, fluid_ideal.cpp) or the
// look-up-table fluid (FluidReal, fluid_real.cpp) with its tables in host memory.
struct HostFluid {
    bool real = false;
    double gamma = 1.4, R = 287.0, cp = 0.0, cv = 0.0;
    LutSet lut{};                 // host tables (real)
    std::vector<double> values;   // their values, all tables back to back
    std::vector<std::pair<std::string, std::string>> header;   // .lut header keys

    // FLUID_MODEL, FLUID_GAMMA, FLUID_R_CONSTANT, FLUID_TABLE_FILE (relative paths from
    // baseDir when it is set, else from the working folder)
    void load(const Config& cfg, const std::string& baseDir);

    // static energy from velocity and total energy: the ideal gas squares |u| after the
    // square root (fluid_ideal.cpp:74-77), the table fluid takes u.u (fluid_base.hpp:79-81)
    double e_u_et(double ux, double uy, double uz, double et) const;
    double p_rho_e(double rho, double e) const;
    double T_rho_e(double rho, double e) const;
    double a_rho_e(double rho, double e) const;
    double s_rho_e(double rho, double e) const;
    double p_rho_T(double rho, double T) const;
    double e_p_rho(double p, double rho) const;
    double total_pressure(double rho, double ux, double uy, double uz, double et) const;
    double total_temperature(double rho, double ux, double uy, double uz, double et) const;
};

// Which fields the volume file holds (Config::getOutputFieldsType, config.cpp:629-648).
enum class OutputFields { Primary, Secondary, TurboBfm };
OutputFields getOutputFields(const Config& cfg);

// Everything the volume file reads, in the reference layout ((i*NJ + j)*NK + k).
struct VolumeData {
    long ni = 0, nj = 0, nk = 0;
    const double *x = nullptr, *y = nullptr, *z = nullptr;   // grid points
    const double* U = nullptr;                               // 5 x N conserved
    const double *finv = nullptr, *fvisc = nullptr;          // 3 x N, xyz triples (may be null)
    const double* dev = nullptr;                             // N (may be null)
    const double *blockage = nullptr, *rpm = nullptr;        // grid fields (BFM)
    const double *radius = nullptr, *theta = nullptr;        // node radius and angle
};

// Output::writeSolution: NI / NJ / NK lines, "x,y,z" and the field names in
// std::map (name) order, one row per node with k fastest, 16 significant digits.
void writeVolumeCsv(const std::string& path, const Config& cfg, const HostFluid& fluid, const VolumeData& d);

// File name of the volume solution at an iteration (Output::getOutputFilename,
// output.cpp:352-364): SOLUTION_NAME, with _<6-digit iteration> when SAVE_UNSTEADY.
std::string volumeFileName(const Config& cfg, long iteration);

// Initial conserved state (5 x N, reference layout) and body forces (3 x N each) of a
// case: RESTART_SOLUTION = yes reads RESTART_SOLUTION_FILEPATH (standard, axisymmetric
// or nearest-neighbour restart), otherwise the uniform state of INIT_*.  periodic: the
// case has a periodic patch pair (Mesh::isPeriodicityActive).  restartForces is set when
// the restart file provided the body forces.
void initialState(const Config& cfg, const std::string& baseDir, const HostFluid& fluid, const GridInput& grid,
                  Topology topo, bool periodic, std::vector<double>& U, std::vector<double>& finv,
                  std::vector<double>& fvisc, bool& restartForces);

}  // namespace tbfm
