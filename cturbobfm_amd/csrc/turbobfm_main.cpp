// `turbobfm input.ini` for the GPU path (replaces src/turbobfm.cpp:7-39 when the
// explicit residual path runs on the device).  Reads the same input.ini, grid CSV and
// boundaries CSV, advances up to N_ITERATIONS iterations (stopping at the reference's
// convergence test), prints the residual table like Solver::printLogResiduals
// (solver.cpp:1473-1484) and writes residuals.csv and the final state.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/tbfm_gpu.h"

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s input.ini [max_iterations]\n", argv[0]);
        return 1;
    }
    tbfm_case* c = nullptr;
    if (tbfm_gpu_case_create_from_ini(argv[1], &c) != TBFM_OK) {
        std::fprintf(stderr, "Solver execution halted: %s\n", tbfm_gpu_last_error(nullptr));
        return 1;
    }
    long ni, nj, nk;
    tbfm_gpu_dims(c, &ni, &nj, &nk);
    long nmax = 0;
    {
        // N_ITERATIONS from input.ini, overridable on the command line
        FILE* f = std::fopen(argv[1], "r");
        char line[4096];
        while (f && std::fgets(line, sizeof line, f)) {
            std::string s(line);
            size_t eq = s.find('=');
            if (eq == std::string::npos) continue;
            std::string k = s.substr(0, eq);
            k.erase(0, k.find_first_not_of(" \t"));
            k.erase(k.find_last_not_of(" \t") + 1);
            if (k == "N_ITERATIONS") nmax = (long)std::atof(s.c_str() + eq + 1);
        }
        if (f) std::fclose(f);
    }
    if (argc >= 3) nmax = std::atol(argv[2]);
    const long chunk = 100;
    std::vector<double> hist;
    long conv = 0, done = 0;
    std::printf("|%14s|%14s|%14s|%14s|%14s|%14s|\n", "Iteration", "rms[Rho]", "rms[RhoU]", "rms[RhoV]", "rms[RhoW]", "rms[RhoE]");
    while (done < nmax && conv == 0) {
        long n = std::min(chunk, nmax - done);
        std::vector<double> lr(5 * n, 0.0);
        int rc = tbfm_gpu_advance(c, n, lr.data(), &conv);
        long rows = conv ? conv - done : n;
        for (long r = 0; r < rows; ++r) {
            std::printf("|%-14ld|%14.6f|%14.6f|%14.6f|%14.6f|%14.6f|\n", done + r + 1, lr[5 * r], lr[5 * r + 1],
                        lr[5 * r + 2], lr[5 * r + 3], lr[5 * r + 4]);
            hist.insert(hist.end(), lr.begin() + 5 * r, lr.begin() + 5 * r + 5);
        }
        if (rc != TBFM_OK) {
            std::fprintf(stderr, "Solver execution halted: %s\n", tbfm_gpu_last_error(c));
            tbfm_gpu_case_destroy(c);
            return 1;
        }
        done += rows;
    }
    if (conv) std::printf("\nConvergence reached at iteration %ld\n", conv);
    FILE* f = std::fopen("residuals.csv", "w");
    if (f) {
        std::fprintf(f, "Iteration,rms[Rho],rms[RhoU],rms[RhoV],rms[RhoW],rms[RhoE]\n");
        for (size_t r = 0; r < hist.size() / 5; ++r)
            std::fprintf(f, "%zu,%.6g,%.6g,%.6g,%.6g,%.6g\n", r + 1, hist[5 * r], hist[5 * r + 1], hist[5 * r + 2],
                         hist[5 * r + 3], hist[5 * r + 4]);
        std::fclose(f);
    }
    const long N = ni * nj * nk;
    std::vector<double> U(5 * N);
    tbfm_gpu_get_state(c, U.data());
    f = std::fopen("solution.csv", "w");
    if (f) {
        std::fprintf(f, "NI=%ld\nNJ=%ld\nNK=%ld\nrho,rhoU,rhoV,rhoW,rhoE\n", ni, nj, nk);
        for (long n = 0; n < N; ++n)
            std::fprintf(f, "%.16e,%.16e,%.16e,%.16e,%.16e\n", U[n], U[N + n], U[2 * N + n], U[3 * N + n], U[4 * N + n]);
        std::fclose(f);
    }
    double turbo[10];
    if (tbfm_gpu_get_turbo(c, turbo) == TBFM_OK)
        std::printf("\nMass Flow [kg/s]: %.6g\nTotal Pressure Ratio [-]: %.6g\nTotal Temperature Ratio [-]: %.6g\n",
                    turbo[6], turbo[7], turbo[8]);
    tbfm_gpu_case_destroy(c);
    return 0;
}
