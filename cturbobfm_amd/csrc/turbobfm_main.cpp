// `turbobfm input.ini` for the GPU path (replaces src/turbobfm.cpp:7-39 when the
// explicit residual path runs on the device).  Reads the same input.ini, grid CSV,
// boundaries CSV and restart file, advances up to N_ITERATIONS iterations (stopping at
// the reference's convergence test), prints the residual table like
// Solver::printLogResiduals (solver.cpp:1473-1484) and writes the reference's files in
// the working folder:
//   Volume_CSV/<SOLUTION_NAME>[_<iteration>].csv  every SAVE_ITERATIONS_INTERVAL
//       iterations, at N_ITERATIONS and at convergence (solver.cpp:1400-1413,
//       output.cpp:367-447);
//   residuals.csv  "Rho,RhoU,RhoV,RhoW,RhoE" and one row of log residuals per iteration,
//       six significant digits (Solver::writeLogResidualsToCsvFile, solver.cpp:2520-2556);
//   turbo.csv  with TURBO_OUTPUT = yes: time [us], mass flow, PRtt, TRtt, ETAtt of every
//       iteration, six significant digits (Solver::writeTurboPerformanceToCsvFile,
//       solver.cpp:2558-2595), the rows paired with the time list the way the reference
//       pairs them between its periodic flushes.
#include <sys/stat.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/tbfm_gpu.h"

namespace {

// KEY = value of input.ini ("" when absent)
std::string iniValue(const char* ini, const char* key) {
    FILE* f = std::fopen(ini, "r");
    char line[4096];
    std::string out;
    while (f && std::fgets(line, sizeof line, f)) {
        std::string s(line);
        if (s.empty() || s[0] == ';' || s[0] == '[') continue;
        size_t eq = s.find('=');
        if (eq == std::string::npos) continue;
        std::string k = s.substr(0, eq);
        k.erase(0, k.find_first_not_of(" \t"));
        k.erase(k.find_last_not_of(" \t") + 1);
        if (k == key) {
            out = s.substr(eq + 1);
            out.erase(0, out.find_first_not_of(" \t"));
            out.erase(out.find_last_not_of(" \t\r\n") + 1);
        }
    }
    if (f) std::fclose(f);
    return out;
}

bool writeVolume(tbfm_case* c, const char* ini, long it) {
    char name[4096];
    if (tbfm_volume_file_name(ini, it, name, sizeof name) != TBFM_OK) {
        std::fprintf(stderr, "Solver execution halted: %s\n", tbfm_gpu_last_error(nullptr));
        return false;
    }
    mkdir("Volume_CSV", 0755);
    const std::string path = std::string("Volume_CSV/") + name + ".csv";
    if (tbfm_gpu_write_solution(c, path.c_str()) != TBFM_OK) {
        std::fprintf(stderr, "Solver execution halted: %s\n", tbfm_gpu_last_error(c));
        return false;
    }
    std::printf("\nSolution written to file: %s\n\n", name);
    return true;
}

bool iniBool(const char* ini, const char* key) {
    std::string v = iniValue(ini, key);
    std::transform(v.begin(), v.end(), v.begin(), [](unsigned char ch) { return (char)std::tolower(ch); });
    return v == "yes" || v == "true" || v == "1";
}

// turbo.csv from the per-iteration rows (time after the iteration, mass flow, PRtt, TRtt,
// ETAtt).  The reference keeps a time list that starts with 0 and grows by one entry per
// iteration, and a list of turbo rows; at every flush (SAVE_ITERATIONS_INTERVAL,
// N_ITERATIONS, HISTORY_BUFFER_SIZE residual rows, convergence) row q is written with
// time entry q and both lists are emptied (solver.cpp:1383-1421, 2558-2595).
void writeTurboCsv(const std::vector<double>& rows, long done, long conv, long nmax, long freq, long bufsize) {
    FILE* f = std::fopen("turbo.csv", "w");
    if (!f) return;
    std::fprintf(f, "Time[\xce\xbcs],Massflow[kg/s],PRtt,TRtt,ETAtt\n");
    std::vector<double> times{0.0};
    std::vector<const double*> pending;
    double cur = 0.0;
    long buffered = 0;
    auto flush = [&]() {
        for (size_t q = 0; q < pending.size(); ++q) {
            const double t = (q < times.size()) ? times[q] : cur;
            const double* r = pending[q];
            std::fprintf(f, "%g,%g,%g,%g,%g\n", t * 1E6, r[1], r[2], r[3], r[4]);
        }
        pending.clear();
        times.clear();
        buffered = 0;
    };
    for (long it = 1; it <= done; ++it) {
        const double* r = &rows[5 * (it - 1)];
        pending.push_back(r);
        cur = r[0];
        times.push_back(cur);
        ++buffered;
        if (conv == it) { flush(); break; }
        if (it % freq == 0 || it == nmax || buffered >= bufsize) flush();
    }
    std::fclose(f);
    std::printf("\nTurbo performance appended to turbo.csv\n\n");
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s input.ini [max_iterations]\n", argv[0]);
        return 1;
    }
    const char* ini = argv[1];
    tbfm_case* c = nullptr;
    if (tbfm_gpu_case_create_from_ini(ini, &c) != TBFM_OK) {
        std::fprintf(stderr, "Solver execution halted: %s\n", tbfm_gpu_last_error(nullptr));
        return 1;
    }
    long nmax = (long)std::atof(iniValue(ini, "N_ITERATIONS").c_str());
    if (argc >= 3) nmax = std::atol(argv[2]);
    const std::string every = iniValue(ini, "SAVE_ITERATIONS_INTERVAL");
    const long interval = every.empty() ? 0 : (long)std::atof(every.c_str());
    if (interval <= 0) {
        std::fprintf(stderr, "Solver execution halted: SAVE_ITERATIONS_INTERVAL must be a positive number\n");
        tbfm_gpu_case_destroy(c);
        return 1;
    }
    std::vector<double> hist;
    long conv = 0, done = 0;
    std::printf("|%14s|%14s|%14s|%14s|%14s|%14s|\n", "Iteration", "rms[Rho]", "rms[RhoU]", "rms[RhoV]", "rms[RhoW]", "rms[RhoE]");
    int status = 0;
    while (done < nmax && conv == 0) {
        // advance to the next output iteration (or the end) in one call
        long n = std::min(interval - done % interval, nmax - done);
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
            status = 1;
            break;
        }
        done += rows;
        if (conv) std::printf("\nConvergence reached at iteration %ld\n", conv);
        if (conv || done % interval == 0 || done == nmax)
            if (!writeVolume(c, ini, done)) { status = 1; break; }
    }
    FILE* f = std::fopen("residuals.csv", "w");
    if (f) {
        std::fprintf(f, "Rho,RhoU,RhoV,RhoW,RhoE\n");
        for (size_t r = 0; r < hist.size() / 5; ++r)
            std::fprintf(f, "%g,%g,%g,%g,%g\n", hist[5 * r], hist[5 * r + 1], hist[5 * r + 2], hist[5 * r + 3],
                         hist[5 * r + 4]);
        std::fclose(f);
        std::printf("\nLog residuals appended to residuals.csv\n\n");
    }
    if (status == 0 && done > 0 && iniBool(ini, "TURBO_OUTPUT")) {
        std::vector<double> rows(5 * done);
        if (tbfm_gpu_get_turbo_history(c, 1, done, rows.data()) == TBFM_OK) {
            const std::string hb = iniValue(ini, "HISTORY_BUFFER_SIZE");
            const long bufsize = hb.empty() ? interval : (long)std::atof(hb.c_str());
            writeTurboCsv(rows, done, conv, nmax, interval, bufsize);
        } else {
            std::fprintf(stderr, "turbo history: %s\n", tbfm_gpu_last_error(c));
        }
    }
    double turbo[10];
    if (status == 0 && tbfm_gpu_get_turbo(c, turbo) == TBFM_OK)
        std::printf("\nMass Flow [kg/s]: %.6g\nTotal Pressure Ratio [-]: %.6g\nTotal Temperature Ratio [-]: %.6g\n",
                    turbo[6], turbo[7], turbo[8]);
    tbfm_gpu_case_destroy(c);
    return status;
}
