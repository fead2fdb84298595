// mph_run.cpp -- command-line host over libmph_hip.so with the reference
// executable's interface (source/main.cpp:313-521):
//
//   mph_run data grid prof_pattern vtk_pattern log [threads] [--dim 2|3]
//           [--rel-residual R] [--device D] [--mixed] [--factored]
//           [--convergence-check]
//
// It reads the same *.data / *.grid files, runs the same time loop, writes the same
// *.prof (every OutputInterval) and *.vtk (every VtkOutputInterval) files named by
// the printf patterns and the step number, and logs the CR summary line of every
// step in the reference's format (main.cpp:4381).  `threads` is accepted and ignored
// (the reference's OpenMP thread count).  --mixed and --factored select the
// mixed-precision solve and the factored particle matrix (mph_params); without
// them the run is the reference's arithmetic.  --convergence-check logs what the
// reference built with CONVERGENCE_CHECK logs: a line per CR iteration
// (main.cpp:4350) before the summary line and the true residual's line after it
// (main.cpp:4396).
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mph_abi.h"

namespace {

FILE* g_log = nullptr;

void logf(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void logf(const char* fmt, ...) {
    va_list a, b;
    va_start(a, fmt);
    va_copy(b, a);
    if (g_log) std::vfprintf(g_log, fmt, a);
    std::vfprintf(stderr, fmt, b);
    va_end(a);
    va_end(b);
    if (g_log) std::fflush(g_log);
}

int die(const char* what) {
    logf("error: %s: %s\n", what, mph_last_error());
    return 1;
}

double scalar(mph_scene* s, const char* n) {
    double v = 0;
    mph_scene_scalar(s, n, &v);
    return v;
}

}  // namespace

int main(int argc, char** argv) {
    std::vector<std::string> pos;
    int dim = 3, device = 0, mixedp = 0, fact = 0, check = 0;
    double relres = 0;
#ifdef MPH_RUN_2D
    dim = 2;
#endif
    for (int k = 1; k < argc; ++k) {
        if (!std::strcmp(argv[k], "--dim") && k + 1 < argc) dim = std::atoi(argv[++k]);
        else if (!std::strcmp(argv[k], "--rel-residual") && k + 1 < argc) relres = std::atof(argv[++k]);
        else if (!std::strcmp(argv[k], "--device") && k + 1 < argc) device = std::atoi(argv[++k]);
        else if (!std::strcmp(argv[k], "--mixed")) mixedp = 1;
        else if (!std::strcmp(argv[k], "--convergence-check")) check = 1;
        else if (!std::strcmp(argv[k], "--factored")) fact = 1;
        else pos.push_back(argv[k]);
    }
    // defaults of main.cpp:58-62
    const std::string data = pos.size() > 0 ? pos[0] : "sample.data";
    const std::string grid = pos.size() > 1 ? pos[1] : "sample.grid";
    const std::string prof = pos.size() > 2 ? pos[2] : "sample%03d.prof";
    const std::string vtk = pos.size() > 3 ? pos[3] : "sample%03d.vtk";
    const std::string logn = pos.size() > 4 ? pos[4] : "sample.log";
    g_log = std::fopen(logn.c_str(), "w");

    mph_params p;
    mph_params_default(&p);
    p.dim = dim;
    if (relres > 0) p.cr_tolerance = relres * relres;
    if (mixedp) p.precision = MPH_PRECISION_MIXED;
    if (fact) p.operator_form = MPH_OPERATOR_FACTORED;
    if (mph_read_data_file(data.c_str(), &p)) return die("reading data file");
    int np = 0;
    if (mph_read_grid_header(grid.c_str(), &np)) return die("reading grid header");
    std::vector<int> type(np);
    std::vector<double> xyz(3 * (size_t)np), vel(3 * (size_t)np);
    mph_grid g{};
    g.np = np;
    if (mph_read_grid_file(grid.c_str(), np, &g.time, &g.spacing, g.domain_min, g.domain_max, type.data(),
                           xyz.data(), vel.data()))
        return die("reading grid file");
    g.type = type.data();
    g.pos = xyz.data();
    g.vel = vel.data();

    mph_scene* s = nullptr;
    if (mph_scene_create(&p, &g, device, &s)) return die("creating the scene");
    if (check && mph_scene_set_option(s, "ConvergenceCheck", 1)) return die("--convergence-check");
    std::vector<double> hist;
    logf("N0a = %e, N0p = %e\n", scalar(s, "N0a"), scalar(s, "N0p"));
    logf("CellCount[DIM]: %d %d %d\n", (int)scalar(s, "CellCount0"), (int)scalar(s, "CellCount1"),
         (int)scalar(s, "CellCount2"));
    // main.cpp:374 writes output.vtk here with uninitialised diagnostic fields; the
    // arrays of this library start at zero instead
    if (mph_scene_write_vtk(s, "output.vtk")) return die("writing output.vtk");

    const double dt = p.dt;
    int iStep = (int)scalar(s, "iStep");
    const auto t0 = std::chrono::steady_clock::now();
    char name[1024];
    for (;;) {
        const double time = scalar(s, "Time");
        if (!(time < p.end_time + 1.0e-5 * dt)) break;
        if (time + 1.0e-5 * dt >= scalar(s, "OutputNext")) {
            std::snprintf(name, sizeof name, prof.c_str(), iStep);
            if (mph_scene_write_prof(s, name)) return die("writing prof file");
            logf("@ Prof Output Time : %e\n", time);
        }
        mph_step_stats st;
        if (mph_scene_step(s, 1, &st)) return die("time step");
        if (check) {  // main.cpp:4350, one line per iteration that did not stop
            size_t bytes = 0;
            if (mph_scene_field_size(s, "ConvergenceHistory", &bytes, nullptr)) return die("convergence record");
            hist.resize(bytes / sizeof(double));
            if (mph_scene_read(s, "ConvergenceHistory", hist.data(), bytes)) return die("convergence record");
            for (size_t k = 0; 2 * k + 1 < hist.size(); ++k)
                logf("line:%d, iter,=%d, rr0=,%e, rr=,%e, rs0=,%e, rs=,%e\n", 4350, (int)k, st.rr0, hist[2 * k],
                     st.rs0, hist[2 * k + 1]);
        }
        logf("line:%d, iter,=%d, rr0=,%e, rr=,%e, rs0=,%e, rs=,%e\n", 4381, st.cr_iterations, st.rr0, st.rr, st.rs0,
             st.rs);
        if (check)  // main.cpp:4396, the true residual
            logf("line:%d, iter,=%d, rr0=,%e, rr=,%e, rs0=,%e, rs=,%e\n", 4396, st.cr_iterations, st.rr0,
                 scalar(s, "CheckRR"), st.rs0, scalar(s, "CheckRS"));
        if (st.vtk_steps) {
            std::snprintf(name, sizeof name, vtk.c_str(), iStep);
            if (mph_scene_write_vtk(s, name)) return die("writing vtk file");
            logf("@ Vtk Output Time : %e\n", time);
        }
        iStep++;
    }
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    logf("main loop: %lf [wall sec]\n", wall);
    mph_scene_destroy(s);
    if (g_log) std::fclose(g_log);
    return 0;
}
