// chaos_regression -- timing harness in the manner of the reference's
// src/main.cu:26-624 (runBifurcationPerformanceTests, runLLEPerformanceTests,
// runBasinsPerformanceTests): for each module, run the file-writing C++
// entry point over a resolution sweep at a fixed modelling time and over a
// modelling-time sweep at a fixed resolution, and append one row per run to
// <results>/<module>_performance_results.csv:
//
//     Test,Parameter,Value,Library,ExecutionTime_ms[,Dimension]
//
// (Test = module, Parameter = Resolution | ModelingTime, Value = the swept
// number; the reference writes the same header but drops the Test field from
// its rows, src/main.cu:79,457,511; the plotting scripts select on Parameter,
// Value, Library and ExecutionTime_ms, scripts/plot_bifurcation_performance.py).
//
// Libraries:
//   hip-per-job  the reference's C++ API (Bifurcation::bifurcation2D,
//                LLE::LLE2D, Basins::basinsOfAttraction_2), one synchronous
//                call per run, map files included, as src/main.cu times it;
//   hip-batched  the same run as a job of the batched engine (submit, drain,
//                chaos_server_write_result), and, per sweep, every run of the
//                sweep submitted at once (Parameter ResolutionSweep /
//                ModelingTimeSweep, Value = jobs, time = makespan; the
//                hip-per-job row of the same name is the sum of its runs).
// The two libraries' map files are compared byte for byte; a difference is
// reported and makes the exit status 3.
//
// Parameter sets (--params):
//   rk4        (default) the RK4 systems the engine runs: LLE on Rossler
//              (a, c) in [0.05,0.35] x [2,10], bifurcation on Lorenz
//              (rho, beta) in [20,200] x [0.5,4], basins on Lorenz ICs
//              (x0, y0) in [-20,20]^2, z0 = 25, rho = 24.5;
//   reference  the sets of src/main.cu:449-624 (Chameleon for bifurcation2D
//              and LLE2D, the BASINS model's parameters for basins); the
//              engine runs RK4 systems only, so these are hip-per-job only.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "LLEHost.h"
#include "basinsHOST.h"
#include "bifurcationHOST.h"
#include "chaoslibrary.h"

namespace {

enum Module { BIF = 0, LLE_M = 1, BASINS = 2 };

struct Suite {
    Module module;
    const char* name;      // Test column and file stem
    const char* results;   // results file name
    int system;
    double CT, TT, NT, h, eps, maxValue;
    int preScaller, writableVar;
    std::vector<double> init, values, ranges;
    int idx[2];
};

// src/main.cu:449-485 (bifurcation), 219-246 (LLE), 26-47 (basins).
std::vector<Suite> reference_suites()
{
    Suite b{BIF, "Bifurcation", "bifurcation_performance_results.csv", CHAOS_SYS_CHAMELEON,
            400, 2000, 0, 0.01, 0.001, 10000, 1, 0,
            {3, 3, 0}, {0.5, 3, 0, 2, 1, 18, 1}, {0, 5, 0, 20}, {3, 5}};
    Suite l{LLE_M, "LLE", "lle_performance_results.csv", CHAOS_SYS_CHAMELEON,
            400, 2000, 0.5, 0.01, 1e-6, 10000, 1, 0,
            {3, 3, 0}, {0.5, 3, 0, 2, 1, 18, 1}, {0, 5, 0, 20}, {3, 5}};
    Suite s{BASINS, "Basins", "basins_performance_results.csv", CHAOS_SYS_BASINS,
            500, 4000, 0, 0.01, 0.05, 100000000, 1, 1,
            {0, 0, 0}, {0.5, 0.1, 1.4, 15.552, 2}, {-6, 6, -6, 6}, {0, 1}};
    return {b, l, s};
}

// The same sweeps (CT, TT, eps, maxValue, preScaller of src/main.cu) on the
// RK4 systems: Lorenz peak-count plane, Rossler LLE plane, Lorenz IC plane.
std::vector<Suite> rk4_suites()
{
    const double beta = 8.0 / 3.0;
    Suite b{BIF, "Bifurcation", "bifurcation_performance_results.csv", CHAOS_SYS_LORENZ_RK4,
            400, 2000, 0, 0.01, 0.001, 10000, 1, 0,
            {1, 1, 1}, {10, 28, beta}, {20, 200, 0.5, 4}, {1, 2}};
    Suite l{LLE_M, "LLE", "lle_performance_results.csv", CHAOS_SYS_ROSSLER_RK4,
            400, 2000, 0.5, 0.01, 1e-6, 10000, 1, 0,
            {0.1, 0.1, 0.1}, {0.2, 0.2, 5.7}, {0.05, 0.35, 2, 10}, {0, 2}};
    Suite s{BASINS, "Basins", "basins_performance_results.csv", CHAOS_SYS_LORENZ_RK4,
            500, 4000, 0, 0.01, 0.05, 100000000, 1, 0,
            {0, 0, 25}, {10, 24.5, beta}, {-20, 20, -20, 20}, {0, 1}};
    return {b, l, s};
}

struct Config {
    std::string params = "rk4";
    std::vector<std::string> suites{"bifurcation", "lle", "basins"};
    bool per_job = true, batched = true;
    std::vector<int> resolutions{100, 200, 300, 400, 500, 600};
    std::vector<int> times{1000, 2000, 3000, 4000, 5000, 6000};
    int fixed_resolution = 300;
    double fixed_time = -1;  // < 0: the suite's own (2000 or 4000)
    double ct = -1;          // < 0: the suite's own recording time
    int semantics = CHAOS_SEM_REFERENCE;
    std::string out = "workspace/regression";
    std::string results = ".";
};

// One run of a sweep.
struct Run {
    const char* parameter;  // Resolution | ModelingTime
    int value;
    int nPts;
    double TT;
};

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

std::vector<int> int_list(const char* s)
{
    std::vector<int> v;
    std::stringstream ss(s);
    std::string t;
    while (std::getline(ss, t, ','))
        if (!t.empty()) v.push_back(std::atoi(t.c_str()));
    return v;
}

std::vector<std::string> str_list(const char* s)
{
    std::vector<std::string> v;
    std::stringstream ss(s);
    std::string t;
    while (std::getline(ss, t, ','))
        if (!t.empty()) v.push_back(t);
    return v;
}

void usage()
{
    std::fprintf(stderr,
                 "usage: chaos_regression [--params rk4|reference] [--suites bifurcation,lle,basins]\n"
                 "       [--libraries hip-per-job,hip-batched] [--resolutions 100,...] [--times 1000,...]\n"
                 "       [--fixed-resolution N] [--fixed-time T] [--ct T] [--semantics reference|config]\n"
                 "       [--out DIR] [--results DIR]\n");
}

bool parse(int argc, char** argv, Config& c)
{
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-h" || a == "--help") return false;
        if (i + 1 >= argc) {
            std::fprintf(stderr, "missing value for %s\n", a.c_str());
            return false;
        }
        const char* v = argv[++i];
        if (a == "--params") c.params = v;
        else if (a == "--suites") c.suites = str_list(v);
        else if (a == "--libraries") {
            c.per_job = c.batched = false;
            for (const auto& l : str_list(v)) {
                if (l == "hip-per-job") c.per_job = true;
                else if (l == "hip-batched") c.batched = true;
                else {
                    std::fprintf(stderr, "unknown library %s\n", l.c_str());
                    return false;
                }
            }
        } else if (a == "--resolutions") c.resolutions = int_list(v);
        else if (a == "--times") c.times = int_list(v);
        else if (a == "--fixed-resolution") c.fixed_resolution = std::atoi(v);
        else if (a == "--fixed-time") c.fixed_time = std::atof(v);
        else if (a == "--ct") c.ct = std::atof(v);
        else if (a == "--semantics") c.semantics = std::string(v) == "config" ? CHAOS_SEM_CONFIG : CHAOS_SEM_REFERENCE;
        else if (a == "--out") c.out = v;
        else if (a == "--results") c.results = v;
        else {
            std::fprintf(stderr, "unknown option %s\n", a.c_str());
            return false;
        }
    }
    if (c.params != "rk4" && c.params != "reference") {
        std::fprintf(stderr, "--params must be rk4 or reference\n");
        return false;
    }
    for (int r : c.resolutions)
        if (r < 2) {
            std::fprintf(stderr, "resolutions must be >= 2\n");
            return false;
        }
    if (c.fixed_resolution < 2) {
        std::fprintf(stderr, "--fixed-resolution must be >= 2\n");
        return false;
    }
    return true;
}

bool engine_runs(int system) { return system == CHAOS_SYS_ROSSLER_RK4 || system == CHAOS_SYS_LORENZ_RK4; }

std::string map_path(const Config& c, const Suite& s, const char* lib, const Run& r)
{
    return c.out + "/" + s.name + "_" + lib + "_" + r.parameter + "_" + std::to_string(r.value) + ".csv";
}

// The files a run writes (basins: the labels file and its four/five side files).
std::vector<std::string> files_of(const Suite& s, const std::string& path, int semantics)
{
    std::vector<std::string> f{path};
    if (s.module == BASINS)
        for (int k = 1; k <= (semantics == CHAOS_SEM_REFERENCE ? 3 : 4); ++k)
            f.push_back(path + "_" + std::to_string(k) + ".csv");
    return f;
}

bool same_bytes(const std::string& a, const std::string& b)
{
    std::ifstream fa(a, std::ios::binary), fb(b, std::ios::binary);
    if (!fa || !fb) return false;
    std::stringstream sa, sb;
    sa << fa.rdbuf();
    sb << fb.rdbuf();
    return sa.str() == sb.str();
}

// hip-per-job: the reference's own C++ call (src/main.cu:494-507, 252-261, 55-69).
void call_api(const Suite& s, const Run& r, double CT, const std::string& path)
{
    const int nic = (int)s.init.size(), nv = (int)s.values.size();
    switch (s.module) {
    case BIF:
        Bifurcation::bifurcation2D(CT, r.nPts, s.h, nic, s.init.data(), s.ranges.data(), s.idx, s.writableVar,
                                   s.maxValue, r.TT, s.values.data(), nv, s.preScaller, s.eps, path);
        break;
    case LLE_M:
        LLE::LLE2D(CT, s.NT, r.nPts, s.h, s.eps, s.init.data(), nic, s.ranges.data(), s.idx, s.writableVar,
                   s.maxValue, r.TT, s.values.data(), nv, path);
        break;
    case BASINS:
        Basins::basinsOfAttraction_2(CT, r.nPts, s.h, nic, s.init.data(), s.ranges.data(), s.idx, s.writableVar,
                                     s.maxValue, r.TT, s.values.data(), nv, s.preScaller, s.eps, path);
        break;
    }
}

chaos_job make_job(const Suite& s, const Run& r, double CT, const chaos_options& o)
{
    chaos_job j;
    std::memset(&j, 0, sizeof j);
    j.kind = s.module == BIF ? CHAOS_JOB_BIF2D : s.module == LLE_M ? CHAOS_JOB_LLE2D : CHAOS_JOB_BASINS2D;
    j.system = s.system;
    j.semantics = o.semantics;
    j.nPts = r.nPts;
    j.tMax = CT;
    j.NT = s.NT;
    j.h = s.h;
    j.eps = s.eps;
    j.maxValue = s.maxValue;
    j.transientTime = r.TT;
    j.preScaller = s.preScaller;
    j.writableVar = s.writableVar;
    j.min_pts = o.min_pts;
    j.amountOfInitialConditions = (int)s.init.size();
    j.seed = o.seed;
    std::copy(s.init.begin(), s.init.end(), j.initialConditions);
    std::copy(s.values.begin(), s.values.end(), j.values);
    j.amountOfValues = (int)s.values.size();
    j.indicesOfMutVars[0] = s.idx[0];
    j.indicesOfMutVars[1] = s.idx[1];
    std::copy(s.ranges.begin(), s.ranges.end(), j.ranges);
    return j;
}

void check(int rc, const char* what)
{
    if (rc != CHAOS_OK) throw std::runtime_error(std::string(what) + ": " + chaos_last_error());
}

// hip-batched: submit every run together, drain, write each job's files.
// Returns the makespan; per-run times are submit -> files written.
double run_batched(void* srv, const Suite& s, const std::vector<Run>& runs, double CT, const chaos_options& o,
                   const Config& c, std::vector<double>& ms, std::vector<double>& steps)
{
    std::vector<long long> ids(runs.size());
    const double t0 = now_ms();
    for (size_t k = 0; k < runs.size(); ++k) {
        chaos_job j = make_job(s, runs[k], CT, o);
        check(chaos_server_submit(srv, &j, &ids[k]), "chaos_server_submit");
    }
    check(chaos_server_drain(srv), "chaos_server_drain");
    for (size_t k = 0; k < runs.size(); ++k) {
        check(chaos_server_write_result(srv, ids[k], map_path(c, s, "hip-batched", runs[k]).c_str(), 0),
              "chaos_server_write_result");
        ms[k] = now_ms() - t0;
        chaos_job_info info;
        check(chaos_server_job_info(srv, ids[k], &info), "chaos_server_job_info");
        steps[k] = info.rk4_steps;
        check(chaos_server_release(srv, ids[k]), "chaos_server_release");
    }
    return now_ms() - t0;
}

// Stats of this thread's C++ API calls (chaos_set_thread_stats).
chaos_stats g_stats;

struct Totals {
    int rows = 0, errors = 0, compared = 0, mismatched = 0;
};

void run_suite(const Config& c, const Suite& base, void* srv, Totals& T)
{
    Suite s = base;
    if (c.ct > 0) s.CT = c.ct;
    if (c.fixed_time >= 0) s.TT = c.fixed_time;
    chaos_stats& stats = g_stats;
    chaos_options o;
    chaos_options_init(&o);
    o.system = s.system;
    o.semantics = c.semantics;
    check(chaos_set_default_options(&o), "chaos_set_default_options");

    // the suite's inputs, one line, for tools that re-run the same cases
    auto list = [](const std::vector<double>& v) {
        std::string t;
        char b[40];
        for (size_t k = 0; k < v.size(); ++k) {
            std::snprintf(b, sizeof b, "%s%.17g", k ? "," : "", v[k]);
            t += b;
        }
        return t;
    };
    std::printf("SUITE %s system=%d semantics=%d CT=%.17g TT=%.17g NT=%.17g h=%.17g eps=%.17g maxValue=%.17g "
                "preScaller=%d writableVar=%d seed=%llu init=%s values=%s ranges=%s idx=%d,%d\n",
                s.name, s.system, c.semantics, s.CT, s.TT, s.NT, s.h, s.eps, s.maxValue, s.preScaller,
                s.writableVar, (unsigned long long)o.seed, list(s.init).c_str(), list(s.values).c_str(),
                list(s.ranges).c_str(), s.idx[0], s.idx[1]);

    const std::string rpath = c.results + "/" + s.results;
    std::ofstream out(rpath);
    if (!out) throw std::runtime_error("cannot open " + rpath);
    const bool bif = s.module == BIF;
    out << "Test,Parameter,Value,Library,ExecutionTime_ms" << (bif ? ",Dimension" : "") << "\n";
    // one CSV row, and a "RUN ..." line on stdout with the RK4 steps the run
    // integrated (all points; 0 when it failed)
    auto row = [&](const char* parameter, int value, const char* lib, double ms, bool ok, double steps) {
        std::printf("RUN %s %s %d %s %.3f %.6e %d\n", s.name, parameter, value, lib, ms, ok ? steps : 0.0,
                    ok ? 1 : 0);
        char buf[64];
        std::snprintf(buf, sizeof buf, "%.3f", ms);
        out << s.name << "," << parameter << "," << value << "," << lib << "," << (ok ? buf : "ERROR")
            << (bif ? ",2D" : "") << "\n";
        out.flush();
        ++T.rows;
        if (!ok) ++T.errors;
    };
    const bool batched = c.batched && srv && engine_runs(s.system);
    if (c.batched && !engine_runs(s.system))
        std::printf("%s: the batched engine runs the RK4 systems only; hip-batched rows skipped\n", s.name);

    std::vector<std::vector<Run>> sweeps(2);
    for (int res : c.resolutions) sweeps[0].push_back({"Resolution", res, res, s.TT});
    for (int t : c.times) sweeps[1].push_back({"ModelingTime", t, c.fixed_resolution, (double)t});
    const char* sweep_names[2] = {"ResolutionSweep", "ModelingTimeSweep"};

    for (int w = 0; w < 2; ++w) {
        const auto& runs = sweeps[w];
        if (runs.empty()) continue;
        double per_job_sum = 0, per_job_steps = 0;
        bool per_job_ok = true;
        for (const Run& r : runs) {
            if (c.per_job) {
                std::printf("%s %s=%d hip-per-job ...\n", s.name, r.parameter, r.value);
                std::fflush(stdout);
                bool ok = true;
                std::memset(&stats, 0, sizeof stats);
                const double t0 = now_ms();
                try {
                    call_api(s, r, s.CT, map_path(c, s, "hip-per-job", r));
                } catch (const std::exception& e) {
                    std::fprintf(stderr, "  error: %s\n", e.what());
                    ok = per_job_ok = false;
                }
                const double ms = now_ms() - t0;
                per_job_sum += ms;
                per_job_steps += stats.rk4_steps;
                row(r.parameter, r.value, "hip-per-job", ms, ok, stats.rk4_steps);
                std::printf("  %.1f ms\n", ms);
            }
            if (batched) {
                std::printf("%s %s=%d hip-batched ...\n", s.name, r.parameter, r.value);
                std::fflush(stdout);
                std::vector<double> ms(1), steps(1);
                bool ok = true;
                try {
                    run_batched(srv, s, {r}, s.CT, o, c, ms, steps);
                } catch (const std::exception& e) {
                    std::fprintf(stderr, "  error: %s\n", e.what());
                    ok = false;
                }
                row(r.parameter, r.value, "hip-batched", ms[0], ok, steps[0]);
                std::printf("  %.1f ms\n", ms[0]);
                if (ok && c.per_job) {
                    const auto a = files_of(s, map_path(c, s, "hip-per-job", r), c.semantics);
                    const auto b = files_of(s, map_path(c, s, "hip-batched", r), c.semantics);
                    for (size_t k = 0; k < a.size(); ++k) {
                        ++T.compared;
                        if (!same_bytes(a[k], b[k])) {
                            ++T.mismatched;
                            std::printf("  MISMATCH %s vs %s\n", a[k].c_str(), b[k].c_str());
                        }
                    }
                }
            }
        }
        if (runs.size() > 1) {
            if (c.per_job)
                row(sweep_names[w], (int)runs.size(), "hip-per-job", per_job_sum, per_job_ok, per_job_steps);
            if (batched) {
                std::printf("%s %s hip-batched (%zu jobs at once) ...\n", s.name, sweep_names[w], runs.size());
                std::fflush(stdout);
                std::vector<double> ms(runs.size()), steps(runs.size());
                bool ok = true;
                double span = 0;
                try {
                    span = run_batched(srv, s, runs, s.CT, o, c, ms, steps);
                } catch (const std::exception& e) {
                    std::fprintf(stderr, "  error: %s\n", e.what());
                    ok = false;
                }
                double total = 0;
                for (double x : steps) total += x;
                row(sweep_names[w], (int)runs.size(), "hip-batched", span, ok, total);
                std::printf("  makespan %.1f ms (per-job sum %.1f ms)\n", span, per_job_sum);
            }
        }
    }
    std::printf("wrote %s\n", rpath.c_str());
}

}  // namespace

int main(int argc, char** argv)
{
    Config c;
    if (!parse(argc, argv, c)) {
        usage();
        return 2;
    }
    int ndev = 0;
    if (chaos_device_count(&ndev) != CHAOS_OK || ndev < 1) {
        std::fprintf(stderr, "chaos_regression: no HIP device\n");
        return 1;
    }
    const std::string mk = "mkdir -p '" + c.out + "' '" + c.results + "'";
    if (std::system(mk.c_str()) != 0) {
        std::fprintf(stderr, "chaos_regression: cannot create %s or %s\n", c.out.c_str(), c.results.c_str());
        return 1;
    }
    const auto all = c.params == "rk4" ? rk4_suites() : reference_suites();
    void* srv = nullptr;
    Totals T;
    try {
        check(chaos_set_thread_stats(&g_stats), "chaos_set_thread_stats");
        if (c.batched) check(chaos_server_create(0, nullptr, &srv), "chaos_server_create");
        for (const auto& name : c.suites) {
            const Suite* s = nullptr;
            for (const auto& x : all)
                if ((name == "bifurcation" && x.module == BIF) || (name == "lle" && x.module == LLE_M) ||
                    (name == "basins" && x.module == BASINS))
                    s = &x;
            if (!s) throw std::runtime_error("unknown suite " + name);
            run_suite(c, *s, srv, T);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "chaos_regression: %s\n", e.what());
        if (srv) chaos_server_destroy(srv);
        return 1;
    }
    if (srv) chaos_server_destroy(srv);
    std::printf("rows %d, errors %d, files compared %d, mismatched %d\n", T.rows, T.errors, T.compared,
                T.mismatched);
    return T.mismatched ? 3 : T.errors ? 4 : 0;
}
