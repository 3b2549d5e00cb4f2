// bal_cli.cpp -- a C++ caller of the drop-in API, built against include/gba
// and linked with libgba_amd.so exactly as a program of the reference would
// be (its benchmark driver benchmarks/bench_bal.cpp:65-136 is the model: same
// options, same gates, same JSON keys).  The tests compile it and compare its
// output with the Python binding's.
//
//   bal_cli FILE [--iterations=N] [--lambda=X] [--json=OUT] [--verbose]
//
// Exit status: 0 solved; 1 usage or I/O error; 2 the solver's initial cost
// disagrees with the reader's model (relative 1e-9); 3 the solve skipped
// non-finite observations.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <string>

#include "gba/bal_problem.hpp"
#include "gba/lm.hpp"

namespace {

struct Options {
    std::string file, json;
    gba::LMOptions lm;
};

bool starts(const std::string& s, const char* key, std::string* rest) {
    const std::string k(key);
    if (s.compare(0, k.size(), k) != 0) return false;
    *rest = s.substr(k.size());
    return true;
}

bool parse(int argc, char** argv, Options* o) {
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        std::string v;
        if (a == "--verbose") o->lm.verbose = true;
        else if (starts(a, "--iterations=", &v)) o->lm.max_iterations = std::atoi(v.c_str());
        else if (starts(a, "--lambda=", &v)) o->lm.initial_lambda = std::strtod(v.c_str(), nullptr);
        else if (starts(a, "--json=", &v)) o->json = v;
        else if (a.rfind("--", 0) != 0 && o->file.empty()) o->file = a;
        else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return false;
        }
    }
    return !o->file.empty();
}

std::string device_name() {
    int dev = 0;
    hipDeviceProp_t prop{};
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
    // the marketing name is empty where the driver's name table is missing
    return prop.name[0] ? prop.name : prop.gcnArchName;
}

}  // namespace

int main(int argc, char** argv) {
    Options opt;
    if (!parse(argc, argv, &opt)) {
        std::fprintf(stderr, "usage: bal_cli FILE [--iterations=N] [--lambda=X] [--json=OUT] [--verbose]\n");
        return 1;
    }
    try {
        gba::BALProblem prob = gba::load_bal(opt.file);
        const double reader_cost = gba::total_squared_error(prob) / 2.0;
        const gba::LMResult res = gba::solve_bal_gpu(prob, opt.lm);
        if (std::fabs(res.initial_cost - reader_cost) > 1e-9 * std::fmax(1.0, std::fabs(reader_cost))) {
            std::fprintf(stderr, "initial cost %.17g differs from the reader's %.17g\n", res.initial_cost,
                         reader_cost);
            return 2;
        }
        if (res.anomalies != 0) {
            std::fprintf(stderr, "%llu non-finite observations were skipped; no result reported\n",
                         res.anomalies);
            return 3;
        }
        const double per_it = res.iterations > 0 ? res.total_time_s / res.iterations : 0.0;
        char out[2048];
        std::snprintf(out, sizeof out,
                      "{\"file\": \"%s\", \"solver\": \"gba-gpu-lm\", \"loss\": \"none\", "
                      "\"num_cameras\": %d, \"num_points\": %d, \"num_observations\": %d, "
                      "\"initial_cost\": %.17g, \"final_cost\": %.17g, \"iterations\": %d, "
                      "\"total_time_s\": %.6f, \"per_iteration_s\": %.6f, \"termination\": \"%s\", "
                      "\"device\": \"%s\"}",
                      opt.file.c_str(), prob.num_cameras, prob.num_points, prob.num_observations, res.initial_cost,
                      res.final_cost, res.iterations, res.total_time_s, per_it, res.termination.c_str(),
                      device_name().c_str());
        std::puts(out);
        if (!opt.json.empty()) {
            FILE* fp = std::fopen(opt.json.c_str(), "w");
            if (!fp || std::fprintf(fp, "%s\n", out) < 0 || std::fclose(fp) != 0) {
                std::fprintf(stderr, "cannot write %s\n", opt.json.c_str());
                return 1;
            }
        }
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
