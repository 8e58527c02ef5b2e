// Command-line driver on the GPU path, with the reference's options
// (mpm_3d/main.cpp:10-32, MPM_Example.cpp:846-863 option registration):
//   -test_number N  -size n  -threads t  -levels l  -cfl c
//   -last_frame f   -o dir   -restart frame   -d 2|3   -seed s   -staged   -v
// Runs MPM_Driver::Execute_Main_Program and writes one frame directory per
// frame; prints the wall time, substeps and particle-substeps/s at the end.
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>

#include "MPM_Driver.h"
#include "Standard_Tests.h"

using namespace mpm_host;

namespace {

struct Options {
    int dim = 3, test_number = 18, size = 32, threads = 1, levels = 1, last_frame = 1, restart = -1;
    double cfl = 0.1;
    uint64_t seed = 0;
    bool staged = false, verbose = false, fluid_source = false;
    std::string output;
};

void usage() {
    std::cerr << "usage: mpm_driver [-d 2|3] [-test_number N] [-size n] [-threads t] [-levels l] [-cfl c]\n"
                 "                  [-last_frame f] [-o dir] [-restart frame] [-seed s] [-staged] [-v]\n"
                 "                  [-fluid_source]\n";
}

bool parse(int argc, char** argv, Options& o) {
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&](const char* what) -> const char* {
            if (i + 1 >= argc) { std::cerr << what << " needs a value\n"; std::exit(2); }
            return argv[++i];
        };
        if (a == "-d") o.dim = std::atoi(next("-d"));
        else if (a == "-test_number") o.test_number = std::atoi(next("-test_number"));
        else if (a == "-size") o.size = std::atoi(next("-size"));
        else if (a == "-threads") o.threads = std::atoi(next("-threads"));
        else if (a == "-levels") o.levels = std::atoi(next("-levels"));
        else if (a == "-cfl") o.cfl = std::atof(next("-cfl"));
        else if (a == "-last_frame") o.last_frame = std::atoi(next("-last_frame"));
        else if (a == "-o") o.output = next("-o");
        else if (a == "-restart") o.restart = std::atoi(next("-restart"));
        else if (a == "-seed") o.seed = std::strtoull(next("-seed"), nullptr, 10);
        else if (a == "-staged") o.staged = true;
        else if (a == "-fluid_source") o.fluid_source = true;
        else if (a == "-v") o.verbose = true;
        else if (a == "-h" || a == "--help") { usage(); std::exit(0); }
        else { std::cerr << "unknown option " << a << "\n"; usage(); return false; }
    }
    if (o.dim != 2 && o.dim != 3) { std::cerr << "-d must be 2 or 3\n"; return false; }
    if (o.levels != 1) { std::cerr << "only -levels 1 is supported (the reference uses one level)\n"; return false; }
    return true;
}

template <int d>
int run(const Options& o) {
    using clock = std::chrono::high_resolution_clock;
    const auto tb = clock::now();
    Standard_Tests<float, d> example;
    example.test_number = o.test_number;
    example.counts = o.size;
    example.threads = o.threads;
    example.levels = o.levels;
    example.last_frame = o.last_frame;
    example.seed = o.seed;
    example.staged = o.staged;
    example.fluid_source_on = o.fluid_source;
    example.cfl = (float)o.cfl;
    example.Parse_Options();
    if (!o.output.empty()) example.output_directory = o.output;
    if (o.restart >= 0) { example.restart = true; example.restart_frame = o.restart; }
    MPM_Driver<float, d> driver(example);
    driver.verbose = o.verbose;
    try {
        driver.Initialize();
        driver.Simulate_To_Frame(example.last_frame);
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    const double secs = std::chrono::duration<double>(clock::now() - tb).count();
    std::cout << "Total duration: " << secs << "\n"
              << "substeps: " << driver.substeps_total << "  particles: " << example.particles.size() << "\n";
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    if (!parse(argc, argv, o)) return 2;
    return o.dim == 3 ? run<3>(o) : run<2>(o);
}
