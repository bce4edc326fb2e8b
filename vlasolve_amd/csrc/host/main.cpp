// vlasolve_hip: command-line counterpart of the reference's vlasolve
// (src/vlasolve.cpp:20-55) running every time step on the GPU.
#include <cstdio>
#include <cstdlib>

#include "vls_host.hpp"

int main(int argc, char **argv) {
  try {
    vlh::Params prm(argc, argv);
    vlh::verbose = prm.get<int>("verbose", "default", 1);
    prm.get<int>("debug", "default", 0);
    prm.get<int>("threads_per_node", "default", 1);
    std::string reload = prm.get<std::string>("reload", "default", "");
    vlh::Solver21 solver;
    solver.init(prm, reload);
    std::string unused = prm.reportUnused();
    if (!unused.empty()) printf("%s", unused.c_str());
    solver.run();
  } catch (const vlh::Failure &e) {
    // the reference prints and exits with -1 (e.g. src/vlasovSolver.hxx:116-120)
    fprintf(stderr, "%s\n", e.what());
    return 255;
  }
  return 0;
}
