// vlasolve with the spherical spline step on the GPU: the reference's own
// driver sequence (parameters, init or -reload, run; src/vlasolve.cpp:20-55)
// around the vlasovSolver21_hip policy of vlasovSolver21_hip.hxx.  Compiled
// against the reference's unmodified headers by integration/Makefile.
#include <stdio.h>

#include <string>

#include "openMP_interface.hxx"
#include "mpiCommunication.hxx"

#define GLOBAL_DEFINITION
#include "global.h"
#undef GLOBAL_DEFINITION

#include "paramsParser.hxx"
#include "solver_traits.hxx"
#include "vlasovSolver21_hip.hxx"

// the (r; v_r, j) spherical grid of the spline solver (src/setup.hxx:7)
typedef vlSolver::traits<1, 2, FLOAT, vlSolver::spline> sphericalTraits;
typedef vlasovSolver21_hip<sphericalTraits::gridHandlerT> hipSolverT;

int main(int argc, char **argv) {
  num_omp_threads = 1;
  verbose = 1;
  debug = 0;
  hipSolverT::mpiComT com(&argc, &argv);
  paramsParser params(argc, argv);
  verbose = params.get<>("verbose", paramsParser::defaultCategory(), verbose);
  debug = params.get<>("debug", paramsParser::defaultCategory(), debug);
  int threads = 1;
#pragma omp parallel
  threads = omp_get_num_threads();
  num_omp_threads = params.get<>("threads_per_node", paramsParser::defaultCategory(), threads);
  omp_set_num_threads(num_omp_threads);
  const std::string reload = params.get<std::string>("reload", paramsParser::defaultCategory(), "");
  hipSolverT solver(com);
  solver.init(params, reload);
  params.reportUnused();
  solver.run();
  return 0;
}
