// fluidsim_legacy -- BGK (LBM3D_Legacy) headless run (main_legacy.cpp:14-23).
#include <cstdio>
#include <exception>

#include "sim_runner.hpp"

int main(int argc, char** argv) {
  DomainConfig dom;
  ObstacleConfig obs;
  RunConfig run;
  parse_common_args(argc, argv, dom, obs, run, nullptr);
  try {
    return run_single_solver<LBM3D_Legacy>(dom, obs, run, "legacy");
  } catch (const std::exception& e) {
    std::fprintf(stderr, "[error] %s\n", e.what());
    return 1;
  }
}
