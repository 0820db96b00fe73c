// fluidsim_home -- HOME (LBM3D_Home) headless run (main_home.cpp:13-20).
#include <cstdio>
#include <exception>

#include "sim_runner.hpp"

int main(int argc, char** argv) {
  DomainConfig dom;
  ObstacleConfig obs;
  RunConfig run;
  parse_common_args(argc, argv, dom, obs, run, nullptr);
  try {
    return run_single_solver<LBM3D_Home>(dom, obs, run, "home");
  } catch (const std::exception& e) {
    std::fprintf(stderr, "[error] %s\n", e.what());
    return 1;
  }
}
