// fluidsim_hybrid -- Hybrid(B0) (LBM3D_Hybrid) headless run (main_hybrid.cpp:18-28).
#include <cstdio>
#include <exception>

#include "sim_runner.hpp"

int main(int argc, char** argv) {
  DomainConfig dom;
  ObstacleConfig obs;
  RunConfig run;
  HybridConfig hyb;
  parse_common_args(argc, argv, dom, obs, run, &hyb);
  try {
    return run_hybrid_b0<LBM3D_Hybrid>(dom, obs, run, hyb);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "[error] %s\n", e.what());
    return 1;
  }
}
