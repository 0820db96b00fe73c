// fluidsim_bench -- lattice-update throughput of the three methods.
//
// The reference's fluidsim_bench is a stub that prints "bench stub" and
// exits 0 (main_bench_experiments.cpp:5-8).  This one measures: with no
// arguments it times a 64^3 periodic Taylor-Green vortex for BGK, HOME and
// Hybrid and exits 0.  It accepts every fluidsim_* flag (domain, --stl,
// --walls yz, --inlet-u, --outlet zero-gradient, --init, --hybrid-band)
// plus
//   --method bgk|home|hybrid|all   (default all)
//   --warmup W                     untimed steps (default 20)
// and --steps K timed steps (default 200).  One line per method:
//   [bench] method=... Nx=.. Ny=.. Nz=.. steps=K ms_per_step=.. MLUPS=..
//           B_per_LUP=.. GBps=.. (algorithmic bytes: 152 BGK/Hybrid, 80 HOME)
#include <chrono>
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>

#include "sim_runner.hpp"

namespace {

template <class Solver>
double time_steps(Solver& s, int warmup, int steps) {
  fs_throw_if(fs_solver_step(s.handle(), warmup), "warmup");
  const auto t0 = std::chrono::steady_clock::now();
  fs_throw_if(fs_solver_step_async(s.handle(), steps), "step");
  fs_throw_if(fs_solver_synchronize(s.handle()), "sync");
  const auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

template <class Solver>
void bench_one(const char* name, double bytes_per_lup, const DomainConfig& dom,
               const ObstacleConfig& obs, const RunConfig& run, const HybridConfig* hyb,
               int warmup) {
  ObstacleData obstacle = build_obstacle(dom, obs);
  const Domain d = to_domain(dom);
  Solver s;
  s.init(d);
  std::vector<unsigned char> band;
  if (hyb) {
    bool special = false;
    build_flags(dom, run, obstacle.solidMask, special);
    band.resize(obstacle.solidMask.size());
    fs_throw_if(fs_band(obstacle.solidMask.data(), d.Nx, d.Ny, d.Nz, hyb->bandD0, band.data()),
                "band");
  }
  setup_solver(s, dom, run, obstacle.solidMask, hyb ? &band : nullptr);
  const double sec = time_steps(s, warmup, run.steps);
  const double cells = double(d.Nx) * d.Ny * d.Nz;
  const double lups = cells * run.steps / sec;
  std::printf(
      "[bench] method=%s Nx=%d Ny=%d Nz=%d steps=%d ms_per_step=%.4f MLUPS=%.1f B_per_LUP=%.0f "
      "GBps=%.1f\n",
      name, d.Nx, d.Ny, d.Nz, run.steps, 1e3 * sec / run.steps, lups / 1e6, bytes_per_lup,
      lups * bytes_per_lup / 1e9);
}

}  // namespace

int main(int argc, char** argv) {
  DomainConfig dom;
  ObstacleConfig obs;
  RunConfig run;
  HybridConfig hyb;
  dom.Nx = dom.Ny = dom.Nz = 64;
  dom.tau = 0.6f;
  run.initMode = "tgv3d";
  run.steps = 200;
  std::string method = "all";
  int warmup = 20;
  bool stepsGiven = false;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--method") && i + 1 < argc) method = argv[++i];
    else if (!std::strcmp(argv[i], "--warmup") && i + 1 < argc) warmup = std::stoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--steps")) stepsGiven = true;
  }
  parse_common_args(argc, argv, dom, obs, run, &hyb);
  if (!stepsGiven) run.steps = 200;
  if (run.steps < 1) run.steps = 1;
  try {
    if (method == "all" || method == "bgk")
      bench_one<LBM3D_Legacy>("bgk", 152.0, dom, obs, run, nullptr, warmup);
    if (method == "all" || method == "home")
      bench_one<LBM3D_Home>("home", 80.0, dom, obs, run, nullptr, warmup);
    if (method == "all" || method == "hybrid")
      bench_one<LBM3D_Hybrid>("hybrid", 152.0, dom, obs, run, &hyb, warmup);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "[error] %s\n", e.what());
    return 1;
  }
  return 0;
}
