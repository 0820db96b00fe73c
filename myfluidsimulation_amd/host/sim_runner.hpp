// sim_runner.hpp -- headless run loops of the drop-in executables (the
// reference's src/apps/sim_runner.hpp contract: same stdout lines, VTK file
// names and cadence -- step 0, then every vtkEvery steps).
//
// The band of the Hybrid method is computed on the GPU (fs_band) with the
// reference rule: fluid cells within 6-neighbour, non-periodic distance d0
// of a solid cell, BFS depth max(1, d0) (sim_runner.hpp:28-91,222-224).
#pragma once

#include <cstdio>
#include <type_traits>
#include <vector>

#include "sim_common.hpp"

template <class Solver>
struct VtkDumper {
  Solver& solver;
  const DomainConfig& dom;
  std::optional<std::filesystem::path> dir;
  std::string tag;
  bool binary;
  float* d_speed = nullptr;

  VtkDumper(Solver& s, const DomainConfig& d, const RunConfig& run, std::string t)
      : solver(s), dom(d), dir(prepare_vtk_dir(run)), tag(std::move(t)), binary(run.vtkBinary) {
    if (dir)
      fs_throw_if(fs_device_malloc((void**)&d_speed,
                                   sizeof(float) * size_t(dom.Nx) * dom.Ny * dom.Nz),
                  "device_malloc");
  }
  ~VtkDumper() {
    if (d_speed) fs_device_free(d_speed);
  }
  void dump(int step) {
    if (!dir) return;
    const int N = dom.Nx * dom.Ny * dom.Nz;
    launch_speed(solver.d_u(), solver.d_v(), solver.d_w(), d_speed, N);
    const std::string r = vtk_path(*dir, "rho_" + tag, step), s = vtk_path(*dir, "speed_" + tag, step),
                      v = vtk_path(*dir, "vel_" + tag, step);
    if (binary) {
      fs_throw_if(fs_write_scalar_vtk_binary(r.c_str(), solver.d_rho(), dom.Nx, dom.Ny, dom.Nz),
                  "write_scalar_vtk_binary");
      fs_throw_if(fs_write_scalar_vtk_binary(s.c_str(), d_speed, dom.Nx, dom.Ny, dom.Nz),
                  "write_scalar_vtk_binary");
      fs_throw_if(fs_write_vector_vtk_binary(v.c_str(), solver.d_u(), solver.d_v(), solver.d_w(),
                                             dom.Nx, dom.Ny, dom.Nz),
                  "write_vector_vtk_binary");
      return;
    }
    write_scalar_vtk(r.c_str(), solver.d_rho(), dom.Nx, dom.Ny, dom.Nz);
    write_scalar_vtk(s.c_str(), d_speed, dom.Nx, dom.Ny, dom.Nz);
    write_vector_vtk(v.c_str(), solver.d_u(), solver.d_v(), solver.d_w(), dom.Nx, dom.Ny, dom.Nz);
  }
};

inline Domain to_domain(const DomainConfig& c) {
  Domain d;
  d.Nx = c.Nx;
  d.Ny = c.Ny;
  d.Nz = c.Nz;
  d.tau = c.tau;
  d.forceX = c.fx;
  d.forceY = c.fy;
  d.forceZ = c.fz;
  return d;
}

// Geometry, flags and initial state shared by both run loops.
template <class Solver>
void setup_solver(Solver& solver, const DomainConfig& domCfg, const RunConfig& runCfg,
                  std::vector<unsigned char>& solid, const std::vector<unsigned char>* band) {
  if constexpr (std::is_same_v<Solver, LBM3D_Home>)
    if (runCfg.bounceHalfway) solver.setBounceRule(FS_BOUNCE_HALFWAY);
  bool special = false;
  std::vector<unsigned char> fl = build_flags(domCfg, runCfg, solid, special);
  if (band)
    for (size_t i = 0; i < fl.size(); ++i)
      if ((*band)[i]) fl[i] |= FS_FLAG_BAND;
  if (special || runCfg.wallsYZ || band)
    solver.setFlags(fl.data());
  else
    solver.setSolidMask(solid.data());
  if (runCfg.inlet) {
    std::vector<float> prof(size_t(domCfg.Ny) * domCfg.Nz, runCfg.inletU);
    solver.setInletProfile(prof.data());
  }
  InitialFields init;
  build_initial_fields(domCfg, runCfg, init);
  solver.reinitEquilibriumFromMacro(init.rho, init.u, init.v, init.w);
}

// run_single_solver<Solver> (sim_runner.hpp:102-192)
template <class Solver>
inline int run_single_solver(const DomainConfig& domCfg, const ObstacleConfig& obsCfg,
                             const RunConfig& runCfg, const char* tag) {
  ObstacleData obstacle = build_obstacle(domCfg, obsCfg);
  const Domain d = to_domain(domCfg);
  Solver solver;
  solver.init(d);
  setup_solver(solver, domCfg, runCfg, obstacle.solidMask, nullptr);

  VtkDumper<Solver> vtk(solver, domCfg, runCfg, tag);
  if (vtk.dir) {
    std::printf("[vtk] %s: dir=%s every=%d\n", tag, vtk.dir->string().c_str(), runCfg.vtkEvery);
    vtk.dump(0);
  }
  std::printf("[run] %s  Nx=%d Ny=%d Nz=%d  tau=%.4g  force=(%.3g,%.3g,%.3g)  steps=%d substeps=%d\n",
              tag, d.Nx, d.Ny, d.Nz, d.tau, d.forceX, d.forceY, d.forceZ, runCfg.steps,
              runCfg.substeps);
  for (int step = 1; step <= runCfg.steps; ++step) {
    solver.step(runCfg.substeps);
    if (vtk.dir && runCfg.vtkEvery > 0 && (step % runCfg.vtkEvery) == 0) vtk.dump(step);
  }
  return 0;
}

// run_hybrid_b0<HybridSolver> (sim_runner.hpp:201-270)
template <class HybridSolver>
inline int run_hybrid_b0(const DomainConfig& domCfg, const ObstacleConfig& obsCfg,
                         const RunConfig& runCfg, const HybridConfig& hybCfg) {
  ObstacleData obstacle = build_obstacle(domCfg, obsCfg);
  const Domain d = to_domain(domCfg);
  // walls are part of the solid set the band is measured from
  bool special = false;
  build_flags(domCfg, runCfg, obstacle.solidMask, special);
  std::vector<unsigned char> band(obstacle.solidMask.size());
  fs_throw_if(fs_band(obstacle.solidMask.data(), d.Nx, d.Ny, d.Nz, hybCfg.bandD0, band.data()),
              "band");

  HybridSolver solver;
  solver.init(d, 0);
  setup_solver(solver, domCfg, runCfg, obstacle.solidMask, &band);

  VtkDumper<HybridSolver> vtk(solver, domCfg, runCfg, "hybrid");
  if (vtk.dir) {
    std::printf("[vtk] hybrid: dir=%s every=%d\n", vtk.dir->string().c_str(), runCfg.vtkEvery);
    vtk.dump(0);
  }
  std::printf(
      "[run] hybrid(B0)  d0=%d  Nx=%d Ny=%d Nz=%d  tau=%.4g  force=(%.3g,%.3g,%.3g)  steps=%d "
      "substeps=%d\n",
      hybCfg.bandD0, d.Nx, d.Ny, d.Nz, d.tau, d.forceX, d.forceY, d.forceZ, runCfg.steps,
      runCfg.substeps);
  for (int step = 1; step <= runCfg.steps; ++step) {
    solver.step(runCfg.substeps);
    if (vtk.dir && runCfg.vtkEvery > 0 && (step % runCfg.vtkEvery) == 0) vtk.dump(step);
  }
  return 0;
}
