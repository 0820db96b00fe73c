// fluidsim_sweep -- the headless Hybrid(B0) band sweep of the reference's
// comparer (main_compare.cpp:505-699, the --hybrid-sweep-max mode, driven by
// scripts/run_b0_sweeps.ps1), on the batched engine: Legacy, HOME and Hybrid
// for every band thickness d0 = 0..max are the domains of one batch and step
// together, instead of one solver after the other.
//
// Same flags as the comparer where they apply (main_compare.cpp:226-301:
// --stl, --force/--force-y/--force-z, --init gauss|slab|none, --amp, --sigma,
// --center, --axis, --width, --amr, --amr-factor, --amr-threshold,
// --hybrid-band, --hybrid-sweep-max, --hybrid-sweep-steps), the same
// 200x150x120 tau 0.58 domain and voxel placement (scale 0.5 about the
// centre), the same stdout lines and CSV columns and row format
// (main_compare.cpp:304-309,384,468,515,630-698).  Extensions: --nx --ny --nz
// --tau, --bounce halfway (HOME bounce-back in the opposite direction instead
// of the reference's pairing table).  The ms_per_step_* columns time each configuration alone on a
// single-domain solver over min(steps, 50) steps (the batch steps them all
// at once, so it has no per-configuration time of its own).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <exception>
#include <string>
#include <vector>

#include "sim_common.hpp"

namespace {

struct SweepConfig {
  int Nx = 200, Ny = 150, Nz = 120;
  float tau = 0.58f, fx = 0.0f, fy = 0.0f, fz = 0.0f;
  std::string stlPath;
  std::string initMode = "none";
  float amp = 1e-3f, sigma = 0.08f, cx = 0.6f, cy = 0.5f, cz = 0.5f;
  int axis = 0;
  float center = 0.5f, width = 0.2f;
  int amrFactor = 1;
  float amrThreshold = 0.5f;
  int hybridBand = 2;
  int sweepMax = -1;
  int sweepSteps = 500;
  // HOME's bounce-back partner: the reference's pairing table, as LBM3D_Home
  // (extension: --bounce halfway)
  int homeBounce = FS_BOUNCE_REFERENCE_TABLE;
};

void parse(int argc, char** argv, SweepConfig& c) {
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    const int left = argc - 1 - i;
    auto F = [&]() { return std::stof(argv[++i]); };
    auto I = [&]() { return std::stoi(argv[++i]); };
    if (a == "--stl" && left >= 1) c.stlPath = argv[++i];
    else if (a == "--force" && left >= 1) c.fx = F();
    else if ((a == "--forcey" || a == "--force-y") && left >= 1) c.fy = F();
    else if ((a == "--forcez" || a == "--force-z") && left >= 1) c.fz = F();
    else if (a == "--init" && left >= 1) c.initMode = argv[++i];
    else if (a == "--amp" && left >= 1) c.amp = F();
    else if (a == "--sigma" && left >= 1) c.sigma = F();
    else if (a == "--center" && left >= 3) {
      c.cx = F();
      c.cy = F();
      c.cz = F();
    } else if (a == "--axis" && left >= 1) {
      const char ax = argv[++i][0];
      c.axis = (ax == 'y') ? 1 : (ax == 'z') ? 2 : 0;
    } else if (a == "--width" && left >= 1) c.width = F();
    else if (a == "--amr") c.amrFactor = std::max(c.amrFactor, 2);
    else if ((a == "--amr-factor" || a == "--amr-refine") && left >= 1) c.amrFactor = I();
    else if (a == "--amr-threshold" && left >= 1) c.amrThreshold = F();
    else if ((a == "--hybrid-band" || a == "--hybrid-d0") && left >= 1)
      c.hybridBand = std::max(0, I());
    else if ((a == "--hybrid-sweep-max" || a == "--hybrid-sweep") && left >= 1)
      c.sweepMax = std::max(0, I());
    else if ((a == "--hybrid-sweep-steps" || a == "--sweep-steps") && left >= 1)
      c.sweepSteps = std::max(0, I());
    // extensions
    else if (a == "--nx" && left >= 1) c.Nx = std::max(8, I());
    else if (a == "--ny" && left >= 1) c.Ny = std::max(8, I());
    else if (a == "--nz" && left >= 1) c.Nz = std::max(1, I());
    else if (a == "--tau" && left >= 1) c.tau = std::max(0.51f, F());
    else if (a == "--bounce" && left >= 1)
      c.homeBounce = std::string(argv[++i]) == "halfway" ? FS_BOUNCE_HALFWAY
                                                          : FS_BOUNCE_REFERENCE_TABLE;
  }
  if (c.amrFactor < 1) c.amrFactor = 1;
  c.amrThreshold = std::clamp(c.amrThreshold, 0.0f, 1.0f);
}

// mean |u| over the near-wall cells and mean |u - base| (float speeds as
// launch_speed computes them), mean rho over the fluid cells
struct Stats {
  double meanU = 0, meanAbsDiff = 0, meanRho = 0;
};

Stats stats(const std::vector<float>& rho, const std::vector<float>& u, const std::vector<float>& v,
            const std::vector<float>& w, const std::vector<int>& near,
            const std::vector<unsigned char>& solid, int nFluid, const std::vector<float>* base,
            std::vector<float>* speedsOut) {
  Stats s;
  double sum = 0, sumAbs = 0;
  if (speedsOut) speedsOut->resize(near.size());
  for (size_t k = 0; k < near.size(); ++k) {
    const int i = near[k];
    const float sp = std::sqrt(u[i] * u[i] + v[i] * v[i] + w[i] * w[i]);
    sum += double(sp);
    if (base) sumAbs += double(std::abs(sp - (*base)[k]));
    if (speedsOut) (*speedsOut)[k] = sp;
  }
  const double cnt = near.empty() ? 1.0 : double(near.size());
  s.meanU = sum / cnt;
  s.meanAbsDiff = sumAbs / cnt;
  double rs = 0;
  for (size_t i = 0; i < rho.size(); ++i)
    if (!solid[i]) rs += double(rho[i]);
  s.meanRho = nFluid > 0 ? rs / double(nFluid) : 0.0;
  return s;
}

double solo_ms_per_step(int method, const fs_domain_desc& dd, const std::vector<unsigned char>& fl,
                        const float* d_rho, int steps, int homeBounce) {
  if (steps <= 0) return 0.0;
  fs_solver* s = nullptr;
  fs_throw_if(fs_solver_create(method, dd.nx, dd.ny, dd.nz, dd.tau, dd.fx, dd.fy, dd.fz, &s),
              "solver_create");
  float ms = 0.0f;
  int rc = method == FS_METHOD_HOME ? fs_solver_set_bounce_rule(s, homeBounce) : FS_OK;
  if (rc == FS_OK) rc = fs_solver_set_flags(s, fl.data());
  if (rc == FS_OK) rc = fs_solver_reinit_from_macro(s, d_rho, nullptr, nullptr, nullptr);
  if (rc == FS_OK) rc = fs_solver_timed_steps(s, steps, &ms);
  fs_solver_destroy(s);
  fs_throw_if(rc, "timed_steps");
  return double(ms) / double(steps);
}

int run(const SweepConfig& c) {
  std::printf(
      "[args] init=%s amp=%.3g sigma=%.3g center=(%.3f,%.3f,%.3f) axis=%d width=%.3f "
      "forceX=%.3g\n",
      c.initMode.c_str(), c.amp, c.sigma, c.cx, c.cy, c.cz, c.axis, c.width, c.fx);
  std::printf("[args] force=(%.3g, %.3g, %.3g)\n", c.fx, c.fy, c.fz);
  std::printf("[amr ] refine=%d threshold=%.2f\n", c.amrFactor, c.amrThreshold);
  const int nx = c.Nx, ny = c.Ny, nz = c.Nz;
  const size_t N = size_t(nx) * ny * nz;

  // obstacle: the comparer stops when the STL cannot be read
  DomainConfig dom;
  dom.Nx = nx;
  dom.Ny = ny;
  dom.Nz = nz;
  ObstacleConfig obs;
  obs.stlPath = c.stlPath;
  obs.amrFactor = c.amrFactor;
  obs.amrThreshold = c.amrThreshold;
  std::vector<unsigned char> solid(N, 0);
  if (!c.stlPath.empty()) {
    int ntri = 0;
    float bmin[3], bmax[3];
    if (fs_load_stl(c.stlPath.c_str(), nullptr, 0, &ntri, bmin, bmax) != FS_OK) {
      std::fprintf(stderr, "Failed to load STL: %s\n", c.stlPath.c_str());
      return 1;
    }
    solid = build_obstacle(dom, obs).solidMask;
  }
  // near-wall fluid cells: a solid 6-neighbour, non-periodic (main_compare.cpp:358-384)
  std::vector<int> near;
  int nFluid = 0;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const size_t i = (size_t(z) * ny + y) * nx + x;
        if (solid[i]) continue;
        ++nFluid;
        const size_t sz = size_t(nx) * ny;
        const bool nb = (x > 0 && solid[i - 1]) || (x < nx - 1 && solid[i + 1]) ||
                        (y > 0 && solid[i - nx]) || (y < ny - 1 && solid[i + nx]) ||
                        (z > 0 && solid[i - sz]) || (z < nz - 1 && solid[i + sz]);
        if (nb) near.push_back(int(i));
      }
  std::printf("[wall] near-wall fluid cells: %zu\n", near.size());

  auto band_of = [&](int d0, int& count) {
    std::vector<unsigned char> band(N, 0);
    if (d0 > 0) fs_throw_if(fs_band(solid.data(), nx, ny, nz, d0, band.data()), "band");
    count = 0;
    for (size_t i = 0; i < N; ++i) count += band[i] ? 1 : 0;
    return band;
  };
  int nLegDefault = 0;
  band_of(c.hybridBand, nLegDefault);
  std::printf("[hybrid] B0 band d0=%d -> legacy cells=%d (fluid only)\n", c.hybridBand,
              nLegDefault);
  if (c.sweepMax < 0) {
    std::printf("[sweep] no --hybrid-sweep-max given: nothing to sweep\n");
    return 0;
  }
  std::printf("[hybrid] B0 sweep enabled: d0=0..%d, steps_per_run=%d\n", c.sweepMax,
              c.sweepSteps);

  // initial density (main_compare.cpp:479-494), equilibrium with u = 0
  InitialFields init;
  RunConfig rc;
  rc.initMode = (c.initMode == "gauss" || c.initMode == "slab") ? c.initMode : "none";
  rc.amp = c.amp;
  rc.sigma = c.sigma;
  rc.cx = c.cx;
  rc.cy = c.cy;
  rc.cz = c.cz;
  rc.slabAxis = c.axis;
  rc.slabCenter = c.center;
  rc.slabWidth = c.width;
  build_initial_fields(dom, rc, init);

  // one batch: Legacy, HOME, Hybrid d0 = 0..max
  const int nd = 2 + c.sweepMax + 1;
  std::vector<fs_domain_desc> descs(nd);
  for (int k = 0; k < nd; ++k)
    descs[k] = fs_domain_desc{k == 0 ? FS_METHOD_BGK : (k == 1 ? FS_METHOD_HOME : FS_METHOD_HYBRID),
                              nx, ny, nz, c.tau, c.fx, c.fy, c.fz};
  std::vector<std::vector<unsigned char>> flags(nd);
  std::vector<int> nLeg(nd, 0);
  for (int k = 0; k < nd; ++k) {
    flags[k].assign(N, 0);
    for (size_t i = 0; i < N; ++i) flags[k][i] = solid[i] ? FS_FLAG_SOLID : 0;
    if (k >= 2) {
      const std::vector<unsigned char> band = band_of(k - 2, nLeg[k]);
      for (size_t i = 0; i < N; ++i)
        if (band[i]) flags[k][i] |= FS_FLAG_BAND;
    }
  }
  fs_batch* b = nullptr;
  fs_throw_if(fs_batch_create(descs.data(), nd, &b), "batch_create");
  struct BatchGuard {
    fs_batch* b;
    ~BatchGuard() { fs_batch_destroy(b); }
  } guard{b};
  fs_throw_if(fs_batch_set_bounce_rule(b, 1, c.homeBounce), "batch_set_bounce_rule");
  for (int k = 0; k < nd; ++k) {
    fs_throw_if(fs_batch_set_flags(b, k, flags[k].data()), "batch_set_flags");
    fs_throw_if(fs_batch_reinit_from_macro(b, k, init.rho, nullptr, nullptr, nullptr),
                "batch_reinit");
  }
  fs_throw_if(fs_batch_step(b, c.sweepSteps), "batch_step");

  std::vector<float> r(N), u(N), v(N), w(N);
  auto fetch = [&](int k) {
    fs_throw_if(fs_batch_copy_macro(b, k, FS_RHO, r.data()), "copy_macro");
    fs_throw_if(fs_batch_copy_macro(b, k, FS_UX, u.data()), "copy_macro");
    fs_throw_if(fs_batch_copy_macro(b, k, FS_UY, v.data()), "copy_macro");
    fs_throw_if(fs_batch_copy_macro(b, k, FS_UZ, w.data()), "copy_macro");
  };
  std::vector<float> baseline;
  fetch(0);
  const Stats L = stats(r, u, v, w, near, solid, nFluid, nullptr, &baseline);
  fetch(1);
  const Stats H = stats(r, u, v, w, near, solid, nFluid, &baseline, nullptr);
  const int timedSteps = std::min(c.sweepSteps, 50);
  const double msL =
      solo_ms_per_step(FS_METHOD_BGK, descs[0], flags[0], init.rho, timedSteps, c.homeBounce);
  const double msH =
      solo_ms_per_step(FS_METHOD_HOME, descs[1], flags[1], init.rho, timedSteps, c.homeBounce);

  std::printf(
      "d0,legacy_cells,legacy_ratio,mean_u_legacy,mean_u_home,mean_u_hybrid,"
      "mean_absdiff_home,mean_absdiff_hybrid,mean_rho_legacy,mean_rho_home,mean_rho_hybrid,"
      "absdiff_rho_home,absdiff_rho_hybrid,ms_per_step_legacy,ms_per_step_home,"
      "ms_per_step_hybrid,amr_factor,amr_threshold,fluid_cells,near_wall_cells\n");
  for (int d0 = 0; d0 <= c.sweepMax; ++d0) {
    const int k = 2 + d0;
    fetch(k);
    const Stats Y = stats(r, u, v, w, near, solid, nFluid, &baseline, nullptr);
    const double msY =
        solo_ms_per_step(FS_METHOD_HYBRID, descs[k], flags[k], init.rho, timedSteps, c.homeBounce);
    const double ratio = nFluid > 0 ? double(nLeg[k]) / double(nFluid) : 0.0;
    std::printf(
        "%d,%d,%.6f,%.9e,%.9e,%.9e,%.9e,%.9e,%.9e,%.9e,%.9e,%.9e,%.9e,%.6f,%.6f,%.6f,%d,%.6f,%d,"
        "%zu\n",
        d0, nLeg[k], ratio, L.meanU, H.meanU, Y.meanU, H.meanAbsDiff, Y.meanAbsDiff, L.meanRho,
        H.meanRho, Y.meanRho, std::abs(H.meanRho - L.meanRho), std::abs(Y.meanRho - L.meanRho),
        msL, msH, msY, c.amrFactor, double(c.amrThreshold), nFluid, near.size());
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  SweepConfig c;
  parse(argc, argv, c);
  try {
    return run(c);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "[error] %s\n", e.what());
    return 1;
  }
}
