// sim_common.hpp -- command line, obstacle and initial-condition helpers of
// the drop-in executables (the reference's src/apps/sim_common.hpp contract).
//
// Flags, defaults, clamps and messages are the reference's
// (sim_common.hpp:53-192, docs/USAGE.md): unknown flags are ignored, a
// missing or unreadable STL prints an error and the run continues without
// an obstacle, and a VTK directory that cannot be created disables output.
// Flags the reference does not have (all optional, default off):
//   --init tgv3d --u0 U      3D Taylor-Green initial rho and u
//   --inlet-u U              equilibrium inlet at x=0, uniform initial flow U
//   --outlet zero-gradient   zero-gradient outlet at x=nx-1
//   --walls yz               solid wall planes at y=0,ny-1 and z=0,nz-1
//   --vtk-format binary      VTK files in the legacy BINARY encoding
//                            (big-endian float32) instead of ASCII
//   --bounce halfway         HOME only: bounce back in the opposite direction
//                            instead of the reference's pairing table
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <optional>
#include <string>
#include <system_error>
#include <vector>

#include "fluidsim.h"
#include "solvers.hpp"

struct DomainConfig {
  int Nx = 200;
  int Ny = 150;
  int Nz = 120;
  float tau = 0.58f;
  float fx = 0.0f;
  float fy = 0.0f;
  float fz = 0.0f;
};

struct ObstacleConfig {
  std::string stlPath;
  float uniformScale = 0.5f;
  float translateX = 0.5f;
  float translateY = 0.5f;
  float translateZ = 0.5f;
  int amrFactor = 1;
  float amrThreshold = 0.5f;
};

struct RunConfig {
  int steps = 200;
  int substeps = 1;
  std::string vtkDir;
  int vtkEvery = 0;
  std::string initMode = "none";  // none | gauss | slab | tgv3d (extension)
  float amp = 1e-3f;
  float sigma = 0.08f;
  float cx = 0.6f, cy = 0.5f, cz = 0.5f;
  int slabAxis = 0;
  float slabCenter = 0.5f;
  float slabWidth = 0.2f;
  // extensions
  float u0 = 0.05f;
  bool inlet = false;
  float inletU = 0.0f;
  bool outlet = false;
  bool wallsYZ = false;
  bool vtkBinary = false;
  bool bounceHalfway = false;  // HOME: opposite-direction bounce-back
};

struct HybridConfig {
  int bandD0 = 2;
};

struct ObstacleData {
  bool hasObstacle = false;
  std::vector<unsigned char> solidMask;  // Nx*Ny*Nz
};

inline void parse_common_args(int argc, char** argv, DomainConfig& dom, ObstacleConfig& obs,
                              RunConfig& run, HybridConfig* hyb) {
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    const int left = argc - 1 - i;  // arguments after this one
    auto I = [&]() { return std::stoi(argv[++i]); };
    auto F = [&]() { return std::stof(argv[++i]); };
    if (a == "--nx" && left >= 1) dom.Nx = I();
    else if (a == "--ny" && left >= 1) dom.Ny = I();
    else if (a == "--nz" && left >= 1) dom.Nz = I();
    else if (a == "--tau" && left >= 1) dom.tau = F();
    else if (a == "--force" && left >= 3) {
      dom.fx = F();
      dom.fy = F();
      dom.fz = F();
    } else if (a == "--force-x" && left >= 1) dom.fx = F();
    else if (a == "--force-y" && left >= 1) dom.fy = F();
    else if (a == "--force-z" && left >= 1) dom.fz = F();
    else if (a == "--stl" && left >= 1) obs.stlPath = argv[++i];
    else if (a == "--voxel-scale" && left >= 1) obs.uniformScale = F();
    else if (a == "--voxel-translate" && left >= 3) {
      obs.translateX = F();
      obs.translateY = F();
      obs.translateZ = F();
    } else if (a == "--amr") obs.amrFactor = std::max(obs.amrFactor, 2);
    else if ((a == "--amr-factor" || a == "--amr-refine") && left >= 1)
      obs.amrFactor = std::max(1, I());
    else if (a == "--amr-threshold" && left >= 1)
      obs.amrThreshold = std::clamp(F(), 0.0f, 1.0f);
    else if (a == "--steps" && left >= 1) run.steps = std::max(0, I());
    else if (a == "--substeps" && left >= 1) run.substeps = std::max(1, I());
    else if (a == "--vtk-dir" && left >= 1) run.vtkDir = argv[++i];
    else if ((a == "--vtk-every" || a == "--vtk-interval") && left >= 1)
      run.vtkEvery = std::max(0, I());
    else if (a == "--init" && left >= 1) run.initMode = argv[++i];
    else if (a == "--amp" && left >= 1) run.amp = F();
    else if (a == "--sigma" && left >= 1) run.sigma = F();
    else if (a == "--center" && left >= 3) {
      run.cx = F();
      run.cy = F();
      run.cz = F();
    } else if (a == "--axis" && left >= 1) {
      const char ax = argv[++i][0];
      run.slabAxis = (ax == 'y') ? 1 : (ax == 'z') ? 2 : 0;
    } else if (a == "--width" && left >= 1) run.slabWidth = F();
    else if (a == "--slab-center" && left >= 1) run.slabCenter = F();
    else if ((a == "--hybrid-band" || a == "--hybrid-d0") && left >= 1) {
      const int v = std::max(0, I());
      if (hyb) hyb->bandD0 = v;
    }
    // ---- extensions
    else if (a == "--u0" && left >= 1) run.u0 = F();
    else if (a == "--inlet-u" && left >= 1) {
      run.inlet = true;
      run.inletU = F();
    } else if (a == "--outlet" && left >= 1) {
      run.outlet = std::string(argv[++i]) == "zero-gradient";
    } else if (a == "--walls" && left >= 1) {
      run.wallsYZ = std::string(argv[++i]) == "yz";
    } else if (a == "--vtk-format" && left >= 1) {
      run.vtkBinary = std::string(argv[++i]) == "binary";
    } else if (a == "--bounce" && left >= 1) {
      run.bounceHalfway = std::string(argv[++i]) == "halfway";
    }
  }
  dom.Nx = std::max(dom.Nx, 8);
  dom.Ny = std::max(dom.Ny, 8);
  dom.Nz = std::max(dom.Nz, 1);
  dom.tau = std::max(dom.tau, 0.51f);
  run.substeps = std::max(run.substeps, 1);
}

// load_obstacle_mesh + build_obstacle (sim_common.hpp:198-238 contract)
inline ObstacleData build_obstacle(const DomainConfig& dom, const ObstacleConfig& obs) {
  ObstacleData out;
  out.solidMask.assign(size_t(dom.Nx) * dom.Ny * dom.Nz, 0);
  if (obs.stlPath.empty()) return out;
  int ntri = 0;
  float bmin[3], bmax[3];
  if (fs_load_stl(obs.stlPath.c_str(), nullptr, 0, &ntri, bmin, bmax) != FS_OK) {
    std::fprintf(stderr, "[error] STL の読み込みに失敗しました: %s\n", obs.stlPath.c_str());
    return out;
  }
  std::vector<float> tris(size_t(ntri) * 9 + 9);
  fs_throw_if(fs_load_stl(obs.stlPath.c_str(), tris.data(), ntri, &ntri, bmin, bmax),
              "load_stl");
  out.hasObstacle = true;
  fs_throw_if(fs_voxelize(tris.data(), ntri, bmin, bmax, dom.Nx, dom.Ny, dom.Nz,
                          obs.uniformScale, obs.translateX, obs.translateY, obs.translateZ,
                          std::max(1, obs.amrFactor), std::clamp(obs.amrThreshold, 0.0f, 1.0f),
                          out.solidMask.data()),
              "voxelize");
  return out;
}

// Extension flags (walls, inlet, outlet) on top of the obstacle mask.
// Returns the flag byte per cell; `special` says whether inlet/outlet cells
// exist (then the solver takes the full flag grid).
inline std::vector<unsigned char> build_flags(const DomainConfig& dom, const RunConfig& run,
                                              std::vector<unsigned char>& solid, bool& special) {
  const int nx = dom.Nx, ny = dom.Ny, nz = dom.Nz;
  auto id = [&](int x, int y, int z) { return (size_t(z) * ny + y) * nx + x; };
  if (run.wallsYZ)
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        if (y == 0 || y == ny - 1 || z == 0 || z == nz - 1)
          for (int x = 0; x < nx; ++x) solid[id(x, y, z)] = 1;
  std::vector<unsigned char> fl(solid.size());
  for (size_t i = 0; i < fl.size(); ++i) fl[i] = solid[i] ? FS_FLAG_SOLID : 0;
  special = false;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y) {
      if (run.inlet && !solid[id(0, y, z)]) {
        fl[id(0, y, z)] |= FS_FLAG_INLET;
        special = true;
      }
      if (run.outlet && nx >= 3 && !solid[id(nx - 1, y, z)] && !solid[id(nx - 2, y, z)]) {
        fl[id(nx - 1, y, z)] |= FS_FLAG_OUTLET;
        special = true;
      }
    }
  return fl;
}

// build_initial_rho_device (sim_common.hpp:242-260) plus the tgv3d / inlet
// extensions, which also give an initial velocity.
struct InitialFields {
  float* rho = nullptr;
  float* u = nullptr;
  float* v = nullptr;
  float* w = nullptr;
  ~InitialFields() {
    for (float* p : {rho, u, v, w})
      if (p) fs_device_free(p);
  }
};

inline void upload(float*& d, const std::vector<float>& h) {
  fs_throw_if(fs_device_malloc((void**)&d, h.size() * sizeof(float)), "device_malloc");
  fs_throw_if(fs_memcpy_htod(d, h.data(), h.size() * sizeof(float)), "memcpy");
}

inline void build_initial_fields(const DomainConfig& dom, const RunConfig& run,
                                 InitialFields& out) {
  const int nx = dom.Nx, ny = dom.Ny, nz = dom.Nz;
  const size_t N = size_t(nx) * ny * nz;
  if (run.initMode == "gauss") {
    fs_throw_if(fs_device_malloc((void**)&out.rho, N * sizeof(float)), "device_malloc");
    make_rho_gauss(out.rho, nx, ny, nz, run.cx, run.cy, run.cz, run.sigma, run.amp);
  } else if (run.initMode == "slab") {
    fs_throw_if(fs_device_malloc((void**)&out.rho, N * sizeof(float)), "device_malloc");
    make_rho_slab(out.rho, nx, ny, nz, run.slabAxis, run.amp, run.slabCenter, run.slabWidth);
  } else if (run.initMode == "tgv3d") {
    // 3D Taylor-Green at cell centres, k = 2 pi / N per axis
    std::vector<float> r(N), u(N), v(N), w(N, 0.0f);
    const double kx = 2 * M_PI / nx, ky = 2 * M_PI / ny, kz = 2 * M_PI / nz, u0 = run.u0;
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          const size_t i = (size_t(z) * ny + y) * nx + x;
          const double X = kx * (x + 0.5), Y = ky * (y + 0.5), Z = kz * (z + 0.5);
          u[i] = float(u0 * std::sin(X) * std::cos(Y) * std::cos(Z));
          v[i] = float(-u0 * std::cos(X) * std::sin(Y) * std::cos(Z));
          r[i] = float(1.0 + (3.0 * u0 * u0 / 16.0) * (std::cos(2 * X) + std::cos(2 * Y)) *
                                 (std::cos(2 * Z) + 2.0));
        }
    upload(out.rho, r);
    upload(out.u, u);
    upload(out.v, v);
    upload(out.w, w);
  } else {
    upload(out.rho, std::vector<float>(N, 1.0f));
  }
  if (run.inlet && !out.u) {
    upload(out.u, std::vector<float>(N, run.inletU));
    upload(out.v, std::vector<float>(N, 0.0f));
    upload(out.w, std::vector<float>(N, 0.0f));
  }
}

// vtk_path / prepare_vtk_dir (sim_common.hpp:264-283)
inline std::string vtk_path(const std::filesystem::path& dir, const std::string& stem, int step) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%s_%06d.vtk", stem.c_str(), step);
  return (dir / buf).string();
}

inline std::optional<std::filesystem::path> prepare_vtk_dir(const RunConfig& run) {
  if (run.vtkDir.empty() || run.vtkEvery <= 0) return std::nullopt;
  std::filesystem::path p(run.vtkDir);
  std::error_code ec;
  std::filesystem::create_directories(p, ec);
  if (ec) {
    std::fprintf(stderr, "[vtk] 出力ディレクトリ作成に失敗: %s (%s)\n", run.vtkDir.c_str(),
                 ec.message().c_str());
    return std::nullopt;
  }
  return p;
}
