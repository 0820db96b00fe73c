// fluidsim_tgv -- Taylor-Green vortex validation harness: LBM3D_Legacy and
// LBM3D_Home from the same initial field, sampled against the analytic decay
// (the reference's src/bench/main_taylor_green.cpp:29-350 contract: flags,
// defaults, stdout lines, CSV columns and VTK names).
//
// Default: the 2D vortex extruded in z, u = u0 sin(kx x) cos(ky y) e^{-nu k^2 t},
// v = -u0 cos(kx x) sin(ky y) e^{-nu k^2 t}, rho = 1 + u0^2 e^{-2 nu k^2 t}/(4 cs2)
// (cos 2kx x + cos 2ky y), energy (u0^2/4) e^{-2 nu k^2 t}
// (main_taylor_green.cpp:68-87,188-190).  --3d (extension): the 3D vortex
// u = u0 sin cos cos, v = -u0 cos sin cos, w = 0 with kz = 2 pi k / Nz,
// rho = 1 + (3 u0^2/16)(cos 2x + cos 2y)(cos 2z + 2) at t = 0, velocity decay
// e^{-nu k^2 t} with k^2 = kx^2 + ky^2 + kz^2 and energy (u0^2/8) e^{-2 nu k^2 t}.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <exception>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "sim_common.hpp"

namespace {

constexpr double kPi = 3.14159265358979323846;

struct TgvConfig {
  int Nx = 128, Ny = 128, Nz = 16;
  int steps = 400;
  int sampleEvery = 10;
  double tau = 0.58;
  double u0 = 0.06;
  double k = 1.0;
  double cs2 = 1.0 / 3.0;
  bool threeD = false;
  std::string vtkDir;
  int vtkEvery = 0;
  std::string csvPath = "out_vtk/tgv_results.csv";
  // derived
  int N = 0;
  double kxW = 0, kyW = 0, kzW = 0, k2 = 0, nu = 0;
};

struct Exact {
  double u, v, w, rho;
};

Exact exact_at(int ix, int iy, int iz, double t, const TgvConfig& c) {
  const double X = (ix + 0.5) * c.kxW, Y = (iy + 0.5) * c.kyW, Z = (iz + 0.5) * c.kzW;
  const double dv = std::exp(-c.nu * c.k2 * t), dr = std::exp(-2.0 * c.nu * c.k2 * t);
  Exact e{};
  if (!c.threeD) {
    e.u = c.u0 * std::sin(X) * std::cos(Y) * dv;
    e.v = -c.u0 * std::cos(X) * std::sin(Y) * dv;
    e.w = 0.0;
    e.rho = 1.0 + (c.u0 * c.u0 * dr / (4.0 * c.cs2)) * (std::cos(2.0 * X) + std::cos(2.0 * Y));
  } else {
    e.u = c.u0 * std::sin(X) * std::cos(Y) * std::cos(Z) * dv;
    e.v = -c.u0 * std::cos(X) * std::sin(Y) * std::cos(Z) * dv;
    e.w = 0.0;
    e.rho = 1.0 + (3.0 * c.u0 * c.u0 * dr / 16.0) * (std::cos(2.0 * X) + std::cos(2.0 * Y)) *
                      (std::cos(2.0 * Z) + 2.0);
  }
  return e;
}

double exact_energy(double t, const TgvConfig& c) {
  return (c.threeD ? 0.125 : 0.25) * c.u0 * c.u0 * std::exp(-2.0 * c.nu * c.k2 * t);
}

struct Sample {
  double energy = 0, l2 = 0, linf = 0;
};

template <class Solver>
Sample sample(Solver& s, const TgvConfig& c, double t, std::vector<float>& u,
              std::vector<float>& v, std::vector<float>& w) {
  const size_t bytes = sizeof(float) * size_t(c.N);
  fs_throw_if(fs_memcpy_dtoh(u.data(), s.d_u(), bytes), "memcpy");
  fs_throw_if(fs_memcpy_dtoh(v.data(), s.d_v(), bytes), "memcpy");
  fs_throw_if(fs_memcpy_dtoh(w.data(), s.d_w(), bytes), "memcpy");
  double e = 0, err2 = 0, emax = 0;
  size_t i = 0;
  for (int z = 0; z < c.Nz; ++z)
    for (int y = 0; y < c.Ny; ++y)
      for (int x = 0; x < c.Nx; ++x, ++i) {
        const Exact a = exact_at(x, y, z, t, c);
        const double du = double(u[i]) - a.u, dv = double(v[i]) - a.v, dw = double(w[i]) - a.w;
        const double d2 = du * du + dv * dv + dw * dw;
        err2 += d2;
        emax = std::max(emax, std::sqrt(d2));
        const double U = u[i], V = v[i], W = w[i];
        e += 0.5 * (U * U + V * V + W * W);
      }
  Sample r;
  r.energy = e / double(c.N);
  r.l2 = std::sqrt(err2 / double(c.N));
  r.linf = emax;
  return r;
}

void parse(int argc, char** argv, TgvConfig& c) {
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    const bool more = i + 1 < argc;
    if (a == "--nx" && more) c.Nx = std::stoi(argv[++i]);
    else if (a == "--ny" && more) c.Ny = std::stoi(argv[++i]);
    else if (a == "--nz" && more) c.Nz = std::stoi(argv[++i]);
    else if (a == "--steps" && more) c.steps = std::stoi(argv[++i]);
    else if (a == "--sample-every" && more) c.sampleEvery = std::stoi(argv[++i]);
    else if (a == "--u0" && more) c.u0 = std::stod(argv[++i]);
    else if (a == "--tau" && more) c.tau = std::stod(argv[++i]);
    else if ((a == "--k" || a == "--wavenumber") && more) c.k = std::stod(argv[++i]);
    else if (a == "--vtk-dir" && more) c.vtkDir = argv[++i];
    else if ((a == "--vtk-every" || a == "--vtk-interval") && more) c.vtkEvery = std::stoi(argv[++i]);
    else if (a == "--csv" && more) c.csvPath = argv[++i];
    else if (a == "--3d") c.threeD = true;
  }
  // clamps of main_taylor_green.cpp:275-281
  if (c.sampleEvery <= 0) c.sampleEvery = 1;
  if (c.vtkEvery < 0) c.vtkEvery = 0;
  if (c.Nz <= 0) c.Nz = 1;
  if (c.Nx <= 0) c.Nx = 16;
  if (c.Ny <= 0) c.Ny = 16;
  if (c.steps < 0) c.steps = 0;
  c.N = c.Nx * c.Ny * c.Nz;
  c.kxW = 2.0 * kPi * c.k / c.Nx;
  c.kyW = 2.0 * kPi * c.k / c.Ny;
  c.kzW = c.threeD ? 2.0 * kPi * c.k / c.Nz : 0.0;
  c.k2 = c.kxW * c.kxW + c.kyW * c.kyW + c.kzW * c.kzW;
  c.nu = (c.tau - 0.5) * c.cs2;
}

int run(const TgvConfig& c) {
  std::cout << "[cfg ] Nx=" << c.Nx << " Ny=" << c.Ny << " Nz=" << c.Nz << " steps=" << c.steps
            << " sample=" << c.sampleEvery << "\n";
  std::cout << "[cfg ] tau=" << c.tau << " nu=" << c.nu << " u0=" << c.u0
            << " kx_wave=" << c.kxW << " ky_wave=" << c.kyW << " k^2=" << c.k2 << "\n";
  if (!c.vtkDir.empty())
    std::cout << "[cfg ] vtk dir=" << c.vtkDir << " every=" << c.vtkEvery << " steps\n";
  std::cout << "[cfg ] csv=" << c.csvPath << "\n";

  Domain d;
  d.Nx = c.Nx;
  d.Ny = c.Ny;
  d.Nz = c.Nz;
  d.tau = float(c.tau);
  LBM3D_Legacy legacy;
  LBM3D_Home home;
  legacy.init(d);
  home.init(d);
  std::vector<unsigned char> mask(size_t(c.N), 0);
  legacy.setSolidMask(mask.data());
  home.setSolidMask(mask.data());
  {
    std::vector<float> r(size_t(c.N)), u(size_t(c.N)), v(size_t(c.N)), w(size_t(c.N), 0.0f);
    size_t i = 0;
    for (int z = 0; z < c.Nz; ++z)
      for (int y = 0; y < c.Ny; ++y)
        for (int x = 0; x < c.Nx; ++x, ++i) {
          const Exact a = exact_at(x, y, z, 0.0, c);
          r[i] = float(a.rho);
          u[i] = float(a.u);
          v[i] = float(a.v);
        }
    InitialFields f;
    upload(f.rho, r);
    upload(f.u, u);
    upload(f.v, v);
    upload(f.w, w);
    legacy.reinitEquilibriumFromMacro(f.rho, f.u, f.v, f.w);
    home.reinitEquilibriumFromMacro(f.rho, f.u, f.v, f.w);
  }

  float* spL = nullptr;
  float* spH = nullptr;
  const bool vtk = !c.vtkDir.empty() && c.vtkEvery > 0;
  if (vtk) {
    fs_throw_if(fs_device_malloc((void**)&spL, sizeof(float) * size_t(c.N)), "device_malloc");
    fs_throw_if(fs_device_malloc((void**)&spH, sizeof(float) * size_t(c.N)), "device_malloc");
  }
  auto dump = [&](int step) {
    if (!vtk) return;
    std::filesystem::create_directories(c.vtkDir);
    const std::filesystem::path dir(c.vtkDir);
    launch_speed(legacy.d_u(), legacy.d_v(), legacy.d_w(), spL, c.N);
    launch_speed(home.d_u(), home.d_v(), home.d_w(), spH, c.N);
    write_scalar_vtk(vtk_path(dir, "rho_legacy", step).c_str(), legacy.d_rho(), c.Nx, c.Ny, c.Nz);
    write_scalar_vtk(vtk_path(dir, "rho_home", step).c_str(), home.d_rho(), c.Nx, c.Ny, c.Nz);
    write_scalar_vtk(vtk_path(dir, "speed_legacy", step).c_str(), spL, c.Nx, c.Ny, c.Nz);
    write_scalar_vtk(vtk_path(dir, "speed_home", step).c_str(), spH, c.Nx, c.Ny, c.Nz);
    write_vector_vtk(vtk_path(dir, "vel_legacy", step).c_str(), legacy.d_u(), legacy.d_v(),
                     legacy.d_w(), c.Nx, c.Ny, c.Nz);
    write_vector_vtk(vtk_path(dir, "vel_home", step).c_str(), home.d_u(), home.d_v(), home.d_w(),
                     c.Nx, c.Ny, c.Nz);
  };

  struct Row {
    int step;
    double t, exact;
    Sample L, H;
  };
  std::vector<Row> rows;
  std::vector<float> u(size_t(c.N)), v(size_t(c.N)), w(size_t(c.N));
  auto sample_now = [&](int step) {
    const double t = double(step);
    Row r{step, t, exact_energy(t, c), sample(legacy, c, t, u, v, w), sample(home, c, t, u, v, w)};
    std::cout << "[step " << std::setw(5) << step << "] "
              << "E_exact=" << std::setw(12) << r.exact << "  E_L=" << std::setw(12) << r.L.energy
              << "  E_H=" << std::setw(12) << r.H.energy << "  L2_L=" << r.L.l2
              << "  L2_H=" << r.H.l2 << "\n";
    if (c.vtkEvery > 0 && step % c.vtkEvery == 0) dump(step);
    rows.push_back(r);
  };

  sample_now(0);
  for (int step = 1; step <= c.steps; ++step) {
    legacy.step(1);
    home.step(1);
    if (step % c.sampleEvery == 0 || step == c.steps) sample_now(step);
  }

  std::ofstream ofs(c.csvPath);
  ofs << "step,time,energy_exact,energy_legacy,energy_home,l2_legacy,l2_home,linf_legacy,"
         "linf_home\n";
  ofs << std::scientific << std::setprecision(8);
  for (const Row& r : rows)
    ofs << r.step << "," << r.t << "," << r.exact << "," << r.L.energy << "," << r.H.energy
        << "," << r.L.l2 << "," << r.H.l2 << "," << r.L.linf << "," << r.H.linf << "\n";
  std::cout << "[write] saved " << rows.size() << " samples to " << c.csvPath << "\n";
  if (spL) fs_device_free(spL);
  if (spH) fs_device_free(spH);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  TgvConfig c;
  parse(argc, argv, c);
  try {
    return run(c);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "[error] %s\n", e.what());
    return 1;
  }
}
