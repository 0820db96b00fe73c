// solvers.hpp -- the reference's Solver classes over the C ABI.
//
// LBM3D_Legacy / LBM3D_Home / LBM3D_Hybrid keep the reference's member
// functions (lbm3d_legacy.hpp:19-72, lbm3d_home.hpp:15-74,
// lbm3d_hybrid.hpp:18-79) so run loops written against the reference compile
// unchanged; each call forwards to libfluidsim.so.  Errors throw
// std::runtime_error, as the reference's CUDA_CHECK does
// (cuda_utils.hpp:8-17).  Unlike the reference classes these are not
// copyable (a copy would double-free the device state).
#pragma once

#include <stdexcept>
#include <string>

#include "fluidsim.h"

// Domain (lbm_common.hpp:6-10)
struct Domain {
  int Nx = 128, Ny = 128, Nz = 96;
  float tau = 0.6f;
  float forceX = 0.0f, forceY = 0.0f, forceZ = 0.0f;
};

inline void fs_throw_if(int rc, const char* what) {
  if (rc != FS_OK) throw std::runtime_error(std::string(what) + ": " + fs_last_error());
}

template <int METHOD>
class FsSolver {
 public:
  FsSolver() = default;
  FsSolver(const FsSolver&) = delete;
  FsSolver& operator=(const FsSolver&) = delete;
  ~FsSolver() { release(); }

  void init(const Domain& d) {
    release();
    fs_throw_if(fs_solver_create(METHOD, d.Nx, d.Ny, d.Nz, d.tau, d.forceX, d.forceY, d.forceZ,
                                 &s_),
                "fs_solver_create");
    Nx_ = d.Nx;
    Ny_ = d.Ny;
    Nz_ = d.Nz;
  }
  void setSolidMask(const unsigned char* h_mask) {
    fs_throw_if(fs_solver_set_solid_mask(s_, h_mask), "setSolidMask");
  }
  // extension: full flag byte (solid / band / inlet / outlet)
  void setFlags(const unsigned char* h_flags) {
    fs_throw_if(fs_solver_set_flags(s_, h_flags), "setFlags");
  }
  void setInletProfile(const float* h_ux) {
    fs_throw_if(fs_solver_set_inlet_profile(s_, h_ux), "setInletProfile");
  }
  void reset() { fs_throw_if(fs_solver_reset(s_), "reset"); }
  void reinitEquilibriumFromMacro(const float* d_rho, const float* d_ux, const float* d_uy,
                                  const float* d_uz) {
    fs_throw_if(fs_solver_reinit_from_macro(s_, d_rho, d_ux, d_uy, d_uz),
                "reinitEquilibriumFromMacro");
  }
  void step(int substeps = 1) { fs_throw_if(fs_solver_step(s_, substeps), "step"); }
  void setForce(float fx, float fy = 0.0f, float fz = 0.0f) {
    fs_throw_if(fs_solver_set_force(s_, fx, fy, fz), "setForce");
  }
  float* d_rho() { return macro(FS_RHO); }
  float* d_u() { return macro(FS_UX); }
  float* d_v() { return macro(FS_UY); }
  float* d_w() { return macro(FS_UZ); }
  int Nx() const { return Nx_; }
  int Ny() const { return Ny_; }
  int Nz() const { return Nz_; }
  int N() const { return Nx_ * Ny_ * Nz_; }
  fs_solver* handle() { return s_; }

 protected:
  void release() {
    if (s_) fs_solver_destroy(s_);
    s_ = nullptr;
  }
  float* macro(int which) {
    float* p = nullptr;
    fs_throw_if(fs_solver_macro_ptr(s_, which, &p), "macro pointer");
    return p;
  }
  fs_solver* s_ = nullptr;
  int Nx_ = 0, Ny_ = 0, Nz_ = 0;
};

class LBM3D_Legacy : public FsSolver<FS_METHOD_BGK> {};

// HOME gathers, so the reference's own bounce-back pairing table (opp19,
// lbm3d_home.cu:27-30, used at :219-226) gives a deterministic update: the
// drop-in class uses it, as the reference does.  setBounceRule
// (FS_BOUNCE_HALFWAY) switches to the mass-conserving opposite direction.
class LBM3D_Home : public FsSolver<FS_METHOD_HOME> {
 public:
  void init(const Domain& d) {
    FsSolver::init(d);
    setBounceRule(FS_BOUNCE_REFERENCE_TABLE);
  }
  void setBounceRule(int rule) {
    fs_throw_if(fs_solver_set_bounce_rule(s_, rule), "setBounceRule");
  }
};

class LBM3D_Hybrid : public FsSolver<FS_METHOD_HYBRID> {
 public:
  // init(const Domain&, int nLegacyCells) -- the count is unused, as in the
  // reference (lbm3d_hybrid.cu:210)
  void init(const Domain& d, int /*nLegacyCells*/ = 0) { FsSolver::init(d); }
  void setLegacyMapping(const unsigned char* h_isLegacy, const int* h_legacySlot_ignored) {
    fs_throw_if(fs_solver_set_legacy_mapping(s_, h_isLegacy, h_legacySlot_ignored),
                "setLegacyMapping");
  }
};
