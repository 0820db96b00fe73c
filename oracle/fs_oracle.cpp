// fs_oracle.cpp -- CPU oracle (test infrastructure only; see fs_oracle.h).
//
// A scalar restatement of the reference algorithms, one function per
// reference routine, in the reference's float expression order.  It is built
// with -ffp-contract=off so every a*b+c rounds twice, as the reference's C++
// expressions are written.  OpenMP splits the cell loops over z-planes; each
// cell's arithmetic is independent of the split, so results do not depend on
// the thread count.
//
// Reference files followed (yagise/MyFluidSimulation):
//   src/lbm/lbm3d_legacy.cu   BGK collide + push stream + halfway bounce-back
//   src/lbm/lbm_reinit.cu     feq re-initialisation (cs2 form)
//   src/lbm/lbm3d_home.cu     HOME moment-only update
//   src/lbm/lbm3d_hybrid.cu   Hybrid(B0) per-cell switch
//   src/apps/sim_runner.hpp   distance-to-solid BFS and band mask
//   src/geom/voxelize.cpp     SAT voxelizer, exterior flood fill, AMR downsample
//   src/geom/stl_loader.cpp   ASCII / binary STL parsing
//   src/moment_vtk_writer.cu  VTK legacy ASCII writer
//   src/fields/*.cu           speed and initial-density profiles
//
// Inlet / outlet cells are a build extension the reference does not have
// (all reference boundaries are periodic).  Their rule is defined here and
// mirrored by the HIP library:
//   * an INLET cell emits feq(rho=1, u=(u_in(y,z),0,0)) in every direction
//     and reports (1, u_in, 0, 0); what streams into it is ignored;
//   * an OUTLET cell (x = nx-1) is a zero-gradient copy of its upstream
//     neighbour x-1: BGK/Hybrid copy the neighbour's populations into the
//     outlet before each step; HOME stores the neighbour's new moments.

#include "fs_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <queue>
#include <sstream>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// D3Q19 velocity set (lbm3d_legacy.cu:16-32)
constexpr int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
// The reference's opp table (lbm3d_legacy.cu:29-32).  For q >= 7 it pairs
// mirror images, not opposites: e_{OPP_REF[7]} = (-1,+1,0) while -e_7 =
// (-1,-1,0).
constexpr int OPP_REF[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
// Opposite directions (e_OPP[q] = -e_q): the halfway bounce-back partner.
constexpr int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};

template <class T>
struct Weights {
  T w[19];
  Weights() {
    w[0] = T(1) / T(3);
    for (int q = 1; q <= 6; ++q) w[q] = T(1) / T(18);
    for (int q = 7; q < 19; ++q) w[q] = T(1) / T(36);
  }
};

// feq (lbm3d_legacy.cu:54-58) == feq19 (lbm3d_hybrid.cu:37-41)
template <class T>
inline T feq_bgk(const T* w, int q, T rho, T ux, T uy, T uz) {
  T eiu = T(CX[q]) * ux + T(CY[q]) * uy + T(CZ[q]) * uz;
  T uu = ux * ux + uy * uy + uz * uz;
  return w[q] * rho * (T(1) + T(3) * eiu + T(4.5) * eiu * eiu - T(1.5) * uu);
}

// kern_reinit_eq body (lbm_reinit.cu:31-49): cs2 form
template <class T>
inline T feq_cs2(const T* w, int q, T r, T u0, T v0, T w0) {
  const T cs2 = T(1) / T(3);
  const T uu = u0 * u0 + v0 * v0 + w0 * w0;
  const T cx = T(CX[q]), cy = T(CY[q]), cz = T(CZ[q]);
  const T cu = cx * u0 + cy * v0 + cz * w0;
  return w[q] * r * (T(1) + cu / cs2 + T(0.5) * (cu * cu) / (cs2 * cs2) - T(0.5) * uu / cs2);
}

// reconstruct_f_post (lbm3d_home.cu:37-81)
template <class T>
inline T home_rec(const T* wt, int q, T rho, T ux, T uy, T uz, T Sxx, T Sxy, T Sxz, T Syy,
                  T Syz, T Szz, T oneMinusOmega) {
  const T cx = T(CX[q]), cy = T(CY[q]), cz = T(CZ[q]);
  const T w = wt[q];
  const T eu = cx * ux + cy * uy + cz * uz;
  const T uu = ux * ux + uy * uy + uz * uz;
  const T feq = w * rho * (T(1) + T(3) * eu + T(4.5) * eu * eu - T(1.5) * uu);
  const T cs2 = T(1) / T(3);
  const T Qxx = cx * cx - cs2;
  const T Qyy = cy * cy - cs2;
  const T Qzz = cz * cz - cs2;
  const T Qxy = cx * cy;
  const T Qxz = cx * cz;
  const T Qyz = cy * cz;
  const T contraction =
      Qxx * Sxx + Qyy * Syy + Qzz * Szz + T(2) * (Qxy * Sxy + Qxz * Sxz + Qyz * Syz);
  const T fneq = T(4.5) * w * contraction;
  return feq + oneMinusOmega * fneq;
}

struct Grid {
  int nx, ny, nz;
  size_t n;
  inline size_t id(int x, int y, int z) const {
    return ((size_t)z * ny + y) * nx + x;
  }
  // periodic neighbour (x+dx, y+dy, z+dz), dx..dz in {-1,0,1}
  inline size_t nb(int x, int y, int z, int dx, int dy, int dz) const {
    int x2 = (x + dx + nx) % nx;
    int y2 = (y + dy + ny) % ny;
    int z2 = (z + dz + nz) % nz;
    return id(x2, y2, z2);
  }
};

struct SolverBase {
  virtual ~SolverBase() {}
  virtual int set_flags(const uint8_t* flags) = 0;
  virtual int set_inlet(const float* ux) = 0;
  virtual int set_omegas(double a, double b, double c) = 0;
  virtual int set_bounce(int rule) = 0;
  virtual int reset() = 0;
  virtual int reinit_f32(const float*, const float*, const float*, const float*) = 0;
  virtual int step(int n) = 0;
  virtual int get_macro(int which, float* o32, double* o64) const = 0;
  virtual size_t state_size() const = 0;
  virtual int get_state(float* o32, double* o64) const = 0;
  virtual int set_state_f32(const float* in) = 0;
};

int validate_flags(const Grid& g, const uint8_t* fl) {
  for (int z = 0; z < g.nz; ++z)
    for (int y = 0; y < g.ny; ++y)
      for (int x = 0; x < g.nx; ++x) {
        const uint8_t f = fl[g.id(x, y, z)];
        if ((f & ORC_INLET) && (f & (ORC_SOLID | ORC_OUTLET))) return -1;
        if (f & ORC_OUTLET) {
          if (f & ORC_SOLID) return -1;
          if (x < 1) return -1;
          const uint8_t u = fl[g.id(x - 1, y, z)];
          if (u & (ORC_SOLID | ORC_INLET | ORC_OUTLET)) return -1;
        }
      }
  return 0;
}

// --------------------------------------------------------------------------
// Populations solvers: Legacy BGK and Hybrid(B0).
// --------------------------------------------------------------------------
template <class T>
struct PopSolver : SolverBase {
  Grid g;
  bool hybrid;
  T tau, fx, fy, fz;
  T omegaL, omegaH, omegaS;
  Weights<T> W;
  std::vector<T> f, fnext, rho, ux, uy, uz;
  std::vector<uint8_t> flags;
  std::vector<T> inlet;  // ny*nz
  bool hasSpecial = false;

  PopSolver(bool hyb, int nx, int ny, int nz, double tau_, double fx_, double fy_,
            double fz_) {
    g = {nx, ny, nz, (size_t)nx * ny * nz};
    hybrid = hyb;
    tau = T(tau_);
    fx = T(fx_);
    fy = T(fy_);
    fz = T(fz_);
    // step(): omega = 1/tau (lbm3d_legacy.cu:213, lbm3d_hybrid.cu:264-267)
    omegaL = omegaH = omegaS = T(1) / tau;
    f.assign(19 * g.n, T(0));
    fnext.assign(19 * g.n, T(0));
    rho.assign(g.n, T(1));
    ux.assign(g.n, T(0));
    uy.assign(g.n, T(0));
    uz.assign(g.n, T(0));
    flags.assign(g.n, 0);
    inlet.assign((size_t)ny * nz, T(0));
    reset();
  }

  int set_flags(const uint8_t* fl) override {
    if (validate_flags(g, fl)) return -1;
    hasSpecial = false;
    for (size_t i = 0; i < g.n; ++i) {
      flags[i] = fl[i] & (ORC_SOLID | ORC_BAND | ORC_INLET | ORC_OUTLET);
      if (flags[i] & (ORC_INLET | ORC_OUTLET)) hasSpecial = true;
    }
    return 0;
  }
  int set_inlet(const float* u) override {
    for (size_t i = 0; i < inlet.size(); ++i) inlet[i] = T(u[i]);
    return 0;
  }
  int set_bounce(int rule) override { return rule == 0 ? 0 : -1; }
  int set_omegas(double a, double b, double c) override {
    omegaL = T(a);
    omegaH = T(b);
    omegaS = T(c);
    return 0;
  }

  // kern_reset (lbm3d_legacy.cu:40-53, lbm3d_hybrid.cu:43-56)
  int reset() override {
    for (size_t id = 0; id < g.n; ++id) {
      rho[id] = T(1);
      ux[id] = uy[id] = uz[id] = T(0);
      for (int q = 0; q < 19; ++q) f[q * g.n + id] = W.w[q] * T(1);
    }
    return 0;
  }

  int reinit_f32(const float* r, const float* u, const float* v, const float* w) override {
    if (!hybrid) {
      // LBM3D_Legacy::reinitEquilibriumFromMacro (lbm3d_legacy.cu:160-208)
      if (!r) return reset();
      for (size_t i = 0; i < g.n; ++i) {
        rho[i] = T(r[i]);
        ux[i] = u ? T(u[i]) : T(0);
        uy[i] = v ? T(v[i]) : T(0);
        uz[i] = w ? T(w[i]) : T(0);
      }
      for (size_t i = 0; i < g.n; ++i)
        for (int q = 0; q < 19; ++q)
          f[q * g.n + i] = feq_cs2(W.w, q, rho[i], ux[i], uy[i], uz[i]);
      fnext = f;
    } else {
      // kern_reinit_equilibrium (lbm3d_hybrid.cu:58-87)
      for (size_t i = 0; i < g.n; ++i) {
        T rr = r ? T(r[i]) : T(1);
        T a = u ? T(u[i]) : T(0);
        T b = v ? T(v[i]) : T(0);
        T c = w ? T(w[i]) : T(0);
        rho[i] = rr;
        ux[i] = a;
        uy[i] = b;
        uz[i] = c;
        for (int q = 0; q < 19; ++q) f[q * g.n + i] = feq_bgk(W.w, q, rr, a, b, c);
      }
    }
    return 0;
  }

  // one time step: kern_collide_stream (lbm3d_legacy.cu:59-114) or
  // kern_collide_stream_b0 (lbm3d_hybrid.cu:89-175), then kern_swap.
  void step1() {
    const size_t N = g.n;
    if (hasSpecial) {
      // outlet: zero-gradient copy of the upstream neighbour before the step
      for (size_t id = 0; id < N; ++id)
        if (flags[id] & ORC_OUTLET)
          for (int q = 0; q < 19; ++q) f[q * N + id] = f[q * N + id - 1];
    }
#pragma omp parallel for schedule(static)
    for (int z = 0; z < g.nz; ++z) {
      for (int y = 0; y < g.ny; ++y) {
        for (int x = 0; x < g.nx; ++x) {
          const size_t id = g.id(x, y, z);
          const uint8_t fl = flags[id];
          if (fl & ORC_SOLID) {
            for (int q = 0; q < 19; ++q) fnext[q * N + id] = f[q * N + id];
            rho[id] = T(1);
            ux[id] = uy[id] = uz[id] = T(0);
            continue;
          }
          T fc[19];
          T r, ux0, uy0, uz0;
          if (fl & ORC_INLET) {
            r = T(1);
            ux0 = inlet[(size_t)z * g.ny + y];
            uy0 = T(0);
            uz0 = T(0);
            for (int q = 0; q < 19; ++q) fc[q] = feq_bgk(W.w, q, r, ux0, uy0, uz0);
          } else {
            r = T(0);
            T jx = T(0), jy = T(0), jz = T(0);
            for (int q = 0; q < 19; ++q) {
              T fq = f[q * N + id];
              r += fq;
              jx += fq * T(CX[q]);
              jy += fq * T(CY[q]);
              jz += fq * T(CZ[q]);
            }
            ux0 = (r > T(0)) ? jx / r : T(0);
            uy0 = (r > T(0)) ? jy / r : T(0);
            uz0 = (r > T(0)) ? jz / r : T(0);
            ux0 += fx;
            uy0 += fy;
            uz0 += fz;
            // an outlet cell collides as its upstream neighbour (same flags)
            const uint8_t cfl = (fl & ORC_OUTLET) ? flags[id - 1] : fl;
            if (!hybrid) {
              const T omega = omegaL;
              for (int q = 0; q < 19; ++q) {
                T fe = feq_bgk(W.w, q, r, ux0, uy0, uz0);
                T fq = f[q * N + id];
                fc[q] = fq + omega * (fe - fq);
              }
            } else {
              T uu = ux0 * ux0 + uy0 * uy0 + uz0 * uz0;
              T feq[19];
              for (int q = 0; q < 19; ++q) {
                T eiu = T(CX[q]) * ux0 + T(CY[q]) * uy0 + T(CZ[q]) * uz0;
                feq[q] = W.w[q] * r * (T(1) + T(3) * eiu + T(4.5) * eiu * eiu - T(1.5) * uu);
              }
              const bool useLegacy = (cfl & ORC_BAND) != 0;
              for (int q = 0; q < 19; ++q) {
                const T fq = f[q * N + id];
                if (useLegacy) {
                  fc[q] = fq + omegaL * (feq[q] - fq);
                } else {
                  T dh = feq[q];
                  T ds = fq - feq[q];
                  fc[q] = fq + omegaH * (dh - fq) + (omegaS - omegaH) * (-ds);
                }
              }
            }
          }
          for (int q = 0; q < 19; ++q) {
            const size_t id2 = g.nb(x, y, z, CX[q], CY[q], CZ[q]);
            if (flags[id2] & ORC_SOLID)
              fnext[OPP[q] * N + id] = fc[q];
            else
              fnext[q * N + id2] = fc[q];
          }
          rho[id] = r;
          ux[id] = ux0;
          uy[id] = uy0;
          uz[id] = uz0;
        }
      }
    }
    f.swap(fnext);  // kern_swap exchanges the buffers element-wise
  }

  int step(int n) override {
    for (int i = 0; i < n; ++i) step1();
    return 0;
  }

  int get_macro(int which, float* o32, double* o64) const override {
    const std::vector<T>* src = which == 0 ? &rho : which == 1 ? &ux : which == 2 ? &uy : &uz;
    for (size_t i = 0; i < g.n; ++i) {
      if (o32) o32[i] = float((*src)[i]);
      if (o64) o64[i] = double((*src)[i]);
    }
    return 0;
  }
  size_t state_size() const override { return 19 * g.n; }
  int get_state(float* o32, double* o64) const override {
    for (size_t i = 0; i < 19 * g.n; ++i) {
      if (o32) o32[i] = float(f[i]);
      if (o64) o64[i] = double(f[i]);
    }
    return 0;
  }
  int set_state_f32(const float* in) override {
    for (size_t i = 0; i < 19 * g.n; ++i) f[i] = T(in[i]);
    return 0;
  }
};

// --------------------------------------------------------------------------
// HOME moment-only solver (lbm3d_home.cu).
// --------------------------------------------------------------------------
template <class T>
struct HomeSolver : SolverBase {
  Grid g;
  T tau, fx, fy, fz, omega;
  int bounce = 0;  // 0: opposite direction, 1: the reference's table
  Weights<T> W;
  std::vector<T> m, mnext;
  std::vector<uint8_t> flags;
  std::vector<T> inlet;

  HomeSolver(int nx, int ny, int nz, double tau_, double fx_, double fy_, double fz_) {
    g = {nx, ny, nz, (size_t)nx * ny * nz};
    tau = T(tau_);
    fx = T(fx_);
    fy = T(fy_);
    fz = T(fz_);
    omega = T(1) / tau;  // lbm3d_home.cu:340
    m.assign(10 * g.n, T(0));
    mnext.assign(10 * g.n, T(0));
    flags.assign(g.n, 0);
    inlet.assign((size_t)ny * nz, T(0));
    reset();
  }
  int set_flags(const uint8_t* fl) override {
    if (validate_flags(g, fl)) return -1;
    for (size_t i = 0; i < g.n; ++i)
      flags[i] = fl[i] & (ORC_SOLID | ORC_BAND | ORC_INLET | ORC_OUTLET);
    return 0;
  }
  int set_inlet(const float* u) override {
    for (size_t i = 0; i < inlet.size(); ++i) inlet[i] = T(u[i]);
    return 0;
  }
  int set_omegas(double, double, double) override { return -1; }
  int set_bounce(int rule) override {
    bounce = rule;
    return 0;
  }

  // kern_reset_mom (lbm3d_home.cu:83-107)
  int reset() override {
    for (size_t id = 0; id < g.n; ++id) {
      for (int k = 0; k < 10; ++k) {
        m[k * g.n + id] = k == 0 ? T(1) : T(0);
        mnext[k * g.n + id] = k == 0 ? T(1) : T(0);
      }
    }
    return 0;
  }
  // kern_reinit_from_macro (lbm3d_home.cu:109-137)
  int reinit_f32(const float* r, const float* u, const float* v, const float* w) override {
    for (size_t id = 0; id < g.n; ++id) {
      T vals[10] = {r ? T(r[id]) : T(1), u ? T(u[id]) : T(0), v ? T(v[id]) : T(0),
                    w ? T(w[id]) : T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
      for (int k = 0; k < 10; ++k) {
        m[k * g.n + id] = vals[k];
        mnext[k * g.n + id] = vals[k];
      }
    }
    return 0;
  }

  // kern_step_moments_only (lbm3d_home.cu:142-286) for the cell at (x,y,z)
  void gather(int x, int y, int z, T* out) const {
    const size_t N = g.n;
    const size_t id = g.id(x, y, z);
    const T* A = m.data();
    const T rhoC = A[0 * N + id], uxC0 = A[1 * N + id], uyC0 = A[2 * N + id],
            uzC0 = A[3 * N + id];
    const T SxxC = A[4 * N + id], SxyC = A[5 * N + id], SxzC = A[6 * N + id];
    const T SyyC = A[7 * N + id], SyzC = A[8 * N + id], SzzC = A[9 * N + id];
    const T oneMinusOmega = T(1) - omega;
    T r = T(0), jx = T(0), jy = T(0), jz = T(0);
    T Mxx = T(0), Mxy = T(0), Mxz = T(0), Myy = T(0), Myz = T(0), Mzz = T(0);
    for (int q = 0; q < 19; ++q) {
      const size_t ids = g.nb(x, y, z, -CX[q], -CY[q], -CZ[q]);
      T fq;
      if (flags[ids] & ORC_SOLID) {
        fq = home_rec(W.w, bounce ? OPP_REF[q] : OPP[q], rhoC, uxC0, uyC0, uzC0, SxxC, SxyC, SxzC, SyyC, SyzC, SzzC,
                      oneMinusOmega);
      } else {
        fq = home_rec(W.w, q, A[0 * N + ids], A[1 * N + ids], A[2 * N + ids], A[3 * N + ids],
                      A[4 * N + ids], A[5 * N + ids], A[6 * N + ids], A[7 * N + ids],
                      A[8 * N + ids], A[9 * N + ids], oneMinusOmega);
      }
      const T cx = T(CX[q]), cy = T(CY[q]), cz = T(CZ[q]);
      r += fq;
      jx += fq * cx;
      jy += fq * cy;
      jz += fq * cz;
      Mxx += fq * cx * cx;
      Mxy += fq * cx * cy;
      Mxz += fq * cx * cz;
      Myy += fq * cy * cy;
      Myz += fq * cy * cz;
      Mzz += fq * cz * cz;
    }
    T ux = (r > T(0)) ? (jx / r) : T(0);
    T uy = (r > T(0)) ? (jy / r) : T(0);
    T uz = (r > T(0)) ? (jz / r) : T(0);
    ux += fx;
    uy += fy;
    uz += fz;
    const T cs2 = T(1) / T(3);
    const T PiEq_xx = r * cs2 + r * ux * ux;
    const T PiEq_xy = r * ux * uy;
    const T PiEq_xz = r * ux * uz;
    const T PiEq_yy = r * cs2 + r * uy * uy;
    const T PiEq_yz = r * uy * uz;
    const T PiEq_zz = r * cs2 + r * uz * uz;
    out[0] = r;
    out[1] = ux;
    out[2] = uy;
    out[3] = uz;
    out[4] = Mxx - PiEq_xx;
    out[5] = Mxy - PiEq_xy;
    out[6] = Mxz - PiEq_xz;
    out[7] = Myy - PiEq_yy;
    out[8] = Myz - PiEq_yz;
    out[9] = Mzz - PiEq_zz;
  }

  void step1() {
    const size_t N = g.n;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < g.nz; ++z) {
      for (int y = 0; y < g.ny; ++y) {
        for (int x = 0; x < g.nx; ++x) {
          const size_t id = g.id(x, y, z);
          const uint8_t fl = flags[id];
          T out[10];
          if (fl & ORC_SOLID) {
            for (int k = 0; k < 10; ++k) out[k] = k == 0 ? T(1) : T(0);
          } else if (fl & ORC_INLET) {
            for (int k = 0; k < 10; ++k) out[k] = T(0);
            out[0] = T(1);
            out[1] = inlet[(size_t)z * g.ny + y];
          } else if (fl & ORC_OUTLET) {
            gather(x - 1, y, z, out);  // zero-gradient copy of x-1
          } else {
            gather(x, y, z, out);
          }
          for (int k = 0; k < 10; ++k) mnext[k * N + id] = out[k];
        }
      }
    }
    m.swap(mnext);  // pointer swap (lbm3d_home.cu:348-351)
  }
  int step(int n) override {
    for (int i = 0; i < n; ++i) step1();
    return 0;
  }
  int get_macro(int which, float* o32, double* o64) const override {
    const T* src = m.data() + (size_t)which * g.n;
    for (size_t i = 0; i < g.n; ++i) {
      if (o32) o32[i] = float(src[i]);
      if (o64) o64[i] = double(src[i]);
    }
    return 0;
  }
  size_t state_size() const override { return 10 * g.n; }
  int get_state(float* o32, double* o64) const override {
    for (size_t i = 0; i < 10 * g.n; ++i) {
      if (o32) o32[i] = float(m[i]);
      if (o64) o64[i] = double(m[i]);
    }
    return 0;
  }
  int set_state_f32(const float* in) override {
    for (size_t i = 0; i < 10 * g.n; ++i) m[i] = T(in[i]);
    return 0;
  }
};

// --------------------------------------------------------------------------
// Geometry: voxelize.cpp restated with a plain float3 (GLM's vec3 arithmetic
// is component-wise; glm::cross(x,y) = (x.y*y.z - y.y*x.z, x.z*y.x - y.z*x.x,
// x.x*y.y - y.x*x.y)).
// --------------------------------------------------------------------------
struct V3 {
  float x, y, z;
};
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
inline V3 vmin3(V3 a, V3 b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
inline V3 vmax3(V3 a, V3 b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }
inline float fmin3(float a, float b, float c) { return std::min(a, std::min(b, c)); }
inline float fmax3(float a, float b, float c) { return std::max(a, std::max(b, c)); }

// triBoxOverlap_SAT (voxelize.cpp:23-81)
bool tri_box_sat(V3 c, V3 half, V3 v0, V3 v1, V3 v2) {
  V3 tv0 = v0 - c, tv1 = v1 - c, tv2 = v2 - c;
  V3 triMin = vmin3(vmin3(tv0, tv1), tv2);
  V3 triMax = vmax3(vmax3(tv0, tv1), tv2);
  if (triMax.x < -half.x || triMin.x > half.x) return false;
  if (triMax.y < -half.y || triMin.y > half.y) return false;
  if (triMax.z < -half.z || triMin.z > half.z) return false;
  V3 e0 = tv1 - tv0, e1 = tv2 - tv1, e2 = tv0 - tv2;
  V3 n = cross(e0, e1);
  float r = half.x * std::fabs(n.x) + half.y * std::fabs(n.y) + half.z * std::fabs(n.z);
  float s = std::fabs(n.x * tv0.x + n.y * tv0.y + n.z * tv0.z);
  if (s > r) return false;
  const V3 axes[3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  const V3 edges[3] = {e0, e1, e2};
  for (int ei = 0; ei < 3; ++ei) {
    const V3 e = edges[ei];
    if (std::fabs(e.x) + std::fabs(e.y) + std::fabs(e.z) < 1e-12f) continue;
    for (int ai = 0; ai < 3; ++ai) {
      V3 L = cross(e, axes[ai]);
      float Llen2 = L.x * L.x + L.y * L.y + L.z * L.z;
      if (Llen2 < 1e-20f) continue;
      float p0 = L.x * tv0.x + L.y * tv0.y + L.z * tv0.z;
      float p1 = L.x * tv1.x + L.y * tv1.y + L.z * tv1.z;
      float p2 = L.x * tv2.x + L.y * tv2.y + L.z * tv2.z;
      float minp = std::min(p0, std::min(p1, p2));
      float maxp = std::max(p0, std::max(p1, p2));
      float rProj = half.x * std::fabs(L.x) + half.y * std::fabs(L.y) + half.z * std::fabs(L.z);
      if (maxp < -rProj || minp > rProj) return false;
    }
  }
  return true;
}

struct Tri {
  V3 v0, v1, v2;
};

// compute_fit_transform + make_transformed_tris (voxelize.cpp:83-113)
std::vector<Tri> transformed_tris(const float* xyz, int ntri, V3 bbmin, V3 bbmax, float scale,
                                  V3 translate) {
  V3 bbsize = bbmax - bbmin;
  float maxDim = std::max(bbsize.x, std::max(bbsize.y, bbsize.z));
  float s = (maxDim > 0.0f) ? (scale / maxDim) : 1.0f;
  V3 center = 0.5f * (bbmin + bbmax);
  V3 trans = translate - s * center;
  std::vector<Tri> tris((size_t)ntri);
  for (int t = 0; t < ntri; ++t) {
    V3 v[3];
    for (int k = 0; k < 3; ++k) {
      const float* p = xyz + 9 * (size_t)t + 3 * k;
      v[k] = s * V3{p[0], p[1], p[2]} + trans;
    }
    tris[t] = {v[0], v[1], v[2]};
  }
  return tris;
}

// voxelize_single_resolution (voxelize.cpp:115-185)
std::vector<uint8_t> voxelize_single(const std::vector<Tri>& tris, int Nx, int Ny, int Nz) {
  const size_t Ntot = (size_t)Nx * Ny * Nz;
  std::vector<uint8_t> surf(Ntot, 0);
  V3 half{0.5f / Nx, 0.5f / Ny, 0.5f / Nz};
  auto idx = [=](int x, int y, int z) { return (size_t)z * Ny * Nx + (size_t)y * Nx + (size_t)x; };
  const float eps = 1e-6f;
  for (size_t it = 0; it < tris.size(); ++it) {
    const Tri& t = tris[it];
    float minx = fmin3(t.v0.x, t.v1.x, t.v2.x) - eps;
    float miny = fmin3(t.v0.y, t.v1.y, t.v2.y) - eps;
    float minz = fmin3(t.v0.z, t.v1.z, t.v2.z) - eps;
    float maxx = fmax3(t.v0.x, t.v1.x, t.v2.x) + eps;
    float maxy = fmax3(t.v0.y, t.v1.y, t.v2.y) + eps;
    float maxz = fmax3(t.v0.z, t.v1.z, t.v2.z) + eps;
    int x0 = std::max(0, (int)std::floor(minx * Nx));
    int y0 = std::max(0, (int)std::floor(miny * Ny));
    int z0 = std::max(0, (int)std::floor(minz * Nz));
    int x1 = std::min(Nx - 1, (int)std::floor(maxx * Nx));
    int y1 = std::min(Ny - 1, (int)std::floor(maxy * Ny));
    int z1 = std::min(Nz - 1, (int)std::floor(maxz * Nz));
    for (int z = z0; z <= z1; ++z)
      for (int y = y0; y <= y1; ++y)
        for (int x = x0; x <= x1; ++x) {
          V3 c{(x + 0.5f) / Nx, (y + 0.5f) / Ny, (z + 0.5f) / Nz};
          if (tri_box_sat(c, half, t.v0, t.v1, t.v2)) surf[idx(x, y, z)] = 1;
        }
  }
  std::vector<uint8_t> vis(Ntot, 0);
  std::queue<size_t> q;
  auto push = [&](int x, int y, int z) {
    if (x < 0 || y < 0 || z < 0 || x >= Nx || y >= Ny || z >= Nz) return;
    size_t i = idx(x, y, z);
    if (vis[i] || surf[i]) return;
    vis[i] = 1;
    q.push(i);
  };
  for (int y = 0; y < Ny; ++y)
    for (int z = 0; z < Nz; ++z) {
      push(0, y, z);
      push(Nx - 1, y, z);
    }
  for (int x = 0; x < Nx; ++x)
    for (int z = 0; z < Nz; ++z) {
      push(x, 0, z);
      push(x, Ny - 1, z);
    }
  for (int x = 0; x < Nx; ++x)
    for (int y = 0; y < Ny; ++y) {
      push(x, y, 0);
      push(x, y, Nz - 1);
    }
  while (!q.empty()) {
    size_t i = q.front();
    q.pop();
    int x = (int)(i % Nx), y = (int)((i / Nx) % Ny), z = (int)(i / ((size_t)Nx * Ny));
    push(x + 1, y, z);
    push(x - 1, y, z);
    push(x, y + 1, z);
    push(x, y - 1, z);
    push(x, y, z + 1);
    push(x, y, z - 1);
  }
  std::vector<uint8_t> solid(Ntot, 0);
  for (size_t i = 0; i < Ntot; ++i)
    if (surf[i] || !vis[i]) solid[i] = 1;
  return solid;
}

// downsample_amr (voxelize.cpp:187-226)
std::vector<uint8_t> downsample(const std::vector<uint8_t>& fine, int refine, int Nx, int Ny,
                                int Nz, float threshold) {
  const int NxF = Nx * refine, NyF = Ny * refine;
  std::vector<uint8_t> coarse((size_t)Nx * Ny * Nz, 0);
  const int refine3 = refine * refine * refine;
  const float thr = std::max(0.0f, std::min(1.0f, threshold));
  for (int cz = 0; cz < Nz; ++cz)
    for (int cy = 0; cy < Ny; ++cy)
      for (int cx = 0; cx < Nx; ++cx) {
        int cnt = 0;
        for (int dz = 0; dz < refine; ++dz)
          for (int dy = 0; dy < refine; ++dy) {
            size_t base = (size_t)(cz * refine + dz) * NyF * NxF +
                          (size_t)(cy * refine + dy) * NxF + (size_t)(cx * refine);
            for (int dx = 0; dx < refine; ++dx) cnt += fine[base + dx];
          }
        float cov = static_cast<float>(cnt) / static_cast<float>(refine3);
        if ((cov > 0.0f) && (cov >= thr)) coarse[((size_t)cz * Ny + cy) * Nx + cx] = 1;
      }
  return coarse;
}

// parse_ascii / parse_binary (stl_loader.cpp:34-100)
bool parse_ascii(std::istream& is, std::vector<float>& xyz) {
  xyz.clear();
  std::string line;
  std::vector<float> tri;
  while (std::getline(is, line)) {
    std::istringstream ss(line);
    std::string tok;
    ss >> tok;
    if (tok == "vertex") {
      float x = 0, y = 0, z = 0;
      ss >> x >> y >> z;
      tri.push_back(x);
      tri.push_back(y);
      tri.push_back(z);
      if (tri.size() == 9) {
        xyz.insert(xyz.end(), tri.begin(), tri.end());
        tri.clear();
      }
    }
  }
  return !xyz.empty();
}
bool parse_binary(std::istream& is, std::vector<float>& xyz) {
  char header[80];
  if (!is.read(header, 80)) return false;
  uint32_t ntri = 0;
  if (!is.read(reinterpret_cast<char*>(&ntri), 4)) return false;
  xyz.clear();
  for (uint32_t i = 0; i < ntri; ++i) {
    char rec[50];
    if (!is.read(rec, 50)) return false;
    float v[12];
    std::memcpy(v, rec, 48);
    xyz.insert(xyz.end(), v + 3, v + 12);
  }
  return true;
}
// load_stl (stl_loader.cpp:102-122)
bool load_stl(const std::string& path, std::vector<float>& xyz) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  char head[6] = {};
  f.read(head, 5);
  f.clear();
  f.seekg(0);
  if (std::string(head) == "solid") {
    std::stringstream ss;
    ss << f.rdbuf();
    std::istringstream is(ss.str());
    if (parse_ascii(is, xyz)) return true;
    std::ifstream fb(path, std::ios::binary);
    if (!fb) return false;
    return parse_binary(fb, xyz);
  }
  return parse_binary(f, xyz);
}

}  // namespace

struct orc_solver {
  SolverBase* impl;
  int method;
};

extern "C" {

int orc_version(void) { return 1; }

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

orc_solver* orc_create(int method, int precision, int nx, int ny, int nz, double tau, double fx,
                       double fy, double fz) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || method < 0 || method > 2) return nullptr;
  SolverBase* impl = nullptr;
  if (method == ORC_HOME) {
    if (precision == 0)
      impl = new HomeSolver<float>(nx, ny, nz, tau, fx, fy, fz);
    else
      impl = new HomeSolver<double>(nx, ny, nz, tau, fx, fy, fz);
  } else {
    const bool hyb = method == ORC_HYBRID;
    if (precision == 0)
      impl = new PopSolver<float>(hyb, nx, ny, nz, tau, fx, fy, fz);
    else
      impl = new PopSolver<double>(hyb, nx, ny, nz, tau, fx, fy, fz);
  }
  return new orc_solver{impl, method};
}

void orc_destroy(orc_solver* s) {
  if (!s) return;
  delete s->impl;
  delete s;
}

int orc_set_flags(orc_solver* s, const uint8_t* fl) { return s->impl->set_flags(fl); }
int orc_set_inlet(orc_solver* s, const float* ux) { return s->impl->set_inlet(ux); }
int orc_set_omegas(orc_solver* s, double a, double b, double c) {
  return s->impl->set_omegas(a, b, c);
}
int orc_set_bounce(orc_solver* s, int rule) { return s->impl->set_bounce(rule); }
int orc_reset(orc_solver* s) { return s->impl->reset(); }

// Writers per fnext slot of one push step (kern_collide_stream's store
// pattern, lbm3d_legacy.cu:92-108) with the given partner table: a valid
// streaming + bounce-back update writes every fluid slot exactly once.
int orc_push_write_counts(const uint8_t* solid, int nx, int ny, int nz, int table,
                          int32_t* counts) {
  Grid g{nx, ny, nz, (size_t)nx * ny * nz};
  const int* opp = table ? OPP_REF : OPP;
  std::memset(counts, 0, sizeof(int32_t) * 19 * g.n);
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const size_t id = g.id(x, y, z);
        if (solid[id]) continue;
        for (int q = 0; q < 19; ++q) {
          const size_t id2 = g.nb(x, y, z, CX[q], CY[q], CZ[q]);
          if (solid[id2])
            counts[opp[q] * g.n + id] += 1;
          else
            counts[q * g.n + id2] += 1;
        }
      }
  return 0;
}
int orc_reinit_f32(orc_solver* s, const float* r, const float* u, const float* v, const float* w) {
  return s->impl->reinit_f32(r, u, v, w);
}
int orc_step(orc_solver* s, int n) { return s->impl->step(n); }
int orc_get_macro_f32(const orc_solver* s, int which, float* out) {
  if (which < 0 || which > 3) return -1;
  return s->impl->get_macro(which, out, nullptr);
}
int orc_get_macro_f64(const orc_solver* s, int which, double* out) {
  if (which < 0 || which > 3) return -1;
  return s->impl->get_macro(which, nullptr, out);
}
size_t orc_state_size(const orc_solver* s) { return s->impl->state_size(); }
int orc_get_state_f32(const orc_solver* s, float* out) { return s->impl->get_state(out, nullptr); }
int orc_get_state_f64(const orc_solver* s, double* out) {
  return s->impl->get_state(nullptr, out);
}
int orc_set_state_f32(orc_solver* s, const float* in) { return s->impl->set_state_f32(in); }

int orc_load_stl(const char* path, float* tri_xyz, int capacity, int* ntri, float* bbmin3,
                 float* bbmax3) {
  std::vector<float> xyz;
  if (!load_stl(path, xyz)) return -1;
  const int nt = (int)(xyz.size() / 9);
  if (ntri) *ntri = nt;
  // bounding box in vertex order (stl_loader.cpp:19-22)
  V3 mn{std::numeric_limits<float>::max(), std::numeric_limits<float>::max(),
        std::numeric_limits<float>::max()};
  V3 mx{-std::numeric_limits<float>::max(), -std::numeric_limits<float>::max(),
        -std::numeric_limits<float>::max()};
  for (size_t i = 0; i + 2 < xyz.size(); i += 3) {
    V3 p{xyz[i], xyz[i + 1], xyz[i + 2]};
    mn = vmin3(mn, p);
    mx = vmax3(mx, p);
  }
  if (bbmin3) {
    bbmin3[0] = mn.x;
    bbmin3[1] = mn.y;
    bbmin3[2] = mn.z;
  }
  if (bbmax3) {
    bbmax3[0] = mx.x;
    bbmax3[1] = mx.y;
    bbmax3[2] = mx.z;
  }
  if (tri_xyz) {
    if (capacity < nt) return -2;
    std::memcpy(tri_xyz, xyz.data(), xyz.size() * sizeof(float));
  }
  return 0;
}

int orc_voxelize(const float* tri_xyz, int ntri, const float* bbmin3, const float* bbmax3,
                 int nx, int ny, int nz, float uniform_scale, float tx, float ty, float tz,
                 int refine, float threshold, uint8_t* mask_out) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return -1;
  auto tris = transformed_tris(tri_xyz, ntri, {bbmin3[0], bbmin3[1], bbmin3[2]},
                               {bbmax3[0], bbmax3[1], bbmax3[2]}, uniform_scale, {tx, ty, tz});
  // voxelize_mesh_to_mask (voxelize.cpp:228-243)
  int r = refine < 1 ? 1 : refine;
  std::vector<uint8_t> out;
  if (r == 1) {
    out = voxelize_single(tris, nx, ny, nz);
  } else {
    auto fine = voxelize_single(tris, nx * r, ny * r, nz * r);
    out = downsample(fine, r, nx, ny, nz, threshold);
  }
  std::memcpy(mask_out, out.data(), out.size());
  return 0;
}

int orc_tri_box_overlap(const float* c, const float* h, const float* a, const float* b,
                        const float* d) {
  return tri_box_sat({c[0], c[1], c[2]}, {h[0], h[1], h[2]}, {a[0], a[1], a[2]},
                     {b[0], b[1], b[2]}, {d[0], d[1], d[2]})
             ? 1
             : 0;
}

// build_distance_to_solid_6n (sim_runner.hpp:28-78)
int orc_distance(const uint8_t* solid, int Nx, int Ny, int Nz, int maxDist, int32_t* dist) {
  const size_t N = (size_t)Nx * Ny * Nz;
  for (size_t i = 0; i < N; ++i) dist[i] = -1;
  if (maxDist <= 0) return 0;
  std::vector<size_t> q;
  for (size_t i = 0; i < N; ++i)
    if (solid[i]) {
      dist[i] = 0;
      q.push_back(i);
    }
  const size_t sy = (size_t)Nx, sz = (size_t)Nx * Ny;
  auto visit = [&](size_t nid, int nd) {
    if (dist[nid] == -1) {
      dist[nid] = nd;
      q.push_back(nid);
    }
  };
  size_t head = 0;
  while (head < q.size()) {
    const size_t id = q[head++];
    const int dcur = dist[id];
    if (dcur >= maxDist) continue;
    const int z = (int)(id / sz);
    const size_t rem = id - (size_t)z * sz;
    const int y = (int)(rem / sy);
    const int x = (int)(rem - (size_t)y * sy);
    const int nd = dcur + 1;
    if (x > 0) visit(id - 1, nd);
    if (x < Nx - 1) visit(id + 1, nd);
    if (y > 0) visit(id - sy, nd);
    if (y < Ny - 1) visit(id + sy, nd);
    if (z > 0) visit(id - sz, nd);
    if (z < Nz - 1) visit(id + sz, nd);
  }
  return 0;
}

// build_legacy_mask_from_distance (sim_runner.hpp:79-91) with the call site's
// maxDist = max(1, d0) (sim_runner.hpp:222)
int orc_band(const uint8_t* solid, int Nx, int Ny, int Nz, int d0, uint8_t* band) {
  const size_t N = (size_t)Nx * Ny * Nz;
  std::vector<int32_t> dist(N);
  orc_distance(solid, Nx, Ny, Nz, std::max(1, d0), dist.data());
  for (size_t i = 0; i < N; ++i) {
    band[i] = 0;
    if (solid[i]) continue;
    if (dist[i] >= 0 && dist[i] <= d0) band[i] = 1;
  }
  return 0;
}

// kern_rho_gauss (init_profiles.cu:7-19); std::exp stands in for __expf
int orc_rho_gauss(float* rho, int Nx, int Ny, int Nz, float cx, float cy, float cz, float sigma,
                  float amp) {
  const size_t N = (size_t)Nx * Ny * Nz;
  for (size_t i = 0; i < N; ++i) {
    int x = (int)(i % Nx), y = (int)((i / Nx) % Ny), z = (int)(i / ((size_t)Nx * Ny));
    float X = (x + 0.5f) / Nx, Y = (y + 0.5f) / Ny, Z = (z + 0.5f) / Nz;
    float dx = X - cx, dy = Y - cy, dz = Z - cz;
    float r2 = dx * dx + dy * dy + dz * dz;
    rho[i] = 1.0f + amp * std::exp(-r2 / (2.f * sigma * sigma));
  }
  return 0;
}
// kern_rho_slab (init_profiles.cu:20-29)
int orc_rho_slab(float* rho, int Nx, int Ny, int Nz, int axis, float amp, float center,
                 float width) {
  const size_t N = (size_t)Nx * Ny * Nz;
  for (size_t i = 0; i < N; ++i) {
    int x = (int)(i % Nx), y = (int)((i / Nx) % Ny), z = (int)(i / ((size_t)Nx * Ny));
    float s = (axis == 0) ? (x + 0.5f) / Nx : (axis == 1) ? (y + 0.5f) / Ny : (z + 0.5f) / Nz;
    float val = (std::fabs(s - center) < 0.5f * width) ? amp : 0.f;
    rho[i] = 1.0f + val;
  }
  return 0;
}
// kern_speed (field_ops.cu:12-18)
int orc_speed(const float* u, const float* v, const float* w, float* out, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    float uu = u[i] * u[i] + v[i] * v[i] + w[i] * w[i];
    out[i] = std::sqrt(uu);
  }
  return 0;
}

// write_scalar_vtk / write_vector_vtk (moment_vtk_writer.cu:9-49)
int orc_write_scalar_vtk(const char* path, const float* h, int Nx, int Ny, int Nz) {
  const size_t N = (size_t)Nx * Ny * Nz;
  std::ofstream f(path);
  if (!f) return -1;
  f << "# vtk DataFile Version 3.0\nscalar\nASCII\n"
    << "DATASET STRUCTURED_POINTS\n"
    << "DIMENSIONS " << Nx << " " << Ny << " " << Nz << "\n"
    << "ORIGIN 0 0 0\n"
    << "SPACING 1 1 1\n"
    << "POINT_DATA " << N << "\n"
    << "SCALARS s float 1\n"
    << "LOOKUP_TABLE default\n";
  for (size_t i = 0; i < N; ++i) f << h[i] << "\n";
  return 0;
}
int orc_write_vector_vtk(const char* path, const float* a, const float* b, const float* c, int Nx,
                         int Ny, int Nz) {
  const size_t N = (size_t)Nx * Ny * Nz;
  std::ofstream f(path);
  if (!f) return -1;
  f << "# vtk DataFile Version 3.0\nvector\nASCII\n"
    << "DATASET STRUCTURED_POINTS\n"
    << "DIMENSIONS " << Nx << " " << Ny << " " << Nz << "\n"
    << "ORIGIN 0 0 0\n"
    << "SPACING 1 1 1\n"
    << "POINT_DATA " << N << "\n"
    << "VECTORS vec float\n";
  for (size_t i = 0; i < N; ++i) f << a[i] << " " << b[i] << " " << c[i] << "\n";
  return 0;
}

}  // extern "C"
