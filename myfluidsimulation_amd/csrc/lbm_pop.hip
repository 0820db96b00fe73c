// lbm_pop.hip -- BGK (LBM3D_Legacy) and Hybrid(B0) (LBM3D_Hybrid) steps on gfx950.
//
// Streaming uses the AA pattern on ONE population buffer, updated in place,
// so each stored value is read once and written once per step (152 B per
// lattice update) instead of the reference's collide into a second buffer
// plus an element-wise swap (lbm3d_legacy.cu:59-118, about 473 B/LUP).
//
//   even step: cell c reads F[q][c], collides, writes f*_q to F[opp q][c]
//   odd  step: cell c reads its incoming f_q from F[opp q][c - e_q] (or
//              F[q][c] when c - e_q is solid: the bounced value), collides,
//              writes f*_q to F[q][c + e_q] (or F[opp q][c] when c + e_q is
//              solid: halfway bounce-back)
//
// Each slot is read and written by the same thread within a step, so the
// in-place update is race free.  After an even step the buffer holds the
// "swapped" representation; after an odd step the reference's natural one.
// The per-cell arithmetic is the reference's (kern_collide_stream,
// lbm3d_legacy.cu:78-108; kern_collide_stream_b0, lbm3d_hybrid.cu:114-168);
// only the storage slots differ, so fields match the reference rule bit for
// bit.
//
// Waves in which no lane touches a solid neighbour (flag bit F_NEAR, derived
// on the device) take a branch without any neighbour flag loads; the test is
// a wave-wide ballot, so the branch is uniform.
#include "d3q19_dev.hpp"

namespace fs {

namespace {

// BGK / hybrid collision in place: f -> post-collision, returns moments.
template <bool HYB>
__device__ __forceinline__ void collide(float (&f)[19], float& r, float& ux, float& uy,
                                        float& uz, const Phys& ph, bool band) {
  float rr = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
  for_q<0, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    rr += f[q];
    acc1<CX[q]>(jx, f[q]);
    acc1<CY[q]>(jy, f[q]);
    acc1<CZ[q]>(jz, f[q]);
  });
  float u0 = (rr > 0.0f) ? jx / rr : 0.0f;
  float v0 = (rr > 0.0f) ? jy / rr : 0.0f;
  float w0 = (rr > 0.0f) ? jz / rr : 0.0f;
  u0 += ph.fx;
  v0 += ph.fy;
  w0 += ph.fz;
  const float uu = u0 * u0 + v0 * v0 + w0 * w0;
  const float uu15 = 1.5f * uu;
  const WRho wr(rr);
  if (!HYB || band) {
    const float om = ph.omegaL;
    for_q<0, 19>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      const float fe = feq_q<q>(wr, u0, v0, w0, uu15);
      f[q] = f[q] + om * (fe - f[q]);
    });
  } else {
    const float oh = ph.omegaH;
    const float osh = ph.omegaS - ph.omegaH;
    for_q<0, 19>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      const float fe = feq_q<q>(wr, u0, v0, w0, uu15);
      const float ds = f[q] - fe;
      f[q] = (f[q] + oh * (fe - f[q])) + osh * (-ds);
    });
  }
  r = rr;
  ux = u0;
  uy = v0;
  uz = w0;
}

__device__ __forceinline__ void inlet_emit(float (&f)[19], float uin) {
  const WRho wr(1.0f);
  const float uu15 = 1.5f * (uin * uin + 0.0f * 0.0f + 0.0f * 0.0f);
  for_q<0, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    f[q] = feq_q<q>(wr, uin, 0.0f, 0.0f, uu15);
  });
}

// odd-step emission from origin cell `c` with neighbour offsets `o`
template <bool CHECK>
__device__ __forceinline__ void emit_odd(float* __restrict__ F, size_t P,
                                         const uint8_t* __restrict__ flags, uint32_t c,
                                         const Nbr& o, const float (&f)[19]) {
  for_q<0, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const uint32_t d = c + nb_off<CX[q], CY[q], CZ[q]>(o);
    if (CHECK && q != 0 && (flags[d] & F_SOLID)) {
      st_pop(&F[size_t(OPP[q]) * P + c], f[q]);
    } else {
      st_pop(&F[size_t(q) * P + d], f[q]);
    }
  });
}

template <bool ODD, bool HYB, bool MACRO>
__global__ void __launch_bounds__(256) k_pop_step(Geo g, Phys ph, PopLaunch a) {
  Cell cl;
  if (!cell_of_thread(g, cl)) return;
  const uint32_t c = cl.c;
  const size_t P = g.P;
  float* __restrict__ F = a.F;
  const uint8_t fl = a.flags[c];
  if (fl & (F_SOLID | F_OUTLET)) {
    // solid: populations frozen, macros (1,0,0,0) (lbm3d_legacy.cu:72-76);
    // outlet: written by its upstream feeder cell
    if (MACRO && (fl & F_SOLID)) {
      a.rho[c] = 1.0f;
      a.ux[c] = 0.0f;
      a.uy[c] = 0.0f;
      a.uz[c] = 0.0f;
    }
    return;
  }
  const Nbr o = nbr_offsets(cl.x, cl.y, cl.z, g);
  const bool anyNear = ODD && (__ballot((fl & F_NEAR) != 0) != 0);
  float f[19];
  float r, ux, uy, uz;
  if (fl & F_INLET) {
    const float uin = a.inlet[size_t(cl.z) * g.ny + cl.y];
    inlet_emit(f, uin);
    r = 1.0f;
    ux = uin;
    uy = 0.0f;
    uz = 0.0f;
  } else {
    if (!ODD) {
      for_q<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f[q] = ld_pop(&F[size_t(q) * P + c]);
      });
    } else if (!anyNear) {
      for_q<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const uint32_t s = c + nb_off<-CX[q], -CY[q], -CZ[q]>(o);
        f[q] = ld_pop(&F[size_t(OPP[q]) * P + s]);
      });
    } else {
      for_q<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const uint32_t s = c + nb_off<-CX[q], -CY[q], -CZ[q]>(o);
        if (q != 0 && (a.flags[s] & F_SOLID))
          f[q] = ld_pop(&F[size_t(q) * P + c]);
        else
          f[q] = ld_pop(&F[size_t(OPP[q]) * P + s]);
      });
    }
    collide<HYB>(f, r, ux, uy, uz, ph, (fl & F_BAND) != 0);
  }

  if (!ODD) {
    for_q<0, 19>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      st_pop(&F[size_t(OPP[q]) * P + c], f[q]);
    });
    if (fl & F_FEEDER) {
      for_q<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        st_pop(&F[size_t(OPP[q]) * P + c + 1], f[q]);
      });
    }
  } else {
    if (!anyNear)
      emit_odd<false>(F, P, a.flags, c, o, f);
    else
      emit_odd<true>(F, P, a.flags, c, o, f);
    if (fl & F_FEEDER) {
      // the outlet cell (x+1 = nx-1) collides exactly like its feeder
      Nbr o1 = o;
      o1.xm = -1;
      o1.xp = -(g.nx - 1);
      emit_odd<true>(F, P, a.flags, c + 1, o1, f);
    }
  }
  if (MACRO) {
    a.rho[c] = r;
    a.ux[c] = ux;
    a.uy[c] = uy;
    a.uz[c] = uz;
    if (fl & F_FEEDER) {
      a.rho[c + 1] = r;
      a.ux[c + 1] = ux;
      a.uy[c + 1] = uy;
      a.uz[c + 1] = uz;
    }
  }
}

// kern_reset (lbm3d_legacy.cu:40-53, lbm3d_hybrid.cu:43-56)
__global__ void __launch_bounds__(256) k_pop_reset(Geo g, PopLaunch a) {
  Cell cl;
  if (!cell_of_thread(g, cl)) return;
  const uint32_t c = cl.c;
  for_q<0, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    a.F[size_t(q) * g.P + c] = Dir<q>::w * 1.0f;
  });
  a.rho[c] = 1.0f;
  a.ux[c] = 0.0f;
  a.uy[c] = 0.0f;
  a.uz[c] = 0.0f;
}

// mode 0: LBM3D_Legacy::reinitEquilibriumFromMacro (lbm3d_legacy.cu:160-208)
//         -> kern_reinit_eq (lbm_reinit.cu:20-50), rho required
// mode 1: kern_reinit_equilibrium (lbm3d_hybrid.cu:58-87)
template <int MODE>
__global__ void __launch_bounds__(256) k_pop_reinit(Geo g, PopLaunch a, const float* rho,
                                                    const float* ux, const float* uy,
                                                    const float* uz) {
  Cell cl;
  if (!cell_of_thread(g, cl)) return;
  const uint32_t c = cl.c;
  const float r = (MODE == 1 && !rho) ? 1.0f : rho[c];
  const float u0 = ux ? ux[c] : 0.0f;
  const float v0 = uy ? uy[c] : 0.0f;
  const float w0 = uz ? uz[c] : 0.0f;
  a.rho[c] = r;
  a.ux[c] = u0;
  a.uy[c] = v0;
  a.uz[c] = w0;
  const WRho wr(r);
  const float uu = u0 * u0 + v0 * v0 + w0 * w0;
  const float uu15 = 1.5f * uu;
  for_q<0, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    a.F[size_t(q) * g.P + c] =
        MODE == 0 ? feq_cs2_q<q>(wr, u0, v0, w0, uu) : feq_q<q>(wr, u0, v0, w0, uu15);
  });
}

// Swapped AA representation -> reference order (incoming populations).
__global__ void __launch_bounds__(256) k_pop_canonical(Geo g, const float* __restrict__ F,
                                                       const uint8_t* __restrict__ flags,
                                                       int swapped, float* __restrict__ out) {
  Cell cl;
  if (!cell_of_thread(g, cl)) return;
  const uint32_t c = cl.c;
  const size_t P = g.P;
  const bool plain = !swapped || (flags[c] & F_SOLID);
  const Nbr o = nbr_offsets(cl.x, cl.y, cl.z, g);
  for_q<0, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    float v;
    if (plain) {
      v = F[size_t(q) * P + c];
    } else {
      const uint32_t s = c + nb_off<-CX[q], -CY[q], -CZ[q]>(o);
      v = (q != 0 && (flags[s] & F_SOLID)) ? F[size_t(q) * P + c] : F[size_t(OPP[q]) * P + s];
    }
    out[size_t(q) * P + c] = v;
  });
}

// Derived flag bits: F_NEAR (a periodic neighbour is solid) and F_FEEDER
// (the cell x+1 is an outlet).
__global__ void __launch_bounds__(256) k_derive_flags(Geo g, const uint8_t* __restrict__ in,
                                                      uint8_t* __restrict__ out) {
  Cell cl;
  if (!cell_of_thread(g, cl)) return;
  const uint32_t c = cl.c;
  uint8_t fl = in[c] & F_PUBLIC;
  const Nbr o = nbr_offsets(cl.x, cl.y, cl.z, g);
  bool near = false;
  for_q<1, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    near |= (in[c + nb_off<CX[q], CY[q], CZ[q]>(o)] & F_SOLID) != 0;
  });
  if (near) fl |= F_NEAR;
  if (cl.x + 1 < g.nx && (in[c + 1] & F_OUTLET)) fl |= F_FEEDER;
  out[c] = fl;
}

inline dim3 lattice_grid(const Geo& g) { return dim3((g.nxy + 255) / 256, g.nz, 1); }

}  // namespace

hipError_t launch_pop_step(const Geo& g, const Phys& ph, const PopLaunch& a, bool odd, bool hybrid,
                           bool write_macro, hipStream_t st) {
  const dim3 grid = lattice_grid(g), block(256);
  const int key = (odd ? 4 : 0) | (hybrid ? 2 : 0) | (write_macro ? 1 : 0);
  switch (key) {
    case 0: hipLaunchKernelGGL((k_pop_step<false, false, false>), grid, block, 0, st, g, ph, a); break;
    case 1: hipLaunchKernelGGL((k_pop_step<false, false, true>), grid, block, 0, st, g, ph, a); break;
    case 2: hipLaunchKernelGGL((k_pop_step<false, true, false>), grid, block, 0, st, g, ph, a); break;
    case 3: hipLaunchKernelGGL((k_pop_step<false, true, true>), grid, block, 0, st, g, ph, a); break;
    case 4: hipLaunchKernelGGL((k_pop_step<true, false, false>), grid, block, 0, st, g, ph, a); break;
    case 5: hipLaunchKernelGGL((k_pop_step<true, false, true>), grid, block, 0, st, g, ph, a); break;
    case 6: hipLaunchKernelGGL((k_pop_step<true, true, false>), grid, block, 0, st, g, ph, a); break;
    default: hipLaunchKernelGGL((k_pop_step<true, true, true>), grid, block, 0, st, g, ph, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_pop_reset(const Geo& g, const PopLaunch& a, hipStream_t st) {
  hipLaunchKernelGGL(k_pop_reset, lattice_grid(g), dim3(256), 0, st, g, a);
  return hipGetLastError();
}

hipError_t launch_pop_reinit(const Geo& g, const PopLaunch& a, int mode, const float* rho,
                             const float* ux, const float* uy, const float* uz, hipStream_t st) {
  if (mode == 0)
    hipLaunchKernelGGL(k_pop_reinit<0>, lattice_grid(g), dim3(256), 0, st, g, a, rho, ux, uy, uz);
  else
    hipLaunchKernelGGL(k_pop_reinit<1>, lattice_grid(g), dim3(256), 0, st, g, a, rho, ux, uy, uz);
  return hipGetLastError();
}

hipError_t launch_pop_canonical(const Geo& g, const float* F, const uint8_t* flags, bool swapped,
                                float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_pop_canonical, lattice_grid(g), dim3(256), 0, st, g, F, flags,
                     swapped ? 1 : 0, out);
  return hipGetLastError();
}

hipError_t launch_derive_flags(const Geo& g, const uint8_t* in, uint8_t* out, hipStream_t st) {
  hipLaunchKernelGGL(k_derive_flags, lattice_grid(g), dim3(256), 0, st, g, in, out);
  return hipGetLastError();
}

}  // namespace fs
