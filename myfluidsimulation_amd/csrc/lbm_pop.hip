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
#include "pop_body.hpp"

namespace fs {

namespace {

template <bool ODD, bool HYB, bool MACRO>
__global__ void __launch_bounds__(256) k_pop_step(Geo g, Phys ph, PopLaunch a) {
  Cell cl;
  if (!cell_of_thread(g, cl)) return;
  pop_cell<ODD, HYB, MACRO>(g, ph, a, cl);
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
// (the cell x+1 is an outlet); and the solid-neighbour mask (bit q: the
// neighbour along e_q is solid), so a step near walls reads one word per
// cell instead of 18 neighbour flag bytes.
__global__ void __launch_bounds__(256) k_derive_flags(Geo g, const uint8_t* __restrict__ in,
                                                      uint8_t* __restrict__ out,
                                                      uint32_t* __restrict__ nbm) {
  Cell cl;
  if (!cell_of_thread(g, cl)) return;
  const uint32_t c = cl.c;
  uint8_t fl = in[c] & F_PUBLIC;
  const Nbr o = nbr_offsets(cl.x, cl.y, cl.z, g);
  uint32_t mask = 0;
  for_q<1, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    if (in[c + nb_off<CX[q], CY[q], CZ[q]>(o)] & F_SOLID) mask |= 1u << q;
  });
  if (mask) fl |= F_NEAR;
  if (nbm) nbm[c] = mask;
  if (cl.x + 1 < g.nx && (in[c + 1] & F_OUTLET)) fl |= F_FEEDER;
  // an outlet (x >= 1) collides as its upstream neighbour, band bit included
  if ((fl & F_OUTLET) && cl.x >= 1) fl = (fl & ~F_BAND) | (in[c - 1] & F_BAND);
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

hipError_t launch_derive_flags(const Geo& g, const uint8_t* in, uint8_t* out, uint32_t* nbm,
                               hipStream_t st) {
  hipLaunchKernelGGL(k_derive_flags, lattice_grid(g), dim3(256), 0, st, g, in, out, nbm);
  return hipGetLastError();
}

}  // namespace fs
