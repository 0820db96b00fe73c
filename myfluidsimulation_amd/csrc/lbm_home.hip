// lbm_home.hip -- HOME moment-only step (LBM3D_Home) on gfx950.
//
// State: 10 moments per cell (rho, u, S = Pi - Pi_eq), two buffers, pointer
// swap per step (lbm3d_home.cu:288-354).  A cell gathers, for each of the 19
// directions, the regularised post-collision population of its upstream
// neighbour, rebuilt from that neighbour's moments (reconstruct_f_post,
// lbm3d_home.cu:37-81), or one of its own populations when the neighbour is
// solid (halfway bounce-back, lbm3d_home.cu:219-225: the opposite direction
// by default, the reference's own pairing table on request), and accumulates
// the new moments (lbm3d_home.cu:238-285).  Per-direction arithmetic is specialised
// at compile time (d3q19_dev.hpp) and keeps the reference's float order.
#include "d3q19_dev.hpp"

namespace fs {

namespace {

__device__ __forceinline__ void load_mom(const float* __restrict__ m, size_t P, uint32_t s,
                                         Mom& v) {
#pragma unroll
  for (int k = 0; k < 10; ++k) v.v[k] = m[size_t(k) * P + s];
}

template <bool CHECK, int BB>
__device__ __forceinline__ void home_gather(const float* __restrict__ m, size_t P,
                                            const uint8_t* __restrict__ flags, uint32_t cc,
                                            const Nbr& o, const Phys& ph, float oneMinusOmega,
                                            float (&out)[10]) {
  Mom mc;
  load_mom(m, P, cc, mc);
  float r = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
  float Mxx = 0.0f, Mxy = 0.0f, Mxz = 0.0f, Myy = 0.0f, Myz = 0.0f, Mzz = 0.0f;
  for_q<0, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int a = CX[q], b = CY[q], c = CZ[q];
    float fq;
    if constexpr (q == 0) {
      fq = home_rec<0>(mc, oneMinusOmega);
    } else {
      const uint32_t s = cc + nb_off<-a, -b, -c>(o);
      if (CHECK && (flags[s] & F_SOLID)) {
        fq = home_rec<BB == 0 ? OPP[q] : OPP_REF[q]>(mc, oneMinusOmega);
      } else {
        Mom ms;
        load_mom(m, P, s, ms);
        fq = home_rec<q>(ms, oneMinusOmega);
      }
    }
    r += fq;
    acc1<a>(jx, fq);
    acc1<b>(jy, fq);
    acc1<c>(jz, fq);
    acc1<a * a>(Mxx, fq);
    acc1<a * b>(Mxy, fq);
    acc1<a * c>(Mxz, fq);
    acc1<b * b>(Myy, fq);
    acc1<b * c>(Myz, fq);
    acc1<c * c>(Mzz, fq);
  });
  float ux = (r > 0.0f) ? (jx / r) : 0.0f;
  float uy = (r > 0.0f) ? (jy / r) : 0.0f;
  float uz = (r > 0.0f) ? (jz / r) : 0.0f;
  ux += ph.fx;
  uy += ph.fy;
  uz += ph.fz;
  constexpr float cs2 = 1.0f / 3.0f;
  const float rcs2 = r * cs2;
  out[0] = r;
  out[1] = ux;
  out[2] = uy;
  out[3] = uz;
  out[4] = Mxx - (rcs2 + (r * ux) * ux);
  out[5] = Mxy - (r * ux) * uy;
  out[6] = Mxz - (r * ux) * uz;
  out[7] = Myy - (rcs2 + (r * uy) * uy);
  out[8] = Myz - (r * uy) * uz;
  out[9] = Mzz - (rcs2 + (r * uz) * uz);
}

template <int BB>
__global__ void __launch_bounds__(256) k_home_step(Geo g, Phys ph, HomeLaunch a) {
  Cell cl;
  if (!cell_of_thread(g, cl)) return;
  const uint32_t c = cl.c;
  const size_t P = g.P;
  const uint8_t fl = a.flags[c];
  float out[10];
  if (fl & F_SOLID) {
    // lbm3d_home.cu:180-188
#pragma unroll
    for (int k = 0; k < 10; ++k) out[k] = k == 0 ? 1.0f : 0.0f;
  } else if (fl & F_INLET) {
#pragma unroll
    for (int k = 0; k < 10; ++k) out[k] = 0.0f;
    out[0] = 1.0f;
    out[1] = a.inlet[size_t(cl.z) * g.ny + cl.y];
  } else {
    // an outlet cell stores the new moments of its upstream neighbour x-1
    const bool outlet = (fl & F_OUTLET) != 0;
    const int xc = outlet ? cl.x - 1 : cl.x;
    const uint32_t cc = outlet ? c - 1 : c;
    const uint8_t flc = outlet ? a.flags[cc] : fl;
    const Nbr o = nbr_offsets(xc, cl.y, cl.z, g);
    const float omm = 1.0f - ph.omegaL;
    if (__ballot((flc & F_NEAR) != 0) == 0)
      home_gather<false, BB>(a.m, P, a.flags, cc, o, ph, omm, out);
    else
      home_gather<true, BB>(a.m, P, a.flags, cc, o, ph, omm, out);
  }
#pragma unroll
  for (int k = 0; k < 10; ++k) a.mn[size_t(k) * P + c] = out[k];
}

// kern_reset_mom / kern_reinit_from_macro (lbm3d_home.cu:83-137): both buffers
__global__ void __launch_bounds__(256) k_home_fill(Geo g, float* m, float* mn, const float* rho,
                                                   const float* ux, const float* uy,
                                                   const float* uz) {
  Cell cl;
  if (!cell_of_thread(g, cl)) return;
  const uint32_t c = cl.c;
  float v[10] = {rho ? rho[c] : 1.0f, ux ? ux[c] : 0.0f, uy ? uy[c] : 0.0f, uz ? uz[c] : 0.0f,
                 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    m[size_t(k) * g.P + c] = v[k];
    mn[size_t(k) * g.P + c] = v[k];
  }
}

inline dim3 lattice_grid(const Geo& g) { return dim3((g.nxy + 255) / 256, g.nz, 1); }

}  // namespace

hipError_t launch_home_step(const Geo& g, const Phys& ph, const HomeLaunch& a, hipStream_t st) {
  if (ph.bounce == 0)
    hipLaunchKernelGGL(k_home_step<0>, lattice_grid(g), dim3(256), 0, st, g, ph, a);
  else
    hipLaunchKernelGGL(k_home_step<1>, lattice_grid(g), dim3(256), 0, st, g, ph, a);
  return hipGetLastError();
}

hipError_t launch_home_fill(const Geo& g, float* m, float* mn, const float* rho, const float* ux,
                            const float* uy, const float* uz, hipStream_t st) {
  hipLaunchKernelGGL(k_home_fill, lattice_grid(g), dim3(256), 0, st, g, m, mn, rho, ux, uy, uz);
  return hipGetLastError();
}

}  // namespace fs
