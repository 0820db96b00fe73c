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
#include "home_pair.hpp"

namespace fs {

namespace {

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

template <int TX, int TY, int ZC, int BB>
__global__ void __launch_bounds__(TX * TY) k_home_tiled(Geo g, Phys ph, HomeLaunch a) {
  __shared__ float fs[19 * TileLds<TX, TY>::PL];
  const int tilesX = (g.nx + TX - 1) / TX;
  // Workgroups the dispatcher sends to one XCD (every 8th) get one
  // contiguous run of tiles, x fastest, then y, then the z-chunk: the tiles
  // an XCD runs at the same time are neighbours in one z-slab, so the halo
  // rows they share are read once into that XCD's L2.
  uint32_t L0 = blockIdx.x;
  const uint32_t total = gridDim.x;
  if ((total & 7u) == 0) L0 = (L0 & 7u) * (total >> 3) + (L0 >> 3);
  const int tilesY = (g.ny + TY - 1) / TY;
  const int bx = int(L0 % uint32_t(tilesX));
  const int rest = int(L0 / uint32_t(tilesX));
  home_tile<TX, TY, ZC, BB>(g, ph, a, bx, rest % tilesY, rest / tilesY, fs);
}

template <int TX, int TY, int ZC, int BB>
__global__ void __launch_bounds__(TX * TY) k_home_tiled2(Geo g, Phys ph, HomeLaunch a) {
  __shared__ float fs[16 * TileLds16<TX, TY>::PL];
  const int tilesX = (g.nx + TX - 1) / TX;
  // the XCD-contiguous tile order of k_home_tiled
  uint32_t L0 = blockIdx.x;
  const uint32_t total = gridDim.x;
  if ((total & 7u) == 0) L0 = (L0 & 7u) * (total >> 3) + (L0 >> 3);
  const int tilesY = (g.ny + TY - 1) / TY;
  const int bx = int(L0 % uint32_t(tilesX));
  const int rest = int(L0 / uint32_t(tilesX));
  home_tile2<TX, TY, ZC, BB>(g, ph, a, bx, rest % tilesY, rest / tilesY, fs);
}

template <int TX, int TY, int ZC, int BB>
__global__ void __launch_bounds__(TX * TY) k_home_tiled3(Geo g, Phys ph, HomeLaunch a) {
  __shared__ float fs[16 * TileLds16<TX, TY>::PL];
  const int tilesX = (g.nx + TX - 1) / TX;
  // the XCD-contiguous tile order of k_home_tiled
  uint32_t L0 = blockIdx.x;
  const uint32_t total = gridDim.x;
  if ((total & 7u) == 0) L0 = (L0 & 7u) * (total >> 3) + (L0 >> 3);
  const int tilesY = (g.ny + TY - 1) / TY;
  const int bx = int(L0 % uint32_t(tilesX));
  const int rest = int(L0 / uint32_t(tilesX));
  home_tile3<TX, TY, ZC, BB>(g, ph, a, bx, rest % tilesY, rest / tilesY, fs);
}

template <int TY, int ZC, int BB>
__global__ void __launch_bounds__(PairLds<TY>::THREADS) k_home_pair(Geo g, Phys ph, HomeLaunch a) {
  __shared__ float lds[PairLds<TY>::FLOATS];
  const int tilesX = (g.nx + 63) / 64;
  // the XCD-contiguous tile order of k_home_tiled
  uint32_t L0 = blockIdx.x;
  const uint32_t total = gridDim.x;
  if ((total & 7u) == 0) L0 = (L0 & 7u) * (total >> 3) + (L0 >> 3);
  const int tilesY = (g.ny + TY - 1) / TY;
  const int bx = int(L0 % uint32_t(tilesX));
  const int rest = int(L0 / uint32_t(tilesX));
  home_pair_tile<TY, ZC, BB>(g, ph, a, bx, rest % tilesY, rest / tilesY, lds);
}

template <int TY, int ZC>
hipError_t launch_pair(const Geo& g, const Phys& ph, const HomeLaunch& a, hipStream_t st) {
  const unsigned blocks =
      unsigned(((g.nx + 63) / 64) * ((g.ny + TY - 1) / TY) * ((g.nz + ZC - 1) / ZC));
  if (ph.bounce == 0)
    hipLaunchKernelGGL((k_home_pair<TY, ZC, 0>), dim3(blocks), dim3(PairLds<TY>::THREADS), 0, st, g, ph,
                       a);
  else
    hipLaunchKernelGGL((k_home_pair<TY, ZC, 1>), dim3(blocks), dim3(PairLds<TY>::THREADS), 0, st, g, ph,
                       a);
  return hipGetLastError();
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

template <int TY, int ZC, int FORM = 1>
hipError_t launch_tiled(const Geo& g, const Phys& ph, const HomeLaunch& a, hipStream_t st) {
  constexpr int TX = 64;
  const unsigned blocks =
      unsigned(((g.nx + TX - 1) / TX) * ((g.ny + TY - 1) / TY) * ((g.nz + ZC - 1) / ZC));
  if constexpr (FORM == 3) {
    if (ph.bounce == 0)
      hipLaunchKernelGGL((k_home_tiled3<TX, TY, ZC, 0>), dim3(blocks), dim3(TX * TY), 0, st, g,
                         ph, a);
    else
      hipLaunchKernelGGL((k_home_tiled3<TX, TY, ZC, 1>), dim3(blocks), dim3(TX * TY), 0, st, g,
                         ph, a);
  } else if constexpr (FORM == 2) {
    if (ph.bounce == 0)
      hipLaunchKernelGGL((k_home_tiled2<TX, TY, ZC, 0>), dim3(blocks), dim3(TX * TY), 0, st, g,
                         ph, a);
    else
      hipLaunchKernelGGL((k_home_tiled2<TX, TY, ZC, 1>), dim3(blocks), dim3(TX * TY), 0, st, g,
                         ph, a);
  } else {
    if (ph.bounce == 0)
      hipLaunchKernelGGL((k_home_tiled<TX, TY, ZC, 0>), dim3(blocks), dim3(TX * TY), 0, st, g, ph,
                         a);
    else
      hipLaunchKernelGGL((k_home_tiled<TX, TY, ZC, 1>), dim3(blocks), dim3(TX * TY), 0, st, g, ph,
                         a);
  }
  return hipGetLastError();
}

}  // namespace

bool home_variant_ok(const Geo& g, int variant) {
  if (variant == kHomePlain) return true;
  if (variant < kHomePlain || variant >= kHomeVariants) return false;
  // the third form and the two-step tile address a moment plane with 32-bit
  // byte offsets
  if (variant >= kHome3T4Z16 && uint64_t(g.P) * sizeof(float) >= (uint64_t(1) << 32)) return false;
  // the tiled kernels take an outlet's moments from the lane at x-1, which
  // must sit in the same 64-wide tile: not so when nx-1 starts a tile
  return (g.nx - 1) % 64 != 0;
}

namespace {

// The one-step default: the plain gather where the grid is too small for the
// tiles, else the third tiled form.
int home_one_step_default(const Geo& g) {
  if (!home_variant_ok(g, kHome2T8Z32)) return kHomePlain;
  // a workgroup walks its whole z-chunk: below 512 tiles of 64x4x16 (about
  // 128^3, measured) too few of them fill the chip and the per-cell gather
  // is faster
  const long cols = (g.nx + 63) / 64;
  const long tiles4 = cols * ((g.ny + 3) / 4) * ((g.nz + 15) / 16);
  if (tiles4 < 512) return kHomePlain;
  // the third tiled form (fewer instructions, bounce-back values from its own
  // emission) with 32-plane chunks where they still leave >= 1024 workgroups
  // (measured against 64-plane chunks, three runs each: 512^3 57.4-57.9
  // against 57.4-58.1 GLUPS, 768x256x256 channel 53.0-54.4 against
  // 53.0-53.2, 256^3 56.3-58.0 against 55.4-57.6), else 64x4x16 tiles
  // (384x128x128 channel 46.8-47.6 GLUPS against 42.7-43.2 for the second
  // form)
  const long tiles8 = cols * ((g.ny + 7) / 8);
  int v = kHome3T4Z16;
  if (g.nz >= 128 && tiles8 * ((g.nz + 31) / 32) >= 1024) v = kHome3T8Z32;
  // 2^30 cells or more: the second form (64-bit addresses) in the same shape
  if (!home_variant_ok(g, v)) v = v == kHome3T8Z32 ? kHome2T8Z32 : kHome2T4Z16;
  return v;
}

}  // namespace

int home_default_variant(const Geo& g) {
  const int v = home_one_step_default(g);
  if (v == kHomePlain || !home_variant_ok(g, kHomeP11Z64)) return v;
  // The two-step tile (1024-thread workgroups, one per CU) where it gives
  // enough of them (measured, two runs each):
  //  * 64x12 columns with 64-plane chunks from 2048 workgroups (512^3: 2752;
  //    71.3 GLUPS against 70.7 for 64x11x64);
  //  * 64x11 columns with 64-plane chunks from 2048 workgroups;
  //  * 64x11 columns with 32-plane chunks from 768 (768x256x256 channel: 2304,
  //    64.4 GLUPS against 61.9 for 64x12x32; 256^3: 768, three full rounds,
  //    65.6 against 57.0 for the third form and 62.0 for 64x12x32).
  const long cols = (g.nx + 63) / 64;
  const long tiles11 = cols * ((g.ny + 10) / 11), tiles12 = cols * ((g.ny + 11) / 12);
  if (tiles12 * ((g.nz + 63) / 64) >= 2048) return kHomeP12Z64;
  if (tiles11 * ((g.nz + 63) / 64) >= 2048) return kHomeP11Z64;
  if (tiles11 * ((g.nz + 31) / 32) >= 768) return kHomeP11Z32;
  return v;
}

int home_single_variant(const Geo& g, int variant) {
  return home_variant_pair(variant) ? home_one_step_default(g) : variant;
}

hipError_t launch_home_pair(const Geo& g, const Phys& ph, const HomeLaunch& a, int variant,
                            hipStream_t st) {
  switch (variant) {
    case kHomeP11Z64: return launch_pair<11, 64>(g, ph, a, st);
    case kHomeP11Z32: return launch_pair<11, 32>(g, ph, a, st);
    case kHomeP10Z64: return launch_pair<10, 64>(g, ph, a, st);
    case kHomeP12Z64: return launch_pair<12, 64>(g, ph, a, st);
    case kHomeP12Z32: return launch_pair<12, 32>(g, ph, a, st);
    default: break;
  }
  return hipErrorInvalidValue;
}

hipError_t launch_home_step(const Geo& g, const Phys& ph, const HomeLaunch& a, int variant,
                            hipStream_t st) {
  if (variant < 0 || !home_variant_ok(g, variant)) variant = home_default_variant(g);
  variant = home_single_variant(g, variant);
  if (!home_variant_ok(g, variant)) variant = kHomePlain;
  switch (variant) {
    case kHomeT4Z16: return launch_tiled<4, 16>(g, ph, a, st);
    case kHomeT8Z16: return launch_tiled<8, 16>(g, ph, a, st);
    case kHomeT8Z32: return launch_tiled<8, 32>(g, ph, a, st);
    case kHome2T4Z16: return launch_tiled<4, 16, 2>(g, ph, a, st);
    case kHome2T8Z16: return launch_tiled<8, 16, 2>(g, ph, a, st);
    case kHome2T8Z32: return launch_tiled<8, 32, 2>(g, ph, a, st);
    case kHome2T16Z32: return launch_tiled<16, 32, 2>(g, ph, a, st);
    case kHome2T8Z64: return launch_tiled<8, 64, 2>(g, ph, a, st);
    case kHome3T4Z16: return launch_tiled<4, 16, 3>(g, ph, a, st);
    case kHome3T8Z32: return launch_tiled<8, 32, 3>(g, ph, a, st);
    case kHome3T8Z64: return launch_tiled<8, 64, 3>(g, ph, a, st);
    default: break;
  }
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
