// lbm_pair.hip -- two BGK (LBM3D_Legacy) or Hybrid(B0) time steps per pass.
//
// The one-step kernels move 19 floats in and 19 out per cell and step (152 B
// per lattice update), so at HBM speed the step time is fixed by bytes.  This
// kernel reads f^t once and writes f^{t+2} once: 76 B per lattice update.
// Step t's post-collision populations live only in LDS.
//
// A workgroup owns a TX x TY column tile (TX = 64 or 32) and a chunk of ZC
// planes and walks the planes z0-1 .. z0+ZC (the tile body is pair_tile in
// pair_body.hpp, shared with the batched kernel):
//   1. every tile cell, and the ring of cells around the tile, reads its f^t
//      (own slots, reference order) and collides (step t); the tile cells put
//      all 19 post-collision values into LDS, the ring cells only the five
//      directions that enter the tile;
//   2. each tile cell picks up the values that stream into its column from
//      LDS: the f^{t+1}_q of the cell at c - e_q arrive from planes z-1 (e_z =
//      +1), z and z+1 (e_z = -1), so they wait in a register ring for one
//      plane; where c - e_q is solid the cell's own f*_{opp q} is taken
//      instead (halfway bounce-back);
//   3. when plane z+1 has been read the cell collides again (step t+1) and
//      pushes f*_q to slot q of c + e_q in the output buffer, or to slot
//      opp(q) of c when c + e_q is solid, as kern_collide_stream does
//      (lbm3d_legacy.cu:78-108, lbm3d_hybrid.cu:114-168).
// The arithmetic per cell is the one-step kernels' (PopCollision), so two
// launches of those equal one launch of this bit for bit.
//
// Boundaries follow the oracle's rule for the channel extension: an inlet
// cell emits feq(1, (u_in, 0, 0)) whatever streams into it (it runs the fluid
// code on zero inputs with PopCollision::set_inlet, which gives exactly that
// feq); an outlet cell collides as its upstream neighbour x-1 (in step 1 it
// reads that cell's f^t, in step 2 it takes the neighbour lane's gathered
// f^{t+1}), so an outlet must not start a tile row (the caller checks).
// Solid cells are never written: their frozen slots must already be equal in
// both buffers.
#include "pair_body.hpp"

namespace fs {

namespace {

template <int TX, int TY, int ZC, bool HYB, bool MACRO>
__global__ void __launch_bounds__(TX * TY) k_pop_pair(Geo g, Phys ph, PairLaunch a) {
  __shared__ float lds[19 * TileLds<TX, TY>::PL];
  // tile order: as k_home_tiled (x fastest, then y, then the z-chunk; each
  // XCD gets one contiguous run)
  const int tilesX = (g.nx + TX - 1) / TX, tilesY = (g.ny + TY - 1) / TY;
  uint32_t L0 = blockIdx.x;
  const uint32_t total = gridDim.x;
  if ((total & 7u) == 0) L0 = (L0 & 7u) * (total >> 3) + (L0 >> 3);
  const int bx = int(L0 % uint32_t(tilesX));
  const int rest = int(L0 / uint32_t(tilesX));
  pair_tile<TX, TY, ZC, HYB, MACRO>(g, ph, a, bx, rest % tilesY, rest / tilesY, lds);
}

template <int TX, int TY, int ZC>
hipError_t launch_pair(const Geo& g, const Phys& ph, const PairLaunch& a, bool hybrid, bool macro,
                       hipStream_t st) {
  const unsigned blocks =
      unsigned(((g.nx + TX - 1) / TX) * ((g.ny + TY - 1) / TY) * ((g.nz + ZC - 1) / ZC));
  const dim3 grid(blocks), block(TX * TY);
  const int key = (hybrid ? 2 : 0) | (macro ? 1 : 0);
  switch (key) {
    case 0: hipLaunchKernelGGL((k_pop_pair<TX, TY, ZC, false, false>), grid, block, 0, st, g, ph, a); break;
    case 1: hipLaunchKernelGGL((k_pop_pair<TX, TY, ZC, false, true>), grid, block, 0, st, g, ph, a); break;
    case 2: hipLaunchKernelGGL((k_pop_pair<TX, TY, ZC, true, false>), grid, block, 0, st, g, ph, a); break;
    default: hipLaunchKernelGGL((k_pop_pair<TX, TY, ZC, true, true>), grid, block, 0, st, g, ph, a); break;
  }
  return hipGetLastError();
}

}  // namespace

int pop_variant_tile_width(int variant) { return variant == kPairX32T8Z16 ? 32 : 64; }

int pop_default_variant(const Geo& g) {
  // A workgroup walks its whole z-chunk, so a small grid gives too few of
  // them to fill 256 CUs: below 512 tiles of 64x4x16 (measured crossover,
  // about 128^3) the one-step kernel, with one workgroup per 256 cells, is
  // faster.
  const long cols = (g.nx + 63) / 64;
  const long tiles4 = cols * ((g.ny + 3) / 4) * ((g.nz + 15) / 16);
  if (tiles4 < 512) return kPopAA;
  // 64x12 tiles (768 threads, one workgroup per CU at three waves per SIMD)
  // re-collide a ring of 156 cells for 768 (64x4: 140 for 256) and re-read
  // the fewest halo rows, but with one workgroup per CU the last round of
  // workgroups must not be nearly empty: 64-plane chunks where they give at
  // least 2048 workgroups (512^3: 2752, 53.3-53.9 GLUPS against 43.3-43.6
  // for 64x4x16), 32-plane chunks from 512 (768x256x256 channel, 2112: 45.5-45.9
  // against 43.5-44.5; 256^3, 704: 42.7-43.1 against 41.0 for 64x8x32)
  const long rows12 = (g.ny + 11) / 12;
  if (cols * rows12 * ((g.nz + 63) / 64) >= 2048) return kPairT12Z64;
  if (cols * rows12 * ((g.nz + 31) / 32) >= 512) return kPairT12Z32;
  // Otherwise 64x4x16 tiles (three workgroups per CU): the 384x128x128
  // channel (264 workgroups of 64x12x32: 28 GLUPS) takes them (39.0).
  return kPairT4Z16;
}

hipError_t launch_pop_pair(const Geo& g, const Phys& ph, const PairLaunch& a, bool hybrid,
                           bool write_macro, int variant, hipStream_t st) {
  if (variant <= kPopAA || variant >= kPopVariants) variant = pop_default_variant(g);
  switch (variant) {
    case kPairT8Z16: return launch_pair<64, 8, 16>(g, ph, a, hybrid, write_macro, st);
    case kPairT8Z32: return launch_pair<64, 8, 32>(g, ph, a, hybrid, write_macro, st);
    case kPairT4Z32: return launch_pair<64, 4, 32>(g, ph, a, hybrid, write_macro, st);
    case kPairX32T8Z16: return launch_pair<32, 8, 16>(g, ph, a, hybrid, write_macro, st);
    case kPairT12Z16: return launch_pair<64, 12, 16>(g, ph, a, hybrid, write_macro, st);
    case kPairT12Z32: return launch_pair<64, 12, 32>(g, ph, a, hybrid, write_macro, st);
    case kPairT12Z64: return launch_pair<64, 12, 64>(g, ph, a, hybrid, write_macro, st);
    default: return launch_pair<64, 4, 16>(g, ph, a, hybrid, write_macro, st);
  }
}

}  // namespace fs
