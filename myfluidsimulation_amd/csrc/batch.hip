// batch.hip -- many independent lattices of uneven size stepped together.
//
// Each domain is described by one entry of a device-resident descriptor
// table (its geometry, relaxation parameters and buffer pointers).  A step is
// one launch for all population domains (BGK and Hybrid) and one for all
// HOME domains.  Each launch runs a persistent grid: every workgroup strides
// over a work list of tiles (a tile belongs to one domain; the list maps
// tile -> domain), so small domains no longer pay a launch each and large
// and small ones share the chip.  The per-tile work is exactly the
// single-domain kernels' (pop_body.hpp, pair_body.hpp, home_body.hpp), so
// every domain's fields equal a single-domain run bit for bit.  Population
// domains advance two steps per launch (k_batch_pair: 64 x 4 tiles where nx
// is a multiple of 64, else 32 x 8 tiles, which waste fewer lanes than
// 64-wide rows on the sweep's other edges of 48..112 cells), or one
// (k_batch_pop) for an odd step count.  HOME domains use the third tiled
// form (home_tile3) on the same two tile shapes (32-plane chunks).
#include "home_body.hpp"
#include "pair_body.hpp"
#include "pop_body.hpp"

namespace fs {

namespace {

// HOME tiles walk 32-plane chunks, population tiles 16-plane ones (seeded
// 64-domain sweep: HOME chunks of 8 / 16 / 32 / 64 / 128 planes 42.1 / 42.9 /
// 43.3 / 41.2 / 39.9 GLUPS with 16-plane population chunks; population
// chunks of 24 or 32 planes 41.4-41.9 / 41.0 with 32-plane HOME chunks)
constexpr int kBatchHomeTX = 32, kBatchHomeTY = 8, kBatchHomeZC = 32;
constexpr int kBatchHomeWideTX = 64, kBatchHomeWideTY = 4;  // nx a multiple of 64
constexpr int kBatchHomeLds =
    16 * (TileLds16<kBatchHomeWideTX, kBatchHomeWideTY>::PL >
                  TileLds16<kBatchHomeTX, kBatchHomeTY>::PL
              ? TileLds16<kBatchHomeWideTX, kBatchHomeWideTY>::PL
              : TileLds16<kBatchHomeTX, kBatchHomeTY>::PL);
static_assert(kBatchHomeWideTX * kBatchHomeWideTY == kBatchHomeTX * kBatchHomeTY, "one block");
// population tiles: 32 x 8, or 64 x 4 for domains whose nx is a multiple of
// 64 (full 64-wide rows); both are 256 threads, so one kernel runs either
constexpr int kBatchPairTX = 32, kBatchPairTY = 8, kBatchPairZC = 16;
constexpr int kBatchWideTX = 64, kBatchWideTY = 4;
constexpr int kBatchPairLds =
    19 * (TileLds<kBatchWideTX, kBatchWideTY>::PL > TileLds<kBatchPairTX, kBatchPairTY>::PL
              ? TileLds<kBatchWideTX, kBatchWideTY>::PL
              : TileLds<kBatchPairTX, kBatchPairTY>::PL);
static_assert(kBatchWideTX * kBatchWideTY == kBatchPairTX * kBatchPairTY, "one block size");

template <bool ODD, bool MACRO>
__global__ void __launch_bounds__(256) k_batch_pop(const BatchPopDomain* __restrict__ doms,
                                                   const uint16_t* __restrict__ tile_dom,
                                                   uint32_t ntiles, int cur) {
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const BatchPopDomain& D = doms[tile_dom[t]];
    const Geo g = D.g;
    PopLaunch a = D.a;
    if (cur) a.F = D.F1;
    const uint32_t local = t - D.tile0;
    const uint32_t z = local / D.tilesPerPlane;
    const uint32_t cp = (local - z * D.tilesPerPlane) * 256u + threadIdx.x;
    if (cp >= g.nxy) continue;
    Cell cl;
    cl.z = int(z);
    const uint32_t y = g.divx.div(cp);
    cl.y = int(y);
    cl.x = int(cp - y * uint32_t(g.nx));
    cl.c = z * g.nxy + cp;
    if (D.hybrid)
      pop_cell<ODD, true, MACRO>(g, D.ph, a, cl);
    else
      pop_cell<ODD, false, MACRO>(g, D.ph, a, cl);
  }
}

template <bool MACRO>
__global__ void __launch_bounds__(kBatchPairTX * kBatchPairTY)
    k_batch_pair(const BatchPopDomain* __restrict__ doms, const uint16_t* __restrict__ tile_dom,
                 uint32_t ntiles, int cur) {
  __shared__ float lds[kBatchPairLds];
  // every thread runs the same trip count: pair_tile synchronises the group
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const BatchPopDomain& D = doms[tile_dom[t]];
    const uint32_t local = t - D.pairTile0;
    const int bz = int(local % uint32_t(D.pairChunks));
    const int rest = int(local / uint32_t(D.pairChunks));
    PairLaunch a;
    a.Fin = cur ? D.F1 : D.a.F;
    a.Fout = cur ? D.a.F : D.F1;
    a.flags = D.a.flags;
    a.nbm = D.a.nbm;
    a.rho = D.a.rho;
    a.ux = D.a.ux;
    a.uy = D.a.uy;
    a.uz = D.a.uz;
    a.inlet = D.a.inlet;
    const int bx = rest % D.pairTilesX, by = rest / D.pairTilesX;
    if (D.pairWide) {
      if (D.hybrid)
        pair_tile<kBatchWideTX, kBatchWideTY, kBatchPairZC, true, MACRO>(D.g, D.ph, a, bx, by, bz,
                                                                         lds);
      else
        pair_tile<kBatchWideTX, kBatchWideTY, kBatchPairZC, false, MACRO>(D.g, D.ph, a, bx, by,
                                                                          bz, lds);
    } else {
      if (D.hybrid)
        pair_tile<kBatchPairTX, kBatchPairTY, kBatchPairZC, true, MACRO>(D.g, D.ph, a, bx, by, bz,
                                                                         lds);
      else
        pair_tile<kBatchPairTX, kBatchPairTY, kBatchPairZC, false, MACRO>(D.g, D.ph, a, bx, by,
                                                                          bz, lds);
    }
  }
}

__global__ void __launch_bounds__(kBatchHomeTX * kBatchHomeTY)
    k_batch_home(const BatchHomeDomain* __restrict__ doms, const uint16_t* __restrict__ tile_dom,
                 uint32_t ntiles, int cur) {
  __shared__ float fs[kBatchHomeLds];
  // every thread runs the same trip count: home_tile synchronises the group
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const BatchHomeDomain& D = doms[tile_dom[t]];
    const uint32_t local = t - D.tile0;
    const int bz = int(local % uint32_t(D.chunks));
    const int rest = int(local / uint32_t(D.chunks));
    HomeLaunch a;
    a.m = cur ? D.m1 : D.m0;
    a.mn = cur ? D.m0 : D.m1;
    a.flags = D.flags;
    a.nbm = D.nbm;
    a.inlet = D.inlet;
    const int bx = rest % D.tilesX, by = rest / D.tilesX;
    if (D.wide) {
      if (D.ph.bounce == 0)
        home_tile3<kBatchHomeWideTX, kBatchHomeWideTY, kBatchHomeZC, 0>(D.g, D.ph, a, bx, by, bz,
                                                                         fs);
      else
        home_tile3<kBatchHomeWideTX, kBatchHomeWideTY, kBatchHomeZC, 1>(D.g, D.ph, a, bx, by, bz,
                                                                         fs);
    } else {
      if (D.ph.bounce == 0)
        home_tile3<kBatchHomeTX, kBatchHomeTY, kBatchHomeZC, 0>(D.g, D.ph, a, bx, by, bz, fs);
      else
        home_tile3<kBatchHomeTX, kBatchHomeTY, kBatchHomeZC, 1>(D.g, D.ph, a, bx, by, bz, fs);
    }
  }
}

}  // namespace

bool batch_home_wide(const Geo& g) { return g.nx % kBatchHomeWideTX == 0; }

void batch_home_tiling(const Geo& g, int& tilesX, int& tilesY, int& chunks) {
  const bool wide = batch_home_wide(g);
  const int tx = wide ? kBatchHomeWideTX : kBatchHomeTX, ty = wide ? kBatchHomeWideTY : kBatchHomeTY;
  tilesX = (g.nx + tx - 1) / tx;
  tilesY = (g.ny + ty - 1) / ty;
  chunks = (g.nz + kBatchHomeZC - 1) / kBatchHomeZC;
}

int batch_home_tile_width() { return kBatchHomeTX; }

bool batch_pair_wide(const Geo& g) { return g.nx % kBatchWideTX == 0; }

void batch_pair_tiling(const Geo& g, int& tilesX, int& tilesY, int& chunks) {
  const bool wide = batch_pair_wide(g);
  const int tx = wide ? kBatchWideTX : kBatchPairTX, ty = wide ? kBatchWideTY : kBatchPairTY;
  tilesX = (g.nx + tx - 1) / tx;
  tilesY = (g.ny + ty - 1) / ty;
  chunks = (g.nz + kBatchPairZC - 1) / kBatchPairZC;
}

int batch_pair_tile_width() { return kBatchPairTX; }

hipError_t launch_batch_pop(const BatchPopDomain* doms, const uint16_t* tile_dom, uint32_t ntiles,
                            int cur, bool odd, bool macro, int blocks, hipStream_t st) {
  if (ntiles == 0) return hipSuccess;
  const unsigned grid = unsigned(blocks <= 0 || uint32_t(blocks) > ntiles ? ntiles : blocks);
  if (odd) {
    if (macro)
      hipLaunchKernelGGL((k_batch_pop<true, true>), dim3(grid), dim3(256), 0, st, doms, tile_dom,
                         ntiles, cur);
    else
      hipLaunchKernelGGL((k_batch_pop<true, false>), dim3(grid), dim3(256), 0, st, doms, tile_dom,
                         ntiles, cur);
  } else {
    if (macro)
      hipLaunchKernelGGL((k_batch_pop<false, true>), dim3(grid), dim3(256), 0, st, doms, tile_dom,
                         ntiles, cur);
    else
      hipLaunchKernelGGL((k_batch_pop<false, false>), dim3(grid), dim3(256), 0, st, doms,
                         tile_dom, ntiles, cur);
  }
  return hipGetLastError();
}

hipError_t launch_batch_pair(const BatchPopDomain* doms, const uint16_t* tile_dom,
                             uint32_t ntiles, int cur, bool macro, int blocks, hipStream_t st) {
  if (ntiles == 0) return hipSuccess;
  const unsigned grid = unsigned(blocks <= 0 || uint32_t(blocks) > ntiles ? ntiles : blocks);
  const dim3 block(kBatchPairTX * kBatchPairTY);
  if (macro)
    hipLaunchKernelGGL(k_batch_pair<true>, dim3(grid), block, 0, st, doms, tile_dom, ntiles, cur);
  else
    hipLaunchKernelGGL(k_batch_pair<false>, dim3(grid), block, 0, st, doms, tile_dom, ntiles, cur);
  return hipGetLastError();
}

hipError_t launch_batch_home(const BatchHomeDomain* doms, const uint16_t* tile_dom,
                             uint32_t ntiles, int cur, int blocks, hipStream_t st) {
  if (ntiles == 0) return hipSuccess;
  const unsigned grid = unsigned(blocks <= 0 || uint32_t(blocks) > ntiles ? ntiles : blocks);
  hipLaunchKernelGGL(k_batch_home, dim3(grid), dim3(kBatchHomeTX * kBatchHomeTY), 0, st, doms,
                     tile_dom, ntiles, cur);
  return hipGetLastError();
}

}  // namespace fs
