// batch.hip -- many independent lattices of uneven size stepped together.
//
// Each domain is described by one entry of a device-resident descriptor
// table (its geometry, relaxation parameters and buffer pointers).  A step is
// one launch for all population domains (BGK and Hybrid) and one for all
// HOME domains.  Each launch runs a persistent grid: every workgroup strides
// over a work list of tiles (a tile belongs to one domain; the list maps
// tile -> domain), so small domains no longer pay a launch each and large
// and small ones share the chip.  The per-tile work is exactly the
// single-domain kernels' (pop_body.hpp, home_body.hpp), so every domain's
// fields equal a single-domain run bit for bit.
#include "home_body.hpp"
#include "pop_body.hpp"

namespace fs {

namespace {

constexpr int kBatchHomeTX = 32, kBatchHomeTY = 8, kBatchHomeZC = 16;

template <bool ODD, bool MACRO>
__global__ void __launch_bounds__(256) k_batch_pop(const BatchPopDomain* __restrict__ doms,
                                                   const uint16_t* __restrict__ tile_dom,
                                                   uint32_t ntiles) {
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const BatchPopDomain& D = doms[tile_dom[t]];
    const Geo g = D.g;
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
      pop_cell<ODD, true, MACRO>(g, D.ph, D.a, cl);
    else
      pop_cell<ODD, false, MACRO>(g, D.ph, D.a, cl);
  }
}

__global__ void __launch_bounds__(kBatchHomeTX * kBatchHomeTY)
    k_batch_home(const BatchHomeDomain* __restrict__ doms, const uint16_t* __restrict__ tile_dom,
                 uint32_t ntiles, int cur) {
  __shared__ float fs[19 * TileLds<kBatchHomeTX, kBatchHomeTY>::PL];
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
    a.inlet = D.inlet;
    if (D.ph.bounce == 0)
      home_tile<kBatchHomeTX, kBatchHomeTY, kBatchHomeZC, 0>(D.g, D.ph, a, rest % D.tilesX,
                                                             rest / D.tilesX, bz, fs);
    else
      home_tile<kBatchHomeTX, kBatchHomeTY, kBatchHomeZC, 1>(D.g, D.ph, a, rest % D.tilesX,
                                                             rest / D.tilesX, bz, fs);
  }
}

}  // namespace

void batch_home_tiling(const Geo& g, int& tilesX, int& tilesY, int& chunks) {
  tilesX = (g.nx + kBatchHomeTX - 1) / kBatchHomeTX;
  tilesY = (g.ny + kBatchHomeTY - 1) / kBatchHomeTY;
  chunks = (g.nz + kBatchHomeZC - 1) / kBatchHomeZC;
}

int batch_home_tile_width() { return kBatchHomeTX; }

hipError_t launch_batch_pop(const BatchPopDomain* doms, const uint16_t* tile_dom, uint32_t ntiles,
                            bool odd, bool macro, int blocks, hipStream_t st) {
  if (ntiles == 0) return hipSuccess;
  const unsigned grid = unsigned(blocks <= 0 || uint32_t(blocks) > ntiles ? ntiles : blocks);
  if (odd) {
    if (macro)
      hipLaunchKernelGGL((k_batch_pop<true, true>), dim3(grid), dim3(256), 0, st, doms, tile_dom,
                         ntiles);
    else
      hipLaunchKernelGGL((k_batch_pop<true, false>), dim3(grid), dim3(256), 0, st, doms, tile_dom,
                         ntiles);
  } else {
    if (macro)
      hipLaunchKernelGGL((k_batch_pop<false, true>), dim3(grid), dim3(256), 0, st, doms, tile_dom,
                         ntiles);
    else
      hipLaunchKernelGGL((k_batch_pop<false, false>), dim3(grid), dim3(256), 0, st, doms,
                         tile_dom, ntiles);
  }
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
