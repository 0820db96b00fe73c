// batch_api.cpp -- the fs_batch_* part of the C ABI (include/fluidsim.h).
//
// A batch owns the device buffers of all its domains in one allocation, the
// device descriptor tables and tile work lists read by batch.hip, and a
// captured hipGraph of STEPS_PER_GRAPH steps that replays the step loop.
// Population domains hold two population buffers and advance two steps per
// launch (an odd step count ends with one in-place AA step); every domain
// steps together, so one buffer index and one step parity describe them all.
// The population domains and the HOME domains never read each other, so a
// run of steps is two independent launch chains: populations on the batch's
// stream, HOME on a second stream forked from it and joined back at the end,
// and the two kinds of domain share the chip instead of taking turns.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "fluidsim.h"
#include "fs_internal.hpp"
#include "solver_impl.hpp"

using fs::fail;


namespace {

constexpr int kStepsPerGraph = 16;  // even number: the graph starts and ends on parity 0

struct Dom {
  fs_domain_desc desc;
  fs::Geo g;
  int pop = -1;   // index in the populations table, or -1
  int home = -1;  // index in the HOME table, or -1
  float* F = nullptr;   // population buffer 0
  float* F1 = nullptr;  // population buffer 1
  float* macro = nullptr;
  float* m[2] = {nullptr, nullptr};
  uint8_t* flags = nullptr;
  uint8_t* stage = nullptr;
  uint32_t* nbm = nullptr;  // solid-neighbour masks
  float* inlet = nullptr;
  std::vector<uint8_t> hflags;
  bool outletTileStart = false;  // population domain: an outlet starts a pair-tile row
};

}  // namespace

struct fs_batch {
  std::vector<Dom> doms;
  std::vector<fs::BatchPopDomain> popTab;
  std::vector<fs::BatchHomeDomain> homeTab;
  uint32_t popTiles = 0, homeTiles = 0, pairTiles = 0;
  char* arena = nullptr;
  fs::BatchPopDomain* d_pop = nullptr;
  fs::BatchHomeDomain* d_home = nullptr;
  uint16_t* d_popTile = nullptr;
  uint16_t* d_pairTile = nullptr;
  uint16_t* d_homeTile = nullptr;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // HOME chain
  hipEvent_t fork = nullptr, join = nullptr;
  int parity = 0;  // shared by every domain: they all step together
  int popCur = 0;  // population buffer holding the state (0: F, 1: F1)
  bool popPair = true;  // population domains take two steps per launch
  int popBlocks = 0, homeBlocks = 0;  // 0 = one workgroup per tile
  bool useGraph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  size_t cells = 0;
};

namespace {

void drop_graph(fs_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

void free_batch(fs_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  (void)hipFree(b->arena);
  (void)hipFree(b->d_pop);
  (void)hipFree(b->d_home);
  (void)hipFree(b->d_popTile);
  (void)hipFree(b->d_pairTile);
  (void)hipFree(b->d_homeTile);
  if (b->side) (void)hipStreamSynchronize(b->side);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->join) (void)hipEventDestroy(b->join);
  if (b->side) (void)hipStreamDestroy(b->side);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int check_dom(const fs_batch* b, int d) {
  if (!b) return fail(FS_ERR_INVALID, "null batch handle");
  if (d < 0 || d >= int(b->doms.size())) return fail(FS_ERR_INVALID, "no domain %d", d);
  return FS_OK;
}

fs::PopLaunch pop_launch(const Dom& D) {
  fs::PopLaunch a;
  a.F = D.F;
  a.flags = D.flags;
  a.nbm = D.nbm;
  a.rho = D.macro;
  a.ux = D.macro + D.g.P;
  a.uy = D.macro + 2 * D.g.P;
  a.uz = D.macro + 3 * D.g.P;
  a.inlet = D.inlet;
  return a;
}

// the population launch arguments on the buffer that holds the state
fs::PopLaunch cur_pop_launch(const fs_batch* b, const Dom& D) {
  fs::PopLaunch a = pop_launch(D);
  if (b->popCur) a.F = D.F1;
  return a;
}

// The two-step kernel never writes solid cells: their frozen slots must be
// equal in both population buffers after any write of the state or flags.
int sync_pop_buffers(fs_batch* b, const Dom& D) {
  if (D.desc.method == fs::kMethodHOME) return FS_OK;
  const float* src = b->popCur ? D.F1 : D.F;
  float* dst = b->popCur ? D.F : D.F1;
  FS_HIP(hipMemcpyAsync(dst, src, sizeof(float) * 19 * D.g.P, hipMemcpyDeviceToDevice,
                        b->stream));
  return FS_OK;
}

int upload_flags(fs_batch* b, Dom& D) {
  FS_HIP(hipMemcpyAsync(D.stage, D.hflags.data(), D.g.n, hipMemcpyHostToDevice, b->stream));
  FS_HIP(fs::launch_derive_flags(D.g, D.stage, D.flags, D.nbm, b->stream));
  if (int rc = sync_pop_buffers(b, D)) return rc;
  FS_HIP(hipStreamSynchronize(b->stream));
  return FS_OK;
}

// n steps: the population chain on the batch's stream (macros written by
// the last step when `macro`), the HOME chain on the side stream
int launch_steps(fs_batch* b, int n, bool macro) {
  if (n <= 0) return FS_OK;
  FS_HIP(hipEventRecord(b->fork, b->stream));
  FS_HIP(hipStreamWaitEvent(b->side, b->fork, 0));
  // populations: two steps per launch from the natural AA representation,
  // one AA step otherwise (the step parity is the AA parity: pairs are even)
  for (int i = 0; i < n;) {
    const bool odd = ((b->parity + i) & 1) != 0;
    if (b->popPair && !odd && n - i >= 2) {
      FS_HIP(fs::launch_batch_pair(b->d_pop, b->d_pairTile, b->pairTiles, b->popCur,
                                   macro && i + 2 == n, b->popBlocks, b->stream));
      b->popCur ^= 1;
      i += 2;
    } else {
      FS_HIP(fs::launch_batch_pop(b->d_pop, b->d_popTile, b->popTiles, b->popCur, odd,
                                  macro && i == n - 1, b->popBlocks, b->stream));
      i += 1;
    }
  }
  for (int i = 0; i < n; ++i)
    FS_HIP(fs::launch_batch_home(b->d_home, b->d_homeTile, b->homeTiles, (b->parity + i) & 1,
                                 b->homeBlocks, b->side));
  FS_HIP(hipEventRecord(b->join, b->side));
  FS_HIP(hipStreamWaitEvent(b->stream, b->join, 0));
  b->parity = (b->parity + n) & 1;
  return FS_OK;
}

int build_graph(fs_batch* b) {
  drop_graph(b);
  FS_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int rc = launch_steps(b, kStepsPerGraph, false);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
  // capture ran nothing: the graph is replayed from (and returns to) parity
  // 0 and population buffer 0
  b->parity = 0;
  b->popCur = 0;
  if (rc != FS_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(FS_ERR_DEVICE, "graph capture: %s", hipGetErrorString(e));
  b->graph = g;
  FS_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return FS_OK;
}

int enqueue(fs_batch* b, int nsteps) {
  if (nsteps < 0) return fail(FS_ERR_INVALID, "nsteps must be >= 0");
  int left = nsteps;
  // whole graphs while more than one graph's worth remains (the last step
  // writes the BGK/Hybrid macros, so it is always a direct launch)
  if (b->useGraph && b->parity == 0 && b->popCur == 0 && left > kStepsPerGraph) {
    if (!b->exec)
      if (int rc = build_graph(b)) return rc;
    while (left > kStepsPerGraph) {
      FS_HIP(hipGraphLaunch(b->exec, b->stream));
      left -= kStepsPerGraph;
    }
  }
  return launch_steps(b, left, true);
}

}  // namespace

extern "C" {

int fs_batch_create(const fs_domain_desc* descs, int ndomains, fs_batch** out) {
  if (!out) return fail(FS_ERR_INVALID, "null output handle");
  *out = nullptr;
  if (!descs || ndomains < 1 || ndomains > 65535)
    return fail(FS_ERR_INVALID, "need 1..65535 domain descriptors");
  if (int rc = fs::runtime_check()) return rc;
  try {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
      return fail(FS_ERR_DEVICE, "no HIP device available");
    fs_batch* b = new fs_batch();
    auto bail = [&](int rc) {
      free_batch(b);
      return rc;
    };
    b->doms.resize(size_t(ndomains));
    size_t bytes = 0;
    auto carve = [&](size_t n) {
      const size_t at = bytes;
      bytes += (n + 255) & ~size_t(255);
      return at;
    };
    std::vector<size_t> offs;  // per domain: F/m0, macro/m1, flags, stage, inlet, nbm, F1
    for (int d = 0; d < ndomains; ++d) {
      const fs_domain_desc& s = descs[d];
      if (s.method < 0 || s.method > 2) return bail(fail(FS_ERR_INVALID, "domain %d: method", d));
      if (s.nx < 2 || s.ny < 2 || s.nz < 1 ||
          (unsigned long long)s.nx * s.ny * s.nz >= (1ull << 31) ||
          // HOME tiles address a moment plane with 32-bit byte offsets
          (s.method == 1 && (unsigned long long)s.nx * s.ny * s.nz >= (1ull << 29)))
        return bail(fail(FS_ERR_INVALID, "domain %d: bad size", d));
      if (!(s.tau > 0.5f) || !std::isfinite(s.tau))
        return bail(fail(FS_ERR_INVALID, "domain %d: tau must be > 0.5", d));
      Dom& D = b->doms[size_t(d)];
      D.desc = s;
      D.g = fs::make_geo(s.nx, s.ny, s.nz);
      D.hflags.assign(D.g.n, 0);
      b->cells += D.g.n;
      const size_t P = D.g.P;
      if (s.method == fs::kMethodHOME) {
        offs.push_back(carve(sizeof(float) * 10 * P));
        offs.push_back(carve(sizeof(float) * 10 * P));
      } else {
        offs.push_back(carve(sizeof(float) * 19 * P));
        offs.push_back(carve(sizeof(float) * 4 * P));
      }
      offs.push_back(carve(P));
      offs.push_back(carve(P));
      offs.push_back(carve(sizeof(float) * size_t(s.ny) * s.nz));
      offs.push_back(carve(sizeof(uint32_t) * P));
      offs.push_back(s.method == fs::kMethodHOME ? 0 : carve(sizeof(float) * 19 * P));
    }
    if (hipStreamCreate(&b->stream) != hipSuccess || hipStreamCreate(&b->side) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess)
      return bail(fail(FS_ERR_DEVICE, "stream"));
    if (hipMalloc((void**)&b->arena, bytes) != hipSuccess)
      return bail(fail(FS_ERR_ALLOC, "batch arena of %zu bytes", bytes));
    std::vector<uint16_t> popTile, pairTile, homeTile;
    for (int d = 0; d < ndomains; ++d) {
      Dom& D = b->doms[size_t(d)];
      const size_t* o = &offs[size_t(d) * 7];
      if (D.desc.method == fs::kMethodHOME) {
        D.m[0] = reinterpret_cast<float*>(b->arena + o[0]);
        D.m[1] = reinterpret_cast<float*>(b->arena + o[1]);
      } else {
        D.F = reinterpret_cast<float*>(b->arena + o[0]);
        D.macro = reinterpret_cast<float*>(b->arena + o[1]);
      }
      D.flags = reinterpret_cast<uint8_t*>(b->arena + o[2]);
      D.stage = reinterpret_cast<uint8_t*>(b->arena + o[3]);
      D.inlet = reinterpret_cast<float*>(b->arena + o[4]);
      D.nbm = reinterpret_cast<uint32_t*>(b->arena + o[5]);
      if (D.desc.method != fs::kMethodHOME) D.F1 = reinterpret_cast<float*>(b->arena + o[6]);
      const float om = 1.0f / D.desc.tau;
      const fs::Phys ph = {om, om, om, D.desc.fx, D.desc.fy, D.desc.fz, FS_BOUNCE_HALFWAY};
      if (D.desc.method == fs::kMethodHOME) {
        fs::BatchHomeDomain h{};
        h.g = D.g;
        h.ph = ph;
        h.m0 = D.m[0];
        h.m1 = D.m[1];
        h.flags = D.flags;
        h.nbm = D.nbm;
        h.inlet = D.inlet;
        h.tile0 = uint32_t(homeTile.size());
        int ty = 0;
        fs::batch_home_tiling(D.g, h.tilesX, ty, h.chunks);
        h.wide = fs::batch_home_wide(D.g) ? 1 : 0;
        const uint32_t nt = uint32_t(h.tilesX) * uint32_t(ty) * uint32_t(h.chunks);
        D.home = int(b->homeTab.size());
        // work-list entries index the HOME descriptor table, not the domain list
        homeTile.insert(homeTile.end(), nt, uint16_t(D.home));
        b->homeTab.push_back(h);
      } else {
        fs::BatchPopDomain p{};
        p.g = D.g;
        p.ph = ph;
        p.a = pop_launch(D);
        p.F1 = D.F1;
        p.hybrid = D.desc.method == fs::kMethodHybrid;
        p.tile0 = uint32_t(popTile.size());
        p.tilesPerPlane = (D.g.nxy + 255) / 256;
        p.pairTile0 = uint32_t(pairTile.size());
        fs::batch_pair_tiling(D.g, p.pairTilesX, p.pairTilesY, p.pairChunks);
        p.pairWide = fs::batch_pair_wide(D.g) ? 1 : 0;
        D.pop = int(b->popTab.size());
        // work-list entries index the populations descriptor table
        popTile.insert(popTile.end(), size_t(p.tilesPerPlane) * D.g.nz, uint16_t(D.pop));
        pairTile.insert(pairTile.end(),
                        size_t(p.pairTilesX) * size_t(p.pairTilesY) * size_t(p.pairChunks),
                        uint16_t(D.pop));
        b->popTab.push_back(p);
      }
    }
    b->popTiles = uint32_t(popTile.size());
    b->pairTiles = uint32_t(pairTile.size());
    b->homeTiles = uint32_t(homeTile.size());
    // every work-list entry must name a descriptor of its own table
    for (uint16_t t : popTile)
      if (t >= b->popTab.size()) return bail(fail(FS_ERR_INVALID, "internal: pop work list"));
    for (uint16_t t : pairTile)
      if (t >= b->popTab.size()) return bail(fail(FS_ERR_INVALID, "internal: pair work list"));
    for (uint16_t t : homeTile)
      if (t >= b->homeTab.size()) return bail(fail(FS_ERR_INVALID, "internal: HOME work list"));
    for (size_t i = 0; i < b->popTab.size(); ++i) {
      const fs::BatchPopDomain& p = b->popTab[i];
      if (p.tile0 + uint64_t(p.tilesPerPlane) * p.g.nz > popTile.size() || popTile[p.tile0] != i)
        return bail(fail(FS_ERR_INVALID, "internal: pop tile range"));
      const uint64_t np = uint64_t(p.pairTilesX) * p.pairTilesY * p.pairChunks;
      if (p.pairTile0 + np > pairTile.size() || pairTile[p.pairTile0] != i)
        return bail(fail(FS_ERR_INVALID, "internal: pair tile range"));
    }
    for (size_t i = 0; i < b->homeTab.size(); ++i) {
      const fs::BatchHomeDomain& h = b->homeTab[i];
      if (h.tile0 >= homeTile.size() || homeTile[h.tile0] != i)
        return bail(fail(FS_ERR_INVALID, "internal: HOME tile range"));
    }
    hipError_t e = hipSuccess;
    if (!b->popTab.empty()) {
      e = hipMalloc((void**)&b->d_pop, sizeof(fs::BatchPopDomain) * b->popTab.size());
      if (e == hipSuccess) e = hipMalloc((void**)&b->d_popTile, sizeof(uint16_t) * popTile.size());
      if (e == hipSuccess)
        e = hipMemcpy(b->d_popTile, popTile.data(), sizeof(uint16_t) * popTile.size(),
                      hipMemcpyHostToDevice);
      if (e == hipSuccess)
        e = hipMalloc((void**)&b->d_pairTile, sizeof(uint16_t) * pairTile.size());
      if (e == hipSuccess)
        e = hipMemcpy(b->d_pairTile, pairTile.data(), sizeof(uint16_t) * pairTile.size(),
                      hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && !b->homeTab.empty()) {
      e = hipMalloc((void**)&b->d_home, sizeof(fs::BatchHomeDomain) * b->homeTab.size());
      if (e == hipSuccess) e = hipMalloc((void**)&b->d_homeTile, sizeof(uint16_t) * homeTile.size());
      if (e == hipSuccess)
        e = hipMemcpy(b->d_homeTile, homeTile.data(), sizeof(uint16_t) * homeTile.size(),
                      hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) return bail(fail(FS_ERR_ALLOC, "batch tables: %s", hipGetErrorString(e)));
    // upload descriptor tables, clear inlets, derive flags, reset every domain
    if (!b->popTab.empty())
      e = hipMemcpy(b->d_pop, b->popTab.data(), sizeof(fs::BatchPopDomain) * b->popTab.size(),
                    hipMemcpyHostToDevice);
    if (e == hipSuccess && !b->homeTab.empty())
      e = hipMemcpy(b->d_home, b->homeTab.data(), sizeof(fs::BatchHomeDomain) * b->homeTab.size(),
                    hipMemcpyHostToDevice);
    if (e != hipSuccess) return bail(fail(FS_ERR_DEVICE, "batch tables: %s", hipGetErrorString(e)));
    for (Dom& D : b->doms) {
      if (hipMemsetAsync(D.inlet, 0, sizeof(float) * size_t(D.g.ny) * D.g.nz, b->stream) !=
          hipSuccess)
        return bail(fail(FS_ERR_DEVICE, "memset"));
      if (int rc = upload_flags(b, D)) return bail(rc);
      hipError_t r = D.desc.method == fs::kMethodHOME
                         ? fs::launch_home_fill(D.g, D.m[0], D.m[1], nullptr, nullptr, nullptr,
                                                nullptr, b->stream)
                         : fs::launch_pop_reset(D.g, pop_launch(D), b->stream);
      if (r != hipSuccess) return bail(fail(FS_ERR_DEVICE, "reset: %s", hipGetErrorString(r)));
      if (int rc = sync_pop_buffers(b, D)) return bail(rc);
    }
    if (hipStreamSynchronize(b->stream) != hipSuccess) return bail(fail(FS_ERR_DEVICE, "sync"));
    *out = b;
    return FS_OK;
  } catch (const std::bad_alloc&) {
    return fail(FS_ERR_ALLOC, "host allocation failed");
  }
}

int fs_batch_destroy(fs_batch* b) {
  free_batch(b);
  return FS_OK;
}

int fs_batch_info(const fs_batch* b, int* ndomains, size_t* cells) {
  if (!b) return fail(FS_ERR_INVALID, "null batch handle");
  if (ndomains) *ndomains = int(b->doms.size());
  if (cells) *cells = b->cells;
  return FS_OK;
}

int fs_batch_set_flags(fs_batch* b, int d, const uint8_t* host_flags) {
  if (int rc = check_dom(b, d)) return rc;
  if (!host_flags) return fail(FS_ERR_INVALID, "null flags");
  Dom& D = b->doms[size_t(d)];
  const fs::Geo& g = D.g;
  for (int z = 0; z < g.nz; ++z)
    for (int y = 0; y < g.ny; ++y) {
      const size_t row = (size_t(z) * g.ny + y) * g.nx;
      for (int x = 0; x < g.nx; ++x) {
        const uint8_t f = host_flags[row + x];
        if ((f & FS_FLAG_INLET) && (f & (FS_FLAG_SOLID | FS_FLAG_OUTLET)))
          return fail(FS_ERR_INVALID, "domain %d: inlet cell is also solid or outlet", d);
        if (f & FS_FLAG_OUTLET) {
          if ((f & FS_FLAG_SOLID) || x < 1 ||
              (host_flags[row + x - 1] & (FS_FLAG_SOLID | FS_FLAG_INLET | FS_FLAG_OUTLET)))
            return fail(FS_ERR_INVALID, "domain %d: outlet cell needs a fluid cell at x-1", d);
          if (D.desc.method == fs::kMethodHOME && x % fs::batch_home_tile_width() == 0)
            return fail(FS_ERR_INVALID,
                        "domain %d: a HOME outlet cell may not start a %d-wide tile (x = %d)", d,
                        fs::batch_home_tile_width(), x);
        }
      }
    }
  bool tileStart = false;
  if (D.desc.method != fs::kMethodHOME)
    for (int z = 0; z < g.nz && !tileStart; ++z)
      for (int y = 0; y < g.ny && !tileStart; ++y) {
        const uint8_t* row = host_flags + (size_t(z) * g.ny + y) * g.nx;
        for (int x = 0; x < g.nx; x += fs::batch_pair_tile_width())
          if (row[x] & FS_FLAG_OUTLET) tileStart = true;
      }
  FS_HIP(hipStreamSynchronize(b->stream));
  D.outletTileStart = tileStart;
  // the two-step kernel takes an outlet's result from the lane at x-1: while
  // any population domain has an outlet that starts a tile row, all of them
  // take one step per launch
  bool pair = true;
  for (const Dom& E : b->doms) pair = pair && !E.outletTileStart;
  if (pair != b->popPair) {
    b->popPair = pair;
    drop_graph(b);
  }
  for (size_t i = 0; i < g.n; ++i) D.hflags[i] = host_flags[i] & fs::F_PUBLIC;
  return upload_flags(b, D);
}

int fs_batch_set_inlet_profile(fs_batch* b, int d, const float* host_ux) {
  if (int rc = check_dom(b, d)) return rc;
  Dom& D = b->doms[size_t(d)];
  const size_t bytes = sizeof(float) * size_t(D.g.ny) * D.g.nz;
  if (host_ux)
    FS_HIP(hipMemcpyAsync(D.inlet, host_ux, bytes, hipMemcpyHostToDevice, b->stream));
  else
    FS_HIP(hipMemsetAsync(D.inlet, 0, bytes, b->stream));
  FS_HIP(hipStreamSynchronize(b->stream));
  return FS_OK;
}

int fs_batch_set_bounce_rule(fs_batch* b, int d, int rule) {
  if (int rc = check_dom(b, d)) return rc;
  if (rule != FS_BOUNCE_HALFWAY && rule != FS_BOUNCE_REFERENCE_TABLE)
    return fail(FS_ERR_INVALID, "unknown bounce rule %d", rule);
  const Dom& D = b->doms[size_t(d)];
  if (D.home < 0) {
    if (rule == FS_BOUNCE_REFERENCE_TABLE)
      return fail(FS_ERR_STATE,
                  "the reference pairing table is not a permutation of the populations; only "
                  "HOME's gather can use it");
    return FS_OK;
  }
  // the kernels read the rule from the device descriptor table on every launch
  b->homeTab[size_t(D.home)].ph.bounce = rule;
  FS_HIP(hipMemcpyAsync(b->d_home + D.home, &b->homeTab[size_t(D.home)],
                        sizeof(fs::BatchHomeDomain), hipMemcpyHostToDevice, b->stream));
  FS_HIP(hipStreamSynchronize(b->stream));
  return FS_OK;
}

int fs_batch_reinit_from_macro(fs_batch* b, int d, const float* rho, const float* ux,
                               const float* uy, const float* uz) {
  if (int rc = check_dom(b, d)) return rc;
  if (b->parity != 0)
    return fail(FS_ERR_STATE, "re-initialise domains after an even number of batch steps");
  Dom& D = b->doms[size_t(d)];
  if (D.desc.method == fs::kMethodHOME) {
    FS_HIP(fs::launch_home_fill(D.g, D.m[0], D.m[1], rho, ux, uy, uz, b->stream));
  } else if (D.desc.method == fs::kMethodBGK && !rho) {
    FS_HIP(fs::launch_pop_reset(D.g, cur_pop_launch(b, D), b->stream));
  } else {
    FS_HIP(fs::launch_pop_reinit(D.g, cur_pop_launch(b, D), D.desc.method == fs::kMethodBGK ? 0 : 1,
                                 rho, ux, uy, uz, b->stream));
  }
  if (int rc = sync_pop_buffers(b, D)) return rc;
  FS_HIP(hipStreamSynchronize(b->stream));
  return FS_OK;
}

int fs_batch_set_launch(fs_batch* b, int pop_blocks, int home_blocks, int use_graph) {
  if (!b) return fail(FS_ERR_INVALID, "null batch handle");
  if (pop_blocks < 0 || home_blocks < 0) return fail(FS_ERR_INVALID, "negative grid size");
  FS_HIP(hipStreamSynchronize(b->stream));
  b->popBlocks = pop_blocks;
  b->homeBlocks = home_blocks;
  b->useGraph = use_graph != 0;
  drop_graph(b);
  return FS_OK;
}

int fs_batch_get_launch(const fs_batch* b, int* pop_blocks, int* home_blocks, int* use_graph) {
  if (!b) return fail(FS_ERR_INVALID, "null batch handle");
  if (pop_blocks) *pop_blocks = b->popBlocks;
  if (home_blocks) *home_blocks = b->homeBlocks;
  if (use_graph) *use_graph = b->useGraph ? 1 : 0;
  return FS_OK;
}

int fs_batch_step(fs_batch* b, int nsteps) {
  if (!b) return fail(FS_ERR_INVALID, "null batch handle");
  if (int rc = enqueue(b, nsteps)) return rc;
  FS_HIP(hipStreamSynchronize(b->stream));
  return FS_OK;
}

int fs_batch_timed_steps(fs_batch* b, int nsteps, float* elapsed_ms) {
  if (!b || !elapsed_ms) return fail(FS_ERR_INVALID, "null argument");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  FS_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(FS_ERR_DEVICE, "hipEventCreate");
  }
  int rc = FS_OK;
  hipError_t e = hipEventRecord(e0, b->stream);
  if (e == hipSuccess) rc = enqueue(b, nsteps);
  if (e == hipSuccess && rc == FS_OK) e = hipEventRecord(e1, b->stream);
  if (e == hipSuccess && rc == FS_OK) e = hipEventSynchronize(e1);
  if (e == hipSuccess && rc == FS_OK) e = hipEventElapsedTime(elapsed_ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != FS_OK) return rc;
  if (e != hipSuccess) return fail(FS_ERR_DEVICE, "timed steps: %s", hipGetErrorString(e));
  return FS_OK;
}

int fs_batch_autotune(fs_batch* b, int trial_steps, int* pop_blocks, int* home_blocks) {
  if (!b) return fail(FS_ERR_INVALID, "null batch handle");
  if (trial_steps < 2 || (trial_steps & 1))
    return fail(FS_ERR_INVALID, "trial_steps must be even and >= 2");
  if (b->parity != 0) return fail(FS_ERR_STATE, "autotune after an even number of steps");
  // candidates: one workgroup per tile, or a persistent grid of k workgroups
  // per CU; each trial runs from a saved copy of every domain's state
  int cus = 0;
  FS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
  const int mults[] = {0, 2, 4, 8, 16};
  size_t bytes = 0;
  for (const Dom& D : b->doms)
    bytes += sizeof(float) * (D.desc.method == fs::kMethodHOME ? 10 : 23) * D.g.P;
  char* saved = nullptr;
  FS_HIP(hipMalloc((void**)&saved, bytes));
  const int cur0 = b->popCur;
  auto copy_state = [&](bool save) -> hipError_t {
    size_t at = 0;
    for (const Dom& D : b->doms) {
      float* src = D.desc.method == fs::kMethodHOME ? D.m[0] : (cur0 ? D.F1 : D.F);
      const size_t n = sizeof(float) * (D.desc.method == fs::kMethodHOME ? 10 : 19) * D.g.P;
      hipError_t e = save ? hipMemcpyAsync(saved + at, src, n, hipMemcpyDeviceToDevice, b->stream)
                          : hipMemcpyAsync(src, saved + at, n, hipMemcpyDeviceToDevice, b->stream);
      if (e != hipSuccess) return e;
      at += n;
      if (D.desc.method != fs::kMethodHOME) {
        const size_t nm = sizeof(float) * 4 * D.g.P;
        e = save ? hipMemcpyAsync(saved + at, D.macro, nm, hipMemcpyDeviceToDevice, b->stream)
                 : hipMemcpyAsync(D.macro, saved + at, nm, hipMemcpyDeviceToDevice, b->stream);
        if (e != hipSuccess) return e;
        at += nm;
      }
    }
    return hipSuccess;
  };
  const int keepPop = b->popBlocks, keepHome = b->homeBlocks;
  const bool keepGraph = b->useGraph;
  int bestPop = 0, bestHome = 0, rc = FS_OK;
  float best = -1.0f;
  hipError_t e = copy_state(true);
  for (int pi = 0; pi < 5 && e == hipSuccess && rc == FS_OK; ++pi)
    for (int hi = 0; hi < 5 && e == hipSuccess && rc == FS_OK; ++hi) {
      if (b->popTab.empty() && pi > 0) continue;
      if (b->homeTab.empty() && hi > 0) continue;
      b->popBlocks = mults[pi] * cus;
      b->homeBlocks = mults[hi] * cus;
      b->useGraph = false;
      float ms = 0.0f;
      rc = enqueue(b, 2);  // warm-up
      if (rc == FS_OK) rc = fs_batch_timed_steps(b, trial_steps, &ms);
      if (rc == FS_OK && (best < 0.0f || ms < best)) {
        best = ms;
        bestPop = b->popBlocks;
        bestHome = b->homeBlocks;
      }
      b->popCur = cur0;
      if (rc == FS_OK) e = copy_state(false);
      b->parity = 0;
    }
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  (void)hipFree(saved);
  b->useGraph = keepGraph;
  drop_graph(b);
  if (rc != FS_OK || e != hipSuccess) {
    b->popBlocks = keepPop;
    b->homeBlocks = keepHome;
    if (rc != FS_OK) return rc;
    return fail(FS_ERR_DEVICE, "autotune: %s", hipGetErrorString(e));
  }
  b->popBlocks = bestPop;
  b->homeBlocks = bestHome;
  if (pop_blocks) *pop_blocks = bestPop;
  if (home_blocks) *home_blocks = bestHome;
  return FS_OK;
}

int fs_batch_copy_macro(fs_batch* b, int d, int which, float* host) {
  if (int rc = check_dom(b, d)) return rc;
  if (which < 0 || which > 3 || !host) return fail(FS_ERR_INVALID, "bad macro selector");
  const Dom& D = b->doms[size_t(d)];
  const float* src = D.desc.method == fs::kMethodHOME
                         ? D.m[b->parity] + size_t(which) * D.g.P
                         : D.macro + size_t(which) * D.g.P;
  FS_HIP(hipMemcpyAsync(host, src, sizeof(float) * D.g.n, hipMemcpyDeviceToHost, b->stream));
  FS_HIP(hipStreamSynchronize(b->stream));
  return FS_OK;
}

int fs_batch_copy_state(fs_batch* b, int d, float* host) {
  if (int rc = check_dom(b, d)) return rc;
  if (!host) return fail(FS_ERR_INVALID, "null output");
  const Dom& D = b->doms[size_t(d)];
  const size_t N = D.g.n, P = D.g.P;
  if (D.desc.method == fs::kMethodHOME) {
    for (int k = 0; k < 10; ++k)
      FS_HIP(hipMemcpyAsync(host + size_t(k) * N, D.m[b->parity] + size_t(k) * P,
                            sizeof(float) * N, hipMemcpyDeviceToHost, b->stream));
  } else {
    float* tmp = nullptr;
    FS_HIP(hipMalloc((void**)&tmp, sizeof(float) * 19 * P));
    hipError_t e = fs::launch_pop_canonical(D.g, b->popCur ? D.F1 : D.F, D.flags, b->parity != 0,
                                            tmp, b->stream);
    for (int q = 0; q < 19 && e == hipSuccess; ++q)
      e = hipMemcpyAsync(host + size_t(q) * N, tmp + size_t(q) * P, sizeof(float) * N,
                         hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(FS_ERR_DEVICE, "copy state: %s", hipGetErrorString(e));
  }
  FS_HIP(hipStreamSynchronize(b->stream));
  return FS_OK;
}

}  // extern "C"
