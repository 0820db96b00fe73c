// geom.hip -- STL voxelizer and Hybrid band construction on gfx950.
//
// Voxelizer (voxelize.cpp:115-243):
//   1. surface: one workgroup per triangle walks the triangle's voxel index
//      box and runs the triangle/box separating-axis test per voxel
//      (triBoxOverlap_SAT, voxelize.cpp:23-81), same float expressions and
//      thresholds, built without contraction, so every decision matches the
//      CPU rule.  Stores of the byte 1 from different triangles may land on
//      one voxel; they write the same value.
//   2. exterior: the reference's 6-neighbour BFS from the six faces is a
//      reachability set, so any order reaches the same fixpoint.  Here it is
//      computed by repeated line sweeps along x, y and z until a full round
//      changes nothing.
//   3. solid = surface or not exterior; optional AMR coverage downsample.
// Band (sim_runner.hpp:28-91): a level-synchronous BFS, one launch per
// distance ring, non-periodic, then band = fluid cells with 1 <= dist <= d0.
#include <mutex>

#include "fluidsim.h"
#include "fs_internal.hpp"

namespace fs {

namespace {

struct F3 {
  float x, y, z;
};
__device__ __forceinline__ F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
  // glm::cross(x, y) = (x.y*y.z - y.y*x.z, x.z*y.x - y.z*x.x, x.x*y.y - y.x*x.y)
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
// std::min / std::max semantics: min(a,b) = (b < a) ? b : a
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }

__device__ bool tri_box(F3 c, F3 half, F3 v0, F3 v1, F3 v2) {
  const F3 tv0 = sub(v0, c), tv1 = sub(v1, c), tv2 = sub(v2, c);
  const float mnx = smin(smin(tv0.x, tv1.x), tv2.x), mxx = smax(smax(tv0.x, tv1.x), tv2.x);
  const float mny = smin(smin(tv0.y, tv1.y), tv2.y), mxy = smax(smax(tv0.y, tv1.y), tv2.y);
  const float mnz = smin(smin(tv0.z, tv1.z), tv2.z), mxz = smax(smax(tv0.z, tv1.z), tv2.z);
  if (mxx < -half.x || mnx > half.x) return false;
  if (mxy < -half.y || mny > half.y) return false;
  if (mxz < -half.z || mnz > half.z) return false;
  const F3 e0 = sub(tv1, tv0), e1 = sub(tv2, tv1), e2 = sub(tv0, tv2);
  const F3 n = cross(e0, e1);
  const float r = half.x * fabsf(n.x) + half.y * fabsf(n.y) + half.z * fabsf(n.z);
  const float s = fabsf(n.x * tv0.x + n.y * tv0.y + n.z * tv0.z);
  if (s > r) return false;
  const F3 edges[3] = {e0, e1, e2};
#pragma unroll
  for (int ei = 0; ei < 3; ++ei) {
    const F3 e = edges[ei];
    if (fabsf(e.x) + fabsf(e.y) + fabsf(e.z) < 1e-12f) continue;
#pragma unroll
    for (int ai = 0; ai < 3; ++ai) {
      const F3 ax = {ai == 0 ? 1.0f : 0.0f, ai == 1 ? 1.0f : 0.0f, ai == 2 ? 1.0f : 0.0f};
      const F3 L = cross(e, ax);
      const float Llen2 = L.x * L.x + L.y * L.y + L.z * L.z;
      if (Llen2 < 1e-20f) continue;
      const float p0 = L.x * tv0.x + L.y * tv0.y + L.z * tv0.z;
      const float p1 = L.x * tv1.x + L.y * tv1.y + L.z * tv1.z;
      const float p2 = L.x * tv2.x + L.y * tv2.y + L.z * tv2.z;
      const float minp = smin(p0, smin(p1, p2));
      const float maxp = smax(p0, smax(p1, p2));
      const float rProj = half.x * fabsf(L.x) + half.y * fabsf(L.y) + half.z * fabsf(L.z);
      if (maxp < -rProj || minp > rProj) return false;
    }
  }
  return true;
}

__global__ void __launch_bounds__(256) k_surface(const TriF* __restrict__ tris, int nx, int ny,
                                                 int nz, uint8_t* __restrict__ surf) {
  const TriF t = tris[blockIdx.x];
  const F3 v0 = {t.v[0], t.v[1], t.v[2]}, v1 = {t.v[3], t.v[4], t.v[5]},
           v2 = {t.v[6], t.v[7], t.v[8]};
  const float eps = 1e-6f;
  // index box (voxelize.cpp:127-139)
  const float minx = smin(v0.x, smin(v1.x, v2.x)) - eps;
  const float miny = smin(v0.y, smin(v1.y, v2.y)) - eps;
  const float minz = smin(v0.z, smin(v1.z, v2.z)) - eps;
  const float maxx = smax(v0.x, smax(v1.x, v2.x)) + eps;
  const float maxy = smax(v0.y, smax(v1.y, v2.y)) + eps;
  const float maxz = smax(v0.z, smax(v1.z, v2.z)) + eps;
  const int x0 = max(0, int(floorf(minx * nx)));
  const int y0 = max(0, int(floorf(miny * ny)));
  const int z0 = max(0, int(floorf(minz * nz)));
  const int x1 = min(nx - 1, int(floorf(maxx * nx)));
  const int y1 = min(ny - 1, int(floorf(maxy * ny)));
  const int z1 = min(nz - 1, int(floorf(maxz * nz)));
  if (x1 < x0 || y1 < y0 || z1 < z0) return;
  const int bx = x1 - x0 + 1, by = y1 - y0 + 1;
  const long long vol = (long long)bx * by * (z1 - z0 + 1);
  const F3 half = {0.5f / nx, 0.5f / ny, 0.5f / nz};
  for (long long v = threadIdx.x; v < vol; v += blockDim.x) {
    const int dx = int(v % bx);
    const long long t2 = v / bx;
    const int dy = int(t2 % by);
    const int dz = int(t2 / by);
    const int x = x0 + dx, y = y0 + dy, z = z0 + dz;
    const F3 c = {(x + 0.5f) / nx, (y + 0.5f) / ny, (z + 0.5f) / nz};
    if (tri_box(c, half, v0, v1, v2)) surf[(size_t(z) * ny + y) * nx + x] = 1;
  }
}

// exterior seeds: non-surface cells on the six faces.  Sixteen cells per
// thread, loaded and stored as one 16-byte vector (the buffers are
// 256-byte aligned; the last thread of an odd-sized grid takes its bytes one
// by one).  Cells are fewer than 2^31, so the index arithmetic is 32-bit.
__global__ void __launch_bounds__(256) k_seed(const uint8_t* __restrict__ surf, int nx, int ny,
                                              int nz, uint8_t* __restrict__ vis) {
  const uint32_t n = uint32_t(nx) * uint32_t(ny) * uint32_t(nz);
  const uint32_t i0 = (blockIdx.x * 256u + threadIdx.x) * 16u;
  if (i0 >= n) return;
  uint32_t row = i0 / uint32_t(nx);
  int x = int(i0 - row * uint32_t(nx));
  int y = int(row % uint32_t(ny)), z = int(row / uint32_t(ny));
  auto seed = [&](uint8_t sv) {
    const bool face = x == 0 || y == 0 || z == 0 || x == nx - 1 || y == ny - 1 || z == nz - 1;
    const uint8_t v = (face && !sv) ? 1 : 0;
    if (++x == nx) {
      x = 0;
      if (++y == ny) y = 0, ++z;
    }
    return v;
  };
  if (i0 + 16u <= n) {
    union {
      uint4 v;
      uint8_t b[16];
    } in, out;
    in.v = *reinterpret_cast<const uint4*>(surf + i0);
#pragma unroll
    for (int j = 0; j < 16; ++j) out.b[j] = seed(in.b[j]);
    *reinterpret_cast<uint4*>(vis + i0) = out.v;
  } else {
    for (uint32_t i = i0; i < n; ++i) vis[i] = seed(surf[i]);
  }
}

// Sets the round's "something changed" flag.  In a round that fills the
// exterior nearly every wavefront reaches new cells, and one atomic each on
// the same word serialises them (with an unconditional atomic the first x
// sweep of a 50 M cell grid took 0.74 ms, the next round, which sets
// nothing, 0.07 ms); the flag only ever goes 0 -> 1, so once a lane sees it
// set there is nothing to do.
__device__ __forceinline__ void mark_changed(int* changed) {
  if (__hip_atomic_load(changed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
    atomicOr(changed, 1);
}

// One sweep per line, forward then backward: a cell becomes reached when the
// cell before it (in the sweep direction) is reached and it is not surface.
// After both passes a line's reached cells are exactly the open runs (maximal
// runs of non-surface cells) that held a reached cell.

// Lines along x, one wavefront per line, 64 cells per step as bit masks
// (bit l = lane l = cell x0 + l).  Within a word, the cells reached from the
// seeds s (a subset of the open cells o) in the direction of higher bits are
// s | (o & ((o + s) ^ o)): the carry of o + s runs from each seed through
// the rest of its open run and stops at the first closed bit.  A reached bit
// 63 carries into bit 0 of the next word; the backward pass is the same on
// bit-reversed words.  The cells of kSweepXBatch words are loaded before the
// first of them is processed, so their loads overlap.  Each lane re-reads
// only its own cells, which it wrote itself in the forward pass.
constexpr int kSweepXBatch = 4;

// one word of the x sweep: cell bytes (vl, sl) of this lane, `in` false for
// lanes past the line end; REV: the backward direction
template <bool REV>
__device__ __forceinline__ void sweep_word(uint8_t vl, uint8_t sl, bool in, int lane,
                                           uint8_t* __restrict__ cell, uint64_t& carry,
                                           bool& any) {
  const uint8_t v = in ? vl : uint8_t(0);
  const uint8_t s = in ? sl : uint8_t(1);  // lanes past the line end are closed
  uint64_t open = __ballot(s == 0), seen = __ballot(v != 0);
  if (REV) {
    open = __builtin_bitreverse64(open);
    seen = __builtin_bitreverse64(seen);
  }
  const uint64_t seeds = (seen | carry) & open;
  uint64_t reach = seeds | (open & ((open + seeds) ^ open));
  carry = reach >> 63;
  if (REV) reach = __builtin_bitreverse64(reach);
  if (in && !v && ((reach >> lane) & 1ull)) {
    *cell = 1;
    any = true;
  }
}

__global__ void __launch_bounds__(256) k_sweep_x(const uint8_t* __restrict__ surf,
                                                 uint8_t* __restrict__ vis, int nx, int nlines,
                                                 int* __restrict__ changed) {
  const int lane = int(threadIdx.x) & 63;
  const long line = long(blockIdx.x) * 4 + (threadIdx.x >> 6);
  // wavefronts past the last line run the same steps with every lane
  // inactive (their ballots are 0; they read cell 0 and write nothing)
  const bool lineOk = line < nlines;
  const size_t base = lineOk ? size_t(line) * size_t(nx) : 0;
  const int chunks = (nx + 63) / 64;
  bool any = false;
  uint64_t carry = 0;
  for (int c0 = 0; c0 < chunks; c0 += kSweepXBatch) {
    uint8_t vl[kSweepXBatch], sl[kSweepXBatch];
#pragma unroll
    for (int j = 0; j < kSweepXBatch; ++j) {
      const int x = (c0 + j) * 64 + lane;
      const size_t i = base + size_t(lineOk && x < nx ? x : 0);  // always a valid cell
      vl[j] = vis[i];
      sl[j] = surf[i];
    }
#pragma unroll
    for (int j = 0; j < kSweepXBatch; ++j) {
      if (c0 + j < chunks) {  // uniform across the wavefront
        const int x = (c0 + j) * 64 + lane;
        const bool in = lineOk && x < nx;
        sweep_word<false>(vl[j], sl[j], in, lane, vis + base + (in ? x : 0), carry, any);
      }
    }
  }
  carry = 0;
  for (int c1 = chunks - 1; c1 >= 0; c1 -= kSweepXBatch) {
    uint8_t vl[kSweepXBatch], sl[kSweepXBatch];
#pragma unroll
    for (int j = 0; j < kSweepXBatch; ++j) {
      const int x = (c1 - j) * 64 + lane;
      const size_t i = base + size_t(lineOk && c1 - j >= 0 && x < nx ? x : 0);
      vl[j] = vis[i];
      sl[j] = surf[i];
    }
#pragma unroll
    for (int j = 0; j < kSweepXBatch; ++j) {
      if (c1 - j >= 0) {
        const int x = (c1 - j) * 64 + lane;
        const bool in = lineOk && x < nx;
        sweep_word<true>(vl[j], sl[j], in, lane, vis + base + (in ? x : 0), carry, any);
      }
    }
  }
  if (any) mark_changed(changed);
}

// Lines along y (axis 1, indexed by (x,z)) or z (axis 2, indexed by (x,y)),
// one thread per line: neighbouring lanes hold neighbouring x, so every
// access is coalesced across the wavefront.  Cells are loaded kSweepBatch at
// a time before any of them is processed, so the loads of a batch overlap
// (a line's cells are its thread's alone: no other thread writes them).
constexpr int kSweepBatch = 8;

__global__ void __launch_bounds__(256) k_sweep_yz(const uint8_t* __restrict__ surf,
                                                  uint8_t* __restrict__ vis, int nx, int ny,
                                                  int nz, int axis, int* __restrict__ changed) {
  const size_t lid = size_t(blockIdx.x) * 256 + threadIdx.x;
  size_t base, stride;
  int len;
  if (axis == 1) {
    if (lid >= size_t(nx) * nz) return;
    const size_t x = lid % nx, z = lid / nx;
    base = z * size_t(nx) * ny + x;
    stride = nx;
    len = ny;
  } else {
    if (lid >= size_t(nx) * ny) return;
    base = lid;
    stride = size_t(nx) * ny;
    len = nz;
  }
  int any = 0;
  uint8_t prev = 0;
  for (int k0 = 0; k0 < len; k0 += kSweepBatch) {
    uint8_t v[kSweepBatch], s[kSweepBatch];
#pragma unroll
    for (int j = 0; j < kSweepBatch; ++j) {
      const bool in = k0 + j < len;
      const size_t i = base + size_t(in ? k0 + j : 0) * stride;  // always a cell of the line
      const uint8_t vl = vis[i], sl = surf[i];
      v[j] = in ? vl : uint8_t(0);
      s[j] = in ? sl : uint8_t(1);
    }
#pragma unroll
    for (int j = 0; j < kSweepBatch; ++j) {
      if (k0 + j < len) {
        if (!v[j] && prev && !s[j]) {
          v[j] = 1;
          vis[base + size_t(k0 + j) * stride] = 1;
          any = 1;
        }
        prev = v[j];
      }
    }
  }
  prev = 0;
  for (int k1 = len - 1; k1 >= 0; k1 -= kSweepBatch) {
    uint8_t v[kSweepBatch], s[kSweepBatch];
#pragma unroll
    for (int j = 0; j < kSweepBatch; ++j) {
      const bool in = k1 - j >= 0;
      const size_t i = base + size_t(in ? k1 - j : 0) * stride;
      const uint8_t vl = vis[i], sl = surf[i];
      v[j] = in ? vl : uint8_t(0);
      s[j] = in ? sl : uint8_t(1);
    }
#pragma unroll
    for (int j = 0; j < kSweepBatch; ++j) {
      if (k1 - j >= 0) {
        if (!v[j] && prev && !s[j]) {
          v[j] = 1;
          vis[base + size_t(k1 - j) * stride] = 1;
          any = 1;
        }
        prev = v[j];
      }
    }
  }
  if (any) mark_changed(changed);
}

// solid = surface or not exterior, sixteen cells per thread
__global__ void __launch_bounds__(256) k_finalize(uint8_t* __restrict__ surf_io,
                                                  const uint8_t* __restrict__ vis, size_t n) {
  const size_t i0 = (size_t(blockIdx.x) * 256 + threadIdx.x) * 16;
  if (i0 >= n) return;
  if (i0 + 16 <= n) {
    union {
      uint4 v;
      uint8_t b[16];
    } s4, v4;
    s4.v = *reinterpret_cast<const uint4*>(surf_io + i0);
    v4.v = *reinterpret_cast<const uint4*>(vis + i0);
#pragma unroll
    for (int j = 0; j < 16; ++j) s4.b[j] = (s4.b[j] || !v4.b[j]) ? 1 : 0;
    *reinterpret_cast<uint4*>(surf_io + i0) = s4.v;
  } else {
    for (size_t i = i0; i < n; ++i) surf_io[i] = (surf_io[i] || !vis[i]) ? 1 : 0;
  }
}

// the same, written as packed bits (bit j of word w: cell 32w + j; bits past
// n are 0): one word per thread
__global__ void __launch_bounds__(256) k_finalize_pack(const uint8_t* __restrict__ surf,
                                                       const uint8_t* __restrict__ vis, size_t n,
                                                       uint32_t* __restrict__ words) {
  const size_t w = size_t(blockIdx.x) * 256 + threadIdx.x;
  const size_t i0 = w * 32;
  if (i0 >= n) return;
  uint32_t bits = 0;
  if (i0 + 32 <= n) {
    union {
      uint4 v[2];
      uint8_t b[32];
    } s4, v4;
    s4.v[0] = *reinterpret_cast<const uint4*>(surf + i0);
    s4.v[1] = *reinterpret_cast<const uint4*>(surf + i0 + 16);
    v4.v[0] = *reinterpret_cast<const uint4*>(vis + i0);
    v4.v[1] = *reinterpret_cast<const uint4*>(vis + i0 + 16);
#pragma unroll
    for (int j = 0; j < 32; ++j) bits |= uint32_t(s4.b[j] || !v4.b[j]) << j;
  } else {
    for (size_t i = i0; i < n; ++i) bits |= uint32_t(surf[i] || !vis[i]) << (i - i0);
  }
  words[w] = bits;
}

// bytes (nonzero = 1) -> packed bits, and back
__global__ void __launch_bounds__(256) k_pack_bits(const uint8_t* __restrict__ in, size_t n,
                                                   uint32_t* __restrict__ words) {
  const size_t w = size_t(blockIdx.x) * 256 + threadIdx.x;
  const size_t i0 = w * 32;
  if (i0 >= n) return;
  uint32_t bits = 0;
  if (i0 + 32 <= n) {
    union {
      uint4 v[2];
      uint8_t b[32];
    } b4;
    b4.v[0] = *reinterpret_cast<const uint4*>(in + i0);
    b4.v[1] = *reinterpret_cast<const uint4*>(in + i0 + 16);
#pragma unroll
    for (int j = 0; j < 32; ++j) bits |= uint32_t(b4.b[j] != 0) << j;
  } else {
    for (size_t i = i0; i < n; ++i) bits |= uint32_t(in[i] != 0) << (i - i0);
  }
  words[w] = bits;
}

__global__ void __launch_bounds__(256) k_unpack_bits(const uint32_t* __restrict__ words, size_t n,
                                                     uint8_t* __restrict__ out) {
  const size_t w = size_t(blockIdx.x) * 256 + threadIdx.x;
  const size_t i0 = w * 32;
  if (i0 >= n) return;
  const uint32_t bits = words[w];
  if (i0 + 32 <= n) {
    union {
      uint4 v[2];
      uint8_t b[32];
    } b4;
#pragma unroll
    for (int j = 0; j < 32; ++j) b4.b[j] = uint8_t((bits >> j) & 1u);
    *reinterpret_cast<uint4*>(out + i0) = b4.v[0];
    *reinterpret_cast<uint4*>(out + i0 + 16) = b4.v[1];
  } else {
    for (size_t i = i0; i < n; ++i) out[i] = uint8_t((bits >> (i - i0)) & 1u);
  }
}

// downsample_amr (voxelize.cpp:187-226)
__global__ void __launch_bounds__(256) k_downsample(const uint8_t* __restrict__ fine, int refine,
                                                    int nx, int ny, int nz, float thr,
                                                    uint8_t* __restrict__ out) {
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  const size_t n = size_t(nx) * ny * nz;
  if (i >= n) return;
  const int cx = int(i % nx), cy = int((i / nx) % ny), cz = int(i / (size_t(nx) * ny));
  const size_t nxf = size_t(nx) * refine, nyf = size_t(ny) * refine;
  int cnt = 0;
  for (int dz = 0; dz < refine; ++dz)
    for (int dy = 0; dy < refine; ++dy) {
      const size_t b = (size_t(cz) * refine + dz) * nyf * nxf + (size_t(cy) * refine + dy) * nxf +
                       size_t(cx) * refine;
      for (int dx = 0; dx < refine; ++dx) cnt += fine[b + dx];
    }
  const float cov = float(cnt) / float(refine * refine * refine);
  out[i] = ((cov > 0.0f) && (cov >= thr)) ? 1 : 0;
}

// band BFS distances: D is uint8_t while the ring count fits below the
// "unreached" value 255 (a quarter of the bytes), int otherwise
template <class D>
struct DistT {
  static constexpr D kUnreached = D(-1);
};

// dist 0 on solid, unreached elsewhere
template <class D>
__global__ void __launch_bounds__(256) k_dist_init(const uint8_t* __restrict__ solid, size_t n,
                                                   D* __restrict__ dist) {
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  dist[i] = solid[i] ? D(0) : DistT<D>::kUnreached;
}
// ring k: unreached cells with a 6-neighbour at k-1 (sim_runner.hpp:57-76);
// cells are fewer than 2^31, so the index arithmetic is 32-bit
template <class D>
__global__ void __launch_bounds__(256) k_dist_ring(D* __restrict__ dist, int nx, int ny, int nz,
                                                   int k) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  const uint32_t n = uint32_t(nx) * uint32_t(ny) * uint32_t(nz);
  if (i >= n) return;
  if (dist[i] != DistT<D>::kUnreached) return;
  const uint32_t row = i / uint32_t(nx);
  const int x = int(i - row * uint32_t(nx));
  const int y = int(row % uint32_t(ny)), z = int(row / uint32_t(ny));
  const uint32_t sy = uint32_t(nx), sz = uint32_t(nx) * uint32_t(ny);
  const D p = D(k - 1);
  bool hit = false;
  if (x > 0) hit |= dist[i - 1] == p;
  if (x < nx - 1) hit |= dist[i + 1] == p;
  if (y > 0) hit |= dist[i - sy] == p;
  if (y < ny - 1) hit |= dist[i + sy] == p;
  if (z > 0) hit |= dist[i - sz] == p;
  if (z < nz - 1) hit |= dist[i + sz] == p;
  if (hit) dist[i] = D(k);
}
template <class D>
__global__ void __launch_bounds__(256) k_band_mask(const uint8_t* __restrict__ solid,
                                                   const D* __restrict__ dist, size_t n, int d0,
                                                   uint8_t* __restrict__ band) {
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const D d = dist[i];
  band[i] = (!solid[i] && d != DistT<D>::kUnreached && int(d) <= d0) ? 1 : 0;
}

// One-byte distances on grids whose rows are whole 16-byte blocks (nx a
// multiple of 16): sixteen cells per thread, loaded and stored as 16-byte
// vectors; a block never straddles a row, so its y/z neighbours are the
// aligned blocks at +-nx and +-nx*ny and its x neighbours the bytes just
// before and after it.
union Cells16 {
  uint4 v;
  uint8_t b[16];
};

__global__ void __launch_bounds__(256) k_dist_init16(const uint8_t* __restrict__ solid,
                                                     uint32_t nblocks, uint8_t* __restrict__ dist) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= nblocks) return;
  Cells16 s, d;
  s.v = reinterpret_cast<const uint4*>(solid)[t];
#pragma unroll
  for (int j = 0; j < 16; ++j) d.b[j] = s.b[j] ? uint8_t(0) : DistT<uint8_t>::kUnreached;
  reinterpret_cast<uint4*>(dist)[t] = d.v;
}

// ring k over 16-cell blocks: every value read is either unchanged by this
// ring or goes from "unreached" to k, neither of which equals k-1, so reading
// the neighbours' blocks while other threads update theirs gives the same
// ring as the one-cell kernel
__global__ void __launch_bounds__(256) k_dist_ring16(uint8_t* __restrict__ dist, int nx, int ny,
                                                     int nz, int k) {
  const uint32_t nb = uint32_t(nx) / 16u * uint32_t(ny) * uint32_t(nz);
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= nb) return;
  constexpr uint8_t U = DistT<uint8_t>::kUnreached;
  Cells16 c;
  c.v = reinterpret_cast<const uint4*>(dist)[t];
  bool any = false;
#pragma unroll
  for (int j = 0; j < 16; ++j) any = any || c.b[j] == U;
  if (!any) return;
  const uint32_t i0 = t * 16u;
  const uint32_t row = i0 / uint32_t(nx);
  const int x0 = int(i0 - row * uint32_t(nx));
  const int y = int(row % uint32_t(ny)), z = int(row / uint32_t(ny));
  const uint32_t by = uint32_t(nx) / 16u, bz = by * uint32_t(ny);
  const uint8_t p = uint8_t(k - 1);
  // a neighbour outside the grid (no periodic wrap) never matches
  Cells16 ym, yp, zm, zp;
  ym.v = yp.v = zm.v = zp.v = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
  const uint4* d4 = reinterpret_cast<const uint4*>(dist);
  if (y > 0) ym.v = d4[t - by];
  if (y < ny - 1) yp.v = d4[t + by];
  if (z > 0) zm.v = d4[t - bz];
  if (z < nz - 1) zp.v = d4[t + bz];
  const uint8_t left = x0 > 0 ? dist[i0 - 1] : U;
  const uint8_t right = x0 + 16 < nx ? dist[i0 + 16] : U;
  Cells16 o = c;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const uint8_t xm = j > 0 ? c.b[j - 1] : left;
    const uint8_t xp = j < 15 ? c.b[j + 1] : right;
    const bool hit = xm == p || xp == p || ym.b[j] == p || yp.b[j] == p || zm.b[j] == p ||
                     zp.b[j] == p;
    if (c.b[j] == U && hit) o.b[j] = uint8_t(k);
  }
  reinterpret_cast<uint4*>(dist)[t] = o.v;
}

__global__ void __launch_bounds__(256) k_band_mask16(const uint8_t* __restrict__ solid,
                                                     const uint8_t* __restrict__ dist,
                                                     uint32_t nblocks, int d0,
                                                     uint8_t* __restrict__ band) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= nblocks) return;
  Cells16 s, d, b;
  s.v = reinterpret_cast<const uint4*>(solid)[t];
  d.v = reinterpret_cast<const uint4*>(dist)[t];
#pragma unroll
  for (int j = 0; j < 16; ++j)
    b.b[j] = (!s.b[j] && d.b[j] != DistT<uint8_t>::kUnreached && int(d.b[j]) <= d0) ? 1 : 0;
  reinterpret_cast<uint4*>(band)[t] = b.v;
}

inline dim3 grid1d(size_t n) { return dim3(unsigned((n + 255) / 256)); }

}  // namespace

hipError_t voxelize_device(const TriF* d_tris, int ntri, int nx, int ny, int nz, uint8_t* d_out,
                           uint8_t* vis, int* changed, int* h_changed, uint32_t* d_words,
                           hipStream_t st) {
  const size_t n = size_t(nx) * ny * nz;
  hipError_t e = hipMemsetAsync(d_out, 0, n, st);
  if (e != hipSuccess) return e;
  if (ntri > 0) {
    hipLaunchKernelGGL(k_surface, dim3(ntri), dim3(256), 0, st, d_tris, nx, ny, nz, d_out);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_seed, grid1d((n + 15) / 16), dim3(256), 0, st, d_out, nx, ny, nz, vis);
  const size_t lines[3] = {size_t(ny) * nz, size_t(nx) * nz, size_t(nx) * ny};
  // every round that changes something grows the set; the set has at most n
  // cells, so the loop ends
  for (size_t round = 0; round <= n; ++round) {
    if ((e = hipMemsetAsync(changed, 0, sizeof(int), st)) != hipSuccess) break;
    hipLaunchKernelGGL(k_sweep_x, dim3(unsigned((lines[0] + 3) / 4)), dim3(256), 0, st, d_out,
                       vis, nx, int(lines[0]), changed);
    for (int ax = 1; ax < 3; ++ax)
      hipLaunchKernelGGL(k_sweep_yz, grid1d(lines[ax]), dim3(256), 0, st, d_out, vis, nx, ny, nz,
                         ax, changed);
    if ((e = hipMemcpyAsync(h_changed, changed, sizeof(int), hipMemcpyDeviceToHost, st)) !=
        hipSuccess)
      break;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) break;
    if (!*h_changed) break;
  }
  if (e == hipSuccess) {
    if (d_words)
      hipLaunchKernelGGL(k_finalize_pack, grid1d((n + 31) / 32), dim3(256), 0, st, d_out, vis, n,
                         d_words);
    else
      hipLaunchKernelGGL(k_finalize, grid1d((n + 15) / 16), dim3(256), 0, st, d_out, vis, n);
    e = hipGetLastError();
  }
  return e;
}

hipError_t pack_bits_device(const uint8_t* d_in, size_t n, uint32_t* d_words, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pack_bits, grid1d((n + 31) / 32), dim3(256), 0, st, d_in, n, d_words);
  return hipGetLastError();
}

hipError_t unpack_bits_device(const uint32_t* d_words, size_t n, uint8_t* d_out, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_unpack_bits, grid1d((n + 31) / 32), dim3(256), 0, st, d_words, n, d_out);
  return hipGetLastError();
}

hipError_t downsample_device(const uint8_t* d_fine, int refine, int nx, int ny, int nz,
                             float threshold, uint8_t* d_out, hipStream_t st) {
  const float thr = fmaxf(0.0f, fminf(1.0f, threshold));
  const size_t n = size_t(nx) * ny * nz;
  hipLaunchKernelGGL(k_downsample, grid1d(n), dim3(256), 0, st, d_fine, refine, nx, ny, nz, thr,
                     d_out);
  return hipGetLastError();
}

namespace {

// rings beyond d0 never enter the band, and no ring exceeds the grid's
// Manhattan diameter
int band_rings(int nx, int ny, int nz, int d0) { return d0 < nx + ny + nz ? d0 : nx + ny + nz; }

template <class D>
hipError_t band_run(const uint8_t* d_solid, int nx, int ny, int nz, int d0, uint8_t* d_band,
                    D* dist, hipStream_t st) {
  const size_t n = size_t(nx) * ny * nz;
  const int rings = band_rings(nx, ny, nz, d0);
  if constexpr (sizeof(D) == 1) {
    // 16-byte blocks need 16-byte aligned buffers (device allocations are;
    // a caller's buffers of fs_band_noalloc may not be)
    const uintptr_t al = reinterpret_cast<uintptr_t>(d_solid) | reinterpret_cast<uintptr_t>(d_band) |
                         reinterpret_cast<uintptr_t>(dist);
    if (nx % 16 == 0 && al % 16 == 0) {
      const uint32_t nb = uint32_t(n / 16);
      hipLaunchKernelGGL(k_dist_init16, grid1d(nb), dim3(256), 0, st, d_solid, nb, dist);
      for (int k = 1; k <= rings; ++k)
        hipLaunchKernelGGL(k_dist_ring16, grid1d(nb), dim3(256), 0, st, dist, nx, ny, nz, k);
      hipLaunchKernelGGL(k_band_mask16, grid1d(nb), dim3(256), 0, st, d_solid, dist, nb, d0,
                         d_band);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL(k_dist_init<D>, grid1d(n), dim3(256), 0, st, d_solid, n, dist);
  for (int k = 1; k <= rings; ++k)
    hipLaunchKernelGGL(k_dist_ring<D>, grid1d(n), dim3(256), 0, st, dist, nx, ny, nz, k);
  hipLaunchKernelGGL(k_band_mask<D>, grid1d(n), dim3(256), 0, st, d_solid, dist, n, d0, d_band);
  return hipGetLastError();
}

bool band_bytes_dist(int nx, int ny, int nz, int d0) { return band_rings(nx, ny, nz, d0) < 255; }

}  // namespace

size_t band_scratch_bytes(int nx, int ny, int nz, int d0) {
  const size_t n = size_t(nx) * ny * nz;
  return band_bytes_dist(nx, ny, nz, d0) ? n : n * sizeof(int);
}

hipError_t band_device(const uint8_t* d_solid, int nx, int ny, int nz, int d0, uint8_t* d_band,
                       uint8_t* d_scratch, hipStream_t st) {
  if (band_bytes_dist(nx, ny, nz, d0))
    return band_run<uint8_t>(d_solid, nx, ny, nz, d0, d_band, d_scratch, st);
  return band_run<int>(d_solid, nx, ny, nz, d0, d_band, reinterpret_cast<int*>(d_scratch), st);
}

// ---------------------------------------------------------------------------
// Band without allocation (fs_band_noalloc): the distance scratch of the BFS
// lives in a file-scope device array of kBandScratchBytes, one copy per GPU,
// held until the process ends.  Host code never passes the array's name to
// a kernel (in host code it is only a stand-in for the symbol): the device
// address of each GPU's copy is looked up once per device with
// hipGetSymbolAddress and kept.  k_dist_init writes every distance a call
// reads, so nothing has to start at zero; calls that overlap on one GPU
// share the array and race.
constexpr size_t kBandScratchBytes = size_t(64) << 20;
__device__ __align__(256) uint8_t scratch[kBandScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

size_t band_noalloc_max_bytes() { return kBandScratchBytes; }

hipError_t band_scratch_address(void** out) {
  *out = nullptr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

}  // namespace fs

extern "C" int fs_band_noalloc(const uint8_t* d_solid, int nx, int ny, int nz, int d0,
                               uint8_t* d_band, void* hip_stream) {
  if (!d_solid || !d_band || nx < 1 || ny < 1 || nz < 1)
    return fs::fail(FS_ERR_INVALID, "bad band arguments");
  const size_t n = size_t(nx) * ny * nz;
  if (n >= (size_t(1) << 31)) return fs::fail(FS_ERR_INVALID, "grid of %zu cells exceeds 2^31", n);
  const int dd = d0 < 0 ? 0 : d0;
  const size_t need = fs::band_scratch_bytes(nx, ny, nz, dd);
  if (need > fs::kBandScratchBytes)
    return fs::fail(FS_ERR_INVALID,
                    "fs_band_noalloc: %dx%dx%d cells with d0 = %d need %zu scratch bytes, more "
                    "than its fixed %zu (use fs_band)",
                    nx, ny, nz, d0, need, fs::kBandScratchBytes);
  if (int rc = fs::runtime_check()) return rc;
  void* scr = nullptr;
  hipError_t e = fs::band_scratch_address(&scr);
  if (e != hipSuccess)
    return fs::fail(FS_ERR_DEVICE, "fs_band_noalloc: scratch address: %s", hipGetErrorString(e));
  e = fs::band_device(d_solid, nx, ny, nz, dd, d_band, static_cast<uint8_t*>(scr),
                      static_cast<hipStream_t>(hip_stream));
  if (e != hipSuccess) return fs::fail(FS_ERR_DEVICE, "fs_band_noalloc: %s", hipGetErrorString(e));
  return FS_OK;
}
