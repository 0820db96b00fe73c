// stage.cpp -- host side of the geometry calls' transfers: pinned staging
// memory, cached device scratch, and multithreaded bit (un)packing.
//
// A solid or band mask is one byte per cell on both sides of the C ABI, but
// it crosses PCIe as packed bits (one eighth of the bytes) through pinned
// host memory: the device packs and unpacks with vector kernels (geom.hip),
// the host with several threads.  The staging buffers and the geometry
// calls' device scratch are kept and grown across calls (one set per
// device, guarded by a mutex that each geometry call holds throughout), so a
// repeated call allocates nothing; fs_geometry_release frees them.
#include <emmintrin.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "fluidsim.h"
#include "fs_internal.hpp"

namespace fs {

namespace {

constexpr int kMaxDevices = 64;
GeomScratch g_scratch[kMaxDevices];

hipError_t grow_device(void** p, size_t* cap, size_t bytes) {
  if (*cap >= bytes && *p) return hipSuccess;
  if (*p) {
    hipError_t e = hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (e != hipSuccess) return e;
  }
  hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 256));
  if (e == hipSuccess) *cap = std::max<size_t>(bytes, 256);
  return e;
}

hipError_t grow_pinned(void** p, size_t* cap, size_t bytes) {
  if (*cap >= bytes && *p) return hipSuccess;
  if (*p) {
    hipError_t e = hipHostFree(*p);
    *p = nullptr;
    *cap = 0;
    if (e != hipSuccess) return e;
  }
  hipError_t e = hipHostMalloc(p, std::max<size_t>(bytes, 256), hipHostMallocDefault);
  if (e == hipSuccess) *cap = std::max<size_t>(bytes, 256);
  return e;
}

// 8 mask bits -> 8 bytes of 0/1 (little endian: bit j -> byte j)
struct ByteLut {
  uint64_t v[256];
  ByteLut() {
    for (int b = 0; b < 256; ++b) {
      uint64_t x = 0;
      for (int j = 0; j < 8; ++j)
        if ((b >> j) & 1) x |= uint64_t(1) << (8 * j);
      v[b] = x;
    }
  }
};
const ByteLut g_lut;

// runs fn(begin_word, end_word) over [0, nwords) on up to 16 threads
template <class Fn>
void parallel_words(size_t nwords, Fn&& fn) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t per = 1 << 16;  // at least 64 Ki words (2 Mi cells) per thread
  const unsigned nt = unsigned(std::min<size_t>({16, hw, (nwords + per - 1) / per}));
  if (nt <= 1) {
    fn(size_t(0), nwords);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt - 1);
  const size_t chunk = (nwords + nt - 1) / nt;
  for (unsigned t = 1; t < nt; ++t) {
    const size_t b = std::min(nwords, t * chunk), e = std::min(nwords, b + chunk);
    th.emplace_back([&fn, b, e] { fn(b, e); });
  }
  fn(size_t(0), std::min(nwords, chunk));
  for (std::thread& t : th) t.join();
}

}  // namespace

GeomScratch* geom_scratch(std::unique_lock<std::mutex>& lock) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  GeomScratch* g = &g_scratch[dev];
  lock = std::unique_lock<std::mutex>(g->mu);
  return g;
}

hipError_t GeomScratch::device(int slot, size_t bytes, void** out) {
  hipError_t e = grow_device(&dev[slot], &devCap[slot], bytes);
  *out = e == hipSuccess ? dev[slot] : nullptr;
  return e;
}

hipError_t GeomScratch::pinned(int slot, size_t bytes, void** out) {
  hipError_t e = grow_pinned(&host[slot], &hostCap[slot], bytes);
  *out = e == hipSuccess ? host[slot] : nullptr;
  return e;
}

void GeomScratch::release() {
  for (int k = 0; k < kSlots; ++k) {
    if (dev[k]) (void)hipFree(dev[k]);
    if (host[k]) (void)hipHostFree(host[k]);
    dev[k] = host[k] = nullptr;
    devCap[k] = hostCap[k] = 0;
  }
}

void unpack_bits_host(const uint32_t* words, size_t n, uint8_t* out) {
  const size_t full = n / 32;
  parallel_words(full, [&](size_t b, size_t e) {
    for (size_t w = b; w < e; ++w) {
      const uint32_t x = words[w];
      uint64_t o[4] = {g_lut.v[x & 0xFF], g_lut.v[(x >> 8) & 0xFF], g_lut.v[(x >> 16) & 0xFF],
                       g_lut.v[x >> 24]};
      std::memcpy(out + w * 32, o, 32);
    }
  });
  for (size_t i = full * 32; i < n; ++i) out[i] = uint8_t((words[full] >> (i - full * 32)) & 1u);
}

void pack_bits_host(const uint8_t* in, size_t n, uint32_t* words) {
  const size_t full = n / 32;
  parallel_words(full, [&](size_t b, size_t e) {
    // 16 bytes at a time: the byte-equal-zero mask, one bit per byte in order
    const __m128i zero = _mm_setzero_si128();
    for (size_t w = b; w < e; ++w) {
      const uint8_t* p = in + w * 32;
      const __m128i lo = _mm_loadu_si128(reinterpret_cast<const __m128i*>(p));
      const __m128i hi = _mm_loadu_si128(reinterpret_cast<const __m128i*>(p + 16));
      const uint32_t z = uint32_t(_mm_movemask_epi8(_mm_cmpeq_epi8(lo, zero))) |
                         (uint32_t(_mm_movemask_epi8(_mm_cmpeq_epi8(hi, zero))) << 16);
      words[w] = ~z;
    }
  });
  if (full * 32 < n) {
    uint32_t x = 0;
    for (size_t i = full * 32; i < n; ++i) x |= uint32_t(in[i] != 0) << (i - full * 32);
    words[full] = x;
  }
}

}  // namespace fs

extern "C" int fs_geometry_release(void) {
  for (int d = 0; d < fs::kMaxDevices; ++d) {
    std::lock_guard<std::mutex> lock(fs::g_scratch[d].mu);
    bool any = false;
    for (int k = 0; k < fs::GeomScratch::kSlots; ++k)
      any = any || fs::g_scratch[d].dev[k] || fs::g_scratch[d].host[k];
    if (!any) continue;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return fs::fail(FS_ERR_DEVICE, "hipGetDevice");
    if (hipSetDevice(d) != hipSuccess) return fs::fail(FS_ERR_DEVICE, "hipSetDevice(%d)", d);
    fs::g_scratch[d].release();
    (void)hipSetDevice(cur);
  }
  return FS_OK;
}
