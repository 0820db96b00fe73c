// probe.hip -- a plain streaming copy, the measured HBM ceiling that the
// lattice kernels are compared with.  Each thread moves four 16-byte
// vectors (all loads issued before the stores); consecutive lanes touch
// consecutive vectors, so every wave reads and writes whole 1 KiB spans.
#include "d3q19_dev.hpp"

namespace fs {

namespace {

constexpr int kVec = 4;

__global__ void __launch_bounds__(256) k_copy(const float4* __restrict__ src,
                                              float4* __restrict__ dst, size_t n4) {
  const size_t base = size_t(blockIdx.x) * (256 * kVec) + threadIdx.x;
  float4 v[kVec];
#pragma unroll
  for (int k = 0; k < kVec; ++k) {
    const size_t i = base + size_t(k) * 256;
    if (i < n4) v[k] = src[i];
  }
#pragma unroll
  for (int k = 0; k < kVec; ++k) {
    const size_t i = base + size_t(k) * 256;
    if (i < n4) dst[i] = v[k];
  }
}

}  // namespace

hipError_t launch_copy_probe(void* dst, const void* src, size_t bytes, hipStream_t st) {
  const size_t n4 = bytes / 16;
  if (n4 == 0) return hipSuccess;
  const size_t blocks = (n4 + 256 * kVec - 1) / (256 * kVec);
  hipLaunchKernelGGL(k_copy, dim3(unsigned(blocks)), dim3(256), 0, st,
                     reinterpret_cast<const float4*>(src), reinterpret_cast<float4*>(dst), n4);
  return hipGetLastError();
}

}  // namespace fs
