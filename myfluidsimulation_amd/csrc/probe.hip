// probe.hip -- a plain streaming copy, the measured HBM ceiling that the
// lattice kernels are compared with (16-byte vector loads and stores,
// grid-stride loop over a grid sized to fill every CU several times).
#include "fs_internal.hpp"

namespace fs {

namespace {

__global__ void __launch_bounds__(256) k_copy(const float4* __restrict__ src,
                                              float4* __restrict__ dst, size_t n4) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n4; i += stride)
    dst[i] = src[i];
}

}  // namespace

hipError_t launch_copy_probe(void* dst, const void* src, size_t bytes, hipStream_t st) {
  const size_t n4 = bytes / 16;
  if (n4 == 0) return hipSuccess;
  size_t blocks = (n4 + 255) / 256;
  if (blocks > 256 * 64) blocks = 256 * 64;
  hipLaunchKernelGGL(k_copy, dim3(unsigned(blocks)), dim3(256), 0, st,
                     reinterpret_cast<const float4*>(src), reinterpret_cast<float4*>(dst), n4);
  return hipGetLastError();
}

}  // namespace fs
