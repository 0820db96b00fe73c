// probe.hip -- a plain streaming copy, the measured HBM ceiling that the
// lattice kernels are compared with.  One 16-byte vector per thread with
// non-temporal loads and stores (the data is touched once, so it need not
// stay in the caches): the fastest of the copy forms measured on MI355X
// (one or four vectors per thread, cached or non-temporal).
#include "d3q19_dev.hpp"

namespace fs {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_copy(const v4f* __restrict__ src, v4f* __restrict__ dst,
                                              size_t n4) {
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (i < n4) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

}  // namespace

hipError_t launch_copy_probe(void* dst, const void* src, size_t bytes, hipStream_t st) {
  const size_t n4 = bytes / 16;
  if (n4 == 0) return hipSuccess;
  const size_t blocks = (n4 + 255) / 256;
  hipLaunchKernelGGL(k_copy, dim3(unsigned(blocks)), dim3(256), 0, st,
                     reinterpret_cast<const v4f*>(src), reinterpret_cast<v4f*>(dst), n4);
  return hipGetLastError();
}

}  // namespace fs
