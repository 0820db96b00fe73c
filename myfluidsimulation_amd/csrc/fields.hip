// fields.hip -- small element-wise kernels behind the field helpers:
// |u| (field_ops.cu:12-28), initial density profiles (init_profiles.cu:7-45)
// and the stand-alone feq re-initialisation (lbm_reinit.cu:20-61).
#include "d3q19_dev.hpp"

namespace fs {

namespace {

constexpr int kBlock = 256;

inline dim3 grid1d(size_t n) { return dim3(unsigned((n + kBlock - 1) / kBlock)); }

__global__ void __launch_bounds__(kBlock) k_speed(const float* __restrict__ u,
                                                  const float* __restrict__ v,
                                                  const float* __restrict__ w,
                                                  float* __restrict__ out, size_t n) {
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const float a = u[i], b = v[i], c = w[i];
  out[i] = sqrtf(a * a + b * b + c * c);
}

// kern_diff / kern_shift (field_ops.cu:20-23,34-37): a - b, a + bias
__global__ void __launch_bounds__(kBlock) k_diff(const float* __restrict__ a,
                                                 const float* __restrict__ b,
                                                 float* __restrict__ out, size_t n) {
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) out[i] = a[i] - b[i];
}

__global__ void __launch_bounds__(kBlock) k_shift(const float* __restrict__ a, float bias,
                                                  float* __restrict__ out, size_t n) {
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) out[i] = a[i] + bias;
}

__global__ void __launch_bounds__(kBlock) k_rho_gauss(float* rho, int nx, int ny, int nz, float cx,
                                                      float cy, float cz, float sigma, float amp) {
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  const size_t n = size_t(nx) * ny * nz;
  if (i >= n) return;
  const int x = int(i % nx), y = int((i / nx) % ny), z = int(i / (size_t(nx) * ny));
  const float X = (x + 0.5f) / nx, Y = (y + 0.5f) / ny, Z = (z + 0.5f) / nz;
  const float dx = X - cx, dy = Y - cy, dz = Z - cz;
  const float r2 = dx * dx + dy * dy + dz * dz;
  rho[i] = 1.0f + amp * expf(-r2 / (2.f * sigma * sigma));
}

__global__ void __launch_bounds__(kBlock) k_rho_slab(float* rho, int nx, int ny, int nz, int axis,
                                                     float amp, float center, float width) {
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  const size_t n = size_t(nx) * ny * nz;
  if (i >= n) return;
  const int x = int(i % nx), y = int((i / nx) % ny), z = int(i / (size_t(nx) * ny));
  const float s = (axis == 0) ? (x + 0.5f) / nx : (axis == 1) ? (y + 0.5f) / ny : (z + 0.5f) / nz;
  const float val = (fabsf(s - center) < 0.5f * width) ? amp : 0.f;
  rho[i] = 1.0f + val;
}

__global__ void __launch_bounds__(kBlock) k_reinit_cs2(const float* rho, const float* ux,
                                                       const float* uy, const float* uz, float* f,
                                                       size_t n, size_t stride) {
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const float r = rho[i];
  const float u0 = ux ? ux[i] : 0.f;
  const float v0 = uy ? uy[i] : 0.f;
  const float w0 = uz ? uz[i] : 0.f;
  const float uu = u0 * u0 + v0 * v0 + w0 * w0;
  const WRho wr(r);
  for_q<0, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    f[size_t(q) * stride + i] = feq_cs2_q<q>(wr, u0, v0, w0, uu);
  });
}

}  // namespace

hipError_t launch_speed_k(const float* u, const float* v, const float* w, float* out, size_t n,
                          hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_speed, grid1d(n), dim3(kBlock), 0, st, u, v, w, out, n);
  return hipGetLastError();
}

hipError_t launch_diff_k(const float* a, const float* b, float* out, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_diff, grid1d(n), dim3(kBlock), 0, st, a, b, out, n);
  return hipGetLastError();
}

hipError_t launch_shift_k(const float* a, float bias, float* out, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_shift, grid1d(n), dim3(kBlock), 0, st, a, bias, out, n);
  return hipGetLastError();
}

hipError_t launch_rho_gauss(float* rho, int nx, int ny, int nz, float cx, float cy, float cz,
                            float sigma, float amp, hipStream_t st) {
  const size_t n = size_t(nx) * ny * nz;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_rho_gauss, grid1d(n), dim3(kBlock), 0, st, rho, nx, ny, nz, cx, cy, cz,
                     sigma, amp);
  return hipGetLastError();
}

hipError_t launch_rho_slab(float* rho, int nx, int ny, int nz, int axis, float amp, float center,
                           float width, hipStream_t st) {
  const size_t n = size_t(nx) * ny * nz;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_rho_slab, grid1d(n), dim3(kBlock), 0, st, rho, nx, ny, nz, axis, amp,
                     center, width);
  return hipGetLastError();
}

hipError_t launch_reinit_cs2(const float* rho, const float* ux, const float* uy, const float* uz,
                             float* f, size_t n, size_t stride, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_reinit_cs2, grid1d(n), dim3(kBlock), 0, st, rho, ux, uy, uz, f, n, stride);
  return hipGetLastError();
}

}  // namespace fs
