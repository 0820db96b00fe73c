// solver_impl.hpp -- the solver handle behind the C ABI, shared by the ABI
// sources (capi.cpp owns it; adjoint_api.cpp drives its kernels).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "fluidsim.h"
#include "fs_internal.hpp"

// Maps a failing HIP call to FS_ERR_DEVICE with the call's text.
#define FS_HIP(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return fs::fail(FS_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

struct fs_solver {
  int method = 0;
  fs::Geo g{};
  fs::Phys ph{};
  float tau = 0.6f;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  // populations (BGK / Hybrid)
  float* F = nullptr;      // the current populations
  float* F2 = nullptr;     // second buffer of the two-step kernels (swapped with F)
  bool f2_synced = false;  // F2's solid-cell slots equal F's
  int pop_variant = -1;    // kernel variant, -1 = default for the grid
  float* macro = nullptr;  // 4 planes: rho, ux, uy, uz
  int parity = 0;          // AA representation: 0 natural, 1 swapped
  // moments (HOME)
  float* mbuf[2] = {nullptr, nullptr};
  int cur = 0;
  int home_variant = -1;  // HOME kernel variant, -1 = default for the grid
  // flags
  uint8_t* flags = nullptr;  // derived flag byte used by the kernels
  uint32_t* nbm = nullptr;   // solid-neighbour masks (bit q: c + e_q is solid)
  uint8_t* stage = nullptr;  // public flag bytes uploaded from the host
  std::vector<uint8_t> hflags;
  // an outlet cell starts a 64-wide (32-wide) row, x % 64 == 0 (x % 32 == 0):
  // the tiled kernels of that width, which take an outlet's result from the
  // lane at x-1, do not apply
  bool outlet_tile_start = false;
  bool outlet_tile_start32 = false;
  // an outlet cell sits on the left ring column of a 64-wide tile (x % 64 ==
  // 63, or x = nx-1 for the first tile) and a cell that reads it across that
  // column (x+1, y +-1, z +-1) is neither inlet nor solid: the two-step HOME
  // tile, which cannot finish such an outlet, does not apply
  bool outlet_ring_read = false;
  float* inlet = nullptr;  // ny*nz
  float* scratch = nullptr;  // 19 planes for state conversion (lazy)

  float* mom_cur() const { return mbuf[cur]; }
};

// Kernel arguments of the population step on the solver's own buffers.
fs::PopLaunch pop_args(fs_solver* s);
// The HOME kernel variant the solver runs (its choice or the grid default;
// the plain gather while an outlet starts a 64-wide tile row).
int solver_home_variant(const fs_solver* s);
// Re-initialises the solver's state from macro fields (device pointers,
// nullable) without synchronising the stream.
int solver_reinit_async(fs_solver* s, const float* rho, const float* ux, const float* uy,
                        const float* uz);
