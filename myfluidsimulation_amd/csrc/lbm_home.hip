// lbm_home.hip -- HOME moment-only step (LBM3D_Home) on gfx950.
//
// State: 10 moments per cell (rho, u, S = Pi - Pi_eq), two buffers, pointer
// swap per step (lbm3d_home.cu:288-354).  A cell gathers, for each of the 19
// directions, the regularised post-collision population of its upstream
// neighbour, rebuilt from that neighbour's moments (reconstruct_f_post,
// lbm3d_home.cu:37-81), or one of its own populations when the neighbour is
// solid (halfway bounce-back, lbm3d_home.cu:219-225: the opposite direction
// by default, the reference's own pairing table on request), and accumulates
// the new moments (lbm3d_home.cu:238-285).  Per-direction arithmetic is specialised
// at compile time (d3q19_dev.hpp) and keeps the reference's float order.
#include "d3q19_dev.hpp"

namespace fs {

namespace {

__device__ __forceinline__ void load_mom(const float* __restrict__ m, size_t P, uint32_t s,
                                         Mom& v) {
#pragma unroll
  for (int k = 0; k < 10; ++k) v.v[k] = m[size_t(k) * P + s];
}

// u = j/r (+F), S = M - Pi_eq (lbm3d_home.cu:252-285)
__device__ __forceinline__ void home_finish(float r, float jx, float jy, float jz, float Mxx,
                                            float Mxy, float Mxz, float Myy, float Myz, float Mzz,
                                            const Phys& ph, float (&out)[10]) {
  float ux = (r > 0.0f) ? (jx / r) : 0.0f;
  float uy = (r > 0.0f) ? (jy / r) : 0.0f;
  float uz = (r > 0.0f) ? (jz / r) : 0.0f;
  ux += ph.fx;
  uy += ph.fy;
  uz += ph.fz;
  constexpr float cs2 = 1.0f / 3.0f;
  const float rcs2 = r * cs2;
  out[0] = r;
  out[1] = ux;
  out[2] = uy;
  out[3] = uz;
  out[4] = Mxx - (rcs2 + (r * ux) * ux);
  out[5] = Mxy - (r * ux) * uy;
  out[6] = Mxz - (r * ux) * uz;
  out[7] = Myy - (rcs2 + (r * uy) * uy);
  out[8] = Myz - (r * uy) * uz;
  out[9] = Mzz - (rcs2 + (r * uz) * uz);
}

template <bool CHECK, int BB>
__device__ __forceinline__ void home_gather(const float* __restrict__ m, size_t P,
                                            const uint8_t* __restrict__ flags, uint32_t cc,
                                            const Nbr& o, const Phys& ph, float oneMinusOmega,
                                            float (&out)[10]) {
  Mom mc;
  load_mom(m, P, cc, mc);
  float r = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
  float Mxx = 0.0f, Mxy = 0.0f, Mxz = 0.0f, Myy = 0.0f, Myz = 0.0f, Mzz = 0.0f;
  for_q<0, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int a = CX[q], b = CY[q], c = CZ[q];
    float fq;
    if constexpr (q == 0) {
      fq = home_rec<0>(mc, oneMinusOmega);
    } else {
      const uint32_t s = cc + nb_off<-a, -b, -c>(o);
      if (CHECK && (flags[s] & F_SOLID)) {
        fq = home_rec<BB == 0 ? OPP[q] : OPP_REF[q]>(mc, oneMinusOmega);
      } else {
        Mom ms;
        load_mom(m, P, s, ms);
        fq = home_rec<q>(ms, oneMinusOmega);
      }
    }
    r += fq;
    acc1<a>(jx, fq);
    acc1<b>(jy, fq);
    acc1<c>(jz, fq);
    acc1<a * a>(Mxx, fq);
    acc1<a * b>(Mxy, fq);
    acc1<a * c>(Mxz, fq);
    acc1<b * b>(Myy, fq);
    acc1<b * c>(Myz, fq);
    acc1<c * c>(Mzz, fq);
  });
  home_finish(r, jx, jy, jz, Mxx, Mxy, Mxz, Myy, Myz, Mzz, ph, out);
}

template <int BB>
__global__ void __launch_bounds__(256) k_home_step(Geo g, Phys ph, HomeLaunch a) {
  Cell cl;
  if (!cell_of_thread(g, cl)) return;
  const uint32_t c = cl.c;
  const size_t P = g.P;
  const uint8_t fl = a.flags[c];
  float out[10];
  if (fl & F_SOLID) {
    // lbm3d_home.cu:180-188
#pragma unroll
    for (int k = 0; k < 10; ++k) out[k] = k == 0 ? 1.0f : 0.0f;
  } else if (fl & F_INLET) {
#pragma unroll
    for (int k = 0; k < 10; ++k) out[k] = 0.0f;
    out[0] = 1.0f;
    out[1] = a.inlet[size_t(cl.z) * g.ny + cl.y];
  } else {
    // an outlet cell stores the new moments of its upstream neighbour x-1
    const bool outlet = (fl & F_OUTLET) != 0;
    const int xc = outlet ? cl.x - 1 : cl.x;
    const uint32_t cc = outlet ? c - 1 : c;
    const uint8_t flc = outlet ? a.flags[cc] : fl;
    const Nbr o = nbr_offsets(xc, cl.y, cl.z, g);
    const float omm = 1.0f - ph.omegaL;
    if (__ballot((flc & F_NEAR) != 0) == 0)
      home_gather<false, BB>(a.m, P, a.flags, cc, o, ph, omm, out);
    else
      home_gather<true, BB>(a.m, P, a.flags, cc, o, ph, omm, out);
  }
#pragma unroll
  for (int k = 0; k < 10; ++k) a.mn[size_t(k) * P + c] = out[k];
}

// ---------------------------------------------------------------------------
// Tiled HOME step.  A workgroup owns a TX x TY column tile of cells and a
// chunk of ZC z-planes.  It walks the source planes z0-1 .. z0+ZC; for each
// plane every cell of the tile computes its 19 post-collision populations
// once (HomeSrc: direction-independent terms shared) into LDS, the halo ring
// computes only the five directions that enter the tile, and each thread
// picks up the 19 values that stream into its column from LDS.  Values for
// output plane o arrive from source planes o-1 (e_z = +1), o (e_z = 0) and
// o+1 (e_z = -1); they wait in registers until o+1 has been read, then the
// cell sums them in the reference's q order.  Per value the arithmetic is
// reconstruct_f_post's, per cell the sums are kern_step_moments_only's
// (lbm3d_home.cu:204-285), so the moments equal the simple kernel's bits.
// ---------------------------------------------------------------------------
// Directions by e_z: +1 {5,11,12,15,16}, 0 {0,1,2,3,4,7,8,9,10},
// -1 {6,13,14,17,18}; kSlot[q] is q's position inside its group.
constexpr int kSlot[19] = {0, 1, 2, 3, 4, 0, 0, 5, 6, 7, 8, 1, 2, 1, 2, 3, 4, 3, 4};

template <int TX, int TY>
struct TileLds {
  static constexpr int LW = TX + 2, LH = TY + 2, PL = LW * LH;
  __device__ static __forceinline__ int at(int q, int row, int col) { return q * PL + row * LW + col; }
};

// halo cell emission: only the directions that point into the tile
template <int TX, int TY, int Q0, int Q1, int Q2, int Q3, int Q4>
__device__ __forceinline__ void emit5(float* fs, const HomeSrc& h, int row, int col) {
  using L = TileLds<TX, TY>;
  fs[L::at(Q0, row, col)] = h.template emit<Q0>();
  fs[L::at(Q1, row, col)] = h.template emit<Q1>();
  fs[L::at(Q2, row, col)] = h.template emit<Q2>();
  fs[L::at(Q3, row, col)] = h.template emit<Q3>();
  fs[L::at(Q4, row, col)] = h.template emit<Q4>();
}

template <int TX, int TY, int ZC, int BB>
__global__ void __launch_bounds__(TX * TY) k_home_tiled(Geo g, Phys ph, HomeLaunch a) {
  using L = TileLds<TX, TY>;
  __shared__ float fs[19 * L::PL];
  const int tid = threadIdx.x;
  const int lx = tid % TX, ly = tid / TX;
  const int tilesX = (g.nx + TX - 1) / TX, tilesY = (g.ny + TY - 1) / TY;
  const int chunks = (g.nz + ZC - 1) / ZC;
  // Workgroups the dispatcher sends to one XCD (every 8th) get one
  // contiguous run of tiles, z-chunk fastest, so the planes and halo rows
  // that neighbouring tiles both read stay in that XCD's L2.
  uint32_t L0 = blockIdx.x;
  const uint32_t total = gridDim.x;
  if ((total & 7u) == 0) L0 = (L0 & 7u) * (total >> 3) + (L0 >> 3);
  const int bz = int(L0 % uint32_t(chunks));
  const int rest = int(L0 / uint32_t(chunks));
  const int bx = rest % tilesX, by = rest / tilesX;
  (void)tilesY;
  const int x0 = bx * TX, y0 = by * TY, z0 = bz * ZC;
  const int TXe = min(TX, g.nx - x0), TYe = min(TY, g.ny - y0), ZCe = min(ZC, g.nz - z0);
  const bool active = lx < TXe && ly < TYe;
  const int x = x0 + lx, y = y0 + ly;
  const size_t P = g.P;
  const float omm = 1.0f - ph.omegaL;

  // halo assignment: bottom row, top row, left column, right column
  const int H = 2 * (TXe + 2) + 2 * TYe;
  int side = -1, hx = 0, hy = 0;
  if (tid < H) {
    if (tid < TXe + 2) {
      side = 0, hx = tid - 1, hy = -1;
    } else if (tid < 2 * (TXe + 2)) {
      side = 1, hx = tid - (TXe + 2) - 1, hy = TYe;
    } else if (tid < 2 * (TXe + 2) + TYe) {
      side = 2, hx = -1, hy = tid - 2 * (TXe + 2);
    } else {
      side = 3, hx = TXe, hy = tid - 2 * (TXe + 2) - TYe;
    }
  }
  const int gx = (x0 + hx + g.nx) % g.nx, gy = (y0 + hy + g.ny) % g.ny;

  float aPrev2[5], aPrev[5], bPrev[9];
  Mom mPrev;
  uint8_t flPrev = 0;
#pragma unroll
  for (int k = 0; k < 5; ++k) aPrev2[k] = aPrev[k] = 0.0f;
#pragma unroll
  for (int k = 0; k < 9; ++k) bPrev[k] = 0.0f;
#pragma unroll
  for (int k = 0; k < 10; ++k) mPrev.v[k] = 0.0f;

  for (int p = z0 - 1; p <= z0 + ZCe; ++p) {
    const int pw = (p + g.nz) % g.nz;
    Mom cur;
    uint8_t flCur = 0;
    if (active) {
      const uint32_t c = (uint32_t(pw) * g.ny + y) * g.nx + x;
#pragma unroll
      for (int k = 0; k < 10; ++k) cur.v[k] = a.m[size_t(k) * P + c];
      flCur = a.flags[c];
      const HomeSrc h(cur, omm);
      for_q<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        fs[L::at(q, ly + 1, lx + 1)] = h.template emit<q>();
      });
    }
    if (side >= 0) {
      const uint32_t c = (uint32_t(pw) * g.ny + gy) * g.nx + gx;
      Mom hm;
#pragma unroll
      for (int k = 0; k < 10; ++k) hm.v[k] = a.m[size_t(k) * P + c];
      const HomeSrc h(hm, omm);
      const int row = hy + 1, col = hx + 1;
      if (side == 0)
        emit5<TX, TY, 3, 7, 8, 15, 17>(fs, h, row, col);  // e_y = +1
      else if (side == 1)
        emit5<TX, TY, 4, 9, 10, 16, 18>(fs, h, row, col);  // e_y = -1
      else if (side == 2)
        emit5<TX, TY, 1, 7, 9, 11, 13>(fs, h, row, col);  // e_x = +1
      else
        emit5<TX, TY, 2, 8, 10, 12, 14>(fs, h, row, col);  // e_x = -1
    }
    __syncthreads();
    if (active) {
      float A[5], B[9], Cd[5];
      for_q<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const float v = fs[L::at(q, ly - CY[q] + 1, lx - CX[q] + 1)];
        if constexpr (CZ[q] > 0) A[kSlot[q]] = v;
        if constexpr (CZ[q] == 0) B[kSlot[q]] = v;
        if constexpr (CZ[q] < 0) Cd[kSlot[q]] = v;
      });
      const int o = p - 1;
      if (o >= z0 && o < z0 + ZCe) {
        const int ow = (o + g.nz) % g.nz;
        const uint32_t co = (uint32_t(ow) * g.ny + y) * g.nx + x;
        const uint8_t fl = flPrev;
        // wave-uniform tests first (every active lane of the wave takes them)
        const bool fluid = (fl & (F_SOLID | F_INLET)) == 0;
        const bool anyNear = __ballot(fluid && (fl & F_NEAR) != 0) != 0;
        const bool outlet = (fl & F_OUTLET) != 0;
        const bool anyOutlet = __ballot(outlet) != 0;
        float out[10];
        if (fl & F_SOLID) {
#pragma unroll
          for (int k = 0; k < 10; ++k) out[k] = k == 0 ? 1.0f : 0.0f;
        } else if (fl & F_INLET) {
#pragma unroll
          for (int k = 0; k < 10; ++k) out[k] = 0.0f;
          out[0] = 1.0f;
          out[1] = a.inlet[size_t(ow) * g.ny + y];
        } else {
          float f[19];
          for_q<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if constexpr (CZ[q] > 0) f[q] = aPrev2[kSlot[q]];
            if constexpr (CZ[q] == 0) f[q] = bPrev[kSlot[q]];
            if constexpr (CZ[q] < 0) f[q] = Cd[kSlot[q]];
          });
          if (anyNear) {
            const Nbr on = nbr_offsets(x, y, ow, g);
            for_q<1, 19>([&](auto qc) {
              constexpr int q = decltype(qc)::value;
              const uint32_t s = co + nb_off<-CX[q], -CY[q], -CZ[q]>(on);
              if (a.flags[s] & F_SOLID) f[q] = home_rec<BB == 0 ? OPP[q] : OPP_REF[q]>(mPrev, omm);
            });
          }
          float r = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
          float Mxx = 0.0f, Mxy = 0.0f, Mxz = 0.0f, Myy = 0.0f, Myz = 0.0f, Mzz = 0.0f;
          for_q<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int ca = CX[q], cb = CY[q], cc = CZ[q];
            const float fq = f[q];
            r += fq;
            acc1<ca>(jx, fq);
            acc1<cb>(jy, fq);
            acc1<cc>(jz, fq);
            acc1<ca * ca>(Mxx, fq);
            acc1<ca * cb>(Mxy, fq);
            acc1<ca * cc>(Mxz, fq);
            acc1<cb * cb>(Myy, fq);
            acc1<cb * cc>(Myz, fq);
            acc1<cc * cc>(Mzz, fq);
          });
          home_finish(r, jx, jy, jz, Mxx, Mxy, Mxz, Myy, Myz, Mzz, ph, out);
        }
        // zero-gradient outlet: the new moments of x-1, the previous lane
        // (the launcher keeps x-1 of an outlet inside the same tile row)
        if (anyOutlet) {
#pragma unroll
          for (int k = 0; k < 10; ++k) {
            const float left = __shfl_up(out[k], 1);
            if (outlet) out[k] = left;
          }
        }
#pragma unroll
        for (int k = 0; k < 10; ++k) a.mn[size_t(k) * P + co] = out[k];
      }
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        aPrev2[k] = aPrev[k];
        aPrev[k] = A[k];
      }
#pragma unroll
      for (int k = 0; k < 9; ++k) bPrev[k] = B[k];
      mPrev = cur;
      flPrev = flCur;
    }
    __syncthreads();
  }
}

// kern_reset_mom / kern_reinit_from_macro (lbm3d_home.cu:83-137): both buffers
__global__ void __launch_bounds__(256) k_home_fill(Geo g, float* m, float* mn, const float* rho,
                                                   const float* ux, const float* uy,
                                                   const float* uz) {
  Cell cl;
  if (!cell_of_thread(g, cl)) return;
  const uint32_t c = cl.c;
  float v[10] = {rho ? rho[c] : 1.0f, ux ? ux[c] : 0.0f, uy ? uy[c] : 0.0f, uz ? uz[c] : 0.0f,
                 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    m[size_t(k) * g.P + c] = v[k];
    mn[size_t(k) * g.P + c] = v[k];
  }
}

inline dim3 lattice_grid(const Geo& g) { return dim3((g.nxy + 255) / 256, g.nz, 1); }

template <int TY, int ZC>
hipError_t launch_tiled(const Geo& g, const Phys& ph, const HomeLaunch& a, hipStream_t st) {
  constexpr int TX = 64;
  const unsigned blocks =
      unsigned(((g.nx + TX - 1) / TX) * ((g.ny + TY - 1) / TY) * ((g.nz + ZC - 1) / ZC));
  if (ph.bounce == 0)
    hipLaunchKernelGGL((k_home_tiled<TX, TY, ZC, 0>), dim3(blocks), dim3(TX * TY), 0, st, g, ph, a);
  else
    hipLaunchKernelGGL((k_home_tiled<TX, TY, ZC, 1>), dim3(blocks), dim3(TX * TY), 0, st, g, ph, a);
  return hipGetLastError();
}

}  // namespace

bool home_variant_ok(const Geo& g, int variant) {
  if (variant == kHomePlain) return true;
  if (variant < kHomePlain || variant >= kHomeVariants) return false;
  // the tiled kernels take an outlet's moments from the lane at x-1, which
  // must sit in the same 64-wide tile: not so when nx-1 starts a tile
  return (g.nx - 1) % 64 != 0;
}

int home_default_variant(const Geo& g) {
  if (!home_variant_ok(g, kHomeT8Z32)) return kHomePlain;
  // deep chunks cut the re-computed halo planes, but only pay when they
  // still leave enough workgroups (>= 4 per CU) to fill the chip
  const long tiles8 = long((g.nx + 63) / 64) * ((g.ny + 7) / 8);
  if (g.nz >= 128 && tiles8 * ((g.nz + 31) / 32) >= 1024) return kHomeT8Z32;
  return kHomeT4Z16;
}

hipError_t launch_home_step(const Geo& g, const Phys& ph, const HomeLaunch& a, int variant,
                            hipStream_t st) {
  if (variant < 0 || !home_variant_ok(g, variant)) variant = home_default_variant(g);
  switch (variant) {
    case kHomeT4Z16: return launch_tiled<4, 16>(g, ph, a, st);
    case kHomeT8Z16: return launch_tiled<8, 16>(g, ph, a, st);
    case kHomeT8Z32: return launch_tiled<8, 32>(g, ph, a, st);
    default: break;
  }
  if (ph.bounce == 0)
    hipLaunchKernelGGL(k_home_step<0>, lattice_grid(g), dim3(256), 0, st, g, ph, a);
  else
    hipLaunchKernelGGL(k_home_step<1>, lattice_grid(g), dim3(256), 0, st, g, ph, a);
  return hipGetLastError();
}

hipError_t launch_home_fill(const Geo& g, float* m, float* mn, const float* rho, const float* ux,
                            const float* uy, const float* uz, hipStream_t st) {
  hipLaunchKernelGGL(k_home_fill, lattice_grid(g), dim3(256), 0, st, g, m, mn, rho, ux, uy, uz);
  return hipGetLastError();
}

}  // namespace fs
