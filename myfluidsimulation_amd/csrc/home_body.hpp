// home_body.hpp -- HOME update bodies shared by the single-domain kernels
// (lbm_home.hip) and the batched kernel (batch.hip): the per-cell gather and
// the tiled plane walk.  See lbm_home.hip for the scheme.
#pragma once

#include "d3q19_dev.hpp"

namespace fs {

namespace {

// New moments are stored non-temporally: nothing reads them in this launch,
// so they should not take L2 lines from the halo rows neighbouring tiles
// read (measured +2-3% on the 512^3 vortex and the 768x256x256 channel).
__device__ __forceinline__ void st_mn(float* p, float v) { __builtin_nontemporal_store(v, p); }

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

// One tile (bx, by) and z-chunk bz of the tiled HOME step; every thread of
// the workgroup must call it (it synchronises the workgroup).  fs: LDS of
// 19 * TileLds<TX,TY>::PL floats.
template <int TX, int TY, int ZC, int BB>
__device__ __forceinline__ void home_tile(const Geo& g, const Phys& ph, const HomeLaunch& a,
                                          int bx, int by, int bz, float* fs) {
  using L = TileLds<TX, TY>;
  const int tid = threadIdx.x;
  const int lx = tid % TX, ly = tid / TX;
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

  // The moments of source plane p+1 are loaded while plane p is computed
  // (one plane of register prefetch), so HBM latency overlaps the
  // reconstruction and the two barriers of each plane.
  Mom nxtC, nxtH;
  uint8_t nxtFl = 0;
  auto load_plane = [&](int p) {
    const int pw = (p + g.nz) % g.nz;
    if (active) {
      const uint32_t c = (uint32_t(pw) * g.ny + y) * g.nx + x;
#pragma unroll
      for (int k = 0; k < 10; ++k) nxtC.v[k] = a.m[size_t(k) * P + c];
      nxtFl = a.flags[c];
    }
    if (side >= 0) {
      const uint32_t c = (uint32_t(pw) * g.ny + gy) * g.nx + gx;
#pragma unroll
      for (int k = 0; k < 10; ++k) nxtH.v[k] = a.m[size_t(k) * P + c];
    }
  };
  load_plane(z0 - 1);

  for (int p = z0 - 1; p <= z0 + ZCe; ++p) {
    const Mom cur = nxtC;
    const uint8_t flCur = nxtFl;
    // an inlet cell finished below (plane p-1) needs its velocity: the load
    // is issued here, so the reconstruction and the barrier hide its latency
    float uinFin = 0.0f;
    if (active && (flPrev & F_INLET)) uinFin = a.inlet[size_t((p - 1 + g.nz) % g.nz) * g.ny + y];
    // likewise a near-wall cell's solid-neighbour mask (one word instead of
    // 18 neighbour flag bytes read where they are needed)
    uint32_t nbmFin = 0u;
    if (active && (flPrev & F_NEAR))
      nbmFin = a.nbm[(uint32_t((p - 1 + g.nz) % g.nz) * g.ny + y) * g.nx + x];
    const Mom hm = nxtH;
    if (p < z0 + ZCe) load_plane(p + 1);
    if (active) {
      const HomeSrc h(cur, omm);
      for_q<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        fs[L::at(q, ly + 1, lx + 1)] = h.template emit<q>();
      });
    }
    if (side >= 0) {
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
    // gather phase: every lane takes part in the wave-wide votes and the
    // shuffle below; lanes outside the tile (or with no output plane yet)
    // vote false and store nothing
    float A[5], B[9], Cd[5];
    if (active) {
      for_q<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const float v = fs[L::at(q, ly - CY[q] + 1, lx - CX[q] + 1)];
        if constexpr (CZ[q] > 0) A[kSlot[q]] = v;
        if constexpr (CZ[q] == 0) B[kSlot[q]] = v;
        if constexpr (CZ[q] < 0) Cd[kSlot[q]] = v;
      });
    }
    const int o = p - 1;
    const bool fin = active && o >= z0 && o < z0 + ZCe;
    const int ow = (o + g.nz) % g.nz;
    const uint32_t co = (uint32_t(ow) * g.ny + (fin ? y : 0)) * g.nx + (fin ? x : 0);
    const uint8_t fl = fin ? flPrev : uint8_t(0);
    const bool fluid = fin && (fl & (F_SOLID | F_INLET)) == 0;
    const bool anyNear = __ballot(fluid && (fl & F_NEAR) != 0) != 0;
    const bool outlet = fin && (fl & F_OUTLET) != 0;
    const bool anyOutlet = __ballot(outlet) != 0;
    float out[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) out[k] = 0.0f;
    if (fin) {
      if (fl & F_SOLID) {
        out[0] = 1.0f;
      } else if (fl & F_INLET) {
        out[0] = 1.0f;
        out[1] = uinFin;
      } else {
        float f[19];
        for_q<0, 19>([&](auto qc) {
          constexpr int q = decltype(qc)::value;
          if constexpr (CZ[q] > 0) f[q] = aPrev2[kSlot[q]];
          if constexpr (CZ[q] == 0) f[q] = bPrev[kSlot[q]];
          if constexpr (CZ[q] < 0) f[q] = Cd[kSlot[q]];
        });
        if (anyNear) {
          // c - e_q is the neighbour along e_opp(q)
          for_q<1, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if ((nbmFin >> OPP[q]) & 1u)
              f[q] = home_rec<BB == 0 ? OPP[q] : OPP_REF[q]>(mPrev, omm);
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
    }
    // zero-gradient outlet: the new moments of x-1, the previous lane (an
    // outlet never starts a tile row: the launchers check it)
    if (anyOutlet) {
#pragma unroll
      for (int k = 0; k < 10; ++k) {
        const float left = __shfl_up(out[k], 1);
        if (outlet) out[k] = left;
      }
    }
    if (fin) {
#pragma unroll
      for (int k = 0; k < 10; ++k) st_mn(a.mn + size_t(k) * P + co, out[k]);
    }
    if (active) {
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

}  // namespace

}  // namespace fs

namespace fs {

namespace {

// ---------------------------------------------------------------------------
// Tiled HOME step, second form (home_tile2).  Same scheme and the same
// per-value and per-cell arithmetic as home_tile, organised for fewer vector
// instructions per lattice update:
//  * the values that wait for the plane above (a three-plane ring), the
//    prefetched moments and the flags live in fixed register slots that
//    rotate once per plane;
//  * the three directions that stay in the cell's own column (rest, +z, -z)
//    never go through LDS: 16 LDS planes instead of 19;
//  * each tile row has one halo role: row 0 the row below the tile, row 1
//    the row above, row 2 the left and right columns (one code path for
//    both, the sign of e_x selected per lane) and the four corners (one
//    direction each); with 64-wide tiles no wave runs two halo branches;
//  * a wave (one 64-wide row or two 32-wide ones) whose 64 output cells are
//    all plain fluid (no solid neighbour, inlet, outlet) takes a path without
//    the per-lane flag handling.
// ---------------------------------------------------------------------------
// LDS plane of each direction with (e_x, e_y) != (0, 0); -1: kept in registers
constexpr int kLds16[19] = {-1, 0, 1, 2, 3, -1, -1, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15};

template <int TX, int TY>
struct TileLds16 {
  static constexpr int LW = TX + 2, LH = TY + 2, PL = LW * LH;
  __device__ static __forceinline__ int at(int slot, int row, int col) {
    return slot * PL + row * LW + col;
  }
};

// emit<Q>'s arithmetic for a direction whose x component is sign sx (a
// lane value): e_x = +1 gives (ux, sxy, sxz) as they are, e_x = -1 their
// negations, which enter every expression exactly as the compile-time signs
// do (negation is exact).  Y, Z: the other two components (compile time).
template <int CY, int CZ>
__device__ __forceinline__ float emit_sx(const HomeSrc& h, float sux, float ssxy, float ssxz) {
  constexpr float w = (CY == 0 && CZ == 0) ? 1.0f / 18.0f : 1.0f / 36.0f;
  constexpr float w45 = 4.5f * w;
  const float wr = (CY == 0 && CZ == 0) ? h.wr1 : h.wr7;
  float eu;
  if constexpr (CY == 0 && CZ == 0)
    eu = sux;
  else if constexpr (CZ == 0)
    eu = sux + sg<CY>(h.uy);
  else
    eu = sux + sg<CZ>(h.uz);
  const float feq = wr * (((1.0f + 3.0f * eu) + (4.5f * eu) * eu) - h.uu15);
  float contraction = (h.pA[0] + (CY ? h.pA[1] : h.pB[1])) + (CZ ? h.pA[2] : h.pB[2]);
  if constexpr (CY != 0) contraction += 2.0f * sg<CY>(ssxy);
  if constexpr (CZ != 0) contraction += 2.0f * sg<CZ>(ssxz);
  return feq + h.omm * (w45 * contraction);
}

// the in-plane diagonal (sx, sy, 0) with both signs lane values
__device__ __forceinline__ float emit_corner(const HomeSrc& h, float sux, float suy, float ssxy) {
  constexpr float w45 = 4.5f * (1.0f / 36.0f);
  const float eu = sux + suy;
  const float feq = h.wr7 * (((1.0f + 3.0f * eu) + (4.5f * eu) * eu) - h.uu15);
  float contraction = (h.pA[0] + h.pA[1]) + h.pB[2];
  contraction += 2.0f * ssxy;
  return feq + h.omm * (w45 * contraction);
}

// emit<Q>'s arithmetic for a direction whose y component is sign sy (a
// lane value), as emit_sx for x: (CX, sy, CZ) with at most one of CX, CZ
// nonzero; suy, ssxy, ssyz are u_y, S_xy, S_yz times sy.
template <int CX, int CZ>
__device__ __forceinline__ float emit_sy(const HomeSrc& h, float suy, float ssxy, float ssyz) {
  constexpr float w = (CX == 0 && CZ == 0) ? 1.0f / 18.0f : 1.0f / 36.0f;
  constexpr float w45 = 4.5f * w;
  const float wr = (CX == 0 && CZ == 0) ? h.wr1 : h.wr7;
  float eu;
  if constexpr (CX == 0 && CZ == 0)
    eu = suy;
  else if constexpr (CZ == 0)
    eu = sg<CX>(h.ux) + suy;
  else
    eu = suy + sg<CZ>(h.uz);
  const float feq = wr * (((1.0f + 3.0f * eu) + (4.5f * eu) * eu) - h.uu15);
  float contraction = ((CX ? h.pA[0] : h.pB[0]) + h.pA[1]) + (CZ ? h.pA[2] : h.pB[2]);
  if constexpr (CX != 0) contraction += 2.0f * sg<CX>(ssxy);
  if constexpr (CZ != 0) contraction += 2.0f * sg<CZ>(ssyz);
  return feq + h.omm * (w45 * contraction);
}

template <int TX, int TY, int ZC, int BB>
__device__ __forceinline__ void home_tile2(const Geo& g, const Phys& ph, const HomeLaunch& a,
                                           int bx, int by, int bz, float* fs) {
  // rows 0, 1 and 2 of the tile carry the halo (one or two rows per wave)
  static_assert(TX == 64 || TX == 32, "a tile row is a wave or half of one");
  static_assert(TY >= 3 && 2 * TY + 4 <= TX, "the columns and corners fit row 2");
  using L = TileLds16<TX, TY>;
  const int tid = threadIdx.x;
  const int lx = tid % TX, ly = tid / TX;
  const int x0 = bx * TX, y0 = by * TY, z0 = bz * ZC;
  const int TXe = min(TX, g.nx - x0), TYe = min(TY, g.ny - y0), ZCe = min(ZC, g.nz - z0);
  const bool active = lx < TXe && ly < TYe;
  const int x = x0 + lx, y = y0 + ly;
  const size_t P = g.P;
  const float omm = 1.0f - ph.omegaL;

  // halo role of this lane: 0 row below, 1 row above, 2 left / right column,
  // 3 corner, -1 none; (hx, hy) tile-relative
  int role = -1, hx = 0, hy = 0, hsgn = 1, hsgny = 1;
  if (ly == 0 && lx < TXe) {
    role = 0, hx = lx, hy = -1;
  } else if (ly == 1 && lx < TXe) {
    role = 1, hx = lx, hy = TYe;
  } else if (ly == 2) {
    if (lx < 2 * TYe) {
      role = 2;
      const bool left = lx < TYe;
      hx = left ? -1 : TXe;
      hy = left ? lx : lx - TYe;
      hsgn = left ? 1 : -1;  // e_x of the directions that enter the tile
    } else if (lx < 2 * TYe + 4) {
      role = 3;
      const int c = lx - 2 * TYe;  // 0 (-1,-1) 1 (TXe,-1) 2 (-1,TYe) 3 (TXe,TYe)
      hx = (c & 1) ? TXe : -1;
      hy = (c & 2) ? TYe : -1;
      hsgn = (c & 1) ? -1 : 1;
      hsgny = (c & 2) ? -1 : 1;
    }
  }
  const int gx = (x0 + hx + g.nx) % g.nx, gy = (y0 + hy + g.ny) % g.ny;
  // corner c's direction (sx, sy, 0) is q = 7 + c; its LDS plane 4 + c
  const int cornerSlot = 4 + ((hsgn < 0) ? 1 : 0) + ((hsgny < 0) ? 2 : 0);
  // the left column writes the planes of q = 1, 7, 9, 11, 13 (0, 4, 6, 8, 10),
  // the right column those of 2, 8, 10, 12, 14 (one further each)
  const int colSlot0 = hsgn > 0 ? 0 : 1;

  Mom Mm[3], Hm[3];
  uint8_t Fl[3] = {0, 0, 0};
  float Ar[3][5], Br[3][9];
#pragma unroll
  for (int s = 0; s < 3; ++s) {
#pragma unroll
    for (int k = 0; k < 10; ++k) Mm[s].v[k] = Hm[s].v[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 5; ++k) Ar[s][k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 9; ++k) Br[s][k] = 0.0f;
  }

  auto load_plane = [&](int p, Mom& mc, uint8_t& fc, Mom& mh) {
    const int pw = (p + g.nz) % g.nz;
    if (active) {
      const uint32_t c = (uint32_t(pw) * g.ny + y) * g.nx + x;
#pragma unroll
      for (int k = 0; k < 10; ++k) mc.v[k] = a.m[size_t(k) * P + c];
      fc = a.flags[c];
    }
    if (role >= 0) {
      const uint32_t c = (uint32_t(pw) * g.ny + gy) * g.nx + gx;
#pragma unroll
      for (int k = 0; k < 10; ++k) mh.v[k] = a.m[size_t(k) * P + c];
    }
  };
  load_plane(z0 - 1, Mm[0], Fl[0], Hm[0]);
  const int last = z0 + ZCe;

  auto step = [&](auto phc, int p) {
    constexpr int PH = decltype(phc)::value;
    constexpr int NX = (PH + 1) % 3, PV = (PH + 2) % 3;
    const Mom& cur = Mm[PH];
    // the output cell of this plane step is (x, y, p-1): its flags, and for
    // an inlet its velocity, for a near-wall cell its solid-neighbour mask
    // (loads issued here, so the reconstruction hides their latency)
    const uint8_t flO = Fl[PV];
    float uinFin = 0.0f;
    if (active && (flO & F_INLET)) uinFin = a.inlet[size_t((p - 1 + g.nz) % g.nz) * g.ny + y];
    uint32_t nbmFin = 0u;
    if (active && (flO & F_NEAR))
      nbmFin = a.nbm[(uint32_t((p - 1 + g.nz) % g.nz) * g.ny + y) * g.nx + x];
    if (p < last) load_plane(p + 1, Mm[NX], Fl[NX], Hm[NX]);
    float f0 = 0.0f, f5 = 0.0f, f6 = 0.0f;
    if (active) {
      const HomeSrc h(cur, omm);
      for_q<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (kLds16[q] >= 0) fs[L::at(kLds16[q], ly + 1, lx + 1)] = h.template emit<q>();
      });
      f0 = h.template emit<0>();
      f5 = h.template emit<5>();
      f6 = h.template emit<6>();
    }
    if (role >= 0) {
      const HomeSrc h(Hm[PH], omm);
      const int row = hy + 1, col = hx + 1;
      if (role == 0) {  // e_y = +1: q = 3, 7, 8, 15, 17
        fs[L::at(kLds16[3], row, col)] = h.template emit<3>();
        fs[L::at(kLds16[7], row, col)] = h.template emit<7>();
        fs[L::at(kLds16[8], row, col)] = h.template emit<8>();
        fs[L::at(kLds16[15], row, col)] = h.template emit<15>();
        fs[L::at(kLds16[17], row, col)] = h.template emit<17>();
      } else if (role == 1) {  // e_y = -1: q = 4, 9, 10, 16, 18
        fs[L::at(kLds16[4], row, col)] = h.template emit<4>();
        fs[L::at(kLds16[9], row, col)] = h.template emit<9>();
        fs[L::at(kLds16[10], row, col)] = h.template emit<10>();
        fs[L::at(kLds16[16], row, col)] = h.template emit<16>();
        fs[L::at(kLds16[18], row, col)] = h.template emit<18>();
      } else if (role == 2) {  // e_x = hsgn: (s,0,0) (s,1,0) (s,-1,0) (s,0,1) (s,0,-1)
        const bool pos = hsgn > 0;
        const float sux = pos ? h.ux : -h.ux, ssxy = pos ? h.sxy : -h.sxy,
                    ssxz = pos ? h.sxz : -h.sxz;
        fs[L::at(colSlot0 + 0, row, col)] = emit_sx<0, 0>(h, sux, ssxy, ssxz);
        fs[L::at(colSlot0 + 4, row, col)] = emit_sx<1, 0>(h, sux, ssxy, ssxz);
        fs[L::at(colSlot0 + 6, row, col)] = emit_sx<-1, 0>(h, sux, ssxy, ssxz);
        fs[L::at(colSlot0 + 8, row, col)] = emit_sx<0, 1>(h, sux, ssxy, ssxz);
        fs[L::at(colSlot0 + 10, row, col)] = emit_sx<0, -1>(h, sux, ssxy, ssxz);
      } else {  // corner: (sx, sy, 0)
        const float sux = hsgn > 0 ? h.ux : -h.ux, suy = hsgny > 0 ? h.uy : -h.uy;
        const float ssxy = (hsgn * hsgny) > 0 ? h.sxy : -h.sxy;
        fs[L::at(cornerSlot, row, col)] = emit_corner(h, sux, suy, ssxy);
      }
    }
    __syncthreads();
    // incoming values of this plane: A (e_z = +1) wait for output p+1, B
    // (e_z = 0) for output p, C (e_z = -1) go to output p-1 now
    float Cd[5];
    if (active) {
      float* A = Ar[PH];
      float* B = Br[PH];
      for_q<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        float v;
        if constexpr (q == 0)
          v = f0;
        else if constexpr (q == 5)
          v = f5;
        else if constexpr (q == 6)
          v = f6;
        else
          v = fs[L::at(kLds16[q], ly - CY[q] + 1, lx - CX[q] + 1)];
        if constexpr (CZ[q] > 0) A[kSlot[q]] = v;
        if constexpr (CZ[q] == 0) B[kSlot[q]] = v;
        if constexpr (CZ[q] < 0) Cd[kSlot[q]] = v;
      });
    }
    const int o = p - 1;
    const bool fin = active && o >= z0 && o < last;
    const float* Ao = Ar[NX];  // from plane o-1 (two steps back)
    const float* Bo = Br[PV];  // from plane o
    auto gather = [&](float (&f)[19]) {
      for_q<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (CZ[q] > 0) f[q] = Ao[kSlot[q]];
        if constexpr (CZ[q] == 0) f[q] = Bo[kSlot[q]];
        if constexpr (CZ[q] < 0) f[q] = Cd[kSlot[q]];
      });
    };
    auto moments = [&](const float (&f)[19], float (&out)[10]) {
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
    };
    const int ow = (o + g.nz) % g.nz;
    const bool plain = fin && (flO & (F_SOLID | F_INLET | F_OUTLET | F_NEAR)) == 0;
    if (__ballot(plain) == ~0ull) {
      // the whole row is plain fluid
      float f[19], out[10];
      gather(f);
      moments(f, out);
      const uint32_t co = (uint32_t(ow) * g.ny + y) * g.nx + x;
#pragma unroll
      for (int k = 0; k < 10; ++k) st_mn(a.mn + size_t(k) * P + co, out[k]);
    } else {
      const uint32_t co = (uint32_t(ow) * g.ny + (fin ? y : 0)) * g.nx + (fin ? x : 0);
      const uint8_t fl = fin ? flO : uint8_t(0);
      const bool fluid = fin && (fl & (F_SOLID | F_INLET)) == 0;
      const bool anyNear = __ballot(fluid && (fl & F_NEAR) != 0) != 0;
      const bool outlet = fin && (fl & F_OUTLET) != 0;
      const bool anyOutlet = __ballot(outlet) != 0;
      float out[10];
#pragma unroll
      for (int k = 0; k < 10; ++k) out[k] = 0.0f;
      if (fin) {
        if (fl & F_SOLID) {
          out[0] = 1.0f;
        } else if (fl & F_INLET) {
          out[0] = 1.0f;
          out[1] = uinFin;
        } else {
          float f[19];
          gather(f);
          if (anyNear) {
            // c - e_q is the neighbour along e_opp(q)
            const Mom& mPrev = Mm[PV];
            for_q<1, 19>([&](auto qc) {
              constexpr int q = decltype(qc)::value;
              if ((nbmFin >> OPP[q]) & 1u)
                f[q] = home_rec<BB == 0 ? OPP[q] : OPP_REF[q]>(mPrev, omm);
            });
          }
          moments(f, out);
        }
      }
      // zero-gradient outlet: the new moments of x-1, the previous lane (an
      // outlet never starts a tile row: the launchers check it)
      if (anyOutlet) {
#pragma unroll
        for (int k = 0; k < 10; ++k) {
          const float left = __shfl_up(out[k], 1);
          if (outlet) out[k] = left;
        }
      }
      if (fin) {
#pragma unroll
        for (int k = 0; k < 10; ++k) st_mn(a.mn + size_t(k) * P + co, out[k]);
      }
    }
    __syncthreads();
  };

  // one loop body on fixed slots (0: this plane, 1: the prefetch, 2: the
  // plane below); the slots rotate after each plane.  (A body unrolled over
  // the three slot phases needs no rotation but took 170 VGPRs, three waves
  // per SIMD, and ran 15% slower.)
  using I0 = std::integral_constant<int, 0>;
  for (int p = z0 - 1; p <= last; ++p) {
    step(I0{}, p);
    const Mom tm = Mm[0];
    Mm[0] = Mm[1];
    Mm[2] = tm;
    Hm[0] = Hm[1];
    const uint8_t tf = Fl[0];
    Fl[0] = Fl[1];
    Fl[2] = tf;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      Ar[1][k] = Ar[2][k];
      Ar[2][k] = Ar[0][k];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) Br[2][k] = Br[0][k];
  }
}

}  // namespace

}  // namespace fs

namespace fs {

namespace {

// ---------------------------------------------------------------------------
// Tiled HOME step, third form (home_tile3).  The scheme, LDS layout, halo
// roles and per-value arithmetic of home_tile2 (so the same bits), with fewer
// vector instructions per lattice update:
//  * a tile cell emits each direction together with its opposite
//    (HomeSrc::emit_pair): the two share e.u, 3*e.u, (4.5*e.u)*e.u and the
//    contraction;
//  * the plane walk is unrolled six planes deep, so every register ring
//    slot (moments: two planes, e_z = 0 values: two, e_z = +1 values and
//    flags: three) is fixed at compile time and no slot is copied;
//  * a near-wall cell reads its own moments back (from the input buffer, the
//    plane the tile read two planes earlier) instead of keeping a third
//    plane of moments in registers;
//  * moments are read at scalar plane bases with 32-bit lane offsets and
//    stored non-temporally.
// The flags of a plane are read with that plane's moments, two planes before
// its cells are finished: read one plane later (with the plane's own step)
// the 512^3 vortex fetched 8.5 instead of 6.9 GB per step from beyond L2.
// ---------------------------------------------------------------------------
// Plane k of a moment buffer read or written at a 32-bit byte offset: the
// plane's base is the same for every lane (scalar registers) and a lane's
// offset is one 32-bit register, so no lane builds 64-bit addresses.  Every
// plane must then be under 4 GiB (grids of fewer than 2^30 cells; the
// launchers check it).
// The empty asm pins the (uniform) base in scalar registers and keeps the
// compiler from folding the lane offset into it first, which would give every
// plane its own 64-bit vector address.
// (The accesses name the global address space: through the opaque base the
// compiler could no longer infer it and would emit flat accesses.)
using gbyte = __attribute__((address_space(1))) char;
__device__ __forceinline__ float ld_b(const float* base, uint32_t boff) {
  asm volatile("" : "+s"(base));
  const gbyte* b = (const gbyte*)(base);
  return *(const __attribute__((address_space(1))) float*)(b + boff);
}
__device__ __forceinline__ void st_b(float* base, uint32_t boff, float v) {
  asm volatile("" : "+s"(base));
  gbyte* b = (gbyte*)(base);
  __builtin_nontemporal_store(v, (__attribute__((address_space(1))) float*)(b + boff));
}

template <int TX, int TY, int ZC, int BB>
__device__ __forceinline__ void home_tile3(const Geo& g, const Phys& ph, const HomeLaunch& a,
                                           int bx, int by, int bz, float* fs) {
  static_assert(TX == 64 || TX == 32, "a tile row is a wave or half of one");
  static_assert(TY >= 3 && 2 * TY + 4 <= TX, "the columns and corners fit row 2");
  using L = TileLds16<TX, TY>;
  const int tid = threadIdx.x;
  const int lx = tid % TX, ly = tid / TX;
  const int x0 = bx * TX, y0 = by * TY, z0 = bz * ZC;
  const int TXe = min(TX, g.nx - x0), TYe = min(TY, g.ny - y0), ZCe = min(ZC, g.nz - z0);
  const bool active = lx < TXe && ly < TYe;
  const int x = x0 + lx, y = y0 + ly;
  const size_t P = g.P;
  const float omm = 1.0f - ph.omegaL;

  // halo role of this lane (as home_tile2): 0 row below, 1 row above, 2 left
  // / right column, 3 corner, -1 none; (hx, hy) tile-relative
  int role = -1, hx = 0, hy = 0, hsgn = 1, hsgny = 1;
  if (ly == 0 && lx < TXe) {
    role = 0, hx = lx, hy = -1;
  } else if (ly == 1 && lx < TXe) {
    role = 1, hx = lx, hy = TYe;
  } else if (ly == 2) {
    if (lx < 2 * TYe) {
      role = 2;
      const bool left = lx < TYe;
      hx = left ? -1 : TXe;
      hy = left ? lx : lx - TYe;
      hsgn = left ? 1 : -1;
    } else if (lx < 2 * TYe + 4) {
      role = 3;
      const int c = lx - 2 * TYe;
      hx = (c & 1) ? TXe : -1;
      hy = (c & 2) ? TYe : -1;
      hsgn = (c & 1) ? -1 : 1;
      hsgny = (c & 2) ? -1 : 1;
    }
  }
  const int gx = (x0 + hx + g.nx) % g.nx, gy = (y0 + hy + g.ny) % g.ny;
  const int cornerSlot = 4 + ((hsgn < 0) ? 1 : 0) + ((hsgny < 0) ? 2 : 0);
  const int colSlot0 = hsgn > 0 ? 0 : 1;
  // plane-independent parts of this lane's cell and halo-cell indices
  const uint32_t cxy = uint32_t(y) * g.nx + x, hxy = uint32_t(gy) * g.nx + gx;
  const uint32_t cxy4 = cxy * 4u, hxy4 = hxy * 4u, plane4 = g.nxy * 4u;

  Mom Mm[2], Hm[2];
  uint8_t Fl[3] = {0, 0, 0};
  float Ar[3][5], Br[2][9];
  // a near-wall cell's own values for the e_z = -1 directions whose source
  // is solid (set with its plane, used when the plane above arrives), and
  // that plane's solid-neighbour mask
  float Cb[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  uint32_t nbPrev = 0u;
#pragma unroll
  for (int s = 0; s < 2; ++s) {
#pragma unroll
    for (int k = 0; k < 10; ++k) Mm[s].v[k] = Hm[s].v[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 9; ++k) Br[s][k] = 0.0f;
  }
#pragma unroll
  for (int s = 0; s < 3; ++s) {
#pragma unroll
    for (int k = 0; k < 5; ++k) Ar[s][k] = 0.0f;
  }

  auto load_plane = [&](int p, Mom& mc, Mom& mh) {
    const uint32_t pw4 = uint32_t((p + g.nz) % g.nz) * plane4;
    if (active) {
#pragma unroll
      for (int k = 0; k < 10; ++k) mc.v[k] = ld_b(a.m + size_t(k) * P, pw4 + cxy4);
    }
    if (role >= 0) {
#pragma unroll
      for (int k = 0; k < 10; ++k) mh.v[k] = ld_b(a.m + size_t(k) * P, pw4 + hxy4);
    }
  };
  load_plane(z0 - 1, Mm[0], Hm[0]);
  if (active) Fl[0] = a.flags[uint32_t((z0 - 1 + g.nz) % g.nz) * g.nxy + cxy];
  const int last = z0 + ZCe;

  // one plane step; K = position of plane p in the six-plane unroll
  auto step = [&](auto kc, int p) {
    constexpr int K = decltype(kc)::value;
    constexpr int S2 = K % 2, N2 = (K + 1) % 2;  // this plane's / the other slot of two
    // slots of three: this plane's, two planes back, one plane back
    constexpr int S3 = K % 3, O3 = (K + 1) % 3, M3 = (K + 2) % 3;
    const Mom& cur = Mm[S2];
    const uint32_t pOut = uint32_t((p - 1 + g.nz) % g.nz);
    // the output cell of this plane step is (x, y, p-1); its flags were read
    // two planes ago; an inlet's velocity and the solid-neighbour mask of
    // this plane's cell are read here, ahead of the reconstruction
    const uint8_t flO = Fl[M3], flP = Fl[S3];
    float uinFin = 0.0f;
    if (active && (flO & F_INLET)) uinFin = a.inlet[size_t(pOut) * g.ny + y];
    uint32_t nbCur = 0u;
    if (active && (flP & F_NEAR)) nbCur = a.nbm[uint32_t((p + g.nz) % g.nz) * g.nxy + cxy];
    if (p < last) load_plane(p + 1, Mm[N2], Hm[N2]);
    if (p < last && active) Fl[O3] = a.flags[uint32_t((p + 1 + g.nz) % g.nz) * g.nxy + cxy];
    float f0 = 0.0f, f5 = 0.0f, f6 = 0.0f;
    if (active) {
      const HomeSrc h(cur, omm);
      auto put = [&](auto qc, float v) {
        constexpr int q = decltype(qc)::value;
        fs[L::at(kLds16[q], ly + 1, lx + 1)] = v;
      };
      using std::integral_constant;
      float u, v;
      h.template emit_pair<1>(u, v);
      put(integral_constant<int, 1>{}, u);
      put(integral_constant<int, 2>{}, v);
      h.template emit_pair<3>(u, v);
      put(integral_constant<int, 3>{}, u);
      put(integral_constant<int, 4>{}, v);
      h.template emit_pair<7>(u, v);
      put(integral_constant<int, 7>{}, u);
      put(integral_constant<int, 10>{}, v);
      h.template emit_pair<8>(u, v);
      put(integral_constant<int, 8>{}, u);
      put(integral_constant<int, 9>{}, v);
      h.template emit_pair<11>(u, v);
      put(integral_constant<int, 11>{}, u);
      put(integral_constant<int, 14>{}, v);
      h.template emit_pair<12>(u, v);
      put(integral_constant<int, 12>{}, u);
      put(integral_constant<int, 13>{}, v);
      h.template emit_pair<15>(u, v);
      put(integral_constant<int, 15>{}, u);
      put(integral_constant<int, 18>{}, v);
      h.template emit_pair<16>(u, v);
      put(integral_constant<int, 16>{}, u);
      put(integral_constant<int, 17>{}, v);
      f0 = h.template emit<0>();
      h.template emit_pair<5>(f5, f6);
    }
    if (role >= 0) {
      const HomeSrc h(Hm[S2], omm);
      const int row = hy + 1, col = hx + 1;
      if (role <= 1) {
        // the row below (e_y = +1: q = 3, 7, 8, 15, 17) or above (e_y = -1:
        // q = 4, 9, 10, 16, 18): one code path, the sign of e_y per lane; the
        // e_y = -1 planes follow the e_y = +1 ones (kLds16)
        const bool pos = role == 0;
        const float suy = pos ? h.uy : -h.uy, ssxy = pos ? h.sxy : -h.sxy,
                    ssyz = pos ? h.syz : -h.syz;
        const int n = pos ? 0 : 1;
        fs[L::at(2 + n, row, col)] = emit_sy<0, 0>(h, suy, ssxy, ssyz);
        fs[L::at(4 + 2 * n, row, col)] = emit_sy<1, 0>(h, suy, ssxy, ssyz);
        fs[L::at(5 + 2 * n, row, col)] = emit_sy<-1, 0>(h, suy, ssxy, ssyz);
        fs[L::at(12 + n, row, col)] = emit_sy<0, 1>(h, suy, ssxy, ssyz);
        fs[L::at(14 + n, row, col)] = emit_sy<0, -1>(h, suy, ssxy, ssyz);
      } else if (role == 2) {  // e_x = hsgn: (s,0,0) (s,1,0) (s,-1,0) (s,0,1) (s,0,-1)
        const bool pos = hsgn > 0;
        const float sux = pos ? h.ux : -h.ux, ssxy = pos ? h.sxy : -h.sxy,
                    ssxz = pos ? h.sxz : -h.sxz;
        fs[L::at(colSlot0 + 0, row, col)] = emit_sx<0, 0>(h, sux, ssxy, ssxz);
        fs[L::at(colSlot0 + 4, row, col)] = emit_sx<1, 0>(h, sux, ssxy, ssxz);
        fs[L::at(colSlot0 + 6, row, col)] = emit_sx<-1, 0>(h, sux, ssxy, ssxz);
        fs[L::at(colSlot0 + 8, row, col)] = emit_sx<0, 1>(h, sux, ssxy, ssxz);
        fs[L::at(colSlot0 + 10, row, col)] = emit_sx<0, -1>(h, sux, ssxy, ssxz);
      } else {  // corner: (sx, sy, 0)
        const float sux = hsgn > 0 ? h.ux : -h.ux, suy = hsgny > 0 ? h.uy : -h.uy;
        const float ssxy = (hsgn * hsgny) > 0 ? h.sxy : -h.sxy;
        fs[L::at(cornerSlot, row, col)] = emit_corner(h, sux, suy, ssxy);
      }
    }
    __syncthreads();
    // incoming values of this plane: A (e_z = +1) wait for output p+1, B
    // (e_z = 0) for output p, C (e_z = -1) go to output p-1 now
    float Cd[5];
    if (active) {
      float* A = Ar[S3];
      float* B = Br[S2];
      for_q<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        float v;
        if constexpr (q == 0)
          v = f0;
        else if constexpr (q == 5)
          v = f5;
        else if constexpr (q == 6)
          v = f6;
        else
          v = fs[L::at(kLds16[q], ly - CY[q] + 1, lx - CX[q] + 1)];
        if constexpr (CZ[q] > 0) A[kSlot[q]] = v;
        if constexpr (CZ[q] == 0) B[kSlot[q]] = v;
        if constexpr (CZ[q] < 0) Cd[kSlot[q]] = v;
      });
    }
    const int o = p - 1;
    const bool fin = active && o >= z0 && o < last;
    const float* Ao = Ar[O3];  // from plane o-1 (two steps back)
    const float* Bo = Br[N2];  // from plane o
    // output o of a near-wall cell: where the source of an e_z = -1
    // direction is solid, its own value (stashed with plane o)
    if (__ballot(nbPrev != 0u) != 0) {
      for_q<1, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (CZ[q] < 0)
          if ((nbPrev >> OPP[q]) & 1u) Cd[kSlot[q]] = Cb[kSlot[q]];
      });
    }
    auto gather = [&](float (&f)[19]) {
      for_q<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (CZ[q] > 0) f[q] = Ao[kSlot[q]];
        if constexpr (CZ[q] == 0) f[q] = Bo[kSlot[q]];
        if constexpr (CZ[q] < 0) f[q] = Cd[kSlot[q]];
      });
    };
    auto moments = [&](const float (&f)[19], float (&out)[10]) {
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
    };
    // every lane sums its gathered values (one copy of the sums in the
    // unrolled body); solid, inlet and outlet cells then replace the result
    float out[10];
    {
      float f[19];
      gather(f);
      moments(f, out);
    }
    const uint8_t fl = fin ? flO : uint8_t(0);
    if (__ballot((fl & (F_SOLID | F_INLET | F_OUTLET)) != 0) != 0) {
      if (fl & (F_SOLID | F_INLET)) {
#pragma unroll
        for (int k = 0; k < 10; ++k) out[k] = 0.0f;
        out[0] = 1.0f;
        if (fl & F_INLET) out[1] = uinFin;
      }
      // zero-gradient outlet: the new moments of x-1, the previous lane (an
      // outlet never starts a tile row: the launchers check it)
      const bool outlet = (fl & F_OUTLET) != 0;
      if (__ballot(outlet) != 0) {
#pragma unroll
        for (int k = 0; k < 10; ++k) {
          const float left = __shfl_up(out[k], 1);
          if (outlet) out[k] = left;
        }
      }
    }
    if (fin) {
      const uint32_t co4 = pOut * plane4 + cxy4;
#pragma unroll
      for (int k = 0; k < 10; ++k) st_b(a.mn + size_t(k) * P, co4, out[k]);
    }
    // the next outputs of a near-wall cell of this plane: where c - e_q is
    // solid the cell takes its own post-collision value in direction OPP[q]
    // (or the reference's pairing), which this plane's emission holds at its
    // own position; e_z = 0 and +1 directions go straight into the ring, the
    // e_z = -1 ones wait for the plane above
    if (__ballot(nbCur != 0u) != 0) {
      float* A = Ar[M3];  // e_z = +1 values of output p (read one plane ago)
      float* B = Br[S2];
      for_q<1, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int bq = BB == 0 ? OPP[q] : OPP_REF[q];
        if ((nbCur >> OPP[q]) & 1u) {
          float v;
          if constexpr (bq == 5)
            v = f5;
          else if constexpr (bq == 6)
            v = f6;
          else
            v = fs[L::at(kLds16[bq], ly + 1, lx + 1)];
          if constexpr (CZ[q] > 0) A[kSlot[q]] = v;
          if constexpr (CZ[q] == 0) B[kSlot[q]] = v;
          if constexpr (CZ[q] < 0) Cb[kSlot[q]] = v;
        }
      });
    }
    nbPrev = nbCur;
    __syncthreads();
  };

  using std::integral_constant;
  for (int p = z0 - 1; p <= last; p += 6) {
    step(integral_constant<int, 0>{}, p);
    if (p + 1 > last) break;
    step(integral_constant<int, 1>{}, p + 1);
    if (p + 2 > last) break;
    step(integral_constant<int, 2>{}, p + 2);
    if (p + 3 > last) break;
    step(integral_constant<int, 3>{}, p + 3);
    if (p + 4 > last) break;
    step(integral_constant<int, 4>{}, p + 4);
    if (p + 5 > last) break;
    step(integral_constant<int, 5>{}, p + 5);
  }
}

}  // namespace

}  // namespace fs
