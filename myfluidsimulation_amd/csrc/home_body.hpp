// home_body.hpp -- HOME update bodies shared by the single-domain kernels
// (lbm_home.hip) and the batched kernel (batch.hip): the per-cell gather and
// the tiled plane walk.  See lbm_home.hip for the scheme.
#pragma once

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
      for (int k = 0; k < 10; ++k) a.mn[size_t(k) * P + co] = out[k];
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
