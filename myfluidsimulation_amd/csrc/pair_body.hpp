// pair_body.hpp -- the two-step BGK / Hybrid tile update shared by the
// single-domain kernel (lbm_pair.hip) and the batched kernel (batch.hip).
// See lbm_pair.hip for the scheme.
#pragma once

#include "home_body.hpp"
#include "pop_body.hpp"

namespace fs {

namespace {

__device__ __forceinline__ int wrap_plane(int p, int nz) {
  return p < 0 ? p + nz : (p >= nz ? p - nz : p);
}

// A ring cell sends into the tile only the directions pointing from its side
// into the tile (five per side); val(q) gives the post-collision value of
// direction q.  Each direction is a compile-time branch on the side, so the
// values never go through a runtime-indexed array.
template <int TX, int TY, class Val>
__device__ __forceinline__ void ring_emit(float* lds, int side, int row, int col, const Val& val) {
  using L = TileLds<TX, TY>;
  for_q<1, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    if constexpr (CX[q] != 0 || CY[q] != 0) {
      bool need = false;
      if constexpr (CY[q] > 0) need = need || side == 0;  // bottom row
      if constexpr (CY[q] < 0) need = need || side == 1;  // top row
      if constexpr (CX[q] > 0) need = need || side == 2;  // left column
      if constexpr (CX[q] < 0) need = need || side == 3;  // right column
      if (need) lds[L::at(q, row, col)] = val(qc);
    }
  });
}

// One TX x TY column tile (bx, by) and z-chunk bz of the two-step update;
// every thread of the workgroup (TX * TY threads) must call it: it
// synchronises the workgroup.  lds: 19 * TileLds<TX,TY>::PL floats.  An
// outlet must not start a TX-wide tile row (the caller checks).
template <int TX, int TY, int ZC, bool HYB, bool MACRO>
__device__ __forceinline__ void pair_tile(const Geo& g, const Phys& ph, const PairLaunch& a,
                                          int bx, int by, int bz, float* lds) {
  constexpr int NT = TX * TY;
  using L = TileLds<TX, TY>;
  constexpr int HMAX = 2 * (TX + 2) + 2 * TY;
  static_assert(NT >= HMAX, "one ring cell per thread at most");
  const int tid = threadIdx.x;
  const int lx = tid % TX, ly = tid / TX;
  const int x0 = bx * TX, y0 = by * TY, z0 = bz * ZC;
  const int TXe = min(TX, g.nx - x0), TYe = min(TY, g.ny - y0), ZCe = min(ZC, g.nz - z0);
  const bool active = lx < TXe && ly < TYe;
  const int x = x0 + lx, y = y0 + ly;
  const size_t P = g.P;
  const int nxy = int(g.nxy);

  // ring cell: bottom row, top row, left column, right column, on the lowest
  // thread ids, so the fewest waves (three) run the ring code
  const int H = 2 * (TXe + 2) + 2 * TYe;
  const int hid = tid;
  int side = -1, hx = 0, hy = 0;
  if (hid < H) {
    if (hid < TXe + 2) {
      side = 0, hx = hid - 1, hy = -1;
    } else if (hid < 2 * (TXe + 2)) {
      side = 1, hx = hid - (TXe + 2) - 1, hy = TYe;
    } else if (hid < 2 * (TXe + 2) + TYe) {
      side = 2, hx = -1, hy = hid - 2 * (TXe + 2);
    } else {
      side = 3, hx = TXe, hy = hid - 2 * (TXe + 2) - TYe;
    }
  }
  const int gx = (x0 + hx + g.nx) % g.nx, gy = (y0 + hy + g.ny) % g.ny;
  const int ownRow = y * g.nx + x, ringRow = gy * g.nx + gx;

  // The tile cells' loads run one plane ahead of their collisions, and the
  // flag bytes two planes ahead, because an outlet cell's populations come
  // from x-1.  Ring cells load when they collide: prefetching them too costs
  // registers that the tile loads need more.
  float nf[19], nh[19];
  float nUin = 0.0f, hUin = 0.0f;  // inlet velocity of an inlet tile / ring cell
  uint32_t nNbm = 0;
  uint8_t fl1 = 0, fl2 = 0, hfl1 = 0, hfl2 = 0;
#pragma unroll
  for (int q = 0; q < 19; ++q) nf[q] = nh[q] = 0.0f;
  auto load_flags = [&](int p, uint8_t& f, uint8_t& hf) {
    const int pw = wrap_plane(p, g.nz);
    if (active) f = a.flags[pw * nxy + ownRow];
    if (side >= 0) hf = a.flags[pw * nxy + ringRow];
  };
  auto load_pops = [&](int p, uint8_t f, uint8_t hf, bool tile, bool ring) {
    const int pw = wrap_plane(p, g.nz);
    if (tile && active) {
      const uint32_t c = uint32_t(pw * nxy + ownRow);
      const uint32_t s = (f & F_OUTLET) ? c - 1 : c;
      if (!(f & F_SOLID)) {
#pragma unroll
        for (int q = 0; q < 19; ++q) nf[q] = a.Fin[size_t(q) * P + s];
      }
      // an inlet's velocity is loaded with the plane, not where it is used
      if (f & F_INLET) nUin = a.inlet[size_t(pw) * g.ny + y];
      nNbm = (f & F_NEAR) ? a.nbm[c] : 0u;
    }
    if (ring && side >= 0) {
      const uint32_t c = uint32_t(pw * nxy + ringRow);
      const uint32_t s = (hf & F_OUTLET) ? c - 1 : c;
      if (!(hf & F_SOLID)) {
#pragma unroll
        for (int q = 0; q < 19; ++q) nh[q] = a.Fin[size_t(q) * P + s];
      }
      if (hf & F_INLET) hUin = a.inlet[size_t(pw) * g.ny + gy];
    }
  };
  uint8_t curFl = 0, hcurFl = 0;
  load_flags(z0 - 1, curFl, hcurFl);
  load_flags(z0, fl1, hfl1);
  load_pops(z0 - 1, curFl, hcurFl, true, false);

  // register ring (values for planes o+1, o) and what plane o needs at its end
  float aPrev2[5], aPrev[5], bPrev[9], cOvPrev[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) aPrev2[k] = aPrev[k] = cOvPrev[k] = 0.0f;
#pragma unroll
  for (int k = 0; k < 9; ++k) bPrev[k] = 0.0f;
  uint8_t flPrev = 0;
  uint32_t nbmPrev = 0;
  float uinPrev = 0.0f;

  for (int p = z0 - 1; p <= z0 + ZCe; ++p) {
    load_pops(p, curFl, hcurFl, false, true);
    const uint32_t curNbm = nNbm;
    const float curUin = nUin;
    // ---- step t on plane p: tile cells ----
    // Inlet lanes run the fluid code on zero inputs with set_inlet moments;
    // the choice is per wave, so waves without an inlet lane pay nothing.
    const bool tileInlet = active && (curFl & F_INLET) != 0;
    const bool tileInletWave = __ballot(tileInlet) != 0;
    if (active && !(curFl & F_SOLID)) {
      if (!tileInletWave) {
        const PopCollision<HYB> pc(pop_sums(nf), ph, (curFl & F_BAND) != 0);
        float post[19];
        pc.post_all(nf, post);
#pragma unroll
        for (int q = 0; q < 19; ++q) lds[L::at(q, ly + 1, lx + 1)] = post[q];
      } else {
        float in[19];
#pragma unroll
        for (int q = 0; q < 19; ++q) in[q] = tileInlet ? 0.0f : nf[q];
        PopCollision<HYB> pc(pop_sums(in), ph, (curFl & F_BAND) != 0);
        if (tileInlet) pc.set_inlet(curUin);
        float post[19];
        pc.post_all(in, post);
#pragma unroll
        for (int q = 0; q < 19; ++q) lds[L::at(q, ly + 1, lx + 1)] = post[q];
      }
    }
    // ---- step t on plane p: ring cells ----
    const bool ringInlet = side >= 0 && (hcurFl & F_INLET) != 0;
    const bool ringInletWave = __ballot(ringInlet) != 0;
    if (side >= 0 && !(hcurFl & F_SOLID)) {
      const int row = hy + 1, col = hx + 1;
      if (!ringInletWave) {
        const PopCollision<HYB> pc(pop_sums(nh), ph, (hcurFl & F_BAND) != 0);
        ring_emit<TX, TY>(lds, side, row, col, [&](auto qc) {
          constexpr int q = decltype(qc)::value;
          return pc.template post<q>(nh[q]);
        });
      } else {
        float in[19];
#pragma unroll
        for (int q = 0; q < 19; ++q) in[q] = ringInlet ? 0.0f : nh[q];
        PopCollision<HYB> pc(pop_sums(in), ph, (hcurFl & F_BAND) != 0);
        if (ringInlet) pc.set_inlet(hUin);
        ring_emit<TX, TY>(lds, side, row, col, [&](auto qc) {
          constexpr int q = decltype(qc)::value;
          return pc.template post<q>(in[q]);
        });
      }
    }
    // ---- prefetch plane p+1 (flags known) and the flags of plane p+2 ----
    uint8_t nextFl = fl1, nextHfl = hfl1;
    if (p < z0 + ZCe) load_pops(p + 1, fl1, hfl1, true, false);
    fl2 = 0, hfl2 = 0;
    if (p + 2 <= z0 + ZCe) load_flags(p + 2, fl2, hfl2);
    __syncthreads();

    // ---- gather: f^{t+1} for planes p+1 (A), p (B), p-1 (C) ----
    float A[5], B[9], Cd[5];
    if (active) {
      for_q<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const float v = lds[L::at(q, ly - CY[q] + 1, lx - CX[q] + 1)];
        if constexpr (CZ[q] > 0) A[kSlot[q]] = v;
        if constexpr (CZ[q] == 0) B[kSlot[q]] = v;
        if constexpr (CZ[q] < 0) Cd[kSlot[q]] = v;
      });
    } else {
#pragma unroll
      for (int k = 0; k < 5; ++k) A[k] = Cd[k] = 0.0f;
#pragma unroll
      for (int k = 0; k < 9; ++k) B[k] = 0.0f;
    }

    // ---- step t+1 on plane o = p-1 ----
    const int o = p - 1;
    const bool fin = active && o >= z0 && o < z0 + ZCe;
    const int ow = wrap_plane(o, g.nz);
    const uint8_t fl = fin ? flPrev : uint8_t(0);
    const bool live = fin && !(fl & F_SOLID);
    const uint32_t co = uint32_t(ow * nxy + ownRow);
    float f[19];
    float r = 1.0f, u = 0.0f, v = 0.0f, w = 0.0f;
    const bool inl = live && (fl & F_INLET) != 0;
    const bool anyInlet = __ballot(inl) != 0;
    if (live) {
      for_q<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (CZ[q] > 0) f[q] = aPrev2[kSlot[q]];
        if constexpr (CZ[q] == 0) f[q] = bPrev[kSlot[q]];
        if constexpr (CZ[q] < 0) f[q] = ((nbmPrev >> OPP[q]) & 1u) ? cOvPrev[kSlot[q]] : Cd[kSlot[q]];
      });
      if (anyInlet) {
#pragma unroll
        for (int q = 0; q < 19; ++q) f[q] = inl ? 0.0f : f[q];
      }
    } else {
#pragma unroll
      for (int q = 0; q < 19; ++q) f[q] = 0.0f;
    }
    // an outlet collides its upstream neighbour's incoming populations (the
    // previous lane's; its derived band bit is that cell's too)
    const bool outlet = fin && (fl & F_OUTLET) != 0;
    if (__ballot(outlet) != 0) {
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        const float left = __shfl_up(f[q], 1);
        if (outlet) f[q] = left;
      }
    }
    if (live && !anyInlet) {
      const PopCollision<HYB> pc(pop_sums(f), ph, (fl & F_BAND) != 0);
      float in[19];
#pragma unroll
      for (int q = 0; q < 19; ++q) in[q] = f[q];
      pc.post_all(in, f);
      r = pc.r;
      u = pc.ux;
      v = pc.uy;
      w = pc.uz;
    } else if (live) {
      PopCollision<HYB> pc(pop_sums(f), ph, (fl & F_BAND) != 0);
      if (inl) pc.set_inlet(uinPrev);
      float in[19];
#pragma unroll
      for (int q = 0; q < 19; ++q) in[q] = f[q];
      pc.post_all(in, f);
      r = pc.r;
      u = pc.ux;
      v = pc.uy;
      w = pc.uz;
    }
    const uint32_t sm = live ? nbmPrev : 0u;
    const bool anyWall = __ballot(sm != 0) != 0;
    if (live) {
      const Nbr on = nbr_offsets(x, y, ow, g);
      for_q<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if (anyWall && q != 0 && ((sm >> q) & 1u))
          a.Fout[size_t(OPP[q]) * P + co] = f[q];
        else
          a.Fout[size_t(q) * P + co + nb_off<CX[q], CY[q], CZ[q]>(on)] = f[q];
      });
    }
    if (MACRO && fin) {
      a.rho[co] = r;
      a.ux[co] = u;
      a.uy[co] = v;
      a.uz[co] = w;
    }

    // ---- bounce-back for the cell on plane p ----
    // where c - e_q is solid, f^{t+1}_q is the cell's own f*_{opp q} (still in
    // LDS): its A values (gathered one plane ago) and B values are patched
    // now, its C values are kept until plane p is finished
    const uint32_t bbm = (active && !(curFl & F_SOLID)) ? curNbm : 0u;
    if (__ballot(bbm != 0) != 0) {
      for_q<1, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if ((bbm >> OPP[q]) & 1u) {
          const float val = lds[L::at(OPP[q], ly + 1, lx + 1)];
          if constexpr (CZ[q] > 0) aPrev[kSlot[q]] = val;
          if constexpr (CZ[q] == 0) B[kSlot[q]] = val;
          if constexpr (CZ[q] < 0) cOvPrev[kSlot[q]] = val;
        }
      });
    }

    // ---- advance the ring ----
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      aPrev2[k] = aPrev[k];
      aPrev[k] = A[k];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) bPrev[k] = B[k];
    flPrev = curFl;
    nbmPrev = curNbm;
    uinPrev = curUin;
    curFl = nextFl;
    hcurFl = nextHfl;
    fl1 = fl2;
    hfl1 = hfl2;
    __syncthreads();
  }
}

}  // namespace

}  // namespace fs
