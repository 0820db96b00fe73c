// home_pair.hpp -- two HOME steps per pass over the moments (k_home_pair).
//
// The one-step HOME kernels read the 10 moments of every cell and write 10
// new ones per step (80 B per lattice update); their speed is set by that
// traffic.  This tile reads m^t once and writes m^{t+2} once (40 B per
// update): the intermediate moments m^{t+1} live only in registers and LDS.
//
// A workgroup owns a 64 x TY column tile T and a chunk of ZC planes.  m^{t+2}
// at a tile cell gathers the post-collision values that m^{t+1} of its 18
// neighbours emit, so m^{t+1} is needed on E = T plus a one-cell ring
// (66 x (TY+2) cells per plane), and m^{t+1} on E gathers what m^t emits on
// F = E plus another ring (68 x (TY+4)), of which the outer ring only emits
// the directions that point into E.  One lane per cell of E or of F's outer
// ring (64 x (TY+5) lanes, one role per wave; pair_wave_role deals the roles
// to the hardware waves so that the four SIMDs carry similar work):
//   roles 0 .. TY+1   rows y0-1 .. y0+TY of E (x0 .. x0+63); rows 1 .. TY are T
//   role  TY+2        E's two side columns (x0-1, x0+64), F's two side
//                     columns (x0-2, x0+65), the four ends of F's rows and
//                     F's four corners (the side wave)
//   roles TY+3, TY+4  F's rows y0-2 and y0+TY+1; where 64 x (TY+5) would pass
//                     1024 threads (TY = 12), role TY+3 holds both rows, one
//                     cell of each per lane
// Each lane walks the source planes p = z0-2 .. z0+ZC+1; per plane:
//   A  every lane emits m^t(p) of its cell into LDS plane set A (E cells and
//      the side wave all 19 directions, F's rows the five that enter E);
//      barrier;
//   B  every E lane gathers m^{t+1}(p-1) (values from planes p-2, p-1, p,
//      the first two kept in registers) in the reference's q order;
//   C  E lanes emit m^{t+1}(p-1) into LDS plane set C (tile cells all
//      directions, E's ring the ones that enter T);  barrier;
//   D  tile lanes gather m^{t+2}(p-2) the same way and store it.
// Per value and per cell the arithmetic is home_tile2's (HomeSrc::emit /
// emit_pair, the q-ordered sums, home_finish), so two steps of this kernel
// give the bits of two one-step launches.
//
// Boundary rules per step, as the one-step kernels apply them: a solid cell
// takes (1, 0...), an inlet (1, u_in, 0...), a zero-gradient outlet the new
// moments of x-1 (the previous lane), and where c - e_q is solid the cell
// takes its own post-collision value in direction OPP[q] (or the reference's
// pairing, BB = 1).  Those own values are the ones the cell emitted into
// LDS (or kept in registers) in stage A / C, so a near-wall cell reads them
// back instead of rebuilding them.
//
// The launcher runs this kernel only where every outlet's upstream cell
// x-1 is in the same E row as the outlet: no outlet at x % 64 == 0 (it
// would start a tile row), and an outlet on E's left column (x % 64 == 63,
// or x = nx-1 for the first tile) only where the cells that read it (x+1,
// y +-1, z +-1) are all inlet or solid, so its value is never used.
#pragma once

#include "home_body.hpp"

namespace fs {

namespace {

template <int TY>
struct PairLds {
  static constexpr int WA = 68, HA = TY + 4, PA = WA * HA;  // stage A: F, offset (2, 2)
  static constexpr int WC = 66, HC = TY + 2, PC = WC * HC;  // stage C: E, offset (1, 1)
  static constexpr int FLOATS = 16 * (PA + PC);
  // F's two rows share one wave (two cells per lane) where a wave each would
  // pass the 1024 threads of a workgroup
  static constexpr bool FOLD = 64 * (TY + 5) > 1024;
  static constexpr int WAVES = FOLD ? TY + 4 : TY + 5;
  static constexpr int THREADS = 64 * WAVES;
  static_assert(THREADS <= 1024, "one workgroup per tile");
  static_assert(4 * (TY + 2) + 8 <= 64, "the side wave holds E's and F's columns and F's ends");
};

// Directions that enter the inner region from one side of a ring:
// 1 e_y = +1, 2 e_y = -1, 3 e_x = +1, 4 e_x = -1, 5..8 one diagonal each
// (7, 8, 9, 10).  Values go to LDS plane kLds16[q] at position pos.
template <int PS>
__device__ __forceinline__ void emit_ring_set(float* L, int pos, const HomeSrc& h, int set) {
  auto put = [&](auto qc) {
    constexpr int q = decltype(qc)::value;
    L[kLds16[q] * PS + pos] = h.template emit<q>();
  };
  using std::integral_constant;
  switch (set) {
    case 1:
      put(integral_constant<int, 3>{}), put(integral_constant<int, 7>{}),
          put(integral_constant<int, 8>{}), put(integral_constant<int, 15>{}),
          put(integral_constant<int, 17>{});
      break;
    case 2:
      put(integral_constant<int, 4>{}), put(integral_constant<int, 9>{}),
          put(integral_constant<int, 10>{}), put(integral_constant<int, 16>{}),
          put(integral_constant<int, 18>{});
      break;
    case 3:
      put(integral_constant<int, 1>{}), put(integral_constant<int, 7>{}),
          put(integral_constant<int, 9>{}), put(integral_constant<int, 11>{}),
          put(integral_constant<int, 13>{});
      break;
    case 4:
      put(integral_constant<int, 2>{}), put(integral_constant<int, 8>{}),
          put(integral_constant<int, 10>{}), put(integral_constant<int, 12>{}),
          put(integral_constant<int, 14>{});
      break;
    case 5: put(integral_constant<int, 7>{}); break;
    case 6: put(integral_constant<int, 8>{}); break;
    case 7: put(integral_constant<int, 9>{}); break;
    case 8: put(integral_constant<int, 10>{}); break;
    default: break;
  }
}

// All 19 directions of a cell: 16 into LDS at pos, the column directions
// (0, 5, 6) returned in registers.
template <int PS>
__device__ __forceinline__ void emit_all(float* L, int pos, const HomeSrc& h, float& r0, float& r5,
                                         float& r6) {
  auto put2 = [&](auto qc) {
    constexpr int q = decltype(qc)::value;
    float u, v;
    h.template emit_pair<q>(u, v);
    L[kLds16[q] * PS + pos] = u;
    L[kLds16[OPP[q]] * PS + pos] = v;
  };
  using std::integral_constant;
  put2(integral_constant<int, 1>{});
  put2(integral_constant<int, 3>{});
  put2(integral_constant<int, 7>{});
  put2(integral_constant<int, 8>{});
  put2(integral_constant<int, 11>{});
  put2(integral_constant<int, 12>{});
  put2(integral_constant<int, 15>{});
  put2(integral_constant<int, 16>{});
  r0 = h.template emit<0>();
  h.template emit_pair<5>(r5, r6);
}

// The role (the wave number of the layout above) of hardware wave hw.  A
// workgroup's waves are dealt to the CU's four SIMDs in turn (wave hw runs on
// SIMD hw % 4), and every stage ends at a barrier, so a stage lasts as long
// as its busiest SIMD.  With TY = 11 the sixteen roles are eleven tile rows
// (stages A to D), E's two ring rows (A, B and five directions of C), the
// side wave (A, B, ten directions of C) and F's two rows (five directions of
// A).  In role order the side wave would join three tile rows on SIMD 1;
// this order gives SIMD 0 the side wave, one F row and two tile rows, and the
// others three tile rows and one E ring row or F row each (measured on the
// 512^3 vortex: 68.7 -> 71.8 GLUPS).
template <int TY>
__device__ __forceinline__ int pair_wave_role(int hw) {
  if constexpr (TY == 11) {
    // hw:      0   1  2  3   4  5  6  7  8  9  10  11 12 13  14  15
    // role:   13   1  2  3  14  5  6  7  8  9  10  11  4  0  12  15
    constexpr uint64_t kRole = 0xFC04BA98765E321DULL;
    return int((kRole >> (4 * hw)) & 15u);
  } else if constexpr (TY == 12) {
    // twelve tile rows, E's two ring rows, the side wave and the folded F
    // rows: three tile rows and one of the other four roles per SIMD
    // hw:      0   1   2   3  4  5 ... 15
    // role:    0  13  14  15  1  2 ... 12
    constexpr uint64_t kRole = 0xCBA987654321FED0ULL;
    return int((kRole >> (4 * hw)) & 15u);
  } else {
    return hw;
  }
}

// Gather rings of one stage (B or D) for one lane: values with e_z = +1 by
// the parity of the plane that sent them (A[K]: read by the output two planes
// later, which has the same parity, so no value moves between slots), e_z = 0
// for the next output (Br), and the own values that replace e_z = -1 values
// of the next output of a near-wall cell (Cb).  Each slot is read for the
// output due before it is refilled.
struct PairRing {
  float A[2][5], Br[9], Cb[5];
};

// A folded F-row lane gathers nothing, so its stage-D ring (24 registers it
// never uses otherwise) holds the moments of its second cell: plane-parity
// buffer B, moment k.  This keeps the two-row lanes within the registers of
// the others.
template <int B>
__device__ __forceinline__ float& fold_slot(PairRing& R, int k) {
  if constexpr (B == 0)
    return k < 9 ? R.Br[k] : R.A[0][0];
  else
    return k < 5 ? R.A[1][k] : R.Cb[k - 5];
}

template <int TY, int ZC, int BB>
__device__ __forceinline__ void home_pair_tile(const Geo& g, const Phys& ph, const HomeLaunch& a,
                                               int bx, int by, int bz, float* lds) {
  constexpr int TX = 64;
  using PL = PairLds<TY>;
  float* LA = lds;
  float* LC = lds + 16 * PL::PA;
  const int tid = threadIdx.x, l = tid & 63;
  const int w = pair_wave_role<TY>(tid >> 6);
  const int x0 = bx * TX, y0 = by * TY, z0 = bz * ZC;
  const int TXe = min(TX, g.nx - x0), TYe = min(TY, g.ny - y0), ZCe = min(ZC, g.nz - z0);
  const float omm = 1.0f - ph.omegaL;
  const size_t P = g.P;

  // ---- role of this lane ----
  // isE: a cell of E (gets m^{t+1}); isT: a tile cell (gets m^{t+2});
  // setA: stage-A emission (0 all directions, 1..8 a ring set, -1 no cell);
  // setC: stage-C emission of an E cell (0 all, 1..4 a ring set)
  bool isE = false, isT = false;
  int xl = 0, yl = 0, setA = -1, setC = -1;
  if (w <= TY + 1) {
    if (w <= TYe + 1 && l < TXe) {
      isE = true;
      xl = l;
      yl = w - 1;
      isT = yl >= 0 && yl < TYe;
      setA = 0;
      setC = isT ? 0 : (yl < 0 ? 1 : 2);
    }
  } else if (w == TY + 2) {
    const int H = TYe + 2;
    if (l < H) {
      isE = true, xl = -1, yl = l - 1, setA = 0, setC = 3;
    } else if (l < 2 * H) {
      isE = true, xl = TXe, yl = l - H - 1, setA = 0, setC = 4;
    } else if (l < 3 * H) {
      xl = -2, yl = l - 2 * H - 1, setA = 3;
    } else if (l < 4 * H) {
      xl = TXe + 1, yl = l - 3 * H - 1, setA = 4;
    } else if (l < 4 * H + 8) {
      const int c = l - 4 * H;
      if (c < 4) {  // the ends of F's rows
        xl = (c & 1) ? TXe : -1;
        yl = (c & 2) ? TYe + 1 : -2;
        setA = (c & 2) ? 2 : 1;
      } else {  // F's corners: (+1,+1) q7, (-1,+1) q8, (+1,-1) q9, (-1,-1) q10
        const int k = c - 4;
        xl = (k & 1) ? TXe + 1 : -2;
        yl = (k & 2) ? TYe + 1 : -2;
        setA = 5 + k;
      }
    }
  } else if (w == TY + 3) {
    if (l < TXe) xl = l, yl = -2, setA = 1;
  } else if (w == TY + 4) {
    if (l < TXe) xl = l, yl = TYe + 1, setA = 2;
  }
  const bool hasCell = setA >= 0;
  const int gx = ((x0 + xl) % g.nx + g.nx) % g.nx;
  const int gy = ((y0 + yl) % g.ny + g.ny) % g.ny;
  const uint32_t cxy = uint32_t(gy) * g.nx + gx;
  const uint32_t cxy4 = cxy * 4u, plane4 = g.nxy * 4u;
  const int posA = (yl + 2) * PL::WA + (xl + 2);
  // folded F rows: the lane's second cell (x0+l, y0+TYe+1), direction set 2
  const bool fold = PL::FOLD && w == TY + 3 && hasCell;
  const uint32_t cxy2_4 =
      (uint32_t(((y0 + TYe + 1) % g.ny + g.ny) % g.ny) * g.nx + uint32_t(gx)) * 4u;
  const int posA2 = (TYe + 3) * PL::WA + (xl + 2);
  const int posC = (yl + 1) * PL::WC + (xl + 1);
  const int last = z0 + ZCe + 1;  // the last source plane

  Mom Mm[2];
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int k = 0; k < 10; ++k) Mm[s].v[k] = 0.0f;
  PairRing RB, RD;
#pragma unroll
  for (int k = 0; k < 5; ++k)
    RB.A[0][k] = RB.A[1][k] = RD.A[0][k] = RD.A[1][k] = RB.Cb[k] = RD.Cb[k] = 0.0f;
#pragma unroll
  for (int k = 0; k < 9; ++k) RB.Br[k] = RD.Br[k] = 0.0f;
  // flags of planes p (byte 0), p-1 (byte 1), p-2 (byte 2) of E cells;
  // solid-neighbour masks of the same planes (near-wall cells only)
  uint32_t flr = 0u, nb0 = 0u, nb1 = 0u, nb2 = 0u;

  auto plane_of = [&](int p) { return uint32_t((p % g.nz + g.nz) % g.nz); };
  auto load_mom_plane = [&](int p, Mom& m) {
    if (hasCell) {
      const uint32_t off = plane_of(p) * plane4 + cxy4;
#pragma unroll
      for (int k = 0; k < 10; ++k) m.v[k] = ld_b(a.m + size_t(k) * P, off);
    }
  };
  auto load_fold = [&](auto bc, int p) {
    constexpr int B = decltype(bc)::value;
    if constexpr (PL::FOLD) {
      if (fold) {
        const uint32_t off = plane_of(p) * plane4 + cxy2_4;
#pragma unroll
        for (int k = 0; k < 10; ++k) fold_slot<B>(RD, k) = ld_b(a.m + size_t(k) * P, off);
      }
    }
  };
  load_fold(std::integral_constant<int, 0>{}, z0 - 2);
  load_mom_plane(z0 - 2, Mm[0]);
  if (isE) flr = a.flags[plane_of(z0 - 2) * g.nxy + cxy];

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

  // One gather stage.  Reads this plane's incoming values from LDS set Lx
  // (layout width W, plane stride PS, own position pos): the e_z = -1 ones
  // finish output plane pOut now, the e_z = 0 / +1 ones go to the ring.
  // nbOut / nbCur: solid-neighbour masks of planes pOut and pOut+1 (the
  // plane just emitted); r0, r5, r6: this lane's own column values of the
  // plane just emitted.  Returns the q-ordered sums of pOut's gathered
  // values (home_finish) in out, computed before the ring slots they come
  // from are refilled.  K: the parity of this plane.  need (the same for the
  // whole workgroup): pOut is an output of this chunk; the first planes of a
  // chunk only fill the ring, and out is left unset.
  auto gather_stage = [&](auto kc, auto wc, auto psc, PairRing& R, const float* Lx, int pos,
                          bool need,
                          bool anyNearOut, bool anyNearCur, uint32_t nbOut, uint32_t nbCur,
                          float r0, float r5, float r6, float (&out)[10]) {
    constexpr int K = decltype(kc)::value;
    constexpr int W = decltype(wc)::value;
    constexpr int PS = decltype(psc)::value;
    auto lds_in = [&](auto qc) {  // value arriving along q from c - e_q (in-plane part)
      constexpr int q = decltype(qc)::value;
      if constexpr (q == 0)
        return r0;
      else if constexpr (q == 5)
        return r5;
      else if constexpr (q == 6)
        return r6;
      else
        return Lx[kLds16[q] * PS + pos - CY[q] * W - CX[q]];
    };
    auto own = [&](auto rc) {  // own post-collision value in direction r
      constexpr int r = decltype(rc)::value;
      if constexpr (r == 0)
        return r0;
      else if constexpr (r == 5)
        return r5;
      else if constexpr (r == 6)
        return r6;
      else
        return Lx[kLds16[r] * PS + pos];
    };
    // output pOut: e_z = +1 from the ring (two planes back), e_z = 0 from
    // the ring (one plane back), e_z = -1 from this plane
    if (need) {
      float f[19];
      for_q<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (CZ[q] > 0) f[q] = R.A[K][kSlot[q]];
        if constexpr (CZ[q] == 0) f[q] = R.Br[kSlot[q]];
        if constexpr (CZ[q] < 0) f[q] = lds_in(qc);
      });
      if (anyNearOut) {
        for_q<1, 19>([&](auto qc) {
          constexpr int q = decltype(qc)::value;
          if constexpr (CZ[q] < 0)
            if ((nbOut >> OPP[q]) & 1u) f[q] = R.Cb[kSlot[q]];
        });
      }
      moments(f, out);
    }
    // ring: this plane's e_z = 0 values (next output), e_z = +1 (the one
    // after, same parity)
    for_q<0, 19>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      if constexpr (CZ[q] > 0) R.A[K][kSlot[q]] = lds_in(qc);
      if constexpr (CZ[q] == 0) R.Br[kSlot[q]] = lds_in(qc);
    });
    // the next output (this plane) of a near-wall cell: where c - e_q is
    // solid it takes its own value in direction OPP[q] (or the reference's
    // pairing), which this plane's emission holds
    if (anyNearCur) {
      for_q<1, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int bq = BB == 0 ? OPP[q] : OPP_REF[q];
        if ((nbCur >> OPP[q]) & 1u) {
          const float v = own(std::integral_constant<int, bq>{});
          if constexpr (CZ[q] > 0) R.A[1 - K][kSlot[q]] = v;
          if constexpr (CZ[q] == 0) R.Br[kSlot[q]] = v;
          if constexpr (CZ[q] < 0) R.Cb[kSlot[q]] = v;
        }
      });
    }
  };

  // the step rules of an output cell on top of its gathered sums: solid
  // (1, 0...), inlet (1, u_in, 0...), outlet (x-1's new moments, the
  // previous lane)
  auto finish_cell = [&](bool fin, uint8_t fl, float uin, float (&out)[10]) {
    const uint8_t f2 = fin ? fl : uint8_t(0);
    if (__ballot((f2 & (F_SOLID | F_INLET | F_OUTLET)) != 0) == 0) return;
    if (f2 & (F_SOLID | F_INLET)) {
#pragma unroll
      for (int k = 0; k < 10; ++k) out[k] = 0.0f;
      out[0] = 1.0f;
      if (f2 & F_INLET) out[1] = uin;
    }
    const bool outlet = (f2 & F_OUTLET) != 0;
    if (__ballot(outlet) != 0) {
#pragma unroll
      for (int k = 0; k < 10; ++k) {
        const float left = __shfl_up(out[k], 1);
        if (outlet) out[k] = left;
      }
    }
  };

  auto iter = [&](auto kc, int p) {
    constexpr int K = decltype(kc)::value;
    const int o1 = p - 1, o2 = p - 2;
    const uint8_t flP = uint8_t(flr), fl1 = uint8_t(flr >> 8), fl2 = uint8_t(flr >> 16);
    // loads for this iteration: solid-neighbour mask of plane p, inlet
    // velocities of the two output planes; then the next plane's moments
    // and flags
    nb0 = (isE && (flP & F_NEAR)) ? a.nbm[plane_of(p) * g.nxy + cxy] : 0u;
    float uin1 = 0.0f, uin2 = 0.0f;
    if (isE && (fl1 & F_INLET)) uin1 = a.inlet[size_t(plane_of(o1)) * g.ny + gy];
    if (isT && (fl2 & F_INLET)) uin2 = a.inlet[size_t(plane_of(o2)) * g.ny + gy];
    uint32_t flNext = 0u;
    if (p < last) {
      load_fold(std::integral_constant<int, 1 - K>{}, p + 1);
      load_mom_plane(p + 1, Mm[1 - K]);
      if (isE) flNext = a.flags[plane_of(p + 1) * g.nxy + cxy];
    }

    // ---- A: emit m^t(p) ----
    // The side wave's lanes all emit every direction: its cells need nine
    // different direction sets, which as divergent branches would cost it
    // 43 directions instead of 19 (the extra values land in LDS slots that
    // no gather reads).  Its SIMD sets the pace of this stage: 512^3 vortex
    // 62.5 -> 68.7 GLUPS.
    float a0 = 0.0f, a5 = 0.0f, a6 = 0.0f;
    if (hasCell) {
      const HomeSrc h(Mm[K], omm);
      if (setA == 0 || w == TY + 2)
        emit_all<PL::PA>(LA, posA, h, a0, a5, a6);
      else
        emit_ring_set<PL::PA>(LA, posA, h, setA);
    }
    if constexpr (PL::FOLD) {
      if (fold) {
        Mom m2;
#pragma unroll
        for (int k = 0; k < 10; ++k) m2.v[k] = fold_slot<K>(RD, k);
        const HomeSrc h2(m2, omm);
        emit_ring_set<PL::PA>(LA, posA2, h2, 2);
      }
    }
    __syncthreads();

    // ---- B: m^{t+1}(o1) on E ----
    const bool finB = isE && o1 >= z0 - 1 && o1 <= z0 + ZCe;
    const bool anyNear1 = __ballot(isE && (fl1 & F_NEAR)) != 0;
    const bool anyNearP = __ballot(isE && (flP & F_NEAR)) != 0;
    // m1 and out below are set only where a lane has an output (isE / isT
    // and need); nothing reads them elsewhere, so they get no default
    float m1[10];
    if (isE) {
      gather_stage(kc, std::integral_constant<int, PL::WA>{},
                   std::integral_constant<int, PL::PA>{}, RB, LA, posA, o1 >= z0 - 1, anyNear1,
                   anyNearP, nb1, nb0, a0, a5, a6, m1);
    }
    finish_cell(finB, fl1, uin1, m1);

    // ---- C: emit m^{t+1}(o1) ----
    float c0 = 0.0f, c5 = 0.0f, c6 = 0.0f;
    if (finB) {
      Mom mm;
#pragma unroll
      for (int k = 0; k < 10; ++k) mm.v[k] = m1[k];
      const HomeSrc h(mm, omm);
      if (setC == 0)
        emit_all<PL::PC>(LC, posC, h, c0, c5, c6);
      else
        emit_ring_set<PL::PC>(LC, posC, h, setC);
    }
    __syncthreads();

    // ---- D: m^{t+2}(o2) on T, stored ----
    {
      const bool finD = isT && o2 >= z0 && o2 < z0 + ZCe;
      const bool anyNear2 = __ballot(isT && (fl2 & F_NEAR)) != 0;
      const bool anyNearC = __ballot(isT && (fl1 & F_NEAR)) != 0;
      float out[10];
      if (isT) {
        gather_stage(kc, std::integral_constant<int, PL::WC>{},
                     std::integral_constant<int, PL::PC>{}, RD, LC, posC, o2 >= z0, anyNear2,
                     anyNearC, nb2, nb1, c0, c5, c6, out);
      }
      finish_cell(finD, fl2, uin2, out);
      if (finD) {
        const uint32_t off = plane_of(o2) * plane4 + cxy4;
#pragma unroll
        for (int k = 0; k < 10; ++k) st_b(a.mn + size_t(k) * P, off, out[k]);
      }
    }
    flr = (flr << 8) | flNext;
    nb2 = nb1;
    nb1 = nb0;
  };

  using std::integral_constant;
  for (int p = z0 - 2; p <= last; p += 2) {
    iter(integral_constant<int, 0>{}, p);
    if (p + 1 > last) break;
    iter(integral_constant<int, 1>{}, p + 1);
  }
}

}  // namespace

}  // namespace fs
