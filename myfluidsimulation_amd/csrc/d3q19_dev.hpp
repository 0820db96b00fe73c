// d3q19_dev.hpp -- per-direction D3Q19 arithmetic for the HIP kernels.
//
// Each helper is specialised on the direction index at compile time, so the
// lattice constants fold away.  Terms that the reference multiplies by a zero
// velocity component are dropped and terms multiplied by +-1 become +-x:
// both are exact in IEEE arithmetic (x*1 == x, x*-1 == -x, y + (+-0) == y for
// any nonzero y), so every stored value keeps the reference's bits while
// the non-zero terms are combined in the reference's order.  The library is
// built with -ffp-contract=off so no multiply-add is fused.
#pragma once

#include <type_traits>

#include "fs_internal.hpp"

namespace fs {

template <int Q>
struct Dir {
  static constexpr int cx = CX[Q], cy = CY[Q], cz = CZ[Q], opp = OPP[Q];
  // w19 (lbm3d_legacy.cu:25-28)
  static constexpr float w = Q == 0 ? 1.0f / 3.0f : (Q <= 6 ? 1.0f / 18.0f : 1.0f / 36.0f);
};

template <int C>
__device__ __forceinline__ float sg(float v) {
  return C > 0 ? v : -v;
}

// e_q . u = cx*ux + cy*uy + cz*uz with the zero terms dropped
template <int Q>
__device__ __forceinline__ float edot(float ux, float uy, float uz) {
  constexpr int a = Dir<Q>::cx, b = Dir<Q>::cy, c = Dir<Q>::cz;
  if constexpr (a == 0 && b == 0 && c == 0) {
    return 0.0f;
  } else if constexpr (b == 0 && c == 0) {
    return sg<a>(ux);
  } else if constexpr (a == 0 && c == 0) {
    return sg<b>(uy);
  } else if constexpr (a == 0 && b == 0) {
    return sg<c>(uz);
  } else if constexpr (c == 0) {
    return sg<a>(ux) + sg<b>(uy);
  } else if constexpr (b == 0) {
    return sg<a>(ux) + sg<c>(uz);
  } else {
    return sg<b>(uy) + sg<c>(uz);
  }
}

// acc += v * c for c in {-1,0,1}
template <int C>
__device__ __forceinline__ void acc1(float& acc, float v) {
  if constexpr (C > 0) acc += v;
  if constexpr (C < 0) acc -= v;
}

// w_q * rho for the three weight classes
struct WRho {
  float w0, w1, w7;
  __device__ __forceinline__ WRho(float r)
      : w0(Dir<0>::w * r), w1(Dir<1>::w * r), w7(Dir<7>::w * r) {}
  template <int Q>
  __device__ __forceinline__ float get() const {
    return Q == 0 ? w0 : (Q <= 6 ? w1 : w7);
  }
};

// feq (lbm3d_legacy.cu:54-58), feq19 (lbm3d_hybrid.cu:37-41) and the
// equilibrium part of reconstruct_f_post (lbm3d_home.cu:57-59):
//   w*rho*(1 + 3eu + 4.5eu*eu - 1.5uu)
template <int Q>
__device__ __forceinline__ float feq_q(const WRho& wr, float ux, float uy, float uz, float uu15) {
  const float eu = edot<Q>(ux, uy, uz);
  return wr.template get<Q>() * (((1.0f + 3.0f * eu) + (4.5f * eu) * eu) - uu15);
}

// feq_q<Q> and feq_q<OPP[Q]> together, with the same bits: e_opp = -e_q, so
// the opposite direction's e.u is -eu (negation is exact), its 3*eu is
// -(3*eu), its (4.5*eu)*eu the same product (round-to-nearest is symmetric),
// and 1 + (-t) is 1 - t
template <int Q>
__device__ __forceinline__ void feq_pair(const WRho& wr, float ux, float uy, float uz, float uu15,
                                         float& fq, float& fo) {
  static_assert(Q != 0, "the rest direction has no partner");
  const float eu = edot<Q>(ux, uy, uz);
  const float t1 = 3.0f * eu;
  const float t2 = (4.5f * eu) * eu;
  const float w = wr.template get<Q>();
  fq = w * (((1.0f + t1) + t2) - uu15);
  fo = w * (((1.0f - t1) + t2) - uu15);
}

// The nine (q, OPP[q]) pairs of D3Q19 by their first member.
constexpr int kPairFirst[9] = {1, 3, 5, 7, 8, 11, 12, 15, 16};

// kern_reinit_eq's cs2 form (lbm_reinit.cu:31-49):
//   w*r*(1 + cu/cs2 + 0.5*(cu*cu)/(cs2*cs2) - 0.5*uu/cs2)
template <int Q>
__device__ __forceinline__ float feq_cs2_q(const WRho& wr, float ux, float uy, float uz,
                                           float uu) {
  constexpr float cs2 = 1.0f / 3.0f;
  constexpr float cs4 = cs2 * cs2;
  const float cu = edot<Q>(ux, uy, uz);
  return wr.template get<Q>() * (((1.0f + cu / cs2) + (0.5f * (cu * cu)) / cs4) - (0.5f * uu) / cs2);
}

// 10 HOME moments of one cell: rho, ux, uy, uz, Sxx, Sxy, Sxz, Syy, Syz, Szz
struct Mom {
  float v[10];
};

// reconstruct_f_post (lbm3d_home.cu:37-81) for direction Q
template <int Q>
__device__ __forceinline__ float home_rec(const Mom& m, float oneMinusOmega) {
  constexpr int a = Dir<Q>::cx, b = Dir<Q>::cy, c = Dir<Q>::cz;
  constexpr float cs2 = 1.0f / 3.0f;
  constexpr float Qxx = float(a * a) - cs2;
  constexpr float Qyy = float(b * b) - cs2;
  constexpr float Qzz = float(c * c) - cs2;
  constexpr float w = Dir<Q>::w;
  constexpr float w45 = 4.5f * w;
  const float rho = m.v[0], ux = m.v[1], uy = m.v[2], uz = m.v[3];
  const float eu = edot<Q>(ux, uy, uz);
  const float uu = ux * ux + uy * uy + uz * uz;
  const float feq = (w * rho) * (((1.0f + 3.0f * eu) + (4.5f * eu) * eu) - 1.5f * uu);
  // (Qxx*Sxx + Qyy*Syy + Qzz*Szz) + 2*(Qxy*Sxy + Qxz*Sxz + Qyz*Syz); at most
  // one off-diagonal Q is nonzero (+-1) for a D3Q19 direction
  float contraction = (Qxx * m.v[4] + Qyy * m.v[7]) + Qzz * m.v[9];
  if constexpr (a != 0 && b != 0) contraction += 2.0f * sg<a * b>(m.v[5]);
  if constexpr (a != 0 && c != 0) contraction += 2.0f * sg<a * c>(m.v[6]);
  if constexpr (b != 0 && c != 0) contraction += 2.0f * sg<b * c>(m.v[8]);
  const float fneq = w45 * contraction;
  return feq + oneMinusOmega * fneq;
}

// One source cell's shared terms of reconstruct_f_post: everything that does
// not depend on the direction (u.u, w*rho per weight class, the two possible
// Q_aa*S_aa products per axis) is computed once; emit<Q>() then finishes
// direction Q with exactly the operations home_rec<Q> performs, so the result
// has the same bits.
struct HomeSrc {
  float ux, uy, uz, uu15, wr0, wr1, wr7;
  float pA[3], pB[3];  // (1 - 1/3)*S_aa and (0 - 1/3)*S_aa for a = x, y, z
  float sxy, sxz, syz, omm;

  __device__ __forceinline__ HomeSrc(const Mom& m, float oneMinusOmega) {
    constexpr float cs2 = 1.0f / 3.0f;
    constexpr float A = 1.0f - cs2, B = 0.0f - cs2;
    const float rho = m.v[0];
    ux = m.v[1];
    uy = m.v[2];
    uz = m.v[3];
    uu15 = 1.5f * (ux * ux + uy * uy + uz * uz);
    wr0 = Dir<0>::w * rho;
    wr1 = Dir<1>::w * rho;
    wr7 = Dir<7>::w * rho;
    pA[0] = A * m.v[4];
    pB[0] = B * m.v[4];
    pA[1] = A * m.v[7];
    pB[1] = B * m.v[7];
    pA[2] = A * m.v[9];
    pB[2] = B * m.v[9];
    sxy = m.v[5];
    sxz = m.v[6];
    syz = m.v[8];
    omm = oneMinusOmega;
  }

  template <int Q>
  __device__ __forceinline__ float emit() const {
    constexpr int a = Dir<Q>::cx, b = Dir<Q>::cy, c = Dir<Q>::cz;
    constexpr float w45 = 4.5f * Dir<Q>::w;
    const float wr = Q == 0 ? wr0 : (Q <= 6 ? wr1 : wr7);
    const float eu = edot<Q>(ux, uy, uz);
    const float feq = wr * (((1.0f + 3.0f * eu) + (4.5f * eu) * eu) - uu15);
    float contraction = ((a ? pA[0] : pB[0]) + (b ? pA[1] : pB[1])) + (c ? pA[2] : pB[2]);
    if constexpr (a != 0 && b != 0) contraction += 2.0f * sg<a * b>(sxy);
    if constexpr (a != 0 && c != 0) contraction += 2.0f * sg<a * c>(sxz);
    if constexpr (b != 0 && c != 0) contraction += 2.0f * sg<b * c>(syz);
    return feq + omm * (w45 * contraction);
  }

  // emit<Q>() and emit<OPP[Q]>() in one go, with the same bits as the two
  // calls: e_opp = -e_q, so the opposite direction's e.u is -eu, its 3*eu is
  // -(3*eu) and its (4.5*eu)*eu is the same product (negation is exact and
  // round-to-nearest is symmetric), 1 + (-t) is 1 - t, and the contraction
  // (even in e) is the same value.
  template <int Q>
  __device__ __forceinline__ void emit_pair(float& fq, float& fo) const {
    static_assert(Q != 0, "the rest direction has no partner");
    constexpr int a = Dir<Q>::cx, b = Dir<Q>::cy, c = Dir<Q>::cz;
    constexpr float w45 = 4.5f * Dir<Q>::w;
    const float wr = Q <= 6 ? wr1 : wr7;
    const float eu = edot<Q>(ux, uy, uz);
    const float t1 = 3.0f * eu;
    const float t2 = (4.5f * eu) * eu;
    float contraction = ((a ? pA[0] : pB[0]) + (b ? pA[1] : pB[1])) + (c ? pA[2] : pB[2]);
    if constexpr (a != 0 && b != 0) contraction += 2.0f * sg<a * b>(sxy);
    if constexpr (a != 0 && c != 0) contraction += 2.0f * sg<a * c>(sxz);
    if constexpr (b != 0 && c != 0) contraction += 2.0f * sg<b * c>(syz);
    const float fneq = omm * (w45 * contraction);
    fq = wr * (((1.0f + t1) + t2) - uu15) + fneq;
    fo = wr * (((1.0f - t1) + t2) - uu15) + fneq;
  }
};

// Compile-time loop over the 19 directions.
template <int Q, int END, class Fn>
__device__ __forceinline__ void for_q(Fn&& fn) {
  if constexpr (Q < END) {
    fn(std::integral_constant<int, Q>{});
    for_q<Q + 1, END>(fn);
  }
}

// Periodic neighbour offsets of cell (x,y,z) in the linear index.
struct Nbr {
  int xm, xp, ym, yp, zm, zp;
};
__device__ __forceinline__ Nbr nbr_offsets(int x, int y, int z, const Geo& g) {
  Nbr o;
  o.xm = (x == 0) ? (g.nx - 1) : -1;
  o.xp = (x == g.nx - 1) ? -(g.nx - 1) : 1;
  o.ym = (y == 0) ? (g.ny - 1) * g.nx : -g.nx;
  o.yp = (y == g.ny - 1) ? -(g.ny - 1) * g.nx : g.nx;
  o.zm = (z == 0) ? int((g.nz - 1) * g.nxy) : -int(g.nxy);
  o.zp = (z == g.nz - 1) ? -int((g.nz - 1) * g.nxy) : int(g.nxy);
  return o;
}
template <int DX, int DY, int DZ>
__device__ __forceinline__ int nb_off(const Nbr& o) {
  int s = 0;
  if constexpr (DX < 0) s += o.xm;
  if constexpr (DX > 0) s += o.xp;
  if constexpr (DY < 0) s += o.ym;
  if constexpr (DY > 0) s += o.yp;
  if constexpr (DZ < 0) s += o.zm;
  if constexpr (DZ > 0) s += o.zp;
  return s;
}

#ifndef FS_NT
#define FS_NT 0
#endif
#ifndef FS_XCD
#define FS_XCD 0
#endif

// Population loads and stores of the streaming kernels (non-temporal when
// FS_NT: every value is touched once per step, so caching it buys nothing).
__device__ __forceinline__ float ld_pop(const float* p) {
#if FS_NT
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}
__device__ __forceinline__ void st_pop(float* p, float v) {
#if FS_NT
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}

// Cell coordinates for the (plane-linear, z) launch grid.  With FS_XCD the
// workgroup order is remapped so the workgroups the dispatcher sends to one
// XCD (every 8th) cover one contiguous slab of cells, keeping the partial
// cache lines shared by neighbouring workgroups inside one L2.
struct Cell {
  uint32_t c;
  int x, y, z;
};
__device__ __forceinline__ bool cell_of_thread(const Geo& g, Cell& cl) {
  uint32_t bx = blockIdx.x, bz = blockIdx.y;
#if FS_XCD
  const uint32_t total = gridDim.x * gridDim.y;
  if ((total & 7u) == 0) {
    const uint32_t L = blockIdx.y * gridDim.x + blockIdx.x;
    const uint32_t logical = (L & 7u) * (total >> 3) + (L >> 3);
    bz = logical / gridDim.x;
    bx = logical - bz * gridDim.x;
  }
#endif
  const uint32_t cp = bx * blockDim.x + threadIdx.x;
  if (cp >= g.nxy) return false;
  cl.z = int(bz);
  const uint32_t y = g.divx.div(cp);
  cl.y = int(y);
  cl.x = int(cp - y * uint32_t(g.nx));
  cl.c = uint32_t(cl.z) * g.nxy + cp;
  return true;
}

}  // namespace fs
