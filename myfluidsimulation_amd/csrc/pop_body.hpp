// pop_body.hpp -- per-cell BGK / Hybrid AA update shared by the single-domain
// kernel (lbm_pop.hip) and the batched kernel (batch.hip).  See lbm_pop.hip
// for the streaming scheme.
#pragma once

#include "d3q19_dev.hpp"

namespace fs {

namespace {

// BGK / hybrid collision in place: f -> post-collision, returns moments.
template <bool HYB>
__device__ __forceinline__ void collide(float (&f)[19], float& r, float& ux, float& uy,
                                        float& uz, const Phys& ph, bool band) {
  float rr = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
  for_q<0, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    rr += f[q];
    acc1<CX[q]>(jx, f[q]);
    acc1<CY[q]>(jy, f[q]);
    acc1<CZ[q]>(jz, f[q]);
  });
  float u0 = (rr > 0.0f) ? jx / rr : 0.0f;
  float v0 = (rr > 0.0f) ? jy / rr : 0.0f;
  float w0 = (rr > 0.0f) ? jz / rr : 0.0f;
  u0 += ph.fx;
  v0 += ph.fy;
  w0 += ph.fz;
  const float uu = u0 * u0 + v0 * v0 + w0 * w0;
  const float uu15 = 1.5f * uu;
  const WRho wr(rr);
  if (!HYB || band) {
    const float om = ph.omegaL;
    for_q<0, 19>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      const float fe = feq_q<q>(wr, u0, v0, w0, uu15);
      f[q] = f[q] + om * (fe - f[q]);
    });
  } else {
    const float oh = ph.omegaH;
    const float osh = ph.omegaS - ph.omegaH;
    for_q<0, 19>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      const float fe = feq_q<q>(wr, u0, v0, w0, uu15);
      const float ds = f[q] - fe;
      f[q] = (f[q] + oh * (fe - f[q])) + osh * (-ds);
    });
  }
  r = rr;
  ux = u0;
  uy = v0;
  uz = w0;
}

__device__ __forceinline__ void inlet_emit(float (&f)[19], float uin) {
  const WRho wr(1.0f);
  const float uu15 = 1.5f * (uin * uin + 0.0f * 0.0f + 0.0f * 0.0f);
  for_q<0, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    f[q] = feq_q<q>(wr, uin, 0.0f, 0.0f, uu15);
  });
}

// odd-step emission from origin cell `c` with neighbour offsets `o`; solid:
// the cell's solid-neighbour mask (bit q: c + e_q is solid)
template <bool CHECK>
__device__ __forceinline__ void emit_odd(float* __restrict__ F, size_t P, uint32_t solid,
                                         uint32_t c, const Nbr& o, const float (&f)[19]) {
  for_q<0, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const uint32_t d = c + nb_off<CX[q], CY[q], CZ[q]>(o);
    if (CHECK && q != 0 && ((solid >> q) & 1u)) {
      st_pop(&F[size_t(OPP[q]) * P + c], f[q]);
    } else {
      st_pop(&F[size_t(q) * P + d], f[q]);
    }
  });
}

// One cell's AA step (even or odd), shared by the single-domain and the
// batched launchers.
template <bool ODD, bool HYB, bool MACRO>
__device__ __forceinline__ void pop_cell(const Geo& g, const Phys& ph, const PopLaunch& a,
                                         const Cell& cl) {
  const uint32_t c = cl.c;
  const size_t P = g.P;
  float* __restrict__ F = a.F;
  const uint8_t fl = a.flags[c];
  if (fl & (F_SOLID | F_OUTLET)) {
    // solid: populations frozen, macros (1,0,0,0) (lbm3d_legacy.cu:72-76);
    // outlet: written by its upstream feeder cell
    if (MACRO && (fl & F_SOLID)) {
      a.rho[c] = 1.0f;
      a.ux[c] = 0.0f;
      a.uy[c] = 0.0f;
      a.uz[c] = 0.0f;
    }
    return;
  }
  const Nbr o = nbr_offsets(cl.x, cl.y, cl.z, g);
  const bool anyNear = ODD && (__ballot((fl & F_NEAR) != 0) != 0);
  float f[19];
  float r, ux, uy, uz;
  if (fl & F_INLET) {
    const float uin = a.inlet[size_t(cl.z) * g.ny + cl.y];
    inlet_emit(f, uin);
    r = 1.0f;
    ux = uin;
    uy = 0.0f;
    uz = 0.0f;
  } else {
    if (!ODD) {
      for_q<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f[q] = ld_pop(&F[size_t(q) * P + c]);
      });
    } else if (!anyNear) {
      for_q<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const uint32_t s = c + nb_off<-CX[q], -CY[q], -CZ[q]>(o);
        f[q] = ld_pop(&F[size_t(OPP[q]) * P + s]);
      });
    } else {
      const uint32_t solid = a.nbm[c];
      for_q<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const uint32_t s = c + nb_off<-CX[q], -CY[q], -CZ[q]>(o);
        // c - e_q is the neighbour along e_opp(q)
        if (q != 0 && ((solid >> OPP[q]) & 1u))
          f[q] = ld_pop(&F[size_t(q) * P + c]);
        else
          f[q] = ld_pop(&F[size_t(OPP[q]) * P + s]);
      });
    }
    collide<HYB>(f, r, ux, uy, uz, ph, (fl & F_BAND) != 0);
  }

  if (!ODD) {
    for_q<0, 19>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      st_pop(&F[size_t(OPP[q]) * P + c], f[q]);
    });
    if (fl & F_FEEDER) {
      for_q<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        st_pop(&F[size_t(OPP[q]) * P + c + 1], f[q]);
      });
    }
  } else {
    if (!anyNear)
      emit_odd<false>(F, P, 0u, c, o, f);
    else
      emit_odd<true>(F, P, a.nbm[c], c, o, f);
    if (fl & F_FEEDER) {
      // the outlet cell (x+1 = nx-1) collides exactly like its feeder
      Nbr o1 = o;
      o1.xm = -1;
      o1.xp = -(g.nx - 1);
      emit_odd<true>(F, P, a.nbm[c + 1], c + 1, o1, f);
    }
  }
  if (MACRO) {
    a.rho[c] = r;
    a.ux[c] = ux;
    a.uy[c] = uy;
    a.uz[c] = uz;
    if (fl & F_FEEDER) {
      a.rho[c + 1] = r;
      a.ux[c + 1] = ux;
      a.uy[c + 1] = uy;
      a.uz[c + 1] = uz;
    }
  }
}

}  // namespace

}  // namespace fs
