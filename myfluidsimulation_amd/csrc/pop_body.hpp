// pop_body.hpp -- per-cell BGK / Hybrid AA update shared by the single-domain
// kernel (lbm_pop.hip) and the batched kernel (batch.hip).  See lbm_pop.hip
// for the streaming scheme.
#pragma once

#include "d3q19_dev.hpp"

namespace fs {

namespace {

// Density and momentum of 19 populations, summed in q order.
struct PopSums {
  float r, jx, jy, jz;
};
__device__ __forceinline__ PopSums pop_sums(const float (&f)[19]) {
  PopSums m{0.0f, 0.0f, 0.0f, 0.0f};
  for_q<0, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    m.r += f[q];
    acc1<CX[q]>(m.jx, f[q]);
    acc1<CY[q]>(m.jy, f[q]);
    acc1<CZ[q]>(m.jz, f[q]);
  });
  return m;
}

// One cell's BGK / Hybrid collision (lbm3d_legacy.cu:78-100,
// lbm3d_hybrid.cu:114-160): the moments and equilibrium terms once, then
// post<q>(f_q) relaxes direction q, so a caller may finish only the
// directions it needs.  u carries the forcing shift.
template <bool HYB>
struct PopCollision {
  float r, ux, uy, uz, uu15;
  WRho wr;
  bool band;
  float om, oh, osh;

  __device__ __forceinline__ PopCollision(const PopSums& m, const Phys& ph, bool bandCell)
      : r(m.r), wr(m.r), band(bandCell) {
    ux = (m.r > 0.0f) ? m.jx / m.r : 0.0f;
    uy = (m.r > 0.0f) ? m.jy / m.r : 0.0f;
    uz = (m.r > 0.0f) ? m.jz / m.r : 0.0f;
    ux += ph.fx;
    uy += ph.fy;
    uz += ph.fz;
    const float uu = ux * ux + uy * uy + uz * uz;
    uu15 = 1.5f * uu;
    om = ph.omegaL;
    oh = ph.omegaH;
    osh = ph.omegaS - ph.omegaH;
  }

  __device__ __forceinline__ float relax(float fq, float fe) const {
    if (!HYB || band) return fq + om * (fe - fq);
    const float ds = fq - fe;
    return (fq + oh * (fe - fq)) + osh * (-ds);
  }

  template <int Q>
  __device__ __forceinline__ float post(float fq) const {
    return relax(fq, feq_q<Q>(wr, ux, uy, uz, uu15));
  }

  // out[q] = post<q>(in[q]) for all 19 directions, each direction's
  // equilibrium computed together with its opposite's (feq_pair: same bits)
  __device__ __forceinline__ void post_all(const float (&in)[19], float (&out)[19]) const {
    out[0] = post<0>(in[0]);
    for_q<0, 9>([&](auto kc) {
      constexpr int q = kPairFirst[decltype(kc)::value], o = OPP[q];
      float eq, eo;
      feq_pair<q>(wr, ux, uy, uz, uu15, eq, eo);
      out[q] = relax(in[q], eq);
      out[o] = relax(in[o], eo);
    });
  }

  // An inlet cell emits feq(1, (u_in, 0, 0)) whatever streams into it
  // (inlet_emit) and reports those moments.  With these moments and omega 1,
  // post<q>(0.0f) = 0 + 1 * (feq - 0) rounds to exactly that feq (u.u and
  // feq are inlet_emit's expressions), so inlet lanes can run the fluid
  // lanes' code with zero inputs instead of a branch of their own.
  __device__ __forceinline__ void set_inlet(float uin) {
    r = 1.0f;
    wr = WRho(1.0f);
    ux = uin;
    uy = 0.0f;
    uz = 0.0f;
    uu15 = 1.5f * (ux * ux + uy * uy + uz * uz);
    om = 1.0f;
    band = true;
  }
};

// BGK / hybrid collision in place: f -> post-collision, returns moments.
template <bool HYB>
__device__ __forceinline__ void collide(float (&f)[19], float& r, float& ux, float& uy,
                                        float& uz, const Phys& ph, bool band) {
  const PopCollision<HYB> pc(pop_sums(f), ph, band);
  const float in[19] = {f[0], f[1], f[2],  f[3],  f[4],  f[5],  f[6],  f[7],  f[8], f[9],
                        f[10], f[11], f[12], f[13], f[14], f[15], f[16], f[17], f[18]};
  pc.post_all(in, f);
  r = pc.r;
  ux = pc.ux;
  uy = pc.uy;
  uz = pc.uz;
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
      // the outlet cell x+1 collides exactly like its feeder; its own x
      // neighbours: the feeder, and x+2 (periodic when x+1 = nx-1)
      Nbr o1 = o;
      o1.xm = -1;
      o1.xp = (cl.x + 2 == g.nx) ? -(g.nx - 1) : 1;
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
