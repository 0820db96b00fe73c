// adjoint.hip -- reverse-mode kernels of the population (BGK / Hybrid) and
// HOME steps, and the momentum-exchange drag they can be differentiated for.
//
// The adjoint of a step is written in the canonical (reference-order)
// representation, whatever AA parity the forward state was in:
//
// Populations.  Step n maps f^n to f^{n+1}: f*(x) = C(f^n(x)) (collision,
// or feq(1, (u_in,0,0)) on an inlet cell, or the feeder's f* on an outlet
// cell), then f^{n+1}_q(x + e_q) = f*_q(x), or f^{n+1}_{opp q}(x) = f*_q(x)
// when x + e_q is solid (halfway bounce-back).  Transposed:
//   lambda*_q(x) = lambda^{n+1}_q(x + e_q)            (a pull along e_q)
//                | lambda^{n+1}_{opp q}(x) + 2 e_qx g_D  (bounce-back link)
//   lambda^n(x)  = J_C(rho, u)^T lambda*(x) (+ the macros' own gradient)
// J_C depends on f only through rho and u, so a recomputed forward segment
// stores 4 floats per cell and step, not 19.
//
// HOME.  Step n maps moments m^n to m^{n+1}: every gathering cell x sums the
// raw moments R = sum_q M_q f_q of f_q = rec_q(m^n(x - e_q)) (or
// rec_{opp q}(m^n(x)) across a wall) and finishes m^{n+1} = fin(R).  The
// adjoint state is dL/dR (10 floats): a cell s collects
//   nu_p(s) = M_p . dL/dR(s + e_p)  (+ bounce-back and drag terms),
//   dL/dm^n(s) = sum_p J_rec_p(m^n(s))^T nu_p,
//   dL/dR^n(s) = J_fin(m^n(s))^T (dL/dm^n(s) + [outlet at s+1] dL/dm^n(s+1)).
#include "d3q19_dev.hpp"

namespace fs {

namespace {

inline dim3 lattice_grid(const Geo& g) { return dim3((g.nxy + 255) / 256, g.nz, 1); }

// ------------------------------------------------------------ populations
// dL/df* of one emitting cell: pull along e_q, bounce-back links read the
// cell's own opposite slot and add the drag weight of a body neighbour.
template <bool CHECK>
__device__ __forceinline__ void adj_pull(const float* __restrict__ L, size_t P,
                                         const uint8_t* __restrict__ flags, uint32_t c,
                                         const Nbr& o, float gd, float (&ls)[19]) {
  for_q<0, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const uint32_t d = c + nb_off<CX[q], CY[q], CZ[q]>(o);
    if constexpr (CHECK && q != 0) {
      const uint8_t fd = flags[d];
      if (fd & F_SOLID) {
        float v = L[size_t(OPP[q]) * P + c];
        if constexpr (CX[q] != 0) {
          if (fd & F_BODY) v += float(2 * CX[q]) * gd;
        }
        ls[q] = v;
        return;
      }
    }
    ls[q] = L[size_t(q) * P + d];
  });
}

// dL/df of f* = f + om (feq(rho, u) - f), u = j/rho + F, plus the gradient
// (gr, gx, gy, gz) of the macros (rho, u) the step reports.
__device__ __forceinline__ void adj_collide(const float (&ls)[19], float r, float ux, float uy,
                                            float uz, const Phys& ph, float om, float gr, float gx,
                                            float gy, float gz, float (&lo)[19]) {
  const float uu15 = 1.5f * (ux * ux + uy * uy + uz * uz);
  float A = 0.0f, S = 0.0f, Bx = 0.0f, By = 0.0f, Bz = 0.0f;
  for_q<0, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const float t = ls[q] * Dir<q>::w;
    const float eu = edot<q>(ux, uy, uz);
    A += t * (((1.0f + 3.0f * eu) + 4.5f * eu * eu) - uu15);  // dfeq_q/drho
    S += t;
    const float k = t * (3.0f + 9.0f * eu);  // e_q part of dfeq_q/du / rho
    acc1<CX[q]>(Bx, k);
    acc1<CY[q]>(By, k);
    acc1<CZ[q]>(Bz, k);
  });
  // sum_q ls_q dfeq_q/du = rho (B - 3 u S); du/df_p = (e_p - ubar)/rho
  const float inv = 1.0f / r;
  const float vx = (om * (r * (Bx - 3.0f * ux * S)) + gx) * inv;
  const float vy = (om * (r * (By - 3.0f * uy * S)) + gy) * inv;
  const float vz = (om * (r * (Bz - 3.0f * uz * S)) + gz) * inv;
  const float ubx = ux - ph.fx, uby = uy - ph.fy, ubz = uz - ph.fz;
  const float base = (om * A + gr) - ((vx * ubx + vy * uby) + vz * ubz);
  const float omm = 1.0f - om;
  for_q<0, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    lo[q] = omm * ls[q] + (base + edot<q>(vx, vy, vz));
  });
}

template <bool HYB>
__global__ void __launch_bounds__(256) k_adj_pop(Geo g, Phys ph, AdjPopLaunch a) {
  Cell cl;
  if (!cell_of_thread(g, cl)) return;
  const uint32_t c = cl.c;
  const size_t P = g.P;
  const uint8_t fl = a.flags[c];
  // solid populations and an outlet's own ones are never read by the step
  const bool dead = (fl & (F_SOLID | F_OUTLET)) != 0;
  const bool anyNear = __ballot(!dead && (fl & F_NEAR) != 0) != 0;
  if (dead) {
#pragma unroll
    for (int q = 0; q < 19; ++q) a.lout[size_t(q) * P + c] = 0.0f;
    return;
  }
  const float gd = a.gdrag ? *a.gdrag : 0.0f;
  const Nbr o = nbr_offsets(cl.x, cl.y, cl.z, g);
  float ls[19];
  if (anyNear)
    adj_pull<true>(a.lin, P, a.flags, c, o, gd, ls);
  else
    adj_pull<false>(a.lin, P, a.flags, c, o, gd, ls);
  float gr = a.grho ? a.grho[c] : 0.0f;
  float gx = a.gux ? a.gux[c] : 0.0f;
  float gy = a.guy ? a.guy[c] : 0.0f;
  float gz = a.guz ? a.guz[c] : 0.0f;
  if (fl & F_FEEDER) {
    // the outlet cell x+1 emits this cell's f* with its own neighbours
    Nbr o1 = o;
    o1.xm = -1;
    o1.xp = (cl.x + 2 == g.nx) ? -(g.nx - 1) : 1;
    float lx[19];
    adj_pull<true>(a.lin, P, a.flags, c + 1, o1, gd, lx);
#pragma unroll
    for (int q = 0; q < 19; ++q) ls[q] += lx[q];
    if (a.grho) gr += a.grho[c + 1];
    if (a.gux) gx += a.gux[c + 1];
    if (a.guy) gy += a.guy[c + 1];
    if (a.guz) gz += a.guz[c + 1];
  }
  if (fl & F_INLET) {
    // f* = feq(1, (u_in, 0, 0)) and the reported u_x = u_in
    if (a.ginlet) {
      const size_t row = size_t(cl.z) * g.ny + cl.y;
      const float uin = a.inlet[row];
      float d = gx;
      for_q<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (CX[q] != 0) {
          const float eu = sg<CX[q]>(uin);
          d += ls[q] * (Dir<q>::w * ((3.0f + 9.0f * eu) * float(CX[q]) - 3.0f * uin));
        } else {
          d += ls[q] * (Dir<q>::w * (-3.0f * uin));
        }
      });
      a.ginlet[row] += d;
    }
#pragma unroll
    for (int q = 0; q < 19; ++q) a.lout[size_t(q) * P + c] = 0.0f;
    return;
  }
  const float om = (HYB && (fl & F_BAND) == 0) ? ph.omegaS : ph.omegaL;
  float lo[19];
  adj_collide(ls, a.rho[c], a.ux[c], a.uy[c], a.uz[c], ph, om, gr, gx, gy, gz, lo);
#pragma unroll
  for (int q = 0; q < 19; ++q) a.lout[size_t(q) * P + c] = lo[q];
}

// dL/d(rho0, u0) of f^0 = feq(rho0, u0)
__global__ void __launch_bounds__(256) k_adj_pop_init(Geo g, const float* __restrict__ lam,
                                                      const float* __restrict__ rho0,
                                                      const float* __restrict__ ux0,
                                                      const float* __restrict__ uy0,
                                                      const float* __restrict__ uz0, float* grho,
                                                      float* gux, float* guy, float* guz) {
  Cell cl;
  if (!cell_of_thread(g, cl)) return;
  const uint32_t c = cl.c;
  const size_t P = g.P;
  const float r = rho0[c], ux = ux0[c], uy = uy0[c], uz = uz0[c];
  const float uu15 = 1.5f * (ux * ux + uy * uy + uz * uz);
  float A = 0.0f, S = 0.0f, Bx = 0.0f, By = 0.0f, Bz = 0.0f;
  for_q<0, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const float t = lam[size_t(q) * P + c] * Dir<q>::w;
    const float eu = edot<q>(ux, uy, uz);
    A += t * (((1.0f + 3.0f * eu) + 4.5f * eu * eu) - uu15);
    S += t;
    const float k = t * (3.0f + 9.0f * eu);
    acc1<CX[q]>(Bx, k);
    acc1<CY[q]>(By, k);
    acc1<CZ[q]>(Bz, k);
  });
  if (grho) grho[c] = A;
  if (gux) gux[c] = r * (Bx - 3.0f * ux * S);
  if (guy) guy[c] = r * (By - 3.0f * uy * S);
  if (guz) guz[c] = r * (Bz - 3.0f * uz * S);
}

// ------------------------------------------------------------------- HOME
// M_Q . l for the raw-moment adjoint l = (r, jx, jy, jz, Mxx, Mxy, Mxz, Myy,
// Myz, Mzz) read at cell x
template <int Q>
__device__ __forceinline__ float proj_at(const float* __restrict__ L, size_t P, uint32_t x) {
  constexpr int a = CX[Q], b = CY[Q], c = CZ[Q];
  float v = L[x];
  if constexpr (a != 0) v += sg<a>(L[P + x]);
  if constexpr (b != 0) v += sg<b>(L[2 * P + x]);
  if constexpr (c != 0) v += sg<c>(L[3 * P + x]);
  if constexpr (a != 0) v += L[4 * P + x];
  if constexpr (a != 0 && b != 0) v += sg<a * b>(L[5 * P + x]);
  if constexpr (a != 0 && c != 0) v += sg<a * c>(L[6 * P + x]);
  if constexpr (b != 0) v += L[7 * P + x];
  if constexpr (b != 0 && c != 0) v += sg<b * c>(L[8 * P + x]);
  if constexpr (c != 0) v += L[9 * P + x];
  return v;
}

template <int Q>
__device__ __forceinline__ float proj_reg(const float (&l)[10]) {
  constexpr int a = CX[Q], b = CY[Q], c = CZ[Q];
  float v = l[0];
  if constexpr (a != 0) v += sg<a>(l[1]);
  if constexpr (b != 0) v += sg<b>(l[2]);
  if constexpr (c != 0) v += sg<c>(l[3]);
  if constexpr (a != 0) v += l[4];
  if constexpr (a != 0 && b != 0) v += sg<a * b>(l[5]);
  if constexpr (a != 0 && c != 0) v += sg<a * c>(l[6]);
  if constexpr (b != 0) v += l[7];
  if constexpr (b != 0 && c != 0) v += sg<b * c>(l[8]);
  if constexpr (c != 0) v += l[9];
  return v;
}

// sum_p nu_p d rec_p(m)/dm, rec_p = feq_p(rho, u) + omm 4.5 w_p Q_p : S
__device__ __forceinline__ void rec_adjoint(const float (&nu)[19], const Mom& m, float omm,
                                            float (&lam)[10]) {
  constexpr float cs2 = 1.0f / 3.0f;
  const float r = m.v[0], ux = m.v[1], uy = m.v[2], uz = m.v[3];
  const float uu15 = 1.5f * (ux * ux + uy * uy + uz * uz);
  float A = 0.0f, S = 0.0f, Bx = 0.0f, By = 0.0f, Bz = 0.0f;
  float Dxx = 0.0f, Dyy = 0.0f, Dzz = 0.0f, Dxy = 0.0f, Dxz = 0.0f, Dyz = 0.0f;
  for_q<0, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int a = CX[q], b = CY[q], c = CZ[q];
    const float t = nu[q] * Dir<q>::w;
    const float eu = edot<q>(ux, uy, uz);
    A += t * (((1.0f + 3.0f * eu) + 4.5f * eu * eu) - uu15);
    S += t;
    const float k = t * (3.0f + 9.0f * eu);
    acc1<a>(Bx, k);
    acc1<b>(By, k);
    acc1<c>(Bz, k);
    if constexpr (a != 0) Dxx += t;
    if constexpr (b != 0) Dyy += t;
    if constexpr (c != 0) Dzz += t;
    acc1<a * b>(Dxy, t);
    acc1<a * c>(Dxz, t);
    acc1<b * c>(Dyz, t);
  });
  const float k45 = 4.5f * omm;
  lam[0] = A;
  lam[1] = r * (Bx - 3.0f * ux * S);
  lam[2] = r * (By - 3.0f * uy * S);
  lam[3] = r * (Bz - 3.0f * uz * S);
  lam[4] = k45 * (Dxx - cs2 * S);
  lam[5] = 2.0f * k45 * Dxy;
  lam[6] = 2.0f * k45 * Dxz;
  lam[7] = k45 * (Dyy - cs2 * S);
  lam[8] = 2.0f * k45 * Dyz;
  lam[9] = k45 * (Dzz - cs2 * S);
}

// dL/dR from dL/dm through m = fin(R): u = j/r + F, S = M - r cs2 I - r u u
__device__ __forceinline__ void finish_adjoint(const Mom& m, const Phys& ph, const float (&l)[10],
                                               float (&R)[10]) {
  constexpr float cs2 = 1.0f / 3.0f;
  const float r = m.v[0], ux = m.v[1], uy = m.v[2], uz = m.v[3];
  const float bx = ux - ph.fx, by = uy - ph.fy, bz = uz - ph.fz;  // j / r
  const float inv = 1.0f / r;
  R[1] = l[1] * inv - ((2.0f * ux * l[4] + uy * l[5]) + uz * l[6]);
  R[2] = l[2] * inv - ((2.0f * uy * l[7] + ux * l[5]) + uz * l[8]);
  R[3] = l[3] * inv - ((2.0f * uz * l[9] + ux * l[6]) + uy * l[8]);
  float r0 = l[0] - ((l[1] * bx + l[2] * by) + l[3] * bz) * inv;
  r0 += l[4] * ((bx * bx - ph.fx * ph.fx) - cs2);
  r0 += l[7] * ((by * by - ph.fy * ph.fy) - cs2);
  r0 += l[9] * ((bz * bz - ph.fz * ph.fz) - cs2);
  r0 += l[5] * (bx * by - ph.fx * ph.fy);
  r0 += l[6] * (bx * bz - ph.fx * ph.fz);
  r0 += l[8] * (by * bz - ph.fy * ph.fz);
  R[0] = r0;
#pragma unroll
  for (int k = 4; k < 10; ++k) R[k] = l[k];
}

__device__ __forceinline__ void load_mom10(const float* __restrict__ m, size_t P, uint32_t s,
                                           Mom& v) {
#pragma unroll
  for (int k = 0; k < 10; ++k) v.v[k] = m[size_t(k) * P + s];
}

// dL/dm^n(s) of a non-solid cell s.  gathers: s is a gathering cell (its
// bounce-back links read its own moments).
template <bool CHECK, int BB>
__device__ __forceinline__ void home_lam(const AdjHomeLaunch& a, size_t P, uint32_t s,
                                         const Nbr& o, const Mom& ms, bool gathers, float gd,
                                         float omm, float (&lam)[10]) {
  float own[10];
#pragma unroll
  for (int k = 0; k < 10; ++k) own[k] = a.lin[size_t(k) * P + s];
  float nu[19];
  nu[0] = own[0];
  for_q<1, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    nu[q] = proj_at<q>(a.lin, P, s + nb_off<CX[q], CY[q], CZ[q]>(o));
  });
  if constexpr (CHECK) {
    for_q<1, 19>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      // arrival direction q across a wall reuses direction p of s itself
      constexpr int p = BB == 0 ? OPP[q] : OPP_REF[q];
      if (gathers && (a.flags[s + nb_off<-CX[q], -CY[q], -CZ[q]>(o)] & F_SOLID))
        nu[p] += proj_reg<q>(own);
      if constexpr (CX[q] != 0) {
        const uint8_t fd = a.flags[s + nb_off<CX[q], CY[q], CZ[q]>(o)];
        if ((fd & F_SOLID) && (fd & F_BODY)) nu[q] += float(2 * CX[q]) * gd;
      }
    });
  }
  rec_adjoint(nu, ms, omm, lam);
}

template <int BB, bool FINAL>
__global__ void __launch_bounds__(256) k_adj_home(Geo g, Phys ph, AdjHomeLaunch a) {
  Cell cl;
  if (!cell_of_thread(g, cl)) return;
  const uint32_t s = cl.c;
  const size_t P = g.P;
  const uint8_t fl = a.flags[s];
  const bool solid = (fl & F_SOLID) != 0;
  const bool anyNear = __ballot(!solid && (fl & F_NEAR) != 0) != 0;
  if (solid) {
    if (FINAL) {
      if (a.grho) a.grho[s] = 0.0f;
      if (a.gux) a.gux[s] = 0.0f;
      if (a.guy) a.guy[s] = 0.0f;
      if (a.guz) a.guz[s] = 0.0f;
    } else {
#pragma unroll
      for (int k = 0; k < 10; ++k) a.lout[size_t(k) * P + s] = 0.0f;
    }
    return;
  }
  const bool gathers = (fl & (F_INLET | F_OUTLET)) == 0;
  const float gd = a.gdrag ? *a.gdrag : 0.0f;
  const float omm = 1.0f - ph.omegaL;
  const Nbr o = nbr_offsets(cl.x, cl.y, cl.z, g);
  Mom ms;
  load_mom10(a.m, P, s, ms);
  float lam[10];
  if (anyNear)
    home_lam<true, BB>(a, P, s, o, ms, gathers, gd, omm, lam);
  else
    home_lam<false, BB>(a, P, s, o, ms, gathers, gd, omm, lam);
  if (FINAL) {
    if (a.grho) a.grho[s] = lam[0];
    if (a.gux) a.gux[s] = lam[1];
    if (a.guy) a.guy[s] = lam[2];
    if (a.guz) a.guz[s] = lam[3];
    return;
  }
  if (!gathers) {
    // inlet moments are (1, u_in, 0, 0, 0...) from step 1 on; an outlet's
    // gradient reaches the step through its feeder x-1
    if ((fl & F_INLET) && a.ginlet) a.ginlet[size_t(cl.z) * g.ny + cl.y] += lam[1];
#pragma unroll
    for (int k = 0; k < 10; ++k) a.lout[size_t(k) * P + s] = 0.0f;
    return;
  }
  if (fl & F_FEEDER) {
    Nbr o1 = o;
    o1.xm = -1;
    o1.xp = (cl.x + 2 == g.nx) ? -(g.nx - 1) : 1;
    Mom mo;
    load_mom10(a.m, P, s + 1, mo);
    float lo[10];
    home_lam<true, BB>(a, P, s + 1, o1, mo, false, gd, omm, lo);
#pragma unroll
    for (int k = 0; k < 10; ++k) lam[k] += lo[k];
  }
  float R[10];
  finish_adjoint(ms, ph, lam, R);
#pragma unroll
  for (int k = 0; k < 10; ++k) a.lout[size_t(k) * P + s] = R[k];
}

__global__ void __launch_bounds__(256)
    k_adj_home_seed(Geo g, Phys ph, const float* __restrict__ m, const uint8_t* __restrict__ flags,
                    const float* grho, const float* gux, const float* guy, const float* guz,
                    float* __restrict__ lout, float* ginlet) {
  Cell cl;
  if (!cell_of_thread(g, cl)) return;
  const uint32_t s = cl.c;
  const size_t P = g.P;
  const uint8_t fl = flags[s];
  if (fl & (F_SOLID | F_INLET | F_OUTLET)) {
    if ((fl & F_INLET) && ginlet && gux) ginlet[size_t(cl.z) * g.ny + cl.y] += gux[s];
#pragma unroll
    for (int k = 0; k < 10; ++k) lout[size_t(k) * P + s] = 0.0f;
    return;
  }
  float l[10];
#pragma unroll
  for (int k = 0; k < 10; ++k) l[k] = 0.0f;
  if (grho) l[0] = grho[s];
  if (gux) l[1] = gux[s];
  if (guy) l[2] = guy[s];
  if (guz) l[3] = guz[s];
  if (fl & F_FEEDER) {
    if (grho) l[0] += grho[s + 1];
    if (gux) l[1] += gux[s + 1];
    if (guy) l[2] += guy[s + 1];
    if (guz) l[3] += guz[s + 1];
  }
  Mom ms;
  load_mom10(m, P, s, ms);
  float R[10];
  finish_adjoint(ms, ph, l, R);
#pragma unroll
  for (int k = 0; k < 10; ++k) lout[size_t(k) * P + s] = R[k];
}

// ------------------------------------------------------------------- drag
constexpr int kRed = 256;

__device__ __forceinline__ void block_sum_store(float v, float* out) {
  __shared__ float red[kRed];
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int h = kRed / 2; h > 0; h >>= 1) {
    if (t < h) red[t] += red[t + h];
    __syncthreads();
  }
  if (t == 0) *out = red[0];
}

__global__ void __launch_bounds__(kRed) k_drag_pop(Geo g, const float* __restrict__ F,
                                                   const DragLink* __restrict__ links,
                                                   uint32_t nlinks, float* partials) {
  float acc = 0.0f;
  for (uint32_t i = blockIdx.x * kRed + threadIdx.x; i < nlinks; i += gridDim.x * kRed) {
    const DragLink L = links[i];
    for_q<1, 19>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      if constexpr (CX[q] != 0) {
        // post-step AA slot of the bounced value f*_q: F[OPP q][cell]
        if (L.q == uint32_t(q)) acc += float(2 * CX[q]) * F[size_t(OPP[q]) * g.P + L.cell];
      }
    });
  }
  block_sum_store(acc, partials + blockIdx.x);
}

__global__ void __launch_bounds__(kRed) k_drag_home(Geo g, Phys ph, const float* __restrict__ m,
                                                    const DragLink* __restrict__ links,
                                                    uint32_t nlinks, float* partials) {
  const float omm = 1.0f - ph.omegaL;
  float acc = 0.0f;
  for (uint32_t i = blockIdx.x * kRed + threadIdx.x; i < nlinks; i += gridDim.x * kRed) {
    const DragLink L = links[i];
    Mom mm;
    load_mom10(m, g.P, L.cell, mm);
    for_q<1, 19>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      if constexpr (CX[q] != 0) {
        if (L.q == uint32_t(q)) acc += float(2 * CX[q]) * home_rec<q>(mm, omm);
      }
    });
  }
  block_sum_store(acc, partials + blockIdx.x);
}

__global__ void __launch_bounds__(kRed) k_drag_sum(const float* __restrict__ partials, int n,
                                                   float* out) {
  float acc = 0.0f;
  for (int i = threadIdx.x; i < n; i += kRed) acc += partials[i];
  block_sum_store(acc, out);
}

__global__ void __launch_bounds__(256) k_fill(float* p, float v, size_t n) {
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace

hipError_t launch_adj_pop(const Geo& g, const Phys& ph, const AdjPopLaunch& a, bool hybrid,
                          hipStream_t st) {
  if (hybrid)
    hipLaunchKernelGGL(k_adj_pop<true>, lattice_grid(g), dim3(256), 0, st, g, ph, a);
  else
    hipLaunchKernelGGL(k_adj_pop<false>, lattice_grid(g), dim3(256), 0, st, g, ph, a);
  return hipGetLastError();
}

hipError_t launch_adj_pop_init(const Geo& g, const float* lam, const float* rho0, const float* ux0,
                               const float* uy0, const float* uz0, float* grho, float* gux,
                               float* guy, float* guz, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_pop_init, lattice_grid(g), dim3(256), 0, st, g, lam, rho0, ux0, uy0,
                     uz0, grho, gux, guy, guz);
  return hipGetLastError();
}

hipError_t launch_adj_home(const Geo& g, const Phys& ph, const AdjHomeLaunch& a, bool final_step,
                           hipStream_t st) {
  const int key = (ph.bounce ? 2 : 0) | (final_step ? 1 : 0);
  switch (key) {
    case 0: hipLaunchKernelGGL((k_adj_home<0, false>), lattice_grid(g), dim3(256), 0, st, g, ph, a); break;
    case 1: hipLaunchKernelGGL((k_adj_home<0, true>), lattice_grid(g), dim3(256), 0, st, g, ph, a); break;
    case 2: hipLaunchKernelGGL((k_adj_home<1, false>), lattice_grid(g), dim3(256), 0, st, g, ph, a); break;
    default: hipLaunchKernelGGL((k_adj_home<1, true>), lattice_grid(g), dim3(256), 0, st, g, ph, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_adj_home_seed(const Geo& g, const Phys& ph, const float* m, const uint8_t* flags,
                                const float* grho, const float* gux, const float* guy,
                                const float* guz, float* lout, float* ginlet, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_home_seed, lattice_grid(g), dim3(256), 0, st, g, ph, m, flags, grho,
                     gux, guy, guz, lout, ginlet);
  return hipGetLastError();
}

hipError_t launch_drag_pop(const Geo& g, const float* F, const DragLink* links, uint32_t nlinks,
                           float* partials, int nblocks, float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_drag_pop, dim3(nblocks), dim3(kRed), 0, st, g, F, links, nlinks, partials);
  hipLaunchKernelGGL(k_drag_sum, dim3(1), dim3(kRed), 0, st, partials, nblocks, out);
  return hipGetLastError();
}

hipError_t launch_drag_home(const Geo& g, const Phys& ph, const float* m, const DragLink* links,
                            uint32_t nlinks, float* partials, int nblocks, float* out,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_drag_home, dim3(nblocks), dim3(kRed), 0, st, g, ph, m, links, nlinks,
                     partials);
  hipLaunchKernelGGL(k_drag_sum, dim3(1), dim3(kRed), 0, st, partials, nblocks, out);
  return hipGetLastError();
}

hipError_t launch_fill(float* p, float v, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}

}  // namespace fs
