// adjoint_api.cpp -- the fs_adjoint_* part of the C ABI (include/fluidsim.h).
//
// A differentiable run of T steps.  The forward pass steps the solver with
// its own kernels and copies its state every S steps (checkpoints).  The
// backward pass walks the segments in reverse: it restores a segment's
// checkpoint, recomputes the segment (the population kernels write the
// collision moments of every step into per-step slots, HOME writes its
// moment states into per-step slots), then runs the adjoint kernels
// (adjoint.hip) down through the segment.  Device memory:
//   checkpoints  ceil(T/S) x (19 | 10) planes
//   step slots   S x 4 planes (populations) | (S-1) x 10 planes (HOME)
//   adjoints     2 x (19 | 10) planes, plus a 19-plane recompute buffer and
//                the 4 initial macro planes for the population methods.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "fluidsim.h"
#include "fs_internal.hpp"
#include "solver_impl.hpp"

using fs::fail;

struct fs_adjoint {
  fs_solver* s = nullptr;
  int T = 0, S = 0, nck = 0;
  bool home = false, hyb = false;
  size_t P = 0;
  float* ck = nullptr;
  std::vector<int> ckParity;
  float* slots = nullptr;
  float* lam[2] = {nullptr, nullptr};
  float* work = nullptr;
  float* init = nullptr;
  uint8_t* aflags = nullptr;
  fs::DragLink* links = nullptr;
  uint32_t nlinks = 0;
  float* partials = nullptr;
  int nblocks = 1;
  size_t bytes = 0;
  bool ready = false;
};

namespace {

constexpr int kMaxDragBlocks = 256;

void free_adjoint(fs_adjoint* a) {
  if (!a) return;
  if (a->s) (void)hipStreamSynchronize(a->s->stream);
  void* bufs[] = {a->ck, a->slots, a->lam[0], a->lam[1], a->work, a->init,
                  a->aflags, a->links, a->partials};
  for (void* p : bufs) (void)hipFree(p);
  delete a;
}

template <class T>
int dev_alloc(fs_adjoint* a, T** p, size_t bytes) {
  if (bytes == 0) bytes = 1;
  hipError_t e = hipMalloc((void**)p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(FS_ERR_ALLOC, "device allocation of %zu bytes failed: %s", bytes,
                hipGetErrorString(e));
  }
  a->bytes += bytes;
  return FS_OK;
}

// Adjoint flag bytes (the solver's derived flags plus F_BODY) and the drag
// link list: (cell, q) for every non-solid cell and direction q with
// e_qx != 0 whose periodic neighbour along e_q is a body cell.
int build_flags_and_links(fs_adjoint* a, const uint8_t* body) {
  fs_solver* s = a->s;
  const fs::Geo& g = s->g;
  std::vector<uint8_t> fl(g.n);
  FS_HIP(hipMemcpy(fl.data(), s->flags, g.n, hipMemcpyDeviceToHost));
  for (uint32_t c = 0; c < g.n; ++c)
    if ((fl[c] & fs::F_SOLID) && (!body || body[c])) fl[c] |= fs::F_BODY;
  std::vector<fs::DragLink> links;
  for (int z = 0; z < g.nz; ++z)
    for (int y = 0; y < g.ny; ++y)
      for (int x = 0; x < g.nx; ++x) {
        const uint32_t c = (uint32_t(z) * g.ny + y) * g.nx + x;
        if (fl[c] & fs::F_SOLID) continue;
        for (int q = 1; q < 19; ++q) {
          if (fs::CX[q] == 0) continue;
          const int xn = (x + fs::CX[q] + g.nx) % g.nx;
          const int yn = (y + fs::CY[q] + g.ny) % g.ny;
          const int zn = (z + fs::CZ[q] + g.nz) % g.nz;
          const uint32_t d = (uint32_t(zn) * g.ny + yn) * g.nx + xn;
          if (fl[d] & fs::F_BODY) links.push_back({c, uint32_t(q)});
        }
      }
  if (int rc = dev_alloc(a, &a->aflags, g.P)) return rc;
  FS_HIP(hipMemcpy(a->aflags, fl.data(), g.n, hipMemcpyHostToDevice));
  a->nlinks = uint32_t(links.size());
  if (int rc = dev_alloc(a, &a->links, sizeof(fs::DragLink) * std::max<size_t>(1, links.size())))
    return rc;
  if (!links.empty())
    FS_HIP(hipMemcpy(a->links, links.data(), sizeof(fs::DragLink) * links.size(),
                     hipMemcpyHostToDevice));
  a->nblocks = int(std::min<uint32_t>(kMaxDragBlocks, std::max<uint32_t>(1, (a->nlinks + 255) / 256)));
  return dev_alloc(a, &a->partials, sizeof(float) * kMaxDragBlocks);
}

// drag of one step over the body links (there is at least one)
int drag_step(fs_adjoint* a, const float* state, float* out) {
  fs_solver* s = a->s;
  if (a->home)
    FS_HIP(fs::launch_drag_home(s->g, s->ph, state, a->links, a->nlinks, a->partials, a->nblocks,
                                out, s->stream));
  else
    FS_HIP(fs::launch_drag_pop(s->g, state, a->links, a->nlinks, a->partials, a->nblocks, out,
                               s->stream));
  return FS_OK;
}

int copy_or_fill(float* dst, const float* src, float v, size_t n, hipStream_t st) {
  if (src) {
    FS_HIP(hipMemcpyAsync(dst, src, sizeof(float) * n, hipMemcpyDeviceToDevice, st));
  } else {
    FS_HIP(fs::launch_fill(dst, v, n, st));
  }
  return FS_OK;
}

// HOME moment state m^{n0 + j} of the segment starting at checkpoint k
float* home_slot(fs_adjoint* a, int k, int j) {
  if (j == 0) return a->ck + size_t(k) * 10 * a->P;
  return a->slots + size_t(j - 1) * 10 * a->P;
}

}  // namespace

extern "C" {

int fs_adjoint_create(fs_solver* s, int T, int segment, const uint8_t* host_body,
                      fs_adjoint** out) {
  if (!out) return fail(FS_ERR_INVALID, "null output handle");
  *out = nullptr;
  if (!s) return fail(FS_ERR_INVALID, "null solver handle");
  if (T < 1) return fail(FS_ERR_INVALID, "T must be >= 1");
  if (segment < 0 || segment > T) return fail(FS_ERR_INVALID, "segment must be in [0, T]");
  if (int rc = fs::runtime_check()) return rc;
  const fs::Geo& g = s->g;
  for (uint32_t c = 0; c < g.n; ++c)
    if ((s->hflags[c] & FS_FLAG_INLET) && (c % uint32_t(g.nx)) != 0)
      return fail(FS_ERR_INVALID, "inlet cell %u is not at x = 0", c);
  fs_adjoint* a = new (std::nothrow) fs_adjoint();
  if (!a) return fail(FS_ERR_ALLOC, "host allocation failed");
  a->s = s;
  a->T = T;
  a->S = segment > 0 ? segment : std::max(1, int(std::ceil(std::sqrt(double(T)))));
  a->nck = (T + a->S - 1) / a->S;
  a->home = s->method == fs::kMethodHOME;
  a->hyb = s->method == fs::kMethodHybrid;
  a->P = g.P;
  const size_t P = g.P;
  auto bail = [&](int rc) {
    free_adjoint(a);
    return rc;
  };
  if (int rc = build_flags_and_links(a, host_body)) return bail(rc);
  const size_t K = a->home ? 10 : 19;
  if (int rc = dev_alloc(a, &a->ck, sizeof(float) * K * P * a->nck)) return bail(rc);
  a->ckParity.assign(a->nck, 0);
  const size_t slotFloats = a->home ? size_t(std::max(0, a->S - 1)) * 10 * P : size_t(a->S) * 4 * P;
  if (int rc = dev_alloc(a, &a->slots, sizeof(float) * slotFloats)) return bail(rc);
  for (float*& l : a->lam)
    if (int rc = dev_alloc(a, &l, sizeof(float) * K * P)) return bail(rc);
  if (!a->home) {
    if (int rc = dev_alloc(a, &a->work, sizeof(float) * 19 * P)) return bail(rc);
    if (int rc = dev_alloc(a, &a->init, sizeof(float) * 4 * P)) return bail(rc);
  }
  *out = a;
  return FS_OK;
}

int fs_adjoint_destroy(fs_adjoint* a) {
  free_adjoint(a);
  return FS_OK;
}

int fs_adjoint_info(const fs_adjoint* a, int* T, int* segment, int* checkpoints,
                    size_t* device_bytes) {
  if (!a) return fail(FS_ERR_INVALID, "null adjoint handle");
  if (T) *T = a->T;
  if (segment) *segment = a->S;
  if (checkpoints) *checkpoints = a->nck;
  if (device_bytes) *device_bytes = a->bytes;
  return FS_OK;
}

int fs_adjoint_forward(fs_adjoint* a, const float* rho0, const float* ux0, const float* uy0,
                       const float* uz0, const float* inlet, float* out_rho, float* out_ux,
                       float* out_uy, float* out_uz, float* drag) {
  if (!a) return fail(FS_ERR_INVALID, "null adjoint handle");
  fs_solver* s = a->s;
  const fs::Geo& g = s->g;
  const size_t P = g.P;
  hipStream_t st = s->stream;
  a->ready = false;
  if (inlet)
    FS_HIP(hipMemcpyAsync(s->inlet, inlet, sizeof(float) * size_t(g.ny) * g.nz,
                          hipMemcpyDeviceToDevice, st));
  if (a->home) {
    if (int rc = solver_reinit_async(s, rho0, ux0, uy0, uz0)) return rc;
  } else {
    // keep (rho0, u0): the backward pass differentiates feq(rho0, u0)
    if (int rc = copy_or_fill(a->init, rho0, 1.0f, g.n, st)) return rc;
    if (int rc = copy_or_fill(a->init + P, ux0, 0.0f, g.n, st)) return rc;
    if (int rc = copy_or_fill(a->init + 2 * P, uy0, 0.0f, g.n, st)) return rc;
    if (int rc = copy_or_fill(a->init + 3 * P, uz0, 0.0f, g.n, st)) return rc;
    if (int rc = solver_reinit_async(s, a->init, a->init + P, a->init + 2 * P, a->init + 3 * P))
      return rc;
  }
  // without body links every step's drag is zero: one fill for all T steps
  if (drag && a->nlinks == 0) FS_HIP(fs::launch_fill(drag, 0.0f, size_t(a->T), st));
  float* const dragSteps = a->nlinks ? drag : nullptr;
  for (int n = 0; n < a->T; ++n) {
    if (n % a->S == 0) {
      const int k = n / a->S;
      if (a->home) {
        FS_HIP(hipMemcpyAsync(a->ck + size_t(k) * 10 * P, s->mom_cur(), sizeof(float) * 10 * P,
                              hipMemcpyDeviceToDevice, st));
      } else {
        FS_HIP(hipMemcpyAsync(a->ck + size_t(k) * 19 * P, s->F, sizeof(float) * 19 * P,
                              hipMemcpyDeviceToDevice, st));
        a->ckParity[k] = s->parity;
      }
    }
    if (a->home) {
      if (dragSteps)
        if (int rc = drag_step(a, s->mom_cur(), dragSteps + n)) return rc;
      const fs::HomeLaunch hl{s->mbuf[s->cur], s->mbuf[1 - s->cur], s->flags, s->nbm, s->inlet};
      FS_HIP(fs::launch_home_step(g, s->ph, hl, solver_home_variant(s), st));
      s->cur = 1 - s->cur;
    } else {
      FS_HIP(fs::launch_pop_step(g, s->ph, pop_args(s), s->parity != 0, a->hyb, n == a->T - 1, st));
      s->parity ^= 1;
      if (dragSteps)
        if (int rc = drag_step(a, s->F, dragSteps + n)) return rc;
    }
  }
  float* outs[4] = {out_rho, out_ux, out_uy, out_uz};
  for (int k = 0; k < 4; ++k) {
    if (!outs[k]) continue;
    const float* src = a->home ? s->mom_cur() + size_t(k) * P : s->macro + size_t(k) * P;
    FS_HIP(hipMemcpyAsync(outs[k], src, sizeof(float) * g.n, hipMemcpyDeviceToDevice, st));
  }
  a->ready = true;
  return FS_OK;
}

int fs_adjoint_backward(fs_adjoint* a, const float* g_rho, const float* g_ux, const float* g_uy,
                        const float* g_uz, const float* g_drag, float* grad_rho0,
                        float* grad_ux0, float* grad_uy0, float* grad_uz0, float* grad_inlet) {
  if (!a) return fail(FS_ERR_INVALID, "null adjoint handle");
  if (!a->ready) return fail(FS_ERR_STATE, "fs_adjoint_backward needs a forward pass first");
  fs_solver* s = a->s;
  const fs::Geo& g = s->g;
  const size_t P = g.P;
  hipStream_t st = s->stream;
  if (grad_inlet)
    FS_HIP(hipMemsetAsync(grad_inlet, 0, sizeof(float) * size_t(g.ny) * g.nz, st));
  int cur = 0;
  if (a->home) {
    FS_HIP(fs::launch_adj_home_seed(g, s->ph, s->mom_cur(), a->aflags, g_rho, g_ux, g_uy, g_uz,
                                    a->lam[0], grad_inlet, st));
  } else {
    FS_HIP(hipMemsetAsync(a->lam[0], 0, sizeof(float) * 19 * P, st));
  }
  const bool anyG = g_rho || g_ux || g_uy || g_uz;
  for (int k = a->nck - 1; k >= 0; --k) {
    const int n0 = k * a->S, n1 = std::min(a->T, n0 + a->S);
    if (a->home) {
      for (int n = n0; n < n1 - 1; ++n) {
        const fs::HomeLaunch hl{home_slot(a, k, n - n0), home_slot(a, k, n - n0 + 1), s->flags,
                                s->nbm, s->inlet};
        FS_HIP(fs::launch_home_step(g, s->ph, hl, solver_home_variant(s), st));
      }
      for (int n = n1 - 1; n >= n0; --n) {
        fs::AdjHomeLaunch al;
        al.lin = a->lam[cur];
        al.lout = n == 0 ? nullptr : a->lam[1 - cur];
        al.m = home_slot(a, k, n - n0);
        al.flags = a->aflags;
        al.gdrag = g_drag ? g_drag + n : nullptr;
        al.ginlet = n >= 1 ? grad_inlet : nullptr;
        al.grho = grad_rho0;
        al.gux = grad_ux0;
        al.guy = grad_uy0;
        al.guz = grad_uz0;
        FS_HIP(fs::launch_adj_home(g, s->ph, al, n == 0, st));
        cur = 1 - cur;
      }
    } else {
      FS_HIP(hipMemcpyAsync(a->work, a->ck + size_t(k) * 19 * P, sizeof(float) * 19 * P,
                            hipMemcpyDeviceToDevice, st));
      int parity = a->ckParity[k];
      fs::PopLaunch pl = pop_args(s);
      pl.F = a->work;
      for (int n = n0; n < n1; ++n) {
        float* slot = a->slots + size_t(n - n0) * 4 * P;
        pl.rho = slot;
        pl.ux = slot + P;
        pl.uy = slot + 2 * P;
        pl.uz = slot + 3 * P;
        FS_HIP(fs::launch_pop_step(g, s->ph, pl, parity != 0, a->hyb, true, st));
        parity ^= 1;
      }
      for (int n = n1 - 1; n >= n0; --n) {
        const float* slot = a->slots + size_t(n - n0) * 4 * P;
        const bool last = n == a->T - 1 && anyG;
        fs::AdjPopLaunch al;
        al.lin = a->lam[cur];
        al.lout = a->lam[1 - cur];
        al.rho = slot;
        al.ux = slot + P;
        al.uy = slot + 2 * P;
        al.uz = slot + 3 * P;
        al.flags = a->aflags;
        al.inlet = s->inlet;
        al.grho = last ? g_rho : nullptr;
        al.gux = last ? g_ux : nullptr;
        al.guy = last ? g_uy : nullptr;
        al.guz = last ? g_uz : nullptr;
        al.gdrag = g_drag ? g_drag + n : nullptr;
        al.ginlet = grad_inlet;
        FS_HIP(fs::launch_adj_pop(g, s->ph, al, a->hyb, st));
        cur = 1 - cur;
      }
    }
  }
  if (!a->home && (grad_rho0 || grad_ux0 || grad_uy0 || grad_uz0))
    FS_HIP(fs::launch_adj_pop_init(g, a->lam[cur], a->init, a->init + P, a->init + 2 * P,
                                   a->init + 3 * P, grad_rho0, grad_ux0, grad_uy0, grad_uz0, st));
  return FS_OK;
}

}  // extern "C"
