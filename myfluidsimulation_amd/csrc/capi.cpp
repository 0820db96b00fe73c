// capi.cpp -- the C ABI declared in include/fluidsim.h.
//
// Owns every device buffer of a solver, validates arguments before touching
// the device, converts HIP errors and C++ exceptions into status codes and a
// per-thread message, and sequences the kernels of lbm_pop.hip, lbm_home.hip,
// fields.hip and geom.hip on the solver's stream.
#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <hip/hip_version.h>
#include <link.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <vector>

#include "fluidsim.h"
#include "fs_internal.hpp"
#include "solver_impl.hpp"

namespace fs {
bool read_stl(const char* path, std::vector<float>& xyz);
bool vtk_scalar(const char* path, const float* h, int nx, int ny, int nz);
bool vtk_vector(const char* path, const float* a, const float* b, const float* c, int nx, int ny,
                int nz);
bool vtk_scalar_binary(const char* path, const float* h, int nx, int ny, int nz);
bool vtk_vector_binary(const char* path, const float* a, const float* b, const float* c, int nx,
                       int ny, int nz);
}  // namespace fs

namespace {
thread_local std::string g_err;
}  // namespace

int fs::fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

namespace {
using fs::fail;

// Runs `body` and maps escaping C++ exceptions to status codes.
template <class Fn>
int guarded(Fn&& body) {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return fail(FS_ERR_ALLOC, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(FS_ERR_INVALID, "%s", e.what());
  } catch (...) {
    return fail(FS_ERR_INVALID, "unknown C++ exception");
  }
}

bool fault_injected(const char* what) {
  const char* v = std::getenv("FS_FAULT_INJECT");
  return v && std::strstr(v, what) != nullptr;
}

// Every loaded object whose file name starts with libamdhip64.so (the HIP
// runtime; a process can map two copies under different names, e.g. one
// bundled with another package beside the system one).
int collect_hip_runtime(struct dl_phdr_info* info, size_t, void* data) {
  const char* name = info->dlpi_name;
  if (!name || !*name) return 0;
  const char* base = std::strrchr(name, '/');
  base = base ? base + 1 : name;
  if (std::strncmp(base, "libamdhip64.so", 14) == 0)
    static_cast<std::vector<std::string>*>(data)->push_back(name);
  return 0;
}

}  // namespace

// The library must run on the HIP runtime it was compiled against, and alone:
// kernels, streams and device pointers of one runtime are meaningless to a
// second runtime mapped into the same process.  Checked on every entry point
// that creates device state or accepts a foreign stream.
int fs::runtime_check() {
  std::vector<std::string> rts;
  dl_iterate_phdr(collect_hip_runtime, &rts);
  if (rts.size() > 1) {
    std::string all;
    for (const std::string& r : rts) all += (all.empty() ? "" : ", ") + r;
    return fail(FS_ERR_DEVICE,
                "%zu HIP runtimes are mapped into this process (%s); libfluidsim refuses to run "
                "beside a second runtime (load only one, e.g. do not import a package that "
                "bundles its own HIP runtime into a process using libfluidsim)",
                rts.size(), all.c_str());
  }
  int v = 0;
  if (hipRuntimeGetVersion(&v) != hipSuccess)
    return fail(FS_ERR_DEVICE, "hipRuntimeGetVersion failed");
  if (v / 100000 != HIP_VERSION / 100000) {
    Dl_info di{};
    const char* where = "?";
    if (dladdr(reinterpret_cast<void*>(&hipRuntimeGetVersion), &di) && di.dli_fname)
      where = di.dli_fname;
    return fail(FS_ERR_DEVICE,
                "libfluidsim was built against HIP %d.%d but is bound to HIP runtime %d.%d (%s)",
                HIP_VERSION_MAJOR, HIP_VERSION_MINOR, v / 10000000, (v / 100000) % 100, where);
  }
  return FS_OK;
}

namespace {

int check_solver(const fs_solver* s) {
  if (!s) return fail(FS_ERR_INVALID, "null solver handle");
  return FS_OK;
}

void free_solver(fs_solver* s) {
  if (!s) return;
  if (s->own) (void)hipStreamSynchronize(s->own);
  (void)hipFree(s->F);
  (void)hipFree(s->F2);
  (void)hipFree(s->macro);
  (void)hipFree(s->mbuf[0]);
  (void)hipFree(s->mbuf[1]);
  (void)hipFree(s->flags);
  (void)hipFree(s->nbm);
  (void)hipFree(s->stage);
  (void)hipFree(s->inlet);
  (void)hipFree(s->scratch);
  if (s->own) (void)hipStreamDestroy(s->own);
  delete s;
}


int upload_flags(fs_solver* s) {
  s->outlet_tile_start = s->outlet_tile_start32 = s->outlet_ring_read = false;
  const int nx = s->g.nx, ny = s->g.ny, nz = s->g.nz;
  const uint8_t* fl = s->hflags.data();
  auto at = [&](int x, int y, int z) {
    return fl[(size_t((z + nz) % nz) * ny + size_t((y + ny) % ny)) * nx + size_t(x % nx)];
  };
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y) {
      const uint8_t* row = fl + (size_t(z) * ny + y) * nx;
      for (int x = 0; x < nx; x += 32)
        if (row[x] & FS_FLAG_OUTLET) {
          s->outlet_tile_start32 = true;
          if (x % 64 == 0) s->outlet_tile_start = true;
        }
      for (int x = 63; x < nx + 63; x += 64) {
        const int xo = x < nx ? x : nx - 1;  // x % 64 == 63, then nx-1
        if (!(row[xo] & FS_FLAG_OUTLET)) continue;
        for (int dz = -1; dz <= 1; ++dz)
          for (int dy = -1; dy <= 1; ++dy)
            if (!(at(xo + 1, y + dy, z + dz) & (FS_FLAG_INLET | FS_FLAG_SOLID)))
              s->outlet_ring_read = true;
      }
    }
  s->f2_synced = false;
  FS_HIP(hipMemcpyAsync(s->stage, s->hflags.data(), s->g.n, hipMemcpyHostToDevice, s->stream));
  FS_HIP(fs::launch_derive_flags(s->g, s->stage, s->flags, s->nbm, s->stream));
  FS_HIP(hipStreamSynchronize(s->stream));
  return FS_OK;
}

int validate_flags(const fs::Geo& g, const uint8_t* fl) {
  for (int z = 0; z < g.nz; ++z)
    for (int y = 0; y < g.ny; ++y) {
      const size_t row = (size_t(z) * g.ny + y) * g.nx;
      for (int x = 0; x < g.nx; ++x) {
        const uint8_t f = fl[row + x];
        if ((f & FS_FLAG_INLET) && (f & (FS_FLAG_SOLID | FS_FLAG_OUTLET)))
          return fail(FS_ERR_INVALID, "inlet cell (%d,%d,%d) is also solid or outlet", x, y, z);
        if (f & FS_FLAG_OUTLET) {
          if (f & FS_FLAG_SOLID)
            return fail(FS_ERR_INVALID, "outlet cell (%d,%d,%d) is solid", x, y, z);
          if (x < 1) return fail(FS_ERR_INVALID, "outlet cell (%d,%d,%d) at x=0", x, y, z);
          if (fl[row + x - 1] & (FS_FLAG_SOLID | FS_FLAG_INLET | FS_FLAG_OUTLET))
            return fail(FS_ERR_INVALID,
                        "outlet cell (%d,%d,%d) needs a fluid, non-inlet, non-outlet cell at x-1",
                        x, y, z);
        }
      }
    }
  return FS_OK;
}

int reset_impl(fs_solver* s) {
  if (s->method == fs::kMethodHOME) {
    FS_HIP(fs::launch_home_fill(s->g, s->mbuf[0], s->mbuf[1], nullptr, nullptr, nullptr, nullptr,
                                s->stream));
    s->cur = 0;
  } else {
    FS_HIP(fs::launch_pop_reset(s->g, pop_args(s), s->stream));
    s->parity = 0;
    s->f2_synced = false;
  }
  FS_HIP(hipStreamSynchronize(s->stream));
  return FS_OK;
}

}  // namespace

fs::PopLaunch pop_args(fs_solver* s) {
  fs::PopLaunch a;
  a.F = s->F;
  a.flags = s->flags;
  a.nbm = s->nbm;
  a.rho = s->macro;
  a.ux = s->macro + s->g.P;
  a.uy = s->macro + 2 * s->g.P;
  a.uz = s->macro + 3 * s->g.P;
  a.inlet = s->inlet;
  return a;
}

int solver_reinit_async(fs_solver* s, const float* rho, const float* ux, const float* uy,
                        const float* uz) {
  if (s->method == fs::kMethodHOME) {
    // kern_reinit_from_macro (lbm3d_home.cu:109-137) into both buffers
    FS_HIP(fs::launch_home_fill(s->g, s->mbuf[0], s->mbuf[1], rho, ux, uy, uz, s->stream));
    s->cur = 0;
  } else if (s->method == fs::kMethodBGK) {
    // a null rho means reset() (lbm3d_legacy.cu:174-178)
    if (!rho) {
      FS_HIP(fs::launch_pop_reset(s->g, pop_args(s), s->stream));
    } else {
      FS_HIP(fs::launch_pop_reinit(s->g, pop_args(s), 0, rho, ux, uy, uz, s->stream));
    }
    s->parity = 0;
    s->f2_synced = false;
  } else {
    FS_HIP(fs::launch_pop_reinit(s->g, pop_args(s), 1, rho, ux, uy, uz, s->stream));
    s->parity = 0;
    s->f2_synced = false;
  }
  return FS_OK;
}

static int solver_pop_variant(const fs_solver* s) {
  const int v = s->pop_variant >= 0 ? s->pop_variant : fs::pop_default_variant(s->g);
  if (v == fs::kPopAA) return v;
  const bool blocked = fs::pop_variant_tile_width(v) == 32 ? s->outlet_tile_start32
                                                            : s->outlet_tile_start;
  return blocked ? fs::kPopAA : v;
}

int solver_home_variant(const fs_solver* s) {
  if (s->outlet_tile_start) return fs::kHomePlain;
  const int v = s->home_variant >= 0 ? s->home_variant : fs::home_default_variant(s->g);
  return fs::home_variant_pair(v) && s->outlet_ring_read ? fs::home_single_variant(s->g, v) : v;
}

namespace {

int enqueue_steps(fs_solver* s, int nsteps) {
  if (nsteps < 0) return fail(FS_ERR_INVALID, "nsteps must be >= 0");
  if (s->method == fs::kMethodHOME) {
    const int variant = solver_home_variant(s);
    int i = 0;
    if (fs::home_variant_pair(variant)) {
      for (; i + 2 <= nsteps; i += 2) {
        fs::HomeLaunch a{s->mbuf[s->cur], s->mbuf[1 - s->cur], s->flags, s->nbm, s->inlet};
        FS_HIP(fs::launch_home_pair(s->g, s->ph, a, variant, s->stream));
        s->cur = 1 - s->cur;
      }
    }
    for (; i < nsteps; ++i) {
      fs::HomeLaunch a{s->mbuf[s->cur], s->mbuf[1 - s->cur], s->flags, s->nbm, s->inlet};
      FS_HIP(fs::launch_home_step(s->g, s->ph, a, variant, s->stream));
      s->cur = 1 - s->cur;
    }
  } else {
    const bool hyb = s->method == fs::kMethodHybrid;
    const int variant = solver_pop_variant(s);
    int i = 0;
    while (i < nsteps) {
      if (variant != fs::kPopAA && s->parity == 0 && nsteps - i >= 2) {
        // two steps F -> F2, then swap the buffers; F2's solid slots must
        // hold F's frozen values, which nothing else writes
        if (!s->f2_synced) {
          FS_HIP(hipMemcpyAsync(s->F2, s->F, sizeof(float) * 19 * s->g.P,
                                hipMemcpyDeviceToDevice, s->stream));
          s->f2_synced = true;
        }
        const fs::PopLaunch pa = pop_args(s);
        const fs::PairLaunch a{s->F, s->F2, pa.flags, pa.nbm, pa.rho, pa.ux, pa.uy, pa.uz,
                               pa.inlet};
        FS_HIP(fs::launch_pop_pair(s->g, s->ph, a, hyb, i + 2 == nsteps, variant, s->stream));
        std::swap(s->F, s->F2);
        i += 2;
      } else {
        FS_HIP(fs::launch_pop_step(s->g, s->ph, pop_args(s), s->parity != 0, hyb,
                                   i == nsteps - 1, s->stream));
        s->parity ^= 1;
        i += 1;
      }
    }
  }
  return FS_OK;
}

float* macro_plane(fs_solver* s, int which) {
  if (s->method == fs::kMethodHOME) return s->mom_cur() + size_t(which) * s->g.P;
  return s->macro + size_t(which) * s->g.P;
}

int ensure_scratch(fs_solver* s) {
  if (s->scratch) return FS_OK;
  FS_HIP(hipMalloc((void**)&s->scratch, sizeof(float) * 19 * s->g.P));
  return FS_OK;
}

}  // namespace

extern "C" {

int fs_version(void) { return FS_ABI_VERSION; }

const char* fs_last_error(void) { return g_err.c_str(); }

int fs_device_count(int* count) {
  if (!count) return fail(FS_ERR_INVALID, "null count");
  *count = 0;
  if (int rc = fs::runtime_check()) return rc;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return FS_OK;
}

int fs_hip_versions(int* built_against, int* runtime) {
  if (built_against) *built_against = HIP_VERSION;
  if (runtime) {
    *runtime = 0;
    FS_HIP(hipRuntimeGetVersion(runtime));
  }
  return FS_OK;
}

int fs_solver_create(int method, int nx, int ny, int nz, float tau, float fx, float fy, float fz,
                     fs_solver** out) {
  return guarded([&]() -> int {
    if (!out) return fail(FS_ERR_INVALID, "null output handle");
    *out = nullptr;
    if (method < 0 || method > 2) return fail(FS_ERR_INVALID, "unknown method %d", method);
    if (nx < 2 || ny < 2 || nz < 1)
      return fail(FS_ERR_INVALID, "grid %dx%dx%d too small (nx,ny >= 2, nz >= 1)", nx, ny, nz);
    const unsigned long long n = (unsigned long long)nx * ny * nz;
    if (n >= (1ull << 31)) return fail(FS_ERR_INVALID, "grid of %llu cells exceeds 2^31", n);
    if (!(tau > 0.5f) || !std::isfinite(tau))
      return fail(FS_ERR_INVALID, "tau must be finite and > 0.5");
    if (fault_injected("alloc")) return fail(FS_ERR_ALLOC, "injected allocation failure");
    if (int rc = fs::runtime_check()) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
      return fail(FS_ERR_DEVICE, "no HIP device available");
    fs_solver* s = new fs_solver();
    s->method = method;
    s->g = fs::make_geo(nx, ny, nz);
    s->tau = tau;
    const float omega = 1.0f / tau;
    s->ph = {omega, omega, omega, fx, fy, fz, FS_BOUNCE_HALFWAY};
    s->hflags.assign(s->g.n, 0);
    auto bail = [&](int code) {
      free_solver(s);
      return code;
    };
    const size_t P = s->g.P;
    if (hipStreamCreate(&s->own) != hipSuccess) return bail(fail(FS_ERR_DEVICE, "stream"));
    s->stream = s->own;
    hipError_t e = hipSuccess;
    if (method == fs::kMethodHOME) {
      e = hipMalloc((void**)&s->mbuf[0], sizeof(float) * 10 * P);
      if (e == hipSuccess) e = hipMalloc((void**)&s->mbuf[1], sizeof(float) * 10 * P);
    } else {
      e = hipMalloc((void**)&s->F, sizeof(float) * 19 * P);
      if (e == hipSuccess) e = hipMalloc((void**)&s->F2, sizeof(float) * 19 * P);
      if (e == hipSuccess) e = hipMalloc((void**)&s->macro, sizeof(float) * 4 * P);
    }
    if (e == hipSuccess) e = hipMalloc((void**)&s->nbm, sizeof(uint32_t) * P);
    if (e == hipSuccess) e = hipMalloc((void**)&s->flags, P);
    if (e == hipSuccess) e = hipMalloc((void**)&s->stage, P);
    if (e == hipSuccess) e = hipMalloc((void**)&s->inlet, sizeof(float) * size_t(ny) * nz);
    if (e != hipSuccess)
      return bail(fail(FS_ERR_ALLOC, "device allocation failed: %s", hipGetErrorString(e)));
    if (hipMemsetAsync(s->inlet, 0, sizeof(float) * size_t(ny) * nz, s->stream) != hipSuccess)
      return bail(fail(FS_ERR_DEVICE, "memset"));
    int rc = upload_flags(s);
    if (rc == FS_OK) rc = reset_impl(s);
    if (rc != FS_OK) return bail(rc);
    *out = s;
    return FS_OK;
  });
}

int fs_solver_destroy(fs_solver* s) {
  if (!s) return FS_OK;
  free_solver(s);
  return FS_OK;
}

int fs_solver_dims(const fs_solver* s, int* nx, int* ny, int* nz, size_t* n) {
  if (int rc = check_solver(s)) return rc;
  if (nx) *nx = s->g.nx;
  if (ny) *ny = s->g.ny;
  if (nz) *nz = s->g.nz;
  if (n) *n = s->g.n;
  return FS_OK;
}

int fs_solver_method(const fs_solver* s, int* method) {
  if (int rc = check_solver(s)) return rc;
  if (!method) return fail(FS_ERR_INVALID, "null method");
  *method = s->method;
  return FS_OK;
}

int fs_solver_set_solid_mask(fs_solver* s, const uint8_t* mask) {
  if (int rc = check_solver(s)) return rc;
  if (!mask) return fail(FS_ERR_INVALID, "null mask");
  return guarded([&]() -> int {
    for (size_t i = 0; i < s->g.n; ++i)
      s->hflags[i] = uint8_t((mask[i] ? FS_FLAG_SOLID : 0) | (s->hflags[i] & FS_FLAG_BAND));
    return upload_flags(s);
  });
}

int fs_solver_set_legacy_mapping(fs_solver* s, const uint8_t* is_legacy, const int*) {
  if (int rc = check_solver(s)) return rc;
  if (s->method != fs::kMethodHybrid)
    return fail(FS_ERR_STATE, "setLegacyMapping applies to the Hybrid method only");
  return guarded([&]() -> int {
    for (size_t i = 0; i < s->g.n; ++i) {
      const bool b = is_legacy && is_legacy[i];
      s->hflags[i] = uint8_t((s->hflags[i] & ~FS_FLAG_BAND) | (b ? FS_FLAG_BAND : 0));
    }
    return upload_flags(s);
  });
}

int fs_solver_set_flags(fs_solver* s, const uint8_t* fl) {
  if (int rc = check_solver(s)) return rc;
  if (!fl) return fail(FS_ERR_INVALID, "null flags");
  return guarded([&]() -> int {
    if (int rc = validate_flags(s->g, fl)) return rc;
    for (size_t i = 0; i < s->g.n; ++i) s->hflags[i] = fl[i] & fs::F_PUBLIC;
    return upload_flags(s);
  });
}

int fs_solver_copy_flags(fs_solver* s, uint8_t* out) {
  if (int rc = check_solver(s)) return rc;
  if (!out) return fail(FS_ERR_INVALID, "null output");
  std::memcpy(out, s->hflags.data(), s->g.n);
  return FS_OK;
}

int fs_solver_set_inlet_profile(fs_solver* s, const float* ux) {
  if (int rc = check_solver(s)) return rc;
  const size_t bytes = sizeof(float) * size_t(s->g.ny) * s->g.nz;
  if (ux)
    FS_HIP(hipMemcpyAsync(s->inlet, ux, bytes, hipMemcpyHostToDevice, s->stream));
  else
    FS_HIP(hipMemsetAsync(s->inlet, 0, bytes, s->stream));
  FS_HIP(hipStreamSynchronize(s->stream));
  return FS_OK;
}

int fs_solver_set_force(fs_solver* s, float fx, float fy, float fz) {
  if (int rc = check_solver(s)) return rc;
  s->ph.fx = fx;
  s->ph.fy = fy;
  s->ph.fz = fz;
  return FS_OK;
}

int fs_solver_set_omegas(fs_solver* s, float oL, float oH, float oS) {
  if (int rc = check_solver(s)) return rc;
  if (s->method != fs::kMethodHybrid)
    return fail(FS_ERR_STATE, "relaxation rates apply to the Hybrid method only");
  s->ph.omegaL = oL;
  s->ph.omegaH = oH;
  s->ph.omegaS = oS;
  return FS_OK;
}

int fs_solver_set_bounce_rule(fs_solver* s, int rule) {
  if (int rc = check_solver(s)) return rc;
  if (rule != FS_BOUNCE_HALFWAY && rule != FS_BOUNCE_REFERENCE_TABLE)
    return fail(FS_ERR_INVALID, "unknown bounce rule %d", rule);
  if (rule == FS_BOUNCE_REFERENCE_TABLE && s->method != fs::kMethodHOME)
    return fail(FS_ERR_STATE,
                "the reference pairing table is not a permutation of the populations; only "
                "HOME's gather can use it");
  s->ph.bounce = rule;
  return FS_OK;
}

int fs_solver_reset(fs_solver* s) {
  if (int rc = check_solver(s)) return rc;
  return reset_impl(s);
}

int fs_solver_reinit_from_macro(fs_solver* s, const float* rho, const float* ux, const float* uy,
                                const float* uz) {
  if (int rc = check_solver(s)) return rc;
  if (int rc = solver_reinit_async(s, rho, ux, uy, uz)) return rc;
  FS_HIP(hipStreamSynchronize(s->stream));
  return FS_OK;
}

int fs_solver_step(fs_solver* s, int nsteps) {
  if (int rc = check_solver(s)) return rc;
  if (int rc = enqueue_steps(s, nsteps)) return rc;
  FS_HIP(hipStreamSynchronize(s->stream));
  return FS_OK;
}

int fs_solver_step_async(fs_solver* s, int nsteps) {
  if (int rc = check_solver(s)) return rc;
  return enqueue_steps(s, nsteps);
}

int fs_solver_set_stream(fs_solver* s, void* st) {
  if (int rc = check_solver(s)) return rc;
  if (int rc = fs::runtime_check()) return rc;
  FS_HIP(hipStreamSynchronize(s->stream));
  s->stream = st ? reinterpret_cast<hipStream_t>(st) : s->own;
  return FS_OK;
}

int fs_solver_get_stream(const fs_solver* s, void** st) {
  if (int rc = check_solver(s)) return rc;
  if (!st) return fail(FS_ERR_INVALID, "null output");
  *st = reinterpret_cast<void*>(s->stream);
  return FS_OK;
}

int fs_solver_synchronize(fs_solver* s) {
  if (int rc = check_solver(s)) return rc;
  FS_HIP(hipStreamSynchronize(s->stream));
  return FS_OK;
}

int fs_solver_timed_steps(fs_solver* s, int nsteps, float* elapsed_ms) {
  if (int rc = check_solver(s)) return rc;
  if (!elapsed_ms) return fail(FS_ERR_INVALID, "null output");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  FS_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(FS_ERR_DEVICE, "hipEventCreate");
  }
  int rc = FS_OK;
  hipError_t e = hipEventRecord(e0, s->stream);
  if (e == hipSuccess) rc = enqueue_steps(s, nsteps);
  if (e == hipSuccess && rc == FS_OK) e = hipEventRecord(e1, s->stream);
  if (e == hipSuccess && rc == FS_OK) e = hipEventSynchronize(e1);
  if (e == hipSuccess && rc == FS_OK) e = hipEventElapsedTime(elapsed_ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != FS_OK) return rc;
  if (e != hipSuccess) return fail(FS_ERR_DEVICE, "timed steps: %s", hipGetErrorString(e));
  return FS_OK;
}

int fs_solver_set_variant(fs_solver* s, int variant) {
  if (int rc = check_solver(s)) return rc;
  if (s->method != fs::kMethodHOME) {
    if (variant < -1 || variant >= fs::kPopVariants)
      return fail(FS_ERR_INVALID, "BGK / Hybrid variant %d out of range", variant);
    s->pop_variant = variant;
    return FS_OK;
  }
  if (variant >= 0 && !fs::home_variant_ok(s->g, variant))
    return fail(FS_ERR_INVALID, "HOME variant %d does not apply to this grid", variant);
  s->home_variant = variant;
  return FS_OK;
}

int fs_solver_get_variant(const fs_solver* s, int* variant) {
  if (int rc = check_solver(s)) return rc;
  if (!variant) return fail(FS_ERR_INVALID, "null output");
  *variant = s->method == fs::kMethodHOME ? solver_home_variant(s) : solver_pop_variant(s);
  return FS_OK;
}

int fs_solver_autotune(fs_solver* s, int trial_steps, int* chosen) {
  if (int rc = check_solver(s)) return rc;
  if (trial_steps < 1) return fail(FS_ERR_INVALID, "trial_steps must be >= 1");
  const bool home = s->method == fs::kMethodHOME;
  const int nvar = home ? fs::kHomeVariants : fs::kPopVariants;
  // the tiled kernels need every outlet away from a tile row start
  if (s->outlet_tile_start) {
    if (home)
      s->home_variant = fs::kHomePlain;
    else
      s->pop_variant = fs::kPopAA;
    if (chosen) *chosen = 0;
    return FS_OK;
  }
  // time every applicable variant from the current state, restoring it after
  // each trial (populations and the macro planes they reported), and keep the
  // fastest; population trials run an even number of steps from the natural
  // representation
  const size_t bytes = sizeof(float) * (home ? 10 : 19) * s->g.P;
  const size_t macroBytes = home ? 0 : sizeof(float) * 4 * s->g.P;
  float* saved = nullptr;
  FS_HIP(hipMalloc((void**)&saved, bytes + macroBytes));
  const int cur0 = s->cur, parity0 = s->parity;
  float* const F0 = s->F;
  float* const F20 = s->F2;
  const int steps = home ? trial_steps : 2 * ((trial_steps + 1) / 2);
  int best = -1;
  float best_ms = 0.0f;
  int rc = FS_OK;
  float* const savedMacro = saved + bytes / sizeof(float);
  hipError_t e = hipMemcpyAsync(saved, home ? s->mbuf[cur0] : s->F, bytes,
                                hipMemcpyDeviceToDevice, s->stream);
  if (e == hipSuccess && !home)
    e = hipMemcpyAsync(savedMacro, s->macro, macroBytes, hipMemcpyDeviceToDevice, s->stream);
  auto restore = [&]() {
    if (home) {
      s->cur = cur0;
      return hipMemcpyAsync(s->mbuf[cur0], saved, bytes, hipMemcpyDeviceToDevice, s->stream);
    }
    s->F = F0;
    s->F2 = F20;
    s->parity = parity0;
    // F2's solid slots still equal F's (no kernel writes a solid slot)
    hipError_t r = hipMemcpyAsync(s->F, saved, bytes, hipMemcpyDeviceToDevice, s->stream);
    if (r == hipSuccess)
      r = hipMemcpyAsync(s->macro, savedMacro, macroBytes, hipMemcpyDeviceToDevice, s->stream);
    return r;
  };
  for (int v = 0; v < nvar && e == hipSuccess && rc == FS_OK; ++v) {
    if (home && !fs::home_variant_ok(s->g, v)) continue;
    if (home)
      s->home_variant = v;
    else
      s->pop_variant = v;
    float ms = 0.0f;
    // warm-up (code load, caches); a swapped AA state first takes one step
    rc = enqueue_steps(s, home ? 1 : (s->parity ? 3 : 2));
    if (rc == FS_OK) rc = fs_solver_timed_steps(s, steps, &ms);
    if (rc == FS_OK && (best < 0 || ms < best_ms)) {
      best = v;
      best_ms = ms;
    }
    e = restore();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(saved);
  if (rc != FS_OK) return rc;
  if (e != hipSuccess) return fail(FS_ERR_DEVICE, "autotune: %s", hipGetErrorString(e));
  if (home)
    s->home_variant = best;
  else
    s->pop_variant = best;
  if (chosen) *chosen = best;
  return FS_OK;
}

int fs_solver_macro_ptr(fs_solver* s, int which, float** d_ptr) {
  if (int rc = check_solver(s)) return rc;
  if (which < 0 || which > 3 || !d_ptr) return fail(FS_ERR_INVALID, "bad macro selector");
  *d_ptr = macro_plane(s, which);
  return FS_OK;
}

int fs_solver_copy_macro(fs_solver* s, int which, float* host) {
  if (int rc = check_solver(s)) return rc;
  if (which < 0 || which > 3 || !host) return fail(FS_ERR_INVALID, "bad macro selector");
  FS_HIP(hipMemcpyAsync(host, macro_plane(s, which), sizeof(float) * s->g.n,
                        hipMemcpyDeviceToHost, s->stream));
  FS_HIP(hipStreamSynchronize(s->stream));
  return FS_OK;
}

int fs_solver_state_size(const fs_solver* s, size_t* nfloats) {
  if (int rc = check_solver(s)) return rc;
  if (!nfloats) return fail(FS_ERR_INVALID, "null output");
  *nfloats = size_t(s->method == fs::kMethodHOME ? 10 : 19) * s->g.n;
  return FS_OK;
}

int fs_solver_copy_state(fs_solver* s, float* host) {
  if (int rc = check_solver(s)) return rc;
  if (!host) return fail(FS_ERR_INVALID, "null output");
  const size_t N = s->g.n, P = s->g.P;
  if (s->method == fs::kMethodHOME) {
    for (int k = 0; k < 10; ++k)
      FS_HIP(hipMemcpyAsync(host + size_t(k) * N, s->mom_cur() + size_t(k) * P,
                            sizeof(float) * N, hipMemcpyDeviceToHost, s->stream));
  } else {
    if (int rc = ensure_scratch(s)) return rc;
    FS_HIP(fs::launch_pop_canonical(s->g, s->F, s->flags, s->parity != 0, s->scratch, s->stream));
    for (int q = 0; q < 19; ++q)
      FS_HIP(hipMemcpyAsync(host + size_t(q) * N, s->scratch + size_t(q) * P, sizeof(float) * N,
                            hipMemcpyDeviceToHost, s->stream));
  }
  FS_HIP(hipStreamSynchronize(s->stream));
  return FS_OK;
}

int fs_solver_load_state(fs_solver* s, const float* host) {
  if (int rc = check_solver(s)) return rc;
  if (!host) return fail(FS_ERR_INVALID, "null input");
  const size_t N = s->g.n, P = s->g.P;
  if (s->method == fs::kMethodHOME) {
    for (int k = 0; k < 10; ++k)
      FS_HIP(hipMemcpyAsync(s->mom_cur() + size_t(k) * P, host + size_t(k) * N,
                            sizeof(float) * N, hipMemcpyHostToDevice, s->stream));
  } else {
    for (int q = 0; q < 19; ++q)
      FS_HIP(hipMemcpyAsync(s->F + size_t(q) * P, host + size_t(q) * N, sizeof(float) * N,
                            hipMemcpyHostToDevice, s->stream));
    s->parity = 0;
    s->f2_synced = false;
  }
  FS_HIP(hipStreamSynchronize(s->stream));
  return FS_OK;
}

// ------------------------------------------------------------------ geometry

int fs_load_stl(const char* path, float* tri_xyz, int capacity, int* ntri, float* bbmin3,
                float* bbmax3) {
  if (!path) return fail(FS_ERR_INVALID, "null path");
  return guarded([&]() -> int {
    std::vector<float> xyz;
    if (!fs::read_stl(path, xyz)) return fail(FS_ERR_IO, "cannot read STL file %s", path);
    const int nt = int(xyz.size() / 9);
    if (ntri) *ntri = nt;
    float mn[3], mx[3];
    for (int k = 0; k < 3; ++k) {
      mn[k] = std::numeric_limits<float>::max();
      mx[k] = -std::numeric_limits<float>::max();
    }
    for (size_t i = 0; i + 2 < xyz.size(); i += 3)
      for (int k = 0; k < 3; ++k) {
        // std::min / std::max argument order of update_bb (stl_loader.cpp:19-22)
        mn[k] = (xyz[i + k] < mn[k]) ? xyz[i + k] : mn[k];
        mx[k] = (mx[k] < xyz[i + k]) ? xyz[i + k] : mx[k];
      }
    if (bbmin3) std::memcpy(bbmin3, mn, sizeof(mn));
    if (bbmax3) std::memcpy(bbmax3, mx, sizeof(mx));
    if (tri_xyz) {
      if (capacity < nt) return fail(FS_ERR_INVALID, "capacity %d < %d triangles", capacity, nt);
      std::memcpy(tri_xyz, xyz.data(), xyz.size() * sizeof(float));
    }
    return FS_OK;
  });
}

int fs_voxelize(const float* tri_xyz, int ntri, const float* bbmin3, const float* bbmax3, int nx,
                int ny, int nz, float uniform_scale, float tx, float ty, float tz, int refine,
                float threshold, uint8_t* mask_out) {
  if (nx < 1 || ny < 1 || nz < 1 || !mask_out || !bbmin3 || !bbmax3 || ntri < 0 ||
      (ntri > 0 && !tri_xyz))
    return fail(FS_ERR_INVALID, "bad voxelize arguments");
  if (int rc = fs::runtime_check()) return rc;
  return guarded([&]() -> int {
    const int r = refine < 1 ? 1 : refine;
    const size_t nf = size_t(nx) * r * size_t(ny) * r * size_t(nz) * r;
    if (size_t(nx) * r >= (1u << 31) || nf >= (size_t(1) << 31))
      return fail(FS_ERR_INVALID, "refined grid of %zu cells exceeds 2^31", nf);
    // compute_fit_transform + make_transformed_tris (voxelize.cpp:83-113)
    float bb[3];
    for (int k = 0; k < 3; ++k) bb[k] = bbmax3[k] - bbmin3[k];
    const float maxDim = std::max(bb[0], std::max(bb[1], bb[2]));
    const float sc = (maxDim > 0.0f) ? (uniform_scale / maxDim) : 1.0f;
    const float tr[3] = {tx, ty, tz};
    float trans[3];
    for (int k = 0; k < 3; ++k) {
      const float center = 0.5f * (bbmin3[k] + bbmax3[k]);
      trans[k] = tr[k] - sc * center;
    }
    std::vector<fs::TriF> tris(static_cast<size_t>(ntri));
    for (int t = 0; t < ntri; ++t)
      for (int j = 0; j < 9; ++j) tris[t].v[j] = sc * tri_xyz[size_t(t) * 9 + j] + trans[j % 3];
    // device scratch and pinned staging kept between calls; the solid mask
    // crosses PCIe as packed bits
    std::unique_lock<std::mutex> lock;
    fs::GeomScratch* gs = fs::geom_scratch(lock);
    if (!gs) return fail(FS_ERR_DEVICE, "no current HIP device");
    const size_t n = size_t(nx) * ny * nz, nw = (n + 31) / 32;
    void *d_tris = nullptr, *d_fine = nullptr, *d_vis = nullptr, *d_chg = nullptr,
         *d_words = nullptr, *d_out = nullptr, *h_words = nullptr, *h_chg = nullptr;
    hipError_t e = gs->device(0, sizeof(fs::TriF) * std::max(1, ntri), &d_tris);
    if (e == hipSuccess) e = gs->device(1, nf, &d_fine);
    if (e == hipSuccess) e = gs->device(2, nf, &d_vis);
    if (e == hipSuccess) e = gs->device(3, sizeof(int), &d_chg);
    if (e == hipSuccess) e = gs->device(4, sizeof(uint32_t) * nw, &d_words);
    if (e == hipSuccess && r > 1) e = gs->device(5, n, &d_out);
    if (e == hipSuccess) e = gs->pinned(0, sizeof(uint32_t) * nw, &h_words);
    if (e == hipSuccess) e = gs->pinned(1, sizeof(int), &h_chg);
    if (e != hipSuccess)
      return fail(FS_ERR_ALLOC, "voxelize scratch allocation failed: %s", hipGetErrorString(e));
    hipStream_t st = nullptr;
    if (ntri > 0)
      e = hipMemcpyAsync(d_tris, tris.data(), sizeof(fs::TriF) * ntri, hipMemcpyHostToDevice, st);
    uint32_t* words = static_cast<uint32_t*>(d_words);
    if (e == hipSuccess)
      e = fs::voxelize_device(static_cast<fs::TriF*>(d_tris), ntri, nx * r, ny * r, nz * r,
                              static_cast<uint8_t*>(d_fine), static_cast<uint8_t*>(d_vis),
                              static_cast<int*>(d_chg), static_cast<int*>(h_chg),
                              r > 1 ? nullptr : words, st);
    if (e == hipSuccess && r > 1) {
      e = fs::downsample_device(static_cast<uint8_t*>(d_fine), r, nx, ny, nz, threshold,
                                static_cast<uint8_t*>(d_out), st);
      if (e == hipSuccess) e = fs::pack_bits_device(static_cast<uint8_t*>(d_out), n, words, st);
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(h_words, words, sizeof(uint32_t) * nw, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(FS_ERR_DEVICE, "voxelize: %s", hipGetErrorString(e));
    fs::unpack_bits_host(static_cast<const uint32_t*>(h_words), n, mask_out);
    return FS_OK;
  });
}

int fs_band(const uint8_t* solid, int nx, int ny, int nz, int d0, uint8_t* band_out) {
  if (!solid || !band_out || nx < 1 || ny < 1 || nz < 1)
    return fail(FS_ERR_INVALID, "bad band arguments");
  const size_t n = size_t(nx) * ny * nz;
  if (n >= (size_t(1) << 31)) return fail(FS_ERR_INVALID, "grid of %zu cells exceeds 2^31", n);
  if (int rc = fs::runtime_check()) return rc;
  return guarded([&]() -> int {
    std::unique_lock<std::mutex> lock;
    fs::GeomScratch* gs = fs::geom_scratch(lock);
    if (!gs) return fail(FS_ERR_DEVICE, "no current HIP device");
    const int dd = d0 < 0 ? 0 : d0;
    const size_t nw = (n + 31) / 32;
    void *d_solid = nullptr, *d_band = nullptr, *d_scr = nullptr, *d_words = nullptr,
         *h_words = nullptr;
    hipError_t e = gs->device(1, n, &d_solid);
    if (e == hipSuccess) e = gs->device(7, n, &d_band);
    if (e == hipSuccess) e = gs->device(6, fs::band_scratch_bytes(nx, ny, nz, dd), &d_scr);
    if (e == hipSuccess) e = gs->device(4, sizeof(uint32_t) * nw, &d_words);
    if (e == hipSuccess) e = gs->pinned(0, sizeof(uint32_t) * nw, &h_words);
    if (e != hipSuccess)
      return fail(FS_ERR_ALLOC, "band scratch allocation failed: %s", hipGetErrorString(e));
    uint32_t* hw = static_cast<uint32_t*>(h_words);
    uint32_t* words = static_cast<uint32_t*>(d_words);
    hipStream_t st = nullptr;
    // the solid mask and the band cross PCIe as packed bits
    fs::pack_bits_host(solid, n, hw);
    e = hipMemcpyAsync(words, hw, sizeof(uint32_t) * nw, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = fs::unpack_bits_device(words, n, static_cast<uint8_t*>(d_solid), st);
    if (e == hipSuccess)
      e = fs::band_device(static_cast<uint8_t*>(d_solid), nx, ny, nz, dd,
                          static_cast<uint8_t*>(d_band), static_cast<uint8_t*>(d_scr), st);
    if (e == hipSuccess) e = fs::pack_bits_device(static_cast<uint8_t*>(d_band), n, words, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(hw, words, sizeof(uint32_t) * nw, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(FS_ERR_DEVICE, "band: %s", hipGetErrorString(e));
    fs::unpack_bits_host(hw, n, band_out);
    return FS_OK;
  });
}

// ------------------------------------------------------------------ fields

int fs_speed(const float* u, const float* v, const float* w, float* out, size_t n) {
  if (n && (!u || !v || !w || !out)) return fail(FS_ERR_INVALID, "null field");
  FS_HIP(fs::launch_speed_k(u, v, w, out, n, nullptr));
  return FS_OK;
}

int fs_diff(const float* a, const float* b, float* out, size_t n) {
  if (n && (!a || !b || !out)) return fail(FS_ERR_INVALID, "null field");
  FS_HIP(fs::launch_diff_k(a, b, out, n, nullptr));
  return FS_OK;
}

int fs_shift(const float* a, float bias, float* out, size_t n) {
  if (n && (!a || !out)) return fail(FS_ERR_INVALID, "null field");
  FS_HIP(fs::launch_shift_k(a, bias, out, n, nullptr));
  return FS_OK;
}

int fs_rho_gauss(float* rho, int nx, int ny, int nz, float cx, float cy, float cz, float sigma,
                 float amp) {
  if (!rho || nx < 1 || ny < 1 || nz < 1) return fail(FS_ERR_INVALID, "bad gauss arguments");
  FS_HIP(fs::launch_rho_gauss(rho, nx, ny, nz, cx, cy, cz, sigma, amp, nullptr));
  return FS_OK;
}

int fs_rho_slab(float* rho, int nx, int ny, int nz, int axis, float amp, float center,
                float width) {
  if (!rho || nx < 1 || ny < 1 || nz < 1) return fail(FS_ERR_INVALID, "bad slab arguments");
  FS_HIP(fs::launch_rho_slab(rho, nx, ny, nz, axis, amp, center, width, nullptr));
  return FS_OK;
}

int fs_reinit_equilibrium(const float* rho, const float* ux, const float* uy, const float* uz,
                          float* f, size_t n) {
  if (n && (!rho || !f)) return fail(FS_ERR_INVALID, "null rho or f");
  FS_HIP(fs::launch_reinit_cs2(rho, ux, uy, uz, f, n, n, nullptr));
  return FS_OK;
}

namespace {

int write_scalar(const char* path, const float* d_field, int nx, int ny, int nz, bool binary) {
  if (!path || !d_field || nx < 1 || ny < 1 || nz < 1)
    return fail(FS_ERR_INVALID, "bad VTK arguments");
  return guarded([&]() -> int {
    const size_t n = size_t(nx) * ny * nz;
    // device -> pinned staging (kept between calls), ordered after earlier
    // work as the reference's cudaMemcpy (moment_vtk_writer.cu:12)
    std::unique_lock<std::mutex> lock;
    fs::GeomScratch* gs = fs::geom_scratch(lock);
    if (!gs) return fail(FS_ERR_DEVICE, "no current HIP device");
    void* stage = nullptr;
    FS_HIP(gs->pinned(2, n * sizeof(float), &stage));
    float* h = static_cast<float*>(stage);
    FS_HIP(hipMemcpyAsync(h, d_field, n * sizeof(float), hipMemcpyDeviceToHost, nullptr));
    FS_HIP(hipStreamSynchronize(nullptr));
    const bool ok = binary ? fs::vtk_scalar_binary(path, h, nx, ny, nz)
                           : fs::vtk_scalar(path, h, nx, ny, nz);
    if (!ok) return fail(FS_ERR_IO, "failed to write VTK file %s", path);
    return FS_OK;
  });
}

int write_vector(const char* path, const float* a, const float* b, const float* c, int nx, int ny,
                 int nz, bool binary) {
  if (!path || !a || !b || !c || nx < 1 || ny < 1 || nz < 1)
    return fail(FS_ERR_INVALID, "bad VTK arguments");
  return guarded([&]() -> int {
    const size_t n = size_t(nx) * ny * nz;
    // the three components -> pinned staging, three asynchronous copies and
    // one wait (moment_vtk_writer.cu:33-35 copies them one after the other)
    std::unique_lock<std::mutex> lock;
    fs::GeomScratch* gs = fs::geom_scratch(lock);
    if (!gs) return fail(FS_ERR_DEVICE, "no current HIP device");
    void* stage = nullptr;
    FS_HIP(gs->pinned(2, 3 * n * sizeof(float), &stage));
    float* h = static_cast<float*>(stage);
    FS_HIP(hipMemcpyAsync(h, a, n * sizeof(float), hipMemcpyDeviceToHost, nullptr));
    FS_HIP(hipMemcpyAsync(h + n, b, n * sizeof(float), hipMemcpyDeviceToHost, nullptr));
    FS_HIP(hipMemcpyAsync(h + 2 * n, c, n * sizeof(float), hipMemcpyDeviceToHost, nullptr));
    FS_HIP(hipStreamSynchronize(nullptr));
    const float *x = h, *y = h + n, *z = h + 2 * n;
    const bool ok = binary ? fs::vtk_vector_binary(path, x, y, z, nx, ny, nz)
                           : fs::vtk_vector(path, x, y, z, nx, ny, nz);
    if (!ok) return fail(FS_ERR_IO, "failed to write VTK file %s", path);
    return FS_OK;
  });
}

}  // namespace

int fs_write_scalar_vtk(const char* path, const float* d_field, int nx, int ny, int nz) {
  return write_scalar(path, d_field, nx, ny, nz, false);
}

int fs_write_vector_vtk(const char* path, const float* a, const float* b, const float* c, int nx,
                        int ny, int nz) {
  return write_vector(path, a, b, c, nx, ny, nz, false);
}

int fs_write_scalar_vtk_binary(const char* path, const float* d_field, int nx, int ny, int nz) {
  return write_scalar(path, d_field, nx, ny, nz, true);
}

int fs_write_vector_vtk_binary(const char* path, const float* a, const float* b, const float* c,
                               int nx, int ny, int nz) {
  return write_vector(path, a, b, c, nx, ny, nz, true);
}

int fs_device_malloc(void** p, size_t bytes) {
  if (!p) return fail(FS_ERR_INVALID, "null output");
  *p = nullptr;
  if (fault_injected("alloc")) return fail(FS_ERR_ALLOC, "injected allocation failure");
  if (int rc = fs::runtime_check()) return rc;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(FS_ERR_ALLOC, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  return FS_OK;
}
int fs_device_free(void* p) {
  FS_HIP(hipFree(p));
  return FS_OK;
}
int fs_memcpy_htod(void* d, const void* h, size_t bytes) {
  FS_HIP(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
  return FS_OK;
}
int fs_memcpy_dtoh(void* h, const void* d, size_t bytes) {
  FS_HIP(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost));
  return FS_OK;
}
int fs_memset_device(void* d, int value, size_t bytes) {
  FS_HIP(hipMemset(d, value, bytes));
  return FS_OK;
}
int fs_device_synchronize(void) {
  FS_HIP(hipDeviceSynchronize());
  return FS_OK;
}

int fs_copy_probe(void* dst, const void* src, size_t bytes, int iters, float* elapsed_ms) {
  if (!dst || !src || !elapsed_ms || iters < 1 || (bytes % 16) != 0 ||
      (reinterpret_cast<uintptr_t>(dst) % 16) != 0 || (reinterpret_cast<uintptr_t>(src) % 16) != 0)
    return fail(FS_ERR_INVALID, "bad copy probe arguments");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  FS_HIP(hipEventCreate(&e0));
  FS_HIP(hipEventCreate(&e1));
  hipError_t e = hipEventRecord(e0, nullptr);
  for (int i = 0; i < iters && e == hipSuccess; ++i) e = fs::launch_copy_probe(dst, src, bytes, nullptr);
  if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(elapsed_ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return fail(FS_ERR_DEVICE, "copy probe: %s", hipGetErrorString(e));
  return FS_OK;
}

// ------------------------------------------------- reference-named helpers

static void die_on(int rc, const char* what) {
  if (rc == FS_OK) return;
  std::fprintf(stderr, "[error] %s: %s\n", what, g_err.c_str());
  std::abort();
}

void write_scalar_vtk(const char* filename, const float* d_field, int Nx, int Ny, int Nz) {
  die_on(fs_write_scalar_vtk(filename, d_field, Nx, Ny, Nz), "write_scalar_vtk");
}
void write_vector_vtk(const char* filename, const float* d_fx, const float* d_fy,
                      const float* d_fz, int Nx, int Ny, int Nz) {
  die_on(fs_write_vector_vtk(filename, d_fx, d_fy, d_fz, Nx, Ny, Nz), "write_vector_vtk");
}
void make_rho_gauss(float* d_rho, int Nx, int Ny, int Nz, float cx, float cy, float cz,
                    float sigma, float amp) {
  die_on(fs_rho_gauss(d_rho, Nx, Ny, Nz, cx, cy, cz, sigma, amp), "make_rho_gauss");
}
void make_rho_slab(float* d_rho, int Nx, int Ny, int Nz, int axis, float amp, float center,
                   float width) {
  die_on(fs_rho_slab(d_rho, Nx, Ny, Nz, axis, amp, center, width), "make_rho_slab");
}
void reinit_equilibrium_from_macro(const float* d_rho, const float* d_ux, const float* d_uy,
                                   const float* d_uz, float* d_f, int N) {
  die_on(fs_reinit_equilibrium(d_rho, d_ux, d_uy, d_uz, d_f, N < 0 ? 0 : size_t(N)),
         "reinit_equilibrium_from_macro");
}
void launch_speed(const float* u, const float* v, const float* w, float* out, int N) {
  die_on(fs_speed(u, v, w, out, N < 0 ? 0 : size_t(N)), "launch_speed");
}
void launch_diff(const float* a, const float* b, float* out, int N) {
  die_on(fs_diff(a, b, out, N < 0 ? 0 : size_t(N)), "launch_diff");
}
void launch_shift(const float* a, float bias, float* out, int N) {
  die_on(fs_shift(a, bias, out, N < 0 ? 0 : size_t(N)), "launch_shift");
}

}  // extern "C"
