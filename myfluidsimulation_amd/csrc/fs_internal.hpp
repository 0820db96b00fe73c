// fs_internal.hpp -- definitions shared by the HIP kernels and the C ABI layer.
//
// Data layout in HBM (one solver):
//   populations  F[q*P + c]   19 planes, two buffers: the two-step kernel reads
//                             one and writes the other, the one-step AA kernel
//                             updates one in place
//   moments      m[k*P + c]   10 planes, two buffers (HOME ping-pong)
//   macros       rho/ux/uy/uz separate planes of P floats (BGK / Hybrid)
//   flags        uint8 per cell
// c = (z*ny + y)*nx + x (x fastest, as lbm3d_legacy.cu:33-35); P is the cell
// count rounded up to 64 floats so every plane starts on a 256-byte boundary.
// Offsets are 64-bit: 19*P exceeds 2^31 at 512^3.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <mutex>

namespace fs {

// D3Q19 velocity set in the reference order (lbm3d_legacy.cu:16-32).
constexpr int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
// Opposite direction, e_OPP[q] = -e_q: the halfway bounce-back partner and
// the AA storage partner.
constexpr int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
// The reference's own table (lbm3d_legacy.cu:29-32, lbm3d_home.cu:27-30,
// lbm3d_hybrid.cu:28-31).  For q >= 7 it pairs mirror images (one velocity
// component flipped, e.g. 7=(+1,+1,0) with 8=(-1,+1,0)), not opposites.
// Only HOME's bounce-back can use it (FS_BOUNCE_REFERENCE_TABLE): there the
// update is a gather and stays deterministic.
constexpr int OPP_REF[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

// Flag byte: the low four bits are the public FS_FLAG_* bits; the upper bits
// are derived on the device after every flag upload.
constexpr uint8_t F_SOLID = 0x01;
constexpr uint8_t F_BAND = 0x02;
constexpr uint8_t F_INLET = 0x04;
constexpr uint8_t F_OUTLET = 0x08;
constexpr uint8_t F_NEAR = 0x10;    // one of the 18 periodic neighbours is solid
constexpr uint8_t F_FEEDER = 0x20;  // cell x+1 is an outlet: emit for it too
constexpr uint8_t F_BODY = 0x40;    // adjoint flags only: a solid whose links count for the drag
constexpr uint8_t F_PUBLIC = 0x0F;

constexpr int kMethodBGK = 0;
constexpr int kMethodHOME = 1;
constexpr int kMethodHybrid = 2;

// Plane stride: the cell count rounded up to 64 floats (every plane starts on
// a 256-byte boundary) plus 320 floats, so that planes of power-of-two grids
// (512^3: 2^29 bytes) do not all start on the same cache set and channel.
// Measured on the 512^3 vortex: BGK 51.9 -> 53.8 GLUPS, HOME +0.5-1%; the
// 768x256x256 channel +2% for both (strides of 64 and 1280 floats alike).
inline size_t padded_cells(size_t n) { return ((n + 63) & ~size_t(63)) + 320; }

// Unsigned 32-bit division by a runtime-invariant divisor with a multiply
// and shifts (Granlund & Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, round-up variant with the add fix-up).
struct FastDiv {
  uint32_t d, m, s1, s2;
  static FastDiv make(uint32_t d) {
    FastDiv f;
    f.d = d;
    uint32_t l = 0;
    while ((uint64_t(1) << l) < d) ++l;
    f.m = uint32_t(((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1);
    f.s1 = l < 1 ? l : 1;
    f.s2 = l > 0 ? l - 1 : 0;
    return f;
  }
  __host__ __device__ inline uint32_t div(uint32_t n) const {
#ifdef __HIP_DEVICE_COMPILE__
    uint32_t t = __umulhi(m, n);
#else
    uint32_t t = uint32_t((uint64_t(m) * n) >> 32);
#endif
    return (t + ((n - t) >> s1)) >> s2;
  }
};

// Grid geometry passed by value to every lattice kernel.
struct Geo {
  int nx, ny, nz;
  uint32_t nxy;   // nx*ny
  uint32_t n;     // cells
  size_t P;       // plane stride in floats
  FastDiv divx;   // c_plane / nx
};

inline Geo make_geo(int nx, int ny, int nz) {
  Geo g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.nxy = uint32_t(nx) * uint32_t(ny);
  g.n = g.nxy * uint32_t(nz);
  g.P = padded_cells(g.n);
  g.divx = FastDiv::make(uint32_t(nx));
  return g;
}

// Relaxation and forcing parameters of one solver.
struct Phys {
  float omegaL, omegaH, omegaS;  // BGK / hybrid-bulk rates (all 1/tau by default)
  float fx, fy, fz;              // velocity shift u_eq = u + F
  int bounce;                    // 0: opposite direction, 1: reference table (HOME)
};

// ---- launchers (defined in the .hip files; all enqueue on `st`) ----
struct PopLaunch {
  float* F;
  const uint8_t* flags;
  const uint32_t* nbm;  // bit q: the periodic neighbour c + e_q is solid
  float *rho, *ux, *uy, *uz;
  const float* inlet;  // ny*nz inlet x-velocity
};
hipError_t launch_pop_step(const Geo& g, const Phys& ph, const PopLaunch& a, bool odd, bool hybrid,
                           bool write_macro, hipStream_t st);
hipError_t launch_pop_reset(const Geo& g, const PopLaunch& a, hipStream_t st);
// mode 0: Legacy cs2 form, macros copied (null u -> 0)
// mode 1: Hybrid feq19 form, macros copied (null rho -> 1, null u -> 0)
hipError_t launch_pop_reinit(const Geo& g, const PopLaunch& a, int mode, const float* rho,
                             const float* ux, const float* uy, const float* uz, hipStream_t st);
// populations in the swapped (odd) AA representation -> reference order
hipError_t launch_pop_canonical(const Geo& g, const float* F, const uint8_t* flags, bool swapped,
                                float* out, hipStream_t st);
// derived flag bytes; nbm (nullable): per cell, bit q set when the periodic
// neighbour c + e_q is solid.  An outlet's derived F_BAND is its upstream
// neighbour's (it collides as that cell).
hipError_t launch_derive_flags(const Geo& g, const uint8_t* in, uint8_t* out, uint32_t* nbm,
                               hipStream_t st);

// Two BGK / Hybrid steps per launch (lbm_pair.hip): Fin holds f^t in the
// reference order, Fout receives f^{t+2}; the solid cells' slots of Fout
// must already equal Fin's (the kernel never writes them).
struct PairLaunch {
  const float* Fin;
  float* Fout;
  const uint8_t* flags;
  const uint32_t* nbm;
  float *rho, *ux, *uy, *uz;
  const float* inlet;
};
// BGK / Hybrid kernel variants: 0 the one-step AA kernel; 1..5 two steps per
// launch on TX x TY column tiles of ZC planes: 64 x (4,16), (8,16), (8,32),
// (4,32), and 32 x (8,16)
constexpr int kPopAA = 0, kPairT4Z16 = 1, kPairT8Z16 = 2, kPairT8Z32 = 3, kPairT4Z32 = 4,
              kPairX32T8Z16 = 5, kPairT12Z16 = 6, kPairT12Z32 = 7, kPairT12Z64 = 8, kPopVariants = 9;
int pop_default_variant(const Geo& g);
// tile row width of a two-step variant: an outlet may not sit at a multiple
int pop_variant_tile_width(int variant);
// variant in 1..kPopVariants-1 (others: the default)
hipError_t launch_pop_pair(const Geo& g, const Phys& ph, const PairLaunch& a, bool hybrid,
                           bool write_macro, int variant, hipStream_t st);

struct HomeLaunch {
  const float* m;
  float* mn;
  const uint8_t* flags;
  const uint32_t* nbm;  // bit q: the periodic neighbour c + e_q is solid
  const float* inlet;
};
// HOME kernel variants: the plain per-cell gather, three tile shapes (64 x TY
// columns, ZC-plane z-chunks) of the first tiled form and the same three of
// the second (home_tile2), three of the third (home_tile3), and five of the
// two-step tile (home_pair_tile: two steps per launch, 64 x TY x ZC)
constexpr int kHomePlain = 0, kHomeT4Z16 = 1, kHomeT8Z16 = 2, kHomeT8Z32 = 3, kHome2T4Z16 = 4,
              kHome2T8Z16 = 5, kHome2T8Z32 = 6, kHome2T16Z32 = 7, kHome2T8Z64 = 8,
              kHome3T4Z16 = 9, kHome3T8Z32 = 10, kHome3T8Z64 = 11, kHomeP11Z64 = 12,
              kHomeP11Z32 = 13, kHomeP10Z64 = 14, kHomeP12Z64 = 15, kHomeP12Z32 = 16,
              kHomeVariants = 17;
bool home_variant_ok(const Geo& g, int variant);
int home_default_variant(const Geo& g);
// two steps per launch
inline bool home_variant_pair(int variant) { return variant >= kHomeP11Z64 && variant < kHomeVariants; }
// the one-step variant that takes a single step for a two-step variant
int home_single_variant(const Geo& g, int variant);
// variant < 0: the default for this grid; a two-step variant takes one step
// with its one-step counterpart
hipError_t launch_home_step(const Geo& g, const Phys& ph, const HomeLaunch& a, int variant,
                            hipStream_t st);
// two steps m -> mn of a two-step variant
hipError_t launch_home_pair(const Geo& g, const Phys& ph, const HomeLaunch& a, int variant,
                            hipStream_t st);
hipError_t launch_home_fill(const Geo& g, float* m, float* mn, const float* rho, const float* ux,
                            const float* uy, const float* uz, hipStream_t st);

// fields
hipError_t launch_speed_k(const float* u, const float* v, const float* w, float* out, size_t n,
                          hipStream_t st);
hipError_t launch_diff_k(const float* a, const float* b, float* out, size_t n, hipStream_t st);
hipError_t launch_shift_k(const float* a, float bias, float* out, size_t n, hipStream_t st);
hipError_t launch_rho_gauss(float* rho, int nx, int ny, int nz, float cx, float cy, float cz,
                            float sigma, float amp, hipStream_t st);
hipError_t launch_rho_slab(float* rho, int nx, int ny, int nz, int axis, float amp, float center,
                           float width, hipStream_t st);
hipError_t launch_reinit_cs2(const float* rho, const float* ux, const float* uy, const float* uz,
                             float* f, size_t n, size_t stride, hipStream_t st);

// batched domains (batch.hip): descriptor table entries
struct BatchPopDomain {
  Geo g;
  Phys ph;
  PopLaunch a;             // a.F: population buffer 0
  float* F1;               // population buffer 1
  int hybrid;
  uint32_t tile0;          // first one-step work-list tile of this domain
  uint32_t tilesPerPlane;  // 256-cell tiles per z-plane
  uint32_t pairTile0;      // first two-step work-list tile
  int pairTilesX, pairTilesY, pairChunks;
  int pairWide;  // two-step tiles 64 x 4 (nx a multiple of 64) instead of 32 x 8
};
struct BatchHomeDomain {
  Geo g;
  Phys ph;
  float* m0;
  float* m1;
  const uint8_t* flags;
  const uint32_t* nbm;
  const float* inlet;
  uint32_t tile0;
  int tilesX, chunks;
  int wide;  // 64 x 4 tiles (nx a multiple of 64) instead of 32 x 8
};
bool batch_home_wide(const Geo& g);
void batch_home_tiling(const Geo& g, int& tilesX, int& tilesY, int& chunks);
int batch_home_tile_width();
// two-step tile shape of a population domain: 64 x 4 when nx is a multiple
// of 64, else 32 x 8 (both 256 threads)
bool batch_pair_wide(const Geo& g);
void batch_pair_tiling(const Geo& g, int& tilesX, int& tilesY, int& chunks);
int batch_pair_tile_width();
// blocks <= 0: one workgroup per tile; otherwise a persistent grid of that size.
// cur: the population buffer (0 or 1) holding the state.
hipError_t launch_batch_pop(const BatchPopDomain* doms, const uint16_t* tile_dom, uint32_t ntiles,
                            int cur, bool odd, bool macro, int blocks, hipStream_t st);
// two steps of every population domain: buffer cur -> buffer 1-cur
hipError_t launch_batch_pair(const BatchPopDomain* doms, const uint16_t* tile_dom,
                             uint32_t ntiles, int cur, bool macro, int blocks, hipStream_t st);
hipError_t launch_batch_home(const BatchHomeDomain* doms, const uint16_t* tile_dom,
                             uint32_t ntiles, int cur, int blocks, hipStream_t st);

// ---- adjoint (adjoint.hip) ----
// Backward through one population step n (BGK / Hybrid), canonical layout:
// lin = dL/df^{n+1}, lout = dL/df^n (19 planes each).
struct AdjPopLaunch {
  const float* lin;
  float* lout;
  const float *rho, *ux, *uy, *uz;      // the step's collision moments (u with the shift F)
  const uint8_t* flags;                 // derived flags with F_BODY
  const float* inlet;                   // ny*nz inlet x-velocity
  const float *grho, *gux, *guy, *guz;  // dL/d(macros written by this step), or null
  const float* gdrag;                   // dL/dD^n (one device float), or null
  float* ginlet;                        // dL/d inlet accumulator (ny*nz), or null
};
hipError_t launch_adj_pop(const Geo& g, const Phys& ph, const AdjPopLaunch& a, bool hybrid,
                          hipStream_t st);
// dL/d(rho0, u0) of the equilibrium initial populations from dL/df^0
hipError_t launch_adj_pop_init(const Geo& g, const float* lam, const float* rho0, const float* ux0,
                               const float* uy0, const float* uz0, float* grho, float* gux,
                               float* guy, float* guz, hipStream_t st);
// Backward through one HOME step n: lin = dL/dR^{n+1} (10 planes: raw moments
// r, j, M of the gathered populations), m = m^n; writes lout = dL/dR^n, or,
// when final (n = 0), dL/dm^0 components rho, u into grho..guz.
struct AdjHomeLaunch {
  const float* lin;
  float* lout;
  const float* m;
  const uint8_t* flags;
  const float* gdrag;
  float* ginlet;
  float *grho, *gux, *guy, *guz;
};
hipError_t launch_adj_home(const Geo& g, const Phys& ph, const AdjHomeLaunch& a, bool final_step,
                           hipStream_t st);
// dL/dR^T from dL/d(end moments rho, u) (nullable planes) and m^T
hipError_t launch_adj_home_seed(const Geo& g, const Phys& ph, const float* m, const uint8_t* flags,
                                const float* grho, const float* gux, const float* guy,
                                const float* guz, float* lout, float* ginlet, hipStream_t st);
// x-momentum handed to body cells in one step, summed over the link list
// (cell, direction q with e_qx != 0, neighbour along q a body cell):
// populations: 2 e_qx f*_q read from the post-step AA slot F[OPP q][cell];
// moments: 2 e_qx f*_q rebuilt from m^n.  Deterministic two-pass sum.
struct DragLink {
  uint32_t cell;
  uint32_t q;
};
hipError_t launch_drag_pop(const Geo& g, const float* F, const DragLink* links, uint32_t nlinks,
                           float* partials, int nblocks, float* out, hipStream_t st);
hipError_t launch_drag_home(const Geo& g, const Phys& ph, const float* m, const DragLink* links,
                            uint32_t nlinks, float* partials, int nblocks, float* out,
                            hipStream_t st);
hipError_t launch_fill(float* p, float v, size_t n, hipStream_t st);

// error reporting shared by the C ABI sources (capi.cpp, batch_api.cpp)
int fail(int code, const char* fmt, ...);
// FS_ERR_DEVICE (with a message) unless exactly one HIP runtime is mapped
// into the process and its major.minor version is the one compiled against
int runtime_check();

// plain streaming copy (bandwidth ceiling probe)
hipError_t launch_copy_probe(void* dst, const void* src, size_t bytes, hipStream_t st);

// geometry
struct TriF {
  float v[9];
};
// surface + exterior fill into d_out (bytes); vis: n bytes of scratch,
// changed: one device int, h_changed: one host int (pinned); with d_words
// non-null the final solid mask is written there as packed bits
// ((n + 31) / 32 words, bit j of word w = cell 32w + j) and d_out keeps the
// surface only.  Enqueued on st; it synchronises st once per sweep round.
hipError_t voxelize_device(const TriF* d_tris, int ntri, int nx, int ny, int nz, uint8_t* d_out,
                           uint8_t* vis, int* changed, int* h_changed, uint32_t* d_words,
                           hipStream_t st);
// bytes (nonzero = 1) <-> packed bits, as above
hipError_t pack_bits_device(const uint8_t* d_in, size_t n, uint32_t* d_words, hipStream_t st);
hipError_t unpack_bits_device(const uint32_t* d_words, size_t n, uint8_t* d_out, hipStream_t st);
hipError_t downsample_device(const uint8_t* d_fine, int refine, int nx, int ny, int nz,
                             float threshold, uint8_t* d_out, hipStream_t st);
// Device scratch and pinned host staging of the geometry calls (stage.cpp),
// one set per device, grown on demand and kept between calls; a call holds
// the mutex (geom_scratch's lock) while it uses them.
struct GeomScratch {
  static constexpr int kSlots = 8;
  std::mutex mu;
  void* dev[kSlots] = {};
  size_t devCap[kSlots] = {};
  void* host[kSlots] = {};
  size_t hostCap[kSlots] = {};
  hipError_t device(int slot, size_t bytes, void** out);
  hipError_t pinned(int slot, size_t bytes, void** out);
  void release();
};
// the current device's set, locked; nullptr if the device is unknown
GeomScratch* geom_scratch(std::unique_lock<std::mutex>& lock);
// host bit (un)packing of masks, bit j of word w = cell 32w + j (threads)
void unpack_bits_host(const uint32_t* words, size_t n, uint8_t* out);
void pack_bits_host(const uint8_t* in, size_t n, uint32_t* words);

// d_scratch: band_scratch_bytes(nx, ny, nz, d0) bytes (one byte per cell
// while at most 254 distance rings are needed, an int per cell beyond)
size_t band_scratch_bytes(int nx, int ny, int nz, int d0);
hipError_t band_device(const uint8_t* d_solid, int nx, int ny, int nz, int d0, uint8_t* d_band,
                       uint8_t* d_scratch, hipStream_t st);

}  // namespace fs
