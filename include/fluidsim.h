/*
 * fluidsim.h -- C ABI of the MI355X (gfx950) D3Q19 lattice-Boltzmann step and
 * STL voxelizer (libfluidsim.so).
 *
 * Plain C: pointers, sizes and int status codes; no C++ or torch types cross
 * this boundary and no C++ exception escapes it.  Every fs_* function returns
 * FS_OK (0) or a negative FS_ERR_* code; fs_last_error() then holds a message
 * (per calling thread).
 *
 * Each declaration cites the reference interface it replaces
 * (yagise/MyFluidSimulation, file:line).  The reference exposes a duck-typed
 * C++ "Solver" class per method (LBM3D_Legacy / LBM3D_Home / LBM3D_Hybrid);
 * fs_solver is that class behind an opaque handle, with the method chosen at
 * creation.  Device pointers handed out by a solver are owned by it and stay
 * valid until fs_solver_destroy (HOME's macro pointers move on every step, as
 * in the reference).  Memory passed in by callers is never freed by the
 * library.
 *
 * Threading: a solver may be used from one host thread at a time.  Work runs
 * on the solver's own HIP stream (default priority); the synchronous calls
 * wait for it before returning, as every reference solver method ends with
 * cudaDeviceSynchronize (lbm3d_legacy.cu:158,207,223).
 */
#ifndef FLUIDSIM_H
#define FLUIDSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FS_ABI_VERSION 1

/* ---- status codes ---- */
#define FS_OK 0
#define FS_ERR_INVALID (-1) /* bad argument or flag grid */
#define FS_ERR_DEVICE (-2)  /* HIP runtime error (no GPU, launch failure) */
#define FS_ERR_ALLOC (-3)   /* device or host allocation failed */
#define FS_ERR_IO (-4)      /* file could not be opened, read or written */
#define FS_ERR_STATE (-5)   /* call not valid for this solver's method */

/* ---- methods: lbm3d_legacy.hpp:19, lbm3d_home.hpp:15, lbm3d_hybrid.hpp:18 ---- */
#define FS_METHOD_BGK 0    /* LBM3D_Legacy: BGK collide + stream */
#define FS_METHOD_HOME 1   /* LBM3D_Home: 10 moments per cell */
#define FS_METHOD_HYBRID 2 /* LBM3D_Hybrid: BGK band + two-rate bulk */

/* ---- per-cell flag byte ---- */
#define FS_FLAG_SOLID 0x01u  /* solid (reference solid mask, lbm3d_legacy.cu:72) */
#define FS_FLAG_BAND 0x02u   /* Hybrid BGK band (reference isLegacy, lbm3d_hybrid.cu:137) */
#define FS_FLAG_INLET 0x04u  /* extension: equilibrium inlet cell */
#define FS_FLAG_OUTLET 0x08u /* extension: zero-gradient outlet cell (copies x-1) */

/* ---- bounce-back partner rule (fs_solver_set_bounce_rule) ----
 * HALFWAY: a population hitting a solid cell returns in the opposite
 *   direction (e = -e_q): mass-conserving halfway bounce-back, all methods.
 * REFERENCE_TABLE: the partner from the reference's opp table
 *   (lbm3d_legacy.cu:29-32), which for q >= 7 pairs mirror images
 *   (e.g. (+1,+1,0) with (-1,+1,0)).  HOME only: its gather stays
 *   deterministic; in the populations solvers' push update that table
 *   writes some slots twice and others never. */
#define FS_BOUNCE_HALFWAY 0
#define FS_BOUNCE_REFERENCE_TABLE 1

/* ---- macro selectors for fs_solver_macro_ptr / fs_solver_copy_macro ---- */
#define FS_RHO 0
#define FS_UX 1
#define FS_UY 2
#define FS_UZ 3

typedef struct fs_solver fs_solver;

/* ABI version (FS_ABI_VERSION) and the last error message of this thread. */
int fs_version(void);
const char* fs_last_error(void);
/* Number of visible HIP devices (0 when none).  Fails with FS_ERR_DEVICE,
 * like every call that creates device state or takes a stream, when the
 * process maps more than one HIP runtime or the one it maps is not the
 * major.minor version the library was built against (see fs_hip_versions). */
int fs_device_count(int* count);
/* Extension (no reference counterpart): the HIP version the library was
 * compiled against and the version of the runtime it is bound to, both as
 * major*10000000 + minor*100000 + patch; either pointer may be NULL. */
int fs_hip_versions(int* built_against, int* runtime);

/* ===================== solver (the reference Solver concept) ============ */

/* LBM3D_*::init(const Domain&) (lbm3d_legacy.cu:141-147, lbm3d_home.cu:301-307,
 * lbm3d_hybrid.cu:210-221) with Domain{Nx,Ny,Nz,tau,forceX..Z}
 * (lbm_common.hpp:6-10): allocates device state and resets it to rho=1,u=0. */
int fs_solver_create(int method, int nx, int ny, int nz, float tau, float fx,
                     float fy, float fz, fs_solver** out);
/* ~LBM3D_*() (lbm3d_legacy.cu:14): frees every device buffer. */
int fs_solver_destroy(fs_solver* s);
/* Nx() Ny() Nz() N() (lbm3d_legacy.hpp:60-64); any pointer may be NULL. */
int fs_solver_dims(const fs_solver* s, int* nx, int* ny, int* nz, size_t* n);
int fs_solver_method(const fs_solver* s, int* method);

/* setSolidMask(const unsigned char*) (lbm3d_legacy.cu:149-151): N host bytes,
 * nonzero = solid.  Clears any inlet/outlet flags; keeps the band. */
int fs_solver_set_solid_mask(fs_solver* s, const uint8_t* host_mask);
/* LBM3D_Hybrid::setLegacyMapping(h_isLegacy, h_legacySlot_ignored)
 * (lbm3d_hybrid.cu:227-235): N host bytes or NULL (= no band cell); the slot
 * array is ignored, as in the reference.  FS_ERR_STATE for other methods. */
int fs_solver_set_legacy_mapping(fs_solver* s, const uint8_t* host_is_legacy,
                                 const int* ignored_slots);
/* Extension: the whole flag byte per cell (FS_FLAG_*).  INLET cells must not
 * be SOLID; OUTLET cells must sit at x >= 1 with a fluid, non-inlet,
 * non-outlet upstream neighbour x-1.  FS_ERR_INVALID otherwise. */
int fs_solver_set_flags(fs_solver* s, const uint8_t* host_flags);
int fs_solver_copy_flags(fs_solver* s, uint8_t* host_flags);
/* Extension: inlet x-velocity per (y,z) (ny*nz floats, index z*ny+y); NULL
 * sets it to zero.  Host pointer. */
int fs_solver_set_inlet_profile(fs_solver* s, const float* host_ux);
/* setForce(fx,fy,fz) (lbm3d_home.hpp:33, lbm3d_hybrid.hpp:39). */
int fs_solver_set_force(fs_solver* s, float fx, float fy, float fz);
/* Hybrid relaxation rates (lbm3d_hybrid.cu:264-267 hard-wires all three to
 * 1/tau, which is the default here).  FS_ERR_STATE for other methods. */
int fs_solver_set_omegas(fs_solver* s, float omega_legacy, float omega_hydro,
                         float omega_shear);

/* Extension: bounce-back partner rule (FS_BOUNCE_*; default HALFWAY).
 * FS_ERR_STATE for REFERENCE_TABLE on BGK / Hybrid. */
int fs_solver_set_bounce_rule(fs_solver* s, int rule);

/* reset() (lbm3d_legacy.cu:153-159, lbm3d_home.cu:313-319,
 * lbm3d_hybrid.cu:237-243). */
int fs_solver_reset(fs_solver* s);
/* reinitEquilibriumFromMacro(d_rho,d_ux,d_uy,d_uz) (lbm3d_legacy.cu:160-208,
 * lbm3d_home.cu:321-333, lbm3d_hybrid.cu:245-257): device pointers of N
 * floats each, any may be NULL with each method's reference meaning. */
int fs_solver_reinit_from_macro(fs_solver* s, const float* d_rho,
                                const float* d_ux, const float* d_uy,
                                const float* d_uz);
/* step(int substeps) (lbm3d_legacy.cu:210-224, lbm3d_home.cu:335-354,
 * lbm3d_hybrid.cu:259-283): nsteps time steps; returns when they are done. */
int fs_solver_step(fs_solver* s, int nsteps);
/* Extension: enqueue nsteps on the solver's stream and return at once. */
int fs_solver_step_async(fs_solver* s, int nsteps);
/* Extension: stream handling.  set_stream(NULL) restores the solver's own
 * stream; the caller keeps ownership of a stream it passes in. */
int fs_solver_set_stream(fs_solver* s, void* hip_stream);
int fs_solver_get_stream(const fs_solver* s, void** hip_stream);
int fs_solver_synchronize(fs_solver* s);
/* Extension: enqueue nsteps between two HIP events recorded on the solver's
 * stream, wait, and return the elapsed device time in milliseconds. */
int fs_solver_timed_steps(fs_solver* s, int nsteps, float* elapsed_ms);

/* Extension: kernel variant selection.  HOME has a plain per-cell gather
 * (variant 0) and tiled kernels (1: 64x4 columns x 16 planes, 2: 64x8 x 16,
 * 3: 64x8 x 32; 4, 5, 6: the same shapes in a second form with fewer
 * instructions per update; 7: 64x16 x 32 and 8: 64x8 x 64 in that form;
 * 9: 64x4 x 16, 10: 64x8 x 32, 11: 64x8 x 64 in a third form; 12: 64x11 x
 * 64, 13: 64x11 x 32, 14: 64x10 x 64, 15: 64x12 x 64, 16: 64x12 x 32 tiles
 * that advance two steps per launch, the default from about 256^3 cells,
 * which run the one-step default where an outlet's downstream neighbour reads
 * it across a tile's left ring column).
 * BGK / Hybrid have a one-step in-place kernel (variant 0) and kernels that
 * advance two steps per launch through LDS (1: 64x4 x 16, 2: 64x8 x 16,
 * 3: 64x8 x 32, 4: 64x4 x 32, 5: 32x8 x 16, 6: 64x12 x 16, 7: 64x12 x 32,
 * 8: 64x12 x 64; a step(n) call with n odd runs one one-step launch).
 * Every variant gives the same bits.
 * -1 restores the per-grid default (the one-step / plain kernels below about
 * 128^3 cells).  While an outlet cell sits at an x that is a multiple of the
 * variant's tile width (64, or 32 for variant 5) the solver runs variant 0
 * (get_variant reports it).
 * fs_solver_autotune times each applicable variant for trial_steps steps
 * from the current state (restored after each trial) and keeps the
 * fastest; *chosen (may be NULL) receives it. */
int fs_solver_set_variant(fs_solver* s, int variant);
int fs_solver_get_variant(const fs_solver* s, int* variant);
int fs_solver_autotune(fs_solver* s, int trial_steps, int* chosen);

/* d_rho() d_u() d_v() d_w() (lbm3d_legacy.cu:229-236, lbm3d_home.cu:355-362):
 * borrowed device pointer to N floats.  BGK and Hybrid report the moments of
 * the populations at the start of the last step (the reference's one-step
 * lag); HOME reports its current moments. */
int fs_solver_macro_ptr(fs_solver* s, int which, float** d_ptr);
int fs_solver_copy_macro(fs_solver* s, int which, float* host_dst);

/* Extension: raw state in the reference layout: populations f[q*N+id]
 * (19N floats, BGK/Hybrid; the incoming populations of the current time
 * level, as the reference's d_f() holds after step) or moments m[k*N+id]
 * (10N floats, HOME; k = rho,ux,uy,uz,Sxx,Sxy,Sxz,Syy,Syz,Szz). */
int fs_solver_state_size(const fs_solver* s, size_t* nfloats);
int fs_solver_copy_state(fs_solver* s, float* host_dst);
int fs_solver_load_state(fs_solver* s, const float* host_src);

/* ===================== batched domains ================================== */
/* Many independent lattices of uneven size stepped together (replaces the
 * reference's one-process-per-configuration sweeps, main_compare.cpp:511-699,
 * scripts/run_b0_sweeps.ps1).  Every domain gives the same bits as a
 * single-domain solver with the same inputs.  Domains share one step
 * counter; the step loop replays a captured hipGraph. */
typedef struct fs_domain_desc {
  int method; /* FS_METHOD_* */
  int nx, ny, nz;
  float tau;
  float fx, fy, fz;
} fs_domain_desc;
typedef struct fs_batch fs_batch;

/* Allocates and resets (rho=1, u=0) every domain.  A domain has fewer than
 * 2^31 cells, a HOME domain fewer than 2^29 (FS_ERR_INVALID otherwise). */
int fs_batch_create(const fs_domain_desc* descs, int ndomains, fs_batch** out);
int fs_batch_destroy(fs_batch* b);
int fs_batch_info(const fs_batch* b, int* ndomains, size_t* cells);
/* Per-domain set-up, as fs_solver_set_flags / _set_inlet_profile /
 * _reinit_from_macro.  A HOME domain's outlet cell may not sit at an x that
 * is a multiple of 32 (FS_ERR_INVALID).  Re-initialisation needs an even
 * number of batch steps since creation (FS_ERR_STATE otherwise). */
int fs_batch_set_flags(fs_batch* b, int domain, const uint8_t* host_flags);
int fs_batch_set_inlet_profile(fs_batch* b, int domain, const float* host_ux);
/* As fs_solver_set_bounce_rule, per domain (HOME domains may take the
 * reference's pairing table, LBM3D_Home's rule, lbm3d_home.cu:27-30). */
int fs_batch_set_bounce_rule(fs_batch* b, int domain, int rule);
int fs_batch_reinit_from_macro(fs_batch* b, int domain, const float* d_rho,
                               const float* d_ux, const float* d_uy,
                               const float* d_uz);
/* nsteps for every domain; synchronous.  timed_steps: HIP events on the
 * batch's stream around the steps, elapsed device milliseconds. */
int fs_batch_step(fs_batch* b, int nsteps);
int fs_batch_timed_steps(fs_batch* b, int nsteps, float* elapsed_ms);
/* Launch shape: workgroups of the populations and HOME launches (0 = one per
 * tile, otherwise a persistent grid that strides over the tile list) and
 * whether the step loop replays a captured hipGraph. */
int fs_batch_set_launch(fs_batch* b, int pop_blocks, int home_blocks, int use_graph);
int fs_batch_get_launch(const fs_batch* b, int* pop_blocks, int* home_blocks, int* use_graph);
/* Times the candidate grid sizes (trial_steps, even, each from the current
 * state, which is restored) and keeps the fastest pair. */
int fs_batch_autotune(fs_batch* b, int trial_steps, int* pop_blocks, int* home_blocks);
/* Results of one domain, as fs_solver_copy_macro / fs_solver_copy_state. */
int fs_batch_copy_macro(fs_batch* b, int domain, int which, float* host_dst);
int fs_batch_copy_state(fs_batch* b, int domain, float* host_dst);

/* ===================== differentiable runs ============================== */
/* Extension with no counterpart in the reference (its solvers have no
 * adjoint): reverse-mode gradients of a T-step run of a solver through
 * hand-written adjoint kernels, with sqrt(T) checkpointing and segment-wise
 * recomputation of the forward trajectory.  Inputs are the initial macro
 * fields (rho0, u0) and the inlet profile; outputs are the solver's macro
 * fields after T steps (as fs_solver_macro_ptr reports them) and the
 * x-momentum D^n handed to "body" solid cells in each step n by halfway
 * bounce-back (momentum exchange, 2 e_qx f*_q summed over the body links).
 * All pointers are device pointers; work is enqueued on the solver's stream
 * (fs_solver_set_stream) and these calls return without waiting.  The
 * solver's flags, parameters and variant must not change while the handle
 * lives.  Inlet cells, if any, must sit at x = 0. */
typedef struct fs_adjoint fs_adjoint;

/* segment: steps between checkpoints (0 = ceil(sqrt(T))).  host_body: N
 * host bytes, nonzero marks the solid cells whose links count toward the
 * drag, or NULL for every solid cell. */
int fs_adjoint_create(fs_solver* s, int T, int segment, const uint8_t* host_body,
                      fs_adjoint** out);
int fs_adjoint_destroy(fs_adjoint* a);
int fs_adjoint_info(const fs_adjoint* a, int* T, int* segment, int* checkpoints,
                    size_t* device_bytes);
/* Re-initialises the solver at equilibrium from (rho0, u0) (N floats each;
 * NULL: rho 1, u 0), sets the inlet profile (ny*nz floats, NULL keeps the
 * solver's), runs T steps storing the checkpoints, copies the end macro
 * fields to out_* (each may be NULL) and D^0..D^{T-1} to drag (T floats,
 * may be NULL). */
int fs_adjoint_forward(fs_adjoint* a, const float* rho0, const float* ux0,
                       const float* uy0, const float* uz0, const float* inlet,
                       float* out_rho, float* out_ux, float* out_uy,
                       float* out_uz, float* drag);
/* Given dL/d(end macro fields) (N floats each) and dL/dD (T floats), any of
 * them NULL for zero, writes dL/d(rho0, u0) (N floats each) and dL/d(inlet)
 * (ny*nz floats), any of them NULL to skip, for the last forward pass.
 * Leaves the solver at its end state; may be called again with other
 * gradients.  FS_ERR_STATE before a forward pass. */
int fs_adjoint_backward(fs_adjoint* a, const float* g_rho, const float* g_ux,
                        const float* g_uy, const float* g_uz, const float* g_drag,
                        float* grad_rho0, float* grad_ux0, float* grad_uy0,
                        float* grad_uz0, float* grad_inlet);

/* ===================== geometry ========================================= */

/* load_stl(path, TriangleMesh&) (stl_loader.cpp:102-122): triangle soup,
 * 9 floats per triangle (v0,v1,v2), and the vertex bounding box.  Call with
 * tri_xyz == NULL to get the count; FS_ERR_IO when the file cannot be read
 * or parsed (the reference's `false`). */
int fs_load_stl(const char* path, float* tri_xyz, int capacity, int* ntri,
                float* bbmin3, float* bbmax3);
/* voxelize_mesh_to_mask(mesh, VoxelParams) (voxelize.cpp:228-243,
 * VoxelParams voxelize.hpp:16-24), on the GPU: nx*ny*nz host bytes out,
 * 1 = solid, x fastest. */
int fs_voxelize(const float* host_tri_xyz, int ntri, const float* bbmin3,
                const float* bbmax3, int nx, int ny, int nz,
                float uniform_scale, float tx, float ty, float tz, int refine,
                float threshold, uint8_t* host_mask_out);
/* build_distance_to_solid_6n + build_legacy_mask_from_distance
 * (sim_runner.hpp:28-91, maxDist = max(1,d0) as at sim_runner.hpp:222), on
 * the GPU: band = fluid cells within 6-neighbour distance d0 of a solid. */
int fs_band(const uint8_t* host_solid, int nx, int ny, int nz, int d0,
            uint8_t* host_band_out);
/* Extension with no counterpart in the reference, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): fs_band on caller-owned device buffers
 * (nx*ny*nz bytes each), enqueued on hip_stream (NULL: the null stream) and
 * returning without waiting.  It calls no hipMalloc / hipFree: the BFS
 * distances live in a fixed 64 MiB device array inside the library (one per
 * GPU, held until the process ends), so grids needing more scratch -- more
 * than 64 Mi cells, or 16 Mi cells when d0 >= 255 needs int distances -- are
 * rejected with FS_ERR_INVALID.  Calls that overlap on one GPU share that
 * array and race: serialise them. */
int fs_band_noalloc(const uint8_t* d_solid, int nx, int ny, int nz, int d0,
                    uint8_t* d_band, void* hip_stream);

/* Extension: the geometry calls (fs_voxelize, fs_band) and the VTK writers
 * keep their device scratch and pinned host staging between calls (one set
 * per device, grown to the largest grid seen); this frees them.  A later
 * call allocates them again.  These calls are serialised per device. */
int fs_geometry_release(void);

/* ===================== fields and output ================================ */

/* launch_speed (field_ops.cu:25-28) */
int fs_speed(const float* d_u, const float* d_v, const float* d_w,
             float* d_out, size_t n);
/* launch_diff / launch_shift (field_ops.cu:20-23,29-32,34-41): out = a - b,
 * out = a + bias (device, n floats) */
int fs_diff(const float* d_a, const float* d_b, float* d_out, size_t n);
int fs_shift(const float* d_a, float bias, float* d_out, size_t n);
/* make_rho_gauss / make_rho_slab (init_profiles.cu:31-45) */
int fs_rho_gauss(float* d_rho, int nx, int ny, int nz, float cx, float cy,
                 float cz, float sigma, float amp);
int fs_rho_slab(float* d_rho, int nx, int ny, int nz, int axis, float amp,
                float center, float width);
/* reinit_equilibrium_from_macro (lbm_reinit.cu:52-61): f[q*n+i] = feq */
int fs_reinit_equilibrium(const float* d_rho, const float* d_ux,
                          const float* d_uy, const float* d_uz, float* d_f,
                          size_t n);
/* write_scalar_vtk / write_vector_vtk (moment_vtk_writer.cu:9-49): device
 * field(s) of nx*ny*nz floats to a legacy ASCII STRUCTURED_POINTS file. */
int fs_write_scalar_vtk(const char* filename, const float* d_field, int nx,
                        int ny, int nz);
int fs_write_vector_vtk(const char* filename, const float* d_fx,
                        const float* d_fy, const float* d_fz, int nx, int ny,
                        int nz);
/* Extension: the same files in the legacy format's BINARY encoding (same
 * header with BINARY, values as big-endian float32): exact values, about a
 * third of the size, no text formatting. */
int fs_write_scalar_vtk_binary(const char* filename, const float* d_field,
                               int nx, int ny, int nz);
int fs_write_vector_vtk_binary(const char* filename, const float* d_fx,
                               const float* d_fy, const float* d_fz, int nx,
                               int ny, int nz);

/* Device memory helpers for callers built without HIP headers. */
int fs_device_malloc(void** d_ptr, size_t bytes);
int fs_device_free(void* d_ptr);
int fs_memcpy_htod(void* d_dst, const void* h_src, size_t bytes);
int fs_memcpy_dtoh(void* h_dst, const void* d_src, size_t bytes);
int fs_memset_device(void* d_dst, int value, size_t bytes);
int fs_device_synchronize(void);
/* Bandwidth probe: `iters` launches of a streaming copy of `bytes` bytes
 * (src -> dst, both device, 16-byte aligned, bytes % 16 == 0), timed with
 * HIP events; *elapsed_ms is the total.  Gives the measured copy ceiling. */
int fs_copy_probe(void* d_dst, const void* d_src, size_t bytes, int iters, float* elapsed_ms);

/* ===================== reference-named entry points ===================== */
/* Same names and signatures as the reference's extern "C" helpers
 * (declared by its consumers at sim_common.hpp:30-46, sim_runner.hpp:24).
 * Like the reference they return nothing; on failure they print the error
 * to stderr and abort, as the reference's throw across extern "C" ends the
 * process (moment_vtk_writer.cu:14,38). */
void write_scalar_vtk(const char* filename, const float* d_field, int Nx,
                      int Ny, int Nz);
void write_vector_vtk(const char* filename, const float* d_fx,
                      const float* d_fy, const float* d_fz, int Nx, int Ny,
                      int Nz);
void make_rho_gauss(float* d_rho, int Nx, int Ny, int Nz, float cx, float cy,
                    float cz, float sigma, float amp);
void make_rho_slab(float* d_rho, int Nx, int Ny, int Nz, int axis, float amp,
                   float center, float width);
void reinit_equilibrium_from_macro(const float* d_rho, const float* d_ux,
                                   const float* d_uy, const float* d_uz,
                                   float* d_f, int N);
void launch_speed(const float* u, const float* v, const float* w, float* out,
                  int N);
void launch_diff(const float* a, const float* b, float* out, int N);
void launch_shift(const float* a, float bias, float* out, int N);

#ifdef __cplusplus
}
#endif
#endif /* FLUIDSIM_H */
