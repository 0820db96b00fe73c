/*
 * fs_oracle.h -- CPU oracle for the D3Q19 LBM step and the STL voxelizer.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a scalar C++ restatement of the
 * reference update rules (the src/lbm, src/geom, src/apps/sim_runner.hpp and
 * src/moment_vtk_writer.cu files of yagise/MyFluidSimulation).  It is the checker
 * the HIP library is compared against.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product path
 * (libfluidsim.so, the fluidsim_* executables) never links or calls it.
 *
 * Every entry point is plain C so ctypes can bind it.  The arithmetic type is
 * chosen per solver (float or double); FP32 solvers reproduce the reference's
 * float expression order exactly (built with -ffp-contract=off).
 */
#ifndef FS_ORACLE_H
#define FS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* methods (same numbering as the product's fs_method) */
#define ORC_BGK 0
#define ORC_HOME 1
#define ORC_HYBRID 2

/* per-cell flag bits (same values as the product's FS_FLAG_*) */
#define ORC_SOLID 0x01u
#define ORC_BAND 0x02u
#define ORC_INLET 0x04u
#define ORC_OUTLET 0x08u

typedef struct orc_solver orc_solver;

int orc_version(void);
int orc_num_threads(void);

/* precision: 0 = float, 1 = double */
orc_solver* orc_create(int method, int precision, int nx, int ny, int nz,
                       double tau, double fx, double fy, double fz);
void orc_destroy(orc_solver* s);

/* flags: nx*ny*nz bytes; only SOLID/BAND/INLET/OUTLET are meaningful */
int orc_set_flags(orc_solver* s, const uint8_t* flags);
/* inlet x-velocity per (y,z): ny*nz floats, index z*ny+y */
int orc_set_inlet(orc_solver* s, const float* ux);
/* hybrid relaxation rates; the reference sets all three to 1/tau */
int orc_set_omegas(orc_solver* s, double omega_legacy, double omega_hydro,
                   double omega_shear);

/* bounce-back partner: 0 opposite direction, 1 the reference's opp table
 * (HOME only; the populations solvers reject 1) */
int orc_set_bounce(orc_solver* s, int rule);
int orc_reset(orc_solver* s);
/* writers per slot of one push step with partner table 0 (opposite) or 1
 * (the reference's); counts is 19*N int32, solid cells write nothing */
int orc_push_write_counts(const uint8_t* solid, int nx, int ny, int nz,
                          int table, int32_t* counts);
/* nullable pointers follow each method's reference semantics */
int orc_reinit_f32(orc_solver* s, const float* rho, const float* ux,
                   const float* uy, const float* uz);
int orc_step(orc_solver* s, int nsteps);

/* which: 0 rho, 1 ux, 2 uy, 3 uz */
int orc_get_macro_f32(const orc_solver* s, int which, float* out);
int orc_get_macro_f64(const orc_solver* s, int which, double* out);
/* populations f[q*N+id] (19N, BGK/Hybrid) or moments m[k*N+id] (10N, HOME) */
size_t orc_state_size(const orc_solver* s);
int orc_get_state_f32(const orc_solver* s, float* out);
int orc_get_state_f64(const orc_solver* s, double* out);
int orc_set_state_f32(orc_solver* s, const float* in);

/* ---- geometry ---- */
/* two-call pattern: tri_xyz == NULL returns the triangle count in *ntri */
int orc_load_stl(const char* path, float* tri_xyz, int capacity, int* ntri,
                 float* bbmin3, float* bbmax3);
int orc_voxelize(const float* tri_xyz, int ntri, const float* bbmin3,
                 const float* bbmax3, int nx, int ny, int nz,
                 float uniform_scale, float tx, float ty, float tz,
                 int refine, float threshold, uint8_t* mask_out);
/* single SAT test (for known-answer tests) */
int orc_tri_box_overlap(const float* c3, const float* half3, const float* v0,
                        const float* v1, const float* v2);
/* distance-to-solid BFS (int32 out, -1 = unreached) and band mask */
int orc_distance(const uint8_t* solid, int nx, int ny, int nz, int max_dist,
                 int32_t* dist_out);
int orc_band(const uint8_t* solid, int nx, int ny, int nz, int d0,
             uint8_t* band_out);

/* ---- fields / output ---- */
int orc_rho_gauss(float* rho, int nx, int ny, int nz, float cx, float cy,
                  float cz, float sigma, float amp);
int orc_rho_slab(float* rho, int nx, int ny, int nz, int axis, float amp,
                 float center, float width);
int orc_speed(const float* u, const float* v, const float* w, float* out,
              size_t n);
int orc_write_scalar_vtk(const char* path, const float* field, int nx, int ny,
                         int nz);
int orc_write_vector_vtk(const char* path, const float* fx, const float* fy,
                         const float* fz, int nx, int ny, int nz);

#ifdef __cplusplus
}
#endif
#endif
