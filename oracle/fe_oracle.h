// fe_oracle.h -- TEST INFRASTRUCTURE: C ABI of the FP64 CPU oracle (fe_oracle.cpp).
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
#ifndef FE_ORACLE_H
#define FE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_PIC = 0, ORC_FLIP = 1, ORC_APIC = 2 };
// FluidMarker values, src/grid_systems/fluid_markers.hpp:8-13
enum { ORC_AIR = 0, ORC_FLUID = 1, ORC_BOUNDRY = 2 };
enum {
    ORC_U = 0, ORC_V = 1, ORC_W = 2,
    ORC_MU = 3, ORC_MV = 4, ORC_MW = 5,
    ORC_RU = 6, ORC_RV = 7, ORC_RW = 8,
    ORC_SDF = 9, ORC_FLAGS = 10, ORC_PRESSURE = 11, ORC_RHS = 12, ORC_COLLIDER_SDF = 13
};
enum { ORC_SOLVER_JACOBI = 0, ORC_SOLVER_CG = 1, ORC_SOLVER_PCG_JACOBI = 2 };

typedef struct orc_solver_desc {
    int32_t kind;
    int32_t max_iterations;
    int32_t check_interval;  // Jacobi residual check interval
    int32_t relative;        // CG: stop on ||r||/||b|| <= tolerance instead of sigma <= tol^2
    int32_t warm_start;      // CG: start from the previous pressure
    int32_t pad;
    double tolerance;
} orc_solver_desc;

typedef struct orc_solve_stats {
    int32_t iterations;
    int32_t pad;
    double rel_residual;
    double abs_residual;
    double ms;
} orc_solve_stats;

typedef struct orc_scene orc_scene;

orc_scene* orc_create(const int n[3], const double origin[3], const double size[3], int method, double blend,
                      double max_cfl, const double gravity[3]);
void orc_destroy(orc_scene* s);
void orc_set_threads(orc_scene* s, int nthreads);
// Pinning tests only: switch on one deliberate restatement mistake so that a
// test can show its check fails on it.  0 (the default) is the oracle proper.
enum {
    ORC_PLANT_NONE = 0,
    ORC_PLANT_APIC_TRANSPOSE = 1,    // G2P writes the affine matrix C transposed
    ORC_PLANT_FLIP_DELTA_SIGN = 2,   // FLIP delta taken as snapshot - new
    ORC_PLANT_EULER_ADVECTION = 3,   // MoveParticles steps with forward Euler
    ORC_PLANT_APIC_FACE_DIMS = 4     // APIC gradient stencil sized by the face array on the normal axis
};
void orc_plant_mistake(orc_scene* s, int which);
// Solver dot products in storage order over 64 fixed blocks (deterministic
// for any thread count) instead of the reference's serial i-outer order: the
// same converged answer, fast enough for tests on 128^3 grids.  Off by default.
void orc_set_fast_dots(orc_scene* s, int on);
// Surface triangulation: the reference's case tables (the default) or the
// closed-surface face rule.
enum { ORC_SURFACE_REFERENCE = 0, ORC_SURFACE_CLOSED = 1 };
void orc_set_surface_mode(orc_scene* s, int mode);
void orc_set_particles(orc_scene* s, const double* x, const double* v, const double* C, int64_t np);
void orc_get_particles(const orc_scene* s, double* x, double* v, double* C);
int64_t orc_num_particles(const orc_scene* s);
int64_t orc_grid_count(orc_scene* s, int what);
int orc_get_grid(orc_scene* s, int what, void* out);
int orc_set_grid(orc_scene* s, int what, const void* in);

void orc_p2g(orc_scene* s);
void orc_build_sdf(orc_scene* s);
void orc_build_markers(orc_scene* s);
void orc_extrapolate_to_air(orc_scene* s);
void orc_apply_bc(orc_scene* s);
void orc_apply_gravity(orc_scene* s, double dt);
void orc_build_system(orc_scene* s);
void orc_apply_pressure(orc_scene* s);
void orc_g2p(orc_scene* s);
void orc_move_particles(orc_scene* s, double dt);
void orc_pressure(orc_scene* s, const orc_solver_desc* d, orc_solve_stats* st);
void orc_substep(orc_scene* s, double dt, const orc_solver_desc* d, orc_solve_stats* st);
// BackwardEulerDiffusionSolver: viscosity > 0 adds ComputeDiffusion to orc_substep;
// orc_diffusion runs it alone (no boundary condition after it); markers of the
// last solve per component, uint8 FluidMarker values, face-array sized.
void orc_set_viscosity(orc_scene* s, double viscosity);
void orc_set_diffusion_solver(orc_scene* s, const orc_solver_desc* d);
void orc_diffusion(orc_scene* s, double dt, orc_solve_stats st[3]);
int orc_diffusion_markers(const orc_scene* s, int c, uint8_t* out);
// MarchingCubesSolver::BuildSurface on the fluid SDF: malloc'ed vertices,
// normals (3 doubles each) and triangles (3 indices); free with orc_free.
int orc_surface(orc_scene* s, double** vertices, double** normals, int32_t** triangles, int64_t* nv, int64_t* nt);
void orc_free(void* p);
// Mesh colliders: Mesh2SDF::Build of a triangle mesh on a node grid (x fastest;
// triangles as 3 vertex indices), and AddCollider of such a TriangleMeshCollider
// (the scene's collider SDF becomes the running minimum of the sampled collider
// SDFs).  Both return -1 for a vertex index out of range.
int orc_mesh_to_sdf(const int size[3], const double origin[3], const double h[3], const double* vertices, int64_t nv,
                    const int32_t* triangles, int64_t nt, double* out);
int orc_add_mesh_collider(orc_scene* s, const int size[3], const double origin[3], const double h[3],
                          const double* vertices, int64_t nv, const int32_t* triangles, int64_t nt);
void orc_closest_on_triangle(const double x[3], const double p0[3], const double p1[3], const double p2[3],
                             double out[3]);
int orc_mesh_inside(const double* vertices, int64_t nv, const int32_t* triangles, int64_t nt, const double x[3]);
// Scene sources.  A VolumeEmitter (its grid must have the scene's size; -2
// otherwise) emits at the start of every orc_substep; InitializeFromTriangleMesh
// writes the mesh SDF into the fluid SDF and appends particles.  The generator
// standing in for std::rand is orc_uniform01.
int orc_add_volume_emitter(orc_scene* s, const int size[3], const double origin[3], const double h[3],
                           const double* vertices, int64_t nv, const int32_t* triangles, int64_t nt, int ppb,
                           const double velocity[3], uint64_t seed);
int orc_initialize_from_mesh(orc_scene* s, const double* vertices, int64_t nv, const int32_t* triangles, int64_t nt,
                             int ppb, uint64_t seed);
double orc_uniform01(uint64_t seed, uint64_t call, uint64_t n, int axis);
void orc_emit(orc_scene* s);  // the emitters' Emitt alone (the first step of orc_substep)
double orc_cfl(const orc_scene* s, double dt);
unsigned orc_num_substeps(const orc_scene* s, double dt);
int orc_advance_frame(orc_scene* s, double frame_dt, const orc_solver_desc* d);
void orc_last_stats(const orc_scene* s, orc_solve_stats* st);

void orc_barycentric(double x, int iLow, int iHigh, int* i, double* f);
void orc_coords_weights(const int size[3], const double origin[3], const double h[3], const double x[3],
                        int idx_out[24], double w_out[8]);
void orc_coords_grad_weights(const int size[3], const double origin[3], const double h[3], const double x[3],
                             int idx_out[24], double g_out[24]);
double orc_trilerp(const double v[8], double fx, double fy, double fz);
double orc_sample_array(const double* data, const int dims[3], const double origin[3], const double h[3],
                        const double p[3]);
void orc_extrapolate_to_region(const double* input, const int* valid, const int dims[3], int iterations,
                               double* output);
int orc_hash_grid_query(const double* pts, int64_t npts, int sx, double spacing, const double q[3], double radius,
                        int64_t* out, int max_out);
double orc_blas_dot(const double* a, const double* b, const int dims[3]);
void orc_solve_rows(const double* Ac, const double* Ar, const double* Au, const double* Af, const double* b,
                    double* x, const int dims[3], const orc_solver_desc* d, orc_solve_stats* st);
void orc_residual_rows(const double* Ac, const double* Ar, const double* Au, const double* Af, const double* b,
                       const double* x, const int dims[3], double* out);
void orc_spmv_rows(const double* Ac, const double* Ar, const double* Au, const double* Af, const double* v,
                   const int dims[3], double* out);

#ifdef __cplusplus
}
#endif

#endif  // FE_ORACLE_H
