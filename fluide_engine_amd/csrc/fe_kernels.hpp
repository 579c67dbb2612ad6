// fe_kernels.hpp -- host launchers of the gfx950 kernels (one scene, one stream).
#pragma once

#include "fe_internal.hpp"

namespace fe {

struct FaceSet {
    float* a[3];
};
struct QuadSet {
    const float4* a[3];
};

// k_particles.hip
hipError_t launch_fill_i32(hipStream_t st, int32_t* a, int32_t v, int64_t n);
hipError_t launch_keys(fe_scene* s, bool histogram);
hipError_t launch_scan(fe_scene* s, const int32_t* in, int32_t* out, int64_t n);
// exclusive scan of n values into out[0..n] with a caller's tile-sum buffer
// of scan_tiles(n) + 1 entries
int64_t scan_tiles(int64_t n);
hipError_t launch_scan_with(hipStream_t st, const int32_t* in, int32_t* out, int64_t n, int32_t* sums);
hipError_t launch_sort(fe_scene* s, bool keys_ready);
hipError_t launch_p2g(fe_scene* s);
hipError_t launch_g2p_move(fe_scene* s, double dt, bool do_g2p, bool do_move, bool emit_keys);
hipError_t launch_quads(fe_scene* s, float* const a[3], float4* const q[3]);
// times the P2G and G2P variants on the scene's particles; keeps the fastest
hipError_t autotune(fe_scene* s, int reps, fe_tuning_times* times);
hipError_t launch_unpermute(fe_scene* s, const float* src, int ncomp, float* dst);
hipError_t launch_deinterleave(fe_scene* s, const float* src, int ncomp, float* dst);
hipError_t launch_iota(fe_scene* s, int32_t* a, int64_t n);
hipError_t launch_cells_by_pid(fe_scene* s, int32_t* out);

// k_grid.hip
// the fluid SDF of the current bins; buckets forces the 8-bucket search (else the
// scene's mode: buckets with a collider or FE_SDF_SEARCH_BUCKETS, the cell scan otherwise)
hipError_t launch_sdf(fe_scene* s, bool buckets = false);
hipError_t launch_markers(fe_scene* s);
hipError_t launch_extrapolate(fe_scene* s);               // ExtrapolateVelocityToAir, depth iterations
hipError_t launch_extrapolate_from(fe_scene* s, const uint8_t* const init[3]);  // from other validity
hipError_t launch_walls(fe_scene* s);                      // wall layers of ConstrainVelocity
hipError_t launch_gravity(fe_scene* s, double dt, bool walls_after);
hipError_t launch_apply_pressure(fe_scene* s, bool walls_after);
hipError_t launch_cfl(fe_scene* s, double dt);  // result in s->cflmax slot s->cfl_slot: FP64 bits, non-finite flag
hipError_t launch_fill_u8(hipStream_t st, uint8_t* a, uint8_t v, int64_t n);
hipError_t launch_fill_f32(hipStream_t st, float* a, float v, int64_t n);

// k_solver.hip
hipError_t launch_pressure_solve(fe_scene* s, const fe_solver_desc* d);  // rhs + solve (no apply)
// a captured substep's part of a CG-family solve (start + `iters` spmv launches + the state slot to the host),
// and what follows it on the host once the head has run
hipError_t capture_pressure_head(fe_scene* s, const fe_solver_desc* d, int iters, CGRun* out);
hipError_t finish_pressure_after_head(fe_scene* s, const fe_solver_desc* d, const CGRun& head, bool* resumed);
hipError_t launch_final_residual(fe_scene* s);  // FP64 ||b - A p||^2 and ||b||^2 (device)
hipError_t read_final_residual(fe_scene* s, double* rr, double* bb);  // waits for them

// k_mg.hip: multigrid V-cycle preconditioner of the pressure CG
hipError_t mg_setup(fe_scene* s);  // coarse markers and row codes from s->flags / s->code
// pconf (nparts blocks of (r.z, r.r), or null): the CG update's partials, with
// which a V-cycle that follows a confirming residual recomputation ends the
// solve at once when the recomputed residual meets a relative stop.
hipError_t mg_apply(fe_scene* s, const float* r, float* z, SolverScalars* st, double* prz, int tag,
                    const double* pconf, int nparts);
float* mg_first_sweep_target(fe_scene* s);  // where the CG kernels write level 0's first sweep
void mg_free(fe_scene* s);

// k_surface.hip: MarchingCubesSolver::BuildSurface on the fluid SDF
hipError_t surface_build(fe_scene* s);  // counts in s->mc (nv3 + caps, nt3 + caps)
hipError_t surface_download(fe_scene* s, double* pos, double* nrm, int32_t* tri);
void surface_free(fe_scene* s);

// k_collider.hip: ApplyBoundryCondition (walls only without a collider) and
// the collision response of moved particles
hipError_t launch_bc(fe_scene* s);
hipError_t launch_collide(fe_scene* s);
hipError_t launch_sdf_into_collider(fe_scene* s);  // ExtrapolateIntoCollider on the fluid SDF
void collider_free(fe_scene* s);
// fe_mesh_sdf_noalloc (fe_noalloc.hip): the limits of one call and the slices
// of the per-device scratch array that hold its buffers.  noalloc_setup runs
// once per device (fe_ctx_create) with that device current.
constexpr int64_t kNaMaxVerts = 65536;
constexpr int64_t kNaMaxTris = 65536;
constexpr int64_t kNaMaxNodes = int64_t(1) << 21;  // e.g. 128^3
struct NoallocSlices {
    double* verts;
    int32_t* tris;
    double* tv;
    int32_t* tb;
    int32_t* bad;
    double* sdf;
};
hipError_t noalloc_setup(int device);
hipError_t noalloc_slices(int device, NoallocSlices* out);
// mesh_prepare on the device: raw vertices and triangles in, packed triangles
// and index boxes out; *bad set on a bad index or coordinate.
hipError_t launch_mesh_prepare(hipStream_t st, const NodeGrid3& G, const double* verts, int64_t nv,
                               const int32_t* tris, int64_t nt, double* tv, int32_t* tb, int32_t* bad);

// Mesh2SDF on the device: mesh_prepare packs the triangles and their index
// boxes on the host (false on a bad index or coordinate), launch_mesh_sdf
// writes the collider grid's SDF, collider_merge folds it into the scene's
// collider SDF (allocated at MAX_DISTANCE on first use).
bool mesh_prepare(const NodeGrid3& G, const double* verts, int64_t nv, const int32_t* tris, int64_t nt,
                  std::vector<double>& tv, std::vector<int32_t>& tb, double& fill);
hipError_t launch_mesh_sdf(hipStream_t st, const NodeGrid3& G, const double* tv_dev, const int32_t* tb_dev,
                           int64_t nt, double fill, double* out);
hipError_t collider_merge(fe_scene* s, const NodeGrid3& G, const double* c_dev);

// k_sources.hip: Emitter::Emitt of every emitter and InitializeParticles from
// the fluid SDF; *over is set (and an error returned) when the particles would
// exceed the capacity
hipError_t launch_emit(fe_scene* s, bool* over);
hipError_t reserve_particles(fe_scene* s, int64_t capacity);

// k_learn.hip: x0 = alpha * the network's prediction in the scene's pressure
hipError_t net_warm_start(fe_net* net, fe_scene* s);
hipError_t launch_initialize_from_sdf(fe_scene* s, int ppb, uint64_t seed, bool* over);
void sources_free(fe_scene* s);

// k_diffusion.hip: BackwardEulerDiffusionSolver (without the trailing boundary condition)
hipError_t launch_diffusion(fe_scene* s, double dt, const fe_solver_desc* d);
void diffusion_free(fe_scene* s);

}  // namespace fe
