/* fluide_hip.h -- C ABI of the MI355X (gfx950) backend for the Fluide_Engine
 * hybrid particle-grid substep.
 *
 * The reference has no C ABI: its hot path sits behind C++ classes
 * (PhysicsAnimation / HybridSimulator / PICSimulator / FLIPSimulator /
 * APICSimulator and the experimental CudaPICSimulator).  These entry points are
 * what those classes (and the PyFluidEngine module) call instead, at scene
 * granularity: the reference's per-stage objects take host Array3s, which a GPU
 * backend cannot sit behind without a device round trip per stage.
 *
 * Conventions: every function returns 0 (FE_OK) or a negative FE_E_* code and
 * never throws; fe_last_error() holds the message.  Grids are x-fastest
 * (i + sx*(j + sy*k)), like Array3 (src/common/array3.hpp:65-85).  Particle
 * vectors are interleaved xyz, like std::vector<Vector3<double>> in
 * ParticleSystem (src/particle_systems/particle_system.hpp:78-79), but FP32.
 * All device memory is owned by the scene; host buffers by the caller.
 * Calls on one scene are synchronous and must come from one host thread at a
 * time; distinct scenes are independent.
 */
#ifndef FLUIDE_HIP_H
#define FLUIDE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum fe_status {
    FE_OK = 0,
    FE_E_INVALID = -1,     /* bad argument, size or state */
    FE_E_HIP = -2,         /* a HIP runtime call failed (message names it) */
    FE_E_NOMEM = -3,       /* device or host allocation failed */
    FE_E_UNSUPPORTED = -4, /* a feature this backend does not implement */
    FE_E_NO_DEVICE = -5,   /* no usable gfx950 device */
    FE_E_NONFINITE = -6    /* NaN or Inf in the state: uploaded particles, or face velocities
                              when a frame computes its CFL number (the state has diverged) */
};

/* Transfer scheme: PICSimulator (src/pic_simulator.hpp:27), FLIPSimulator
 * (src/flip_simulator.hpp:7), APICSimulator (src/apic_simulator.hpp:8). */
enum fe_method { FE_PIC = 0, FE_FLIP = 1, FE_APIC = 2 };

/* Cell markers: FluidMarker (src/grid_systems/fluid_markers.hpp:8-13). */
enum fe_marker { FE_AIR = 0, FE_FLUID = 1, FE_BOUNDRY = 2 };

/* Linear solver for the pressure projection.
 * JACOBI: JacobiIterationSolver / CudaJacobiIterationSolver semantics
 *   (src/linear_system/jacobi_iteration_solver.cpp:14-37): max_iterations sweeps,
 *   stop when ||b-Ax||_2 < tolerance, checked every check_interval sweeps.
 * CG: ConjugateGradientSolver (src/linear_system/conjugate_gradient_solver.cpp:14-55),
 *   identity preconditioner, residual replacement every 50 iterations or when
 *   sigma grows (and, for the FP32 vectors, whenever ||r|| fell 100x since the
 *   last replacement); stop on sigma <= tol^2, or ||r||/||b|| <= tol when
 *   relative, accepted only on a replaced residual.
 * PCG_JACOBI: CG with the diagonal preconditioner (no reference counterpart).
 * MGPCG: CG preconditioned by one geometric multigrid V-cycle per iteration
 *   (no reference counterpart; same stopping rules and final residual as CG).
 * Vectors are FP32; the solution and every dot product and norm are FP64. */
enum fe_solver_kind { FE_SOLVER_JACOBI = 0, FE_SOLVER_CG = 1, FE_SOLVER_PCG_JACOBI = 2, FE_SOLVER_MGPCG = 3 };

typedef struct fe_solver_desc {
    int32_t kind;
    int32_t max_iterations;
    int32_t check_interval; /* Jacobi residual check interval (reference default 5) */
    int32_t relative;       /* CG: relative stopping rule */
    int32_t warm_start;     /* FE_WARM_START_*: 1 = from the previous pressure (reference behaviour) */
    int32_t reserved;
    double tolerance;
} fe_solver_desc;

typedef struct fe_solve_stats {
    int32_t iterations;
    int32_t converged;
    double rel_residual; /* ||b - A p||_2 / ||b||_2 on the final p, in FP64 */
    double abs_residual;
    double ms;           /* device time of build + solve + apply, HIP events */
} fe_solve_stats;

/* Scene description: HybridSimulator(gridSize, domain) (src/hybrid_simulator.cpp:4-25)
 * plus the simulator settings the reference keeps in members. */
typedef struct fe_scene_desc {
    int32_t n[3];         /* cells per axis */
    int32_t method;       /* fe_method */
    double origin[3];     /* BoundingBox3D origin */
    double size[3];       /* BoundingBox3D size; h = size / n */
    double blend;         /* FLIP PIC weight (_blendingFactor, src/flip_simulator.cpp:5) */
    double max_cfl;       /* _maxCfl (src/hybrid_simulator.cpp:21) */
    double gravity[3];    /* DirectionalField strength (src/forces/directional_field.cpp:13-16) */
    double viscosity;     /* >= 0; > 0 adds BackwardEulerDiffusionSolver to the substep */
    double density;       /* unused by the reference's pressure solve (kept for the API) */
    int64_t max_particles;
} fe_scene_desc;

/* Grid fields for fe_grid_download / fe_grid_upload. */
enum fe_grid_field {
    FE_GRID_U = 0,       /* float, (nx+1)*ny*nz */
    FE_GRID_V = 1,       /* float, nx*(ny+1)*nz */
    FE_GRID_W = 2,       /* float, nx*ny*(nz+1) */
    FE_GRID_MU = 3,      /* uint8 face markers of u (Fluid3::xMarkers) */
    FE_GRID_MV = 4,
    FE_GRID_MW = 5,
    FE_GRID_SU = 6,      /* float, FLIP snapshot of u (FLIPSimulator::_velocityResiduals) */
    FE_GRID_SV = 7,
    FE_GRID_SW = 8,
    FE_GRID_SDF = 9,     /* double, fluid SDF at cell centres (Fluid3::sdf) */
    FE_GRID_FLAGS = 10,  /* uint8 cell markers (fe_marker) */
    FE_GRID_PRESSURE = 11, /* double, pressure (LinearSystem::x) */
    FE_GRID_RHS = 12,    /* float, divergence right-hand side (LinearSystem::b) */
    FE_GRID_DMU = 13,    /* uint8 fe_marker of the u faces in the last diffusion solve
                            (BackwardEulerDiffusionSolver::_fluidMarkers); download only */
    FE_GRID_DMV = 14,
    FE_GRID_DMW = 15,
    FE_GRID_COLLIDER_SDF = 16 /* double, nx*ny*nz: the collider SDF at the nodes origin + h*(i,j,k)
                                 (ColliderCollection::BuildSdf, src/3d/collider_collection.cpp:33-50);
                                 uploading it adds the collider to the scene */
};

/* Stages of one substep, for fe_run_stage (PICSimulator::OnAdvanceTimeStep,
 * src/pic_simulator.cpp:111-237). */
enum fe_stage {
    FE_STAGE_SORT = 0,          /* bin particles by cell (no reference counterpart) */
    FE_STAGE_P2G = 1,           /* TransferParticles2Grid */
    FE_STAGE_SDF = 2,           /* BuildSignedDistanceField (+ cell markers) */
    FE_STAGE_EXTRAPOLATE = 3,   /* ExtrapolateVelocityToAir */
    FE_STAGE_BC = 4,            /* ApplyBoundryCondition (walls; the collider branches with a collider) */
    FE_STAGE_GRAVITY = 5,       /* ExternalForce::ApplyExternalForce */
    FE_STAGE_PRESSURE = 6,      /* SinglePhasePressureSolver::Solve */
    FE_STAGE_G2P = 7,           /* TransferGrid2Particles */
    FE_STAGE_MOVE = 8,          /* MoveParticles */
    FE_STAGE_MARKERS = 9,       /* SinglePhasePressureSolver::BuildMarkers only */
    FE_STAGE_DIFFUSION = 10,    /* BackwardEulerDiffusionSolver::Solve (ComputeDiffusion
                                   without its boundary condition); solver NULL = the scene's */
    FE_STAGE_EMIT = 11          /* Emitter::Emitt of every emitter (first step of a substep) */
};

/* Initial guess of a pressure solve (fe_solver_desc.warm_start). */
enum fe_warm_start {
    FE_WARM_START_ZERO = 0,
    FE_WARM_START_PREVIOUS = 1, /* the previous solve's pressure (Array3::Resize keeps it, src/common/array3.hpp:189-197) */
    FE_WARM_START_LEARNED = 2   /* alpha * the warm-start network's prediction from (b, markers), alpha minimising
                                   ||b - alpha A p||; needs fe_scene_set_warm_start_net.  No reference counterpart. */
};

typedef struct fe_ctx fe_ctx;
typedef struct fe_scene fe_scene;

/* Context: one HIP device.  device = -1 picks the current device. */
int fe_ctx_create(int device, fe_ctx** out);
int fe_ctx_destroy(fe_ctx* ctx);
const char* fe_last_error(const fe_ctx* ctx);
/* Name of the device the context runs on ("gfx950..."), for logs. */
int fe_ctx_device_name(const fe_ctx* ctx, char* buf, size_t len);

/* Scene: device state of HybridSimulator (src/hybrid_simulator.cpp:4-25). */
int fe_scene_create(fe_ctx* ctx, const fe_scene_desc* desc, fe_scene** out);
int fe_scene_destroy(fe_scene* scene);
int fe_scene_sync(fe_scene* scene);
/* Simulator settings changed after construction: SetBlendingFactor
 * (src/flip_simulator.hpp:14-16), SetMaxClf, SetViscosity, SetDensity
 * (src/hybrid_simulator.hpp:37-40) and the DirectionalField forces of
 * AddExternalForce, summed (src/pic_simulator.cpp:41-44).  viscosity > 0 runs
 * ComputeDiffusion in every substep (src/pic_simulator.cpp:214-222). */
int fe_scene_set_params(fe_scene* scene, double blend, double max_cfl, const double gravity[3], double viscosity,
                        double density);

/* The linear solver of the scene's BackwardEulerDiffusionSolver, fixed in the
 * reference to CudaJacobiIterationSolver(1000, 5, 1e-14)
 * (src/fluid_solvers/backward_euler_diffusion_solver.cpp:10), which is the
 * default here; kind must be FE_SOLVER_JACOBI (else FE_E_UNSUPPORTED).
 * Replaces PICSimulator::SetDiffusionSolver (src/pic_simulator.hpp:36) for the
 * one diffusion solver the reference has. */
int fe_scene_set_diffusion_solver(fe_scene* scene, const fe_solver_desc* solver);
/* Per component (u, v, w) of the last diffusion solve: Jacobi sweeps run,
 * converged flag and the last checked ||b - A x||_2 (-1 when none was checked). */
int fe_last_diffusion_stats(const fe_scene* scene, fe_solve_stats stats[3]);

/* Particles: ParticleSystem "postion"/"velocity" attributes
 * (src/hybrid_simulator.hpp:50-51) and the APIC affine rows _cX/_cY/_cZ
 * (src/apic_simulator.hpp:21-23; C is 9 floats per particle: cx, cy, cz rows,
 * or NULL for zero).  Upload replaces all particles (FE_E_NONFINITE, and
 * nothing replaced, if a value is NaN or Inf); download returns them in
 * upload order. */
int fe_particles_upload(fe_scene* scene, const float* x, const float* v, const float* C, int64_t count);
int fe_particles_download(fe_scene* scene, float* x, float* v, float* C, int64_t count);
int fe_particles_count(const fe_scene* scene, int64_t* count);

/* One substep: PICSimulator::OnAdvanceTimeStep (src/pic_simulator.cpp:111-154). */
int fe_substep(fe_scene* scene, double dt, const fe_solver_desc* solver, fe_solve_stats* stats);

/* One frame: PhysicsAnimation::AdvanceTimeStep (src/physics_animation.cpp:64-92),
 * CFL-adaptive substeps; substeps_out may be NULL.  FE_E_NONFINITE when a
 * substep's CFL evaluation finds a NaN or Inf face velocity (the reference has
 * no such check and would go on with the diverged state). */
int fe_advance_frame(fe_scene* scene, double frame_dt, const fe_solver_desc* solver, int32_t* substeps_out);

/* Batches of independent scenes (no reference counterpart; the reference runs
 * one simulator per process).  A batch lists scenes of one context (not owned;
 * they must outlive it); fe_batch_advance_frame advances every scene by one
 * frame (fe_advance_frame) using `workers` host threads, largest scenes first,
 * each scene on its own stream.  substeps and ms (may be NULL) receive per
 * scene, in list order, the substep count and the wall time of its frame.
 * Returns the first failing scene's code. */
typedef struct fe_batch fe_batch;

/* Substeps as captured HIP graphs (no reference counterpart).  With enable = 1
 * each substep of fe_substep / fe_advance_frame / fe_batch_advance_frame that
 * qualifies (no emitter, collider or viscosity, a CG-family solver without the
 * learned warm start, the kernel timer off, after one eager substep) runs as
 * two graph launches instead of about 130 kernel launches: the stages up to
 * the pressure solve with the solve's first iterations, then the pressure
 * application and the advection.  The host adds iterations between them only
 * when the solve has not stopped.  Results are the eager substep's, bit for
 * bit.  In graph substeps fe_solve_stats.ms is the whole substep's device time
 * and fe_stage_times reports zeros.  Graphs are captured on first use and
 * again when a setting they hold changes; enable = 0 releases them.  Counters:
 * graph substeps run, captures made, substeps that needed host iterations
 * (pointers may be NULL). */
int fe_scene_set_graph(fe_scene* scene, int32_t enable);
int fe_scene_graph_stats(const fe_scene* scene, int64_t* substeps, int64_t* captures, int64_t* resumed);

int fe_batch_create(fe_ctx* ctx, fe_scene* const* scenes, int32_t count, int32_t workers, fe_batch** out);
int fe_batch_destroy(fe_batch* batch);
int fe_batch_advance_frame(fe_batch* batch, double frame_dt, const fe_solver_desc* solver, int32_t* substeps,
                           double* ms);

/* Colliders (AddCollider, src/pic_simulator.cpp:46-49): upload the collision
 * SDF with fe_grid_upload(FE_GRID_COLLIDER_SDF).  With a collider, cells outside
 * the fluid whose centre samples it negative are BOUNDRY, the boundary
 * condition runs the collider branches of ConstrainVelocity
 * (src/fluid_solvers/blocked_boundry_condition_solver.cpp:18-291) and moved
 * particles inside it are pushed to its zero level set along its gradient with
 * the inward normal velocity removed.  fe_scene_clear_colliders removes it. */
int fe_scene_clear_colliders(fe_scene* scene);

/* AddCollider(TriangleMeshCollider(gridSize, gridOrigin, gridSpacing, mesh))
 * (src/pic_simulator.cpp:66-69, src/3d/triangle_mesh_collider.cpp:5-9): the
 * collider's SDF is Mesh2SDF::Build (src/3d/mesh_2_sdf.cpp:14-95) of the mesh
 * on its own node grid (grid_size nodes, x fastest; computed on the device),
 * and the scene's collider SDF becomes the running minimum of the colliders'
 * SDFs sampled at its nodes (ColliderCollection::BuildSdf,
 * src/3d/collider_collection.cpp:33-50).  vertices: 3 doubles each;
 * triangles: 3 vertex indices each (the reference's point1Idx..point3Idx).
 * Synchronous.  FE_E_INVALID for an index out of range, a grid size < 1, a
 * spacing <= 0, or a vertex coordinate / spacing[0] beyond the int range. */
int fe_scene_add_mesh_collider(fe_scene* scene, const int32_t grid_size[3], const double grid_origin[3],
                               const double grid_spacing[3], const double* vertices, int64_t vertex_count,
                               const int32_t* triangles, int64_t triangle_count);
/* Mesh2SDF::Build alone: the mesh's SDF on a node grid into sdf (host memory,
 * grid_size[0]*grid_size[1]*grid_size[2] doubles, x fastest).  Synchronous. */
int fe_mesh_sdf(fe_ctx* ctx, const int32_t grid_size[3], const double grid_origin[3], const double grid_spacing[3],
                const double* vertices, int64_t vertex_count, const int32_t* triangles, int64_t triangle_count,
                double* sdf);
/* fe_mesh_sdf without device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (serving loops, harnesses
 * that forbid hipMalloc), e.g. to rebuild a moving collider's SDF every frame.
 * Same arguments and results as fe_mesh_sdf, bit for bit.  Its device buffers
 * live in one static array per GPU, held until the process ends: at most
 * 65536 vertices, 65536 triangles and 2^21 grid nodes per call (FE_E_INVALID
 * beyond, before any device work).  Synchronous, on the null stream.  Calls
 * that overlap on one GPU share that array and race: serialise them. */
int fe_mesh_sdf_noalloc(fe_ctx* ctx, const int32_t grid_size[3], const double grid_origin[3],
                        const double grid_spacing[3], const double* vertices, int64_t vertex_count,
                        const int32_t* triangles, int64_t triangle_count, double* sdf);

/* Candidate search of the fluid SDF (BuildSignedDistanceField,
 * src/pic_simulator.cpp:452-475).  FE_SDF_SEARCH_CELLS (the default) scans the
 * cells within the particle radius r of each centre: every negative value is
 * the reference's, positive ones may be smaller where the reference's buckets
 * miss particles between r and 2r.  FE_SDF_SEARCH_BUCKETS takes exactly the
 * reference's candidates (8 hash buckets of width 2r,
 * src/particle_systems/hash_grid_particle_searcher.cpp:154-204), so every
 * value is the reference's, at about three times the cost; it is used
 * whenever the scene has a collider (the extrapolation into the collider
 * averages positive values). */
enum fe_sdf_search { FE_SDF_SEARCH_CELLS = 0, FE_SDF_SEARCH_BUCKETS = 1 };
int fe_scene_set_sdf_search(fe_scene* scene, int32_t mode);

/* Scene sources (PICSimulator::InitializeFromTriangleMesh and AddEmitter,
 * src/pic_simulator.cpp:33-49).  Both put particles_per_block particles into
 * each picked cell at pos + h * u per axis, pos = the cell's corner, u
 * uniform; u comes from a counter-based generator (SplitMix64) keyed by
 * (seed, call, particle, axis) in place of the reference's std::rand(), so
 * runs repeat.  Particles are appended in the reference's ForEachIndex order
 * (i outer, k inner).  The particle storage grows past max_particles when they
 * need it (the reference's particle system grows without bound).
 *
 * fe_scene_initialize_from_mesh: Mesh2SDF of the mesh into the fluid SDF
 * (cell-centred grid), then InitializeParticles (src/pic_simulator.cpp:521-548):
 * a cell is picked when that SDF sampled at its corner is negative; the new
 * particles are at rest.
 *
 * fe_scene_add_volume_emitter: VolumeEmitter(mesh, resolution, domain,
 * particlesPerBlock, velocity, variance) (src/emmiter/volume_emmiter.cpp:12-22;
 * the variance is unused there too).  Its SDF is Mesh2SDF on the node grid of
 * the domain at that resolution, which must equal the scene's (Emitt reads the
 * fluid and collider SDFs at the emitter's indices).  At the start of every
 * substep (src/pic_simulator.cpp:156-163) a node emits when emitterSdf <= 0,
 * fluidSdf >= 0 and colliderSdf > 0 there, the new particles moving with
 * velocity.  Emitt runs before BuildCollider, whose SDF starts at 0: nothing is
 * emitted in a scene's first substep. */
int fe_scene_initialize_from_mesh(fe_scene* scene, const double* vertices, int64_t vertex_count,
                                  const int32_t* triangles, int64_t triangle_count, int32_t particles_per_block,
                                  uint64_t seed);
int fe_scene_add_volume_emitter(fe_scene* scene, const int32_t resolution[3], const double domain_origin[3],
                                const double domain_size[3], const double* vertices, int64_t vertex_count,
                                const int32_t* triangles, int64_t triangle_count, int32_t particles_per_block,
                                const double velocity[3], uint64_t seed);
/* VolumeEmitter::InitializeFromSdf (src/emmiter/volume_emmiter.cpp:50-56): the
 * emitter's SDF given directly (resolution[0]*[1]*[2] doubles, x fastest). */
int fe_scene_add_volume_emitter_sdf(fe_scene* scene, const int32_t resolution[3], const double domain_origin[3],
                                    const double domain_size[3], const double* sdf, int32_t particles_per_block,
                                    const double velocity[3], uint64_t seed);
int fe_scene_clear_emitters(fe_scene* scene);
/* Particle storage: grow it to at least capacity particles (the current ones
 * are kept; no counterpart in the reference, whose particle system grows by
 * itself), and read the current capacity. */
int fe_scene_reserve_particles(fe_scene* scene, int64_t capacity);
int fe_scene_particle_capacity(const fe_scene* scene, int64_t* capacity);

/* Kernel variants and their autotuner (no reference counterpart).  P2G runs
 * as a gather over 2x2x2 face tiles (FE_P2G_TILES, four lanes per tile), a
 * gather with one thread per face (FE_P2G_FACES) or a scatter over 8x8x8
 * face blocks (FE_P2G_SCATTER, 64-bit integer LDS sums); all three add the
 * same fixed-point terms, so their sums are equal.  G2P and advection sample the face arrays directly
 * (FE_G2P_PAIRS) or from a quad layout built first (FE_G2P_QUADS).  The
 * variants of each stage give bit-identical results, so the choice only
 * changes speed.  fe_scene_autotune times every variant `reps` times on the
 * scene's current particles (it sorts them if needed and leaves the face
 * velocities, markers and FLIP snapshot holding their P2G, which the next
 * substep recomputes anyway; the particles are untouched), keeps the fastest
 * of each stage and reports the choice and the median times in ms (times may
 * be NULL: p2g_ms[variant], g2p_ms[variant]).  Defaults: SCATTER, QUADS. */
enum fe_p2g_variant { FE_P2G_TILES = 0, FE_P2G_FACES = 1, FE_P2G_SCATTER = 2 };
enum fe_g2p_variant { FE_G2P_QUADS = 0, FE_G2P_PAIRS = 1 };
/* The multigrid V-cycle of FE_SOLVER_MGPCG on the levels above its
 * one-workgroup tail: one descent and one ascent kernel per level over LDS
 * tiles with recomputed halos (FE_MG_FUSED, the default), or a kernel per
 * sweep, residual, restriction and prolongation (FE_MG_SPLIT).  Both apply the
 * same preconditioner (the same values per cell); they sum r.z over
 * different blocks, so FP64 roundings of the CG scalars may differ.  Not
 * autotuned. */
enum fe_mg_variant { FE_MG_FUSED = 0, FE_MG_SPLIT = 1 };
typedef struct fe_tuning {
    int32_t p2g; /* fe_p2g_variant */
    int32_t g2p; /* fe_g2p_variant */
    int32_t mg;  /* fe_mg_variant */
} fe_tuning;
typedef struct fe_tuning_times {
    double p2g_ms[3];
    double g2p_ms[2];
} fe_tuning_times;
int fe_scene_get_tuning(const fe_scene* scene, fe_tuning* out);
int fe_scene_set_tuning(fe_scene* scene, const fe_tuning* tuning);
int fe_scene_autotune(fe_scene* scene, int32_t reps, fe_tuning* out, fe_tuning_times* times);

/* Learned warm start (no reference counterpart).  A 4-level convolutional
 * encoder-decoder (3x3x3 kernels, 16-64 channels, BF16 weights and activations
 * with FP32 accumulation on the matrix cores) maps a pressure system's
 * right-hand side and cell markers to a pressure shape; a solve with
 * warm_start = FE_WARM_START_LEARNED starts CG from alpha times it, alpha
 * minimising ||b - alpha A p||, so the initial residual never exceeds ||b||,
 * and stops at the usual tolerance.  The network is trained on the device
 * (Adam on FP32 master weights) on samples taken from solved scenes: each
 * sample is the scene's last system (b, markers) and its solution divided by
 * its RMS over FLUID cells; the loss is the mean squared error over FLUID
 * cells.  Grids need every axis a multiple of 8.
 * Parameters are FP32 in this order: for each convolution (enc0 16->16,
 * enc1 16->32, enc2 32->48, enc3 48->64, dec2 112->48, dec1 80->32,
 * dec0 48->16) W[t][ci][co] (t = (dz+1)*9 + (dy+1)*3 + (dx+1)) then its bias,
 * then the 1x1 head's 16 weights and bias. */
typedef struct fe_net fe_net;
int fe_net_create(fe_ctx* ctx, uint64_t seed, fe_net** net);
int fe_net_destroy(fe_net* net);
int fe_net_param_count(int64_t* count);
int fe_net_get_params(fe_net* net, float* params, int64_t count);
int fe_net_set_params(fe_net* net, const float* params, int64_t count);
int fe_net_add_sample(fe_net* net, fe_scene* scene); /* the scene's last pressure system and solution */
int fe_net_sample_count(const fe_net* net, int64_t* count);
int fe_net_clear_samples(fe_net* net);
/* steps Adam steps, one sample each (picked by seed); losses (steps floats) may be NULL */
int fe_net_train(fe_net* net, int32_t steps, double learning_rate, uint64_t seed, float* losses);
/* Training loss.  FE_NET_LOSS_ENERGY (default): the A-norm error left after
 * the best scaling of the prediction p, relative to the solution x's,
 * 1 - (b.p)^2 / ((p.A.p)(b.x)) over FLUID cells -- what a CG warm start needs;
 * FE_NET_LOSS_MSE: the mean over FLUID cells of (p - x / rms(x))^2. */
enum fe_net_loss { FE_NET_LOSS_MSE = 0, FE_NET_LOSS_ENERGY = 1 };
int fe_net_set_loss(fe_net* net, int32_t mode);
/* One forward and backward pass on host arrays (x fastest): b, markers and the
 * target (the solution x for the energy loss, the normalised one for MSE) on a
 * grid of spacing h (NULL: 1/n per axis); the loss, the gradient of every
 * parameter (fe_net_param_count floats, may be NULL) and the prediction
 * (n[0]*n[1]*n[2] floats, may be NULL). */
int fe_net_loss_and_grads(fe_net* net, const int32_t n[3], const double h[3], const float* b, const uint8_t* flags,
                          const float* target, double* loss, float* grads, float* prediction);
/* The network used by solves with FE_WARM_START_LEARNED (NULL detaches); the
 * scene keeps a pointer, so the network must outlive the scenes that use it
 * (or be detached first).  Scenes advanced on several host threads may share
 * one network (its warm start is serialised).  fe_scene_warm_start_scale: the
 * scale alpha of the scene's last learned warm start. */
int fe_scene_set_warm_start_net(fe_scene* scene, fe_net* net);
/* How alpha is chosen: FE_NET_SCALE_ENERGY (default) minimises the A-norm of
 * the initial error (alpha = b.p / p.A.p, the quantity CG reduces);
 * FE_NET_SCALE_RESIDUAL minimises ||b - alpha A p||. */
enum fe_net_scale { FE_NET_SCALE_RESIDUAL = 0, FE_NET_SCALE_ENERGY = 1 };
int fe_net_set_scale_mode(fe_net* net, int32_t mode);
int fe_scene_warm_start_scale(const fe_scene* scene, double* alpha);

/* Fluid surface: HybridSimulator::GetSurface -> MarchingCubesSolver::BuildSurface
 * (src/3d/marching_cubes_solver.cpp:25-51) on the current fluid SDF, iso value 0.
 * fe_surface_build extracts the mesh on the device and returns its counts;
 * fe_surface_download copies it out: vertices and per-vertex normals as
 * 3 doubles each, triangles as 3 vertex indices (counter-clockwise seen from
 * outside the fluid).  Cube vertices come first, then the caps on the six
 * boundary node planes.  Grids need fewer than 2^31 node edges. */
int fe_surface_build(fe_scene* scene, int64_t* vertex_count, int64_t* triangle_count);
int fe_surface_download(fe_scene* scene, double* vertices, double* normals, int32_t* triangles);
/* Triangulation of the surface: FE_SURFACE_REFERENCE (the default) uses the
 * reference's case tables for cubes and boundary squares
 * (src/3d/marching_cubes_solver.cpp:652-952), so every cube and cap square is
 * triangulated as the reference does, ambiguous cases and their holes
 * included; FE_SURFACE_CLOSED derives every case from one face rule, which
 * closes the surface everywhere.  Both place the same vertices.  In both,
 * when the scene's last fluid SDF came from the cell scan
 * (FE_SDF_SEARCH_CELLS), fe_surface_build first rebuilds it from the same
 * particles with the reference's 8-bucket search, so the positive values that
 * place the vertices are the reference's (the markers do not change). */
enum fe_surface_mode { FE_SURFACE_REFERENCE = 0, FE_SURFACE_CLOSED = 1 };
int fe_scene_set_surface_mode(fe_scene* scene, int32_t mode);

/* PICSimulator::Cfl (src/pic_simulator.cpp:80-97).  NaN face velocities are
 * skipped by the maximum, as by the reference's std::max; any NaN or Inf face
 * value makes the call return FE_E_NONFINITE (cfl_out still holds the value). */
int fe_cfl(fe_scene* scene, double dt, double* cfl_out);

/* Statistics of the most recent pressure solve (SinglePhasePressureSolver
 * keeps no statistics; this is the backend's report). */
int fe_last_solve_stats(const fe_scene* scene, fe_solve_stats* stats);

/* Run one stage of the substep on the current state (tests, profiling). */
int fe_run_stage(fe_scene* scene, int32_t stage, double dt, const fe_solver_desc* solver, fe_solve_stats* stats);

/* Grid state transfer; bytes must equal the field's size. */
int fe_grid_download(fe_scene* scene, int32_t field, void* dst, size_t bytes);
int fe_grid_upload(fe_scene* scene, int32_t field, const void* src, size_t bytes);
int fe_grid_bytes(const fe_scene* scene, int32_t field, size_t* bytes);

/* Cell index of each particle (i + nx*(j + ny*k)) as the sort bins them, in
 * upload order.  Bit-exact with floor((p - origin)/h) in FP64 clamped to the grid. */
int fe_particle_cells(fe_scene* scene, int32_t* cells, int64_t count);

/* Analytic scene seeding (the reference's InitializeParticles,
 * src/pic_simulator.cpp:521-548, draws std::rand(); this is the deterministic
 * replacement): every cell whose centre lies inside the union of the boxes
 * and spheres gets 2x2x2 particles, one per sub-cell at
 * cell_lo + (s + U)*h/2 with U from splitmix64(seed, cell, sub-cell).
 * Host only (no GPU).  Call with x == NULL to get the count. */
typedef struct fe_seed_shape {
    int32_t kind;       /* 0 = box [lo, hi), 1 = sphere (centre, radius in lo[0..2], hi[0]) */
    int32_t reserved;
    double lo[3];
    double hi[3];
} fe_seed_shape;
int fe_seed_particles(const fe_scene_desc* desc, const fe_seed_shape* shapes, int32_t nshapes, uint64_t seed,
                      float* x, int64_t capacity, int64_t* count);

/* Kernel timing of the last fe_substep / fe_run_stage, in ms per stage
 * (index = fe_stage), from HIP events on the scene's stream. */
int fe_stage_times(const fe_scene* scene, double* ms, int32_t count);

/* Per-kernel timer (no reference counterpart; for roofline measurement):
 * HIP events on the scene's stream around every launch of one kernel.
 * fe_kernel_timer(scene, FE_KERNEL_NONE) switches it off. */
enum fe_kernel_id {
    FE_KERNEL_NONE = 0,
    FE_KERNEL_CG_SPMV = 1,    /* d = z + beta d, q = A d, d.q */
    FE_KERNEL_CG_UPDATE = 2,  /* x += alpha d, r -= alpha q, z = M^-1 r, r.z */
    FE_KERNEL_P2G = 3,        /* per-face gather (3 launches per substep, one per component) */
    FE_KERNEL_G2P_MOVE = 4,   /* fused G2P + advection */
    FE_KERNEL_SDF = 5,        /* fluid SDF */
    FE_KERNEL_JACOBI = 6,     /* one Jacobi sweep */
    FE_KERNEL_MG_SMOOTH = 7,  /* FE_MG_SPLIT: smoothing sweep on the finest level (3 per V-cycle) */
    FE_KERNEL_MG_DOWN = 8,    /* FE_MG_FUSED: finest level's descent (sweep, residual, restriction) */
    FE_KERNEL_MG_UP = 9       /* FE_MG_FUSED: finest level's ascent (prolongation, two sweeps, r.z) */
};
int fe_kernel_timer(fe_scene* scene, int32_t kernel_id);
/* Total device ms and launch count since the last reset (reset != 0 clears). */
int fe_kernel_timer_read(fe_scene* scene, double* total_ms, int64_t* launches, int32_t reset);

/* One plain device copy of `bytes` bytes (a multiple of 16) between two
 * scratch buffers the call allocates and frees; *ms = its duration.  No
 * reference counterpart: the known-size kernel that hardware byte counters
 * (FETCH_SIZE, WRITE_SIZE) are scaled by. */
int fe_copy_probe(fe_ctx* ctx, int64_t bytes, double* ms);

/* Library version string. */
const char* fe_version(void);

#ifdef __cplusplus
}
#endif

#endif /* FLUIDE_HIP_H */
