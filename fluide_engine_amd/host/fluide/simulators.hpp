// simulators.hpp -- the reference's simulator classes over the gfx950 backend.
//
// HipPICSimulator / HipFLIPSimulator / HipAPICSimulator keep the method set of
// PhysicsAnimation (src/physics_animation.hpp:6-38), HybridSimulator
// (src/hybrid_simulator.hpp:12-52), PICSimulator (src/pic_simulator.hpp:29-43)
// and FLIPSimulator (src/flip_simulator.hpp:14-16), so code written against
// CudaPICSimulator / CudaFLIPSimulator / CudaAPICSimulator
// (src/cuda_pic_simulator.hpp:29-58) compiles against these through the
// Cuda* aliases at the end.  The state lives on the GPU in one fe_scene
// (include/fluide_hip.h); host copies are made only by the explicit
// particle/grid accessors.
//
// Triangle-mesh scenes (InitializeFromTriangleMesh), TriangleMeshColliders and
// VolumeEmitters run on the device; external forces other than a uniform
// DirectionalField are not on this backend (Unsupported).  GetSurface extracts
// the surface on the GPU; OBJManager and SimulationRunner complete the
// reference's frame loop.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "fluide_hip.h"

namespace fluide {

// src/common/vector3.hpp (the members the simulators use)
template <typename T>
struct Vector3 {
    T x, y, z;
    Vector3(T v = T(0)) : x(v), y(v), z(v) {}
    Vector3(T x_, T y_, T z_) : x(x_), y(y_), z(z_) {}
};

// src/3d/bounding_box_3d.hpp
class BoundingBox3D {
  public:
    BoundingBox3D(Vector3<double> origin = 0, Vector3<double> size = 1) : _origin(origin), _size(size) {}
    bool IsInside(Vector3<double> p) const;
    Vector3<double> Corner(size_t idx) const;
    Vector3<double> GetOrigin() const { return _origin; }
    Vector3<double> GetSize() const { return _size; }
    void SetOrigin(Vector3<double> origin) { _origin = origin; }
    void SetSize(Vector3<double> size) { _size = size; }

  private:
    Vector3<double> _origin;
    Vector3<double> _size;
};

// src/common/frame.hpp: frame index and interval; time = index * interval.
class Frame {
  public:
    Frame() = default;
    explicit Frame(double timeIntervalInSeconds) : _timeIntervalInSeconds(timeIntervalInSeconds) {}
    void Advance() { ++_index; }
    void Advance(int numberOfFrames) { _index += numberOfFrames; }
    Frame& operator++() {
        Advance();
        return *this;
    }
    Frame operator++(int) {
        Frame r = *this;
        Advance();
        return r;
    }
    unsigned long GetIndex() const { return _index; }
    double GetTimeIntervalInSeconds() const { return _timeIntervalInSeconds; }
    double GetTimeInSeconds() const { return _index * _timeIntervalInSeconds; }

  private:
    unsigned long _index = 0;
    double _timeIntervalInSeconds = 1 / 60.0f;
};

// src/forces/external_force.hpp, src/forces/directional_field.hpp
class ExternalForce {
  public:
    virtual ~ExternalForce() = default;
    virtual Vector3<double> Sample(const Vector3<double>& position) const = 0;
};

class DirectionalField : public ExternalForce {
  public:
    explicit DirectionalField(Vector3<double> strength) : _strength(strength) {}
    Vector3<double> Sample(const Vector3<double>&) const override { return _strength; }
    void SetStrength(Vector3<double> strength) { _strength = strength; }

  private:
    Vector3<double> _strength;
};

// src/forces/point_field.hpp: strength / max(|x - origin|, 1) towards the
// origin.  Not uniform, so AddExternalForce rejects it unless its strength is 0.
class PointField : public ExternalForce {
  public:
    PointField(Vector3<double> origin = 0, double strength = 0, double strengthFallOff = 1)
        : _origin(origin), _strength(strength), _strengthFallOff(strengthFallOff) {}
    Vector3<double> Sample(const Vector3<double>& position) const override;
    void SetPosition(Vector3<double> origin) { _origin = origin; }
    void SetStrength(double strength) { _strength = strength; }
    void SetStrengthFallOff(double strengthFallOff) { _strengthFallOff = strengthFallOff; }
    Vector3<double> GetOrigin() const { return _origin; }
    double GetStrength() const { return _strength; }
    double GetStrengthFallOff() const { return _strengthFallOff; }

  private:
    Vector3<double> _origin;
    double _strength, _strengthFallOff;
};

// Host copy of a cell-centred scalar grid (src/3d/scalar_grid3d.hpp): x-fastest.
struct ScalarGrid3D {
    Vector3<size_t> size;
    Vector3<double> origin, spacing;
    std::vector<double> data;
    double operator()(size_t i, size_t j, size_t k) const { return data[i + size.x * (j + size.y * k)]; }
    Vector3<size_t> GetSize() const { return size; }
    Vector3<double> GetOrigin() const { return origin; }
    Vector3<double> GetGridSpacing() const { return spacing; }
};

// src/3d/triangle_3d.hpp
struct Triangle3D {
    size_t point1Idx = 0, point2Idx = 1, point3Idx = 2, normalIdx = 0;
    Triangle3D() = default;
    Triangle3D(size_t p1, size_t p2, size_t p3, size_t n) : point1Idx(p1), point2Idx(p2), point3Idx(p3), normalIdx(n) {}
};
using Triangle3D_t = Triangle3D;

// src/3d/triangle_mesh.hpp (the members the surface and OBJ paths use)
class TriangleMesh {
  public:
    void Clear() {
        _vertices.clear();
        _normals.clear();
        _triangles.clear();
        _objectName.clear();
    }
    void AddVertex(const Vector3<double>& v) { _vertices.push_back(v); }
    void AddNormal(const Vector3<double>& n) { _normals.push_back(n); }
    void AddTriangle(const Triangle3D_t& t) { _triangles.push_back(t); }
    const std::vector<Vector3<double>>& GetVerticies() const { return _vertices; }
    const std::vector<Vector3<double>>& GetNormals() const { return _normals; }
    const std::vector<Triangle3D_t>& GetTriangles() const { return _triangles; }
    std::vector<Vector3<double>>& GetVerticies() { return _vertices; }
    std::vector<Vector3<double>>& GetNormals() { return _normals; }
    std::vector<Triangle3D_t>& GetTriangles() { return _triangles; }
    std::string GetObjectName() const { return _objectName; }
    void SetObjectName(const std::string& name) { _objectName = name; }

  private:
    std::string _objectName;
    std::vector<Vector3<double>> _vertices, _normals;
    std::vector<Triangle3D_t> _triangles;
};

// OBJ files in the reference's layout (src/3d/obj_manager.cpp): header,
// object name, "v", "vn", "s off", "f a//n b//n c//n" (1-based).
class OBJManager {
  public:
    void Save(const std::string& path, const TriangleMesh& mesh) const;
    void Load(const std::string& path, TriangleMesh& mesh) const;
};

// Pressure-solver selection (SetPressureSolver / SetLinearSystemSolver,
// src/fluid_solvers/single_phase_pressure_solver.hpp:21).  The default is the
// reference's CudaJacobiIterationSolver(1000, 5, 1e-14)
// (src/fluid_solvers/single_phase_pressure_solver.cpp:11).
struct PressureSolverSettings {
    int kind = FE_SOLVER_JACOBI;  // fe_solver_kind
    int maxIterations = 1000;
    int checkInterval = 5;
    double tolerance = 1e-14;
    bool relative = false;
    bool warmStart = true;
};

// The linear solver of BackwardEulerDiffusionSolver, which the reference fixes
// to CudaJacobiIterationSolver(1000, 5, 1e-14)
// (src/fluid_solvers/backward_euler_diffusion_solver.cpp:10): Jacobi only.
struct DiffusionSolverSettings {
    int maxIterations = 1000;
    int checkInterval = 5;
    double tolerance = 1e-14;
};

// The reference's solver objects (SetPressureSolver / SetDiffusionSolver take
// them as shared_ptrs, src/pic_simulator.hpp:41-43).  Each linear solver names
// the device solve it stands for; the iteration itself runs in the scene.
class LinearSystemSolver {
  public:
    virtual ~LinearSystemSolver() = default;
    virtual PressureSolverSettings Settings() const = 0;
};

// src/linear_system/jacobi_iteration_solver.hpp: at most maxNumberOfIterations
// sweeps, ||b - A x||_2 < tolerance checked every residualCheckInterval sweeps.
class JacobiIterationSolver : public LinearSystemSolver {
  public:
    JacobiIterationSolver(size_t maxNumberOfIterations, size_t residualCheckInterval, double tolerance);
    PressureSolverSettings Settings() const override { return _s; }

  private:
    PressureSolverSettings _s;
};

// src/linear_system/cuda_jacobi_iteration_solver.hpp: the same sweeps and stop.
class CudaJacobiIterationSolver : public JacobiIterationSolver {
  public:
    using JacobiIterationSolver::JacobiIterationSolver;
};

// src/linear_system/conjugate_gradient_solver.hpp: unpreconditioned CG from the
// previous solution, stopping when r.r <= tolerance^2 or after maxNumberOfIterations.
class ConjugateGradientSolver : public LinearSystemSolver {
  public:
    ConjugateGradientSolver(size_t maxNumberOfIterations, double tolerance);
    PressureSolverSettings Settings() const override { return _s; }

  protected:
    PressureSolverSettings _s;
};

// src/linear_system/cuda_conjugate_gradient_solver.hpp: the same iteration.
class CudaConjugateGradientSolver : public ConjugateGradientSolver {
  public:
    using ConjugateGradientSolver::ConjugateGradientSolver;
};

// Added for the GPU backend (no reference class): CG preconditioned by one
// geometric multigrid V-cycle, stopping on ||b - A x||_2 <= tolerance ||b||_2
// when relative (the default) or r.r <= tolerance^2.
class MultigridPCGSolver : public ConjugateGradientSolver {
  public:
    MultigridPCGSolver(size_t maxNumberOfIterations, double tolerance, bool relative = true);
};

// src/fluid_solvers/pressure_solver.hpp
class PressureSolver {
  public:
    virtual ~PressureSolver() = default;
};

// src/fluid_solvers/single_phase_pressure_solver.hpp: its linear solver
// defaults to CudaJacobiIterationSolver(1000, 5, 1e-14)
// (src/fluid_solvers/single_phase_pressure_solver.cpp:11).
class SinglePhasePressureSolver : public PressureSolver {
  public:
    SinglePhasePressureSolver();
    void SetLinearSystemSolver(const std::shared_ptr<LinearSystemSolver>& solver);
    const std::shared_ptr<LinearSystemSolver>& GetLinearSystemSolver() const { return _solver; }

  private:
    std::shared_ptr<LinearSystemSolver> _solver;
};

// src/fluid_solvers/diffusion_solver.hpp
class DiffusionSolver {
  public:
    virtual ~DiffusionSolver() = default;
};

// src/fluid_solvers/backward_euler_diffusion_solver.hpp: implicit viscosity
// with the fixed CudaJacobiIterationSolver(1000, 5, 1e-14)
// (src/fluid_solvers/backward_euler_diffusion_solver.cpp:10).
class BackwardEulerDiffusionSolver : public DiffusionSolver {
  public:
    DiffusionSolverSettings Settings() const { return DiffusionSolverSettings(); }
};

// src/fluid_solvers/forward_euler_diffusion_solver.hpp: not on this backend
// (SetDiffusionSolver rejects it).
class ForwardEulerDiffusionSolver : public DiffusionSolver {};

class Error : public std::runtime_error {
  public:
    Error(const std::string& what, int code) : std::runtime_error(what), code(code) {}
    int code;
};

class Unsupported : public std::logic_error {
  public:
    using std::logic_error::logic_error;
};

// One fe_ctx per process and device, shared by the simulators on it.
// TriangleMeshCollider (src/3d/triangle_mesh_collider.hpp:11-30): a closed
// triangle mesh whose SDF is built by Mesh2SDF on the node grid (gridSize,
// gridOrigin, gridSpacing).  The SDF is computed on the device when the
// collider is added to a simulator (fe_scene_add_mesh_collider).
// src/3d/collider.hpp.  The colliders of this backend are sampled into the
// scene's collider SDF on the device; their host-side queries are not kept.
class Collider {
  public:
    virtual ~Collider() = default;
};

class TriangleMeshCollider : public Collider {
  public:
    TriangleMeshCollider(const Vector3<size_t>& gridSize, const Vector3<double>& gridOrigin,
                         const Vector3<double>& gridSpacing, const TriangleMesh& mesh);
    const Vector3<size_t>& GetGridSize() const { return _size; }
    const Vector3<double>& GetGridOrigin() const { return _origin; }
    const Vector3<double>& GetGridSpacing() const { return _spacing; }
    const std::vector<double>& Vertices() const { return _vertices; }    // 3 per vertex
    const std::vector<int32_t>& Triangles() const { return _triangles; }  // 3 vertex indices per triangle

  private:
    Vector3<size_t> _size;
    Vector3<double> _origin, _spacing;
    std::vector<double> _vertices;
    std::vector<int32_t> _triangles;
};

// VolumeEmitter (src/emmiter/volume_emmiter.hpp:13-36): particles per block
// with a start velocity in the nodes of the domain grid at `resolution` where
// the emitter's SDF is <= 0 (Mesh2SDF of the mesh, or given directly by
// InitializeFromSdf).  velocityVariance is kept but unused, as in the
// reference.  The seed keys the particle generator (no reference counterpart).
// src/emmiter/emmiter.hpp
class Emitter {
  public:
    virtual ~Emitter() = default;
};

class VolumeEmitter : public Emitter {
  public:
    VolumeEmitter(const TriangleMesh& mesh, Vector3<size_t> resolution, BoundingBox3D domain,
                  const size_t& particlesPerBlock, const Vector3<double>& velocity,
                  const Vector3<double>& velocityVariance);
    void InitializeFromTriangleMesh(const TriangleMesh& mesh);
    void InitializeFromSdf(const ScalarGrid3D& sdf);
    void SetSeed(uint64_t seed) { _seed = seed; }

    // read by the simulator
    const Vector3<size_t>& Resolution() const { return _resolution; }
    const BoundingBox3D& Domain() const { return _domain; }
    size_t ParticlesPerBlock() const { return _particlesPerBlock; }
    const Vector3<double>& Velocity() const { return _velocity; }
    uint64_t Seed() const { return _seed; }
    bool HasSdf() const { return !_sdf.empty(); }
    const std::vector<double>& Sdf() const { return _sdf; }  // x fastest
    const std::vector<double>& Vertices() const { return _vertices; }
    const std::vector<int32_t>& Triangles() const { return _triangles; }

  private:
    Vector3<size_t> _resolution;
    BoundingBox3D _domain;
    size_t _particlesPerBlock;
    Vector3<double> _velocity, _velocityVariance;
    uint64_t _seed = 0;
    std::vector<double> _vertices, _sdf;
    std::vector<int32_t> _triangles;
};

// src/3d/collider_collection.hpp: the colliders of a scene on the node grid
// (size, gridSpacing, origin).  SetColliders hands them to a simulator, whose
// device builds the minimum SDF over them (BuildSdf is that step, done there).
class ColliderCollection {
  public:
    ColliderCollection(const Vector3<size_t>& size, const Vector3<double>& gridSpacing, const Vector3<double>& origin)
        : _size(size), _gridSpacing(gridSpacing), _origin(origin) {}
    void AddCollider(std::shared_ptr<Collider> collider);
    void BuildSdf() {}
    std::vector<std::shared_ptr<Collider>>& GetColliders() { return _colliders; }
    const std::vector<std::shared_ptr<Collider>>& GetColliders() const { return _colliders; }
    Vector3<size_t> GetSize() const { return _size; }
    Vector3<double> GetGridSpacing() const { return _gridSpacing; }
    Vector3<double> GetOrigin() const { return _origin; }

  private:
    std::vector<std::shared_ptr<Collider>> _colliders;
    Vector3<size_t> _size;
    Vector3<double> _gridSpacing, _origin;
};

class Device;

class HipHybridSimulator {
  public:
    HipHybridSimulator(const Vector3<size_t>& gridSize, const BoundingBox3D& domain, int method);
    virtual ~HipHybridSimulator();
    HipHybridSimulator(const HipHybridSimulator&) = delete;
    HipHybridSimulator& operator=(const HipHybridSimulator&) = delete;

    // PhysicsAnimation (src/physics_animation.cpp): frame -> AdvanceTimeStep ->
    // CFL-limited substeps (fe_advance_frame).
    void AdvanceSingleFrame();
    void Update(const Frame& frame);
    Frame GetCurrentFrame() const { return _frame; }
    double GetCurrentTimeInSeconds() const { return _time; }
    unsigned int GetNumberOfSubTimeSteps() const { return _substeps; }
    void SetCurrentFrame(const Frame& frame) { _frame = frame; }

    // HybridSimulator
    double Cfl(double timeIntervalInSeconds) const;
    void AddExternalForce(std::shared_ptr<ExternalForce> force);
    double MaxCfl() const { return _maxCfl; }
    double GetViscosity() const { return _viscosity; }
    double GetDensity() const { return _density; }
    size_t GetParticlesPerBlock() const { return _particlesPerBlock; }
    Vector3<double> GetOrigin() const { return _domain.GetOrigin(); }
    Vector3<size_t> GetResolution() const { return _resolution; }
    Vector3<double> GetGridSpacing() const;
    const ScalarGrid3D& GetFluidSdf() const;  // downloads the current SDF
    size_t GetNumberOfParticles() const;
    void SetMaxClf(double maxCfl);
    void SetViscosity(double viscosity);
    void SetDensity(double density);
    void SetParticlesPerBlock(size_t particlesPerBlock) { _particlesPerBlock = particlesPerBlock; }

    // PICSimulator::SetPressureSolver (src/pic_simulator.hpp:42): a
    // SinglePhasePressureSolver and its linear solver, or the settings directly.
    void SetPressureSolver(std::shared_ptr<PressureSolver> pressureSolver);
    void SetPressureSolver(const PressureSolverSettings& settings);
    PressureSolverSettings GetPressureSolver() const { return _solver; }
    // SetDiffusionSolver (src/pic_simulator.hpp:41) for the reference's
    // BackwardEulerDiffusionSolver; it runs when viscosity > 0.
    void SetDiffusionSolver(std::shared_ptr<DiffusionSolver> diffusionSolver);
    void SetDiffusionSolver(const DiffusionSolverSettings& settings);
    DiffusionSolverSettings GetDiffusionSolver() const { return _diffusion; }

    // HybridSimulator::AddCollider (src/hybrid_simulator.hpp:35,
    // src/pic_simulator.cpp:66-69): the scene's collider SDF becomes the
    // minimum over its colliders (ColliderCollection::BuildSdf).
    // TriangleMeshCollider is the collider type of this backend.
    void AddCollider(std::shared_ptr<Collider> collider);
    // PICSimulator::SetColliders (src/pic_simulator.hpp:43): the collection's
    // colliders replace the simulator's.
    void SetColliders(std::shared_ptr<ColliderCollection> colliders);
    size_t NumberOfColliders() const { return _colliders.size(); }

    // PICSimulator::AddEmitter (src/pic_simulator.cpp:46-49): emits at the
    // start of every substep from the second on.  VolumeEmitter is the emitter
    // type of this backend.
    void AddEmitter(std::shared_ptr<Emitter> emitter);
    size_t NumberOfEmitters() const { return _emitters.size(); }
    // PICSimulator::InitializeFromTriangleMesh (src/pic_simulator.cpp:33-39):
    // Mesh2SDF of the mesh into the fluid SDF, then GetParticlesPerBlock()
    // particles in each cell whose corner samples it negative (appended).
    void InitializeFromTriangleMesh(const TriangleMesh& mesh, uint64_t seed = 0);
    // HybridSimulator::GetSurface: MarchingCubesSolver::BuildSurface on the
    // current fluid SDF (extracted on the GPU), appended to mesh; each
    // triangle's normal index is its first vertex's (per-vertex normals).
    void GetSurface(TriangleMesh& mesh) const;

    // Added for the GPU backend: particle and grid transfer, analytic scenes.
    void InitializeParticles(const std::vector<fe_seed_shape>& shapes, uint64_t seed);
    void SetParticles(const float* x, const float* v, const float* C, int64_t count);  // xyz-interleaved
    void GetParticles(std::vector<float>* x, std::vector<float>* v, std::vector<float>* C) const;
    size_t GridBytes(int field) const;
    void GetGrid(int field, void* dst, size_t bytes) const;
    void SetGrid(int field, const void* src, size_t bytes);
    fe_solve_stats LastSolveStats() const;
    std::vector<fe_solve_stats> LastDiffusionStats() const;  // u, v, w of the last diffusion solve
    int Method() const { return _method; }
    fe_scene* Handle() const { return _scene; }

  protected:
    void SetBlend(double blend);
    double Blend() const { return _blend; }

  private:
    void EnsureScene(int64_t capacity) const;
    void PushParams() const;
    void PushDiffusionSolver() const;
    void PushCollider(const TriangleMeshCollider& c) const;
    void PushEmitter(const VolumeEmitter& e) const;
    fe_scene_desc Desc(int64_t capacity) const;
    void Check(int rc, const char* what) const;

    Vector3<size_t> _resolution;
    BoundingBox3D _domain;
    int _method;
    double _maxCfl = 3;  // src/hybrid_simulator.cpp:21
    double _viscosity = 0;
    double _density = 1;
    size_t _particlesPerBlock = 16;
    double _blend = 0.0;
    Vector3<double> _gravity = 0;
    PressureSolverSettings _solver;
    DiffusionSolverSettings _diffusion;
    std::vector<std::shared_ptr<TriangleMeshCollider>> _colliders;
    std::vector<std::shared_ptr<VolumeEmitter>> _emitters;
    Frame _frame;
    double _time = 0.0;
    unsigned int _substeps = 0;
    std::shared_ptr<Device> _device;
    mutable fe_scene* _scene = nullptr;
    mutable ScalarGrid3D _sdf;
};

// FLIPSimulator and APICSimulator derive from PICSimulator, as in the
// reference (src/flip_simulator.hpp, src/apic_simulator.hpp), so a
// shared_ptr<FLIPSimulator> passes where a shared_ptr<PICSimulator> is taken.
class HipPICSimulator : public HipHybridSimulator {
  public:
    HipPICSimulator(const Vector3<size_t>& gridSize, const BoundingBox3D& domain)
        : HipHybridSimulator(gridSize, domain, FE_PIC) {}

  protected:
    HipPICSimulator(const Vector3<size_t>& gridSize, const BoundingBox3D& domain, int method)
        : HipHybridSimulator(gridSize, domain, method) {}
};

class HipFLIPSimulator : public HipPICSimulator {
  public:
    HipFLIPSimulator(const Vector3<size_t>& gridSize, const BoundingBox3D& domain)
        : HipPICSimulator(gridSize, domain, FE_FLIP) {
        SetBlend(0.05);  // _blendingFactor(0.05), src/flip_simulator.cpp:5
    }
    void SetBlendingFactor(double factor) { SetBlend(factor); }
    double GetBlendingFactor() const { return Blend(); }
};

class HipAPICSimulator : public HipPICSimulator {
  public:
    HipAPICSimulator(const Vector3<size_t>& gridSize, const BoundingBox3D& domain)
        : HipPICSimulator(gridSize, domain, FE_APIC) {}
};

// src/simulation_stats.hpp
struct SimulationStats {
    std::string simulatorType;
    size_t numberOfIterations = 0;
    Vector3<size_t> gridSize = 0;
    size_t numberOfParticles = 0;
    double timeIntervalInSeconds = 0;
    double simulationTimeInSeconds = 0;
    std::vector<double> iterationTimeInSeconds;
    std::vector<double> cflPerIteration;
    void PrintStats() const;
    void Clear() { *this = SimulationStats(); }
};

// SimulationRunner::RunSimulation (src/simulation_runner.cpp:17-55): per frame
// AdvanceSingleFrame, GetSurface and an OBJ file
// <cache>/simulated_frames/<name>_<i>.obj; the log goes to <cache>/log.txt.
class SimulationRunner {
  public:
    void RunSimulation(const std::shared_ptr<HipHybridSimulator>& simulator, double timeIntervalInSeconds,
                       size_t numberOfIterations, const std::string& simulationOutputName,
                       const std::string& cacheFolderPath);
    SimulationStats GetStats() const { return _stats; }

  private:
    SimulationStats _stats;
};

// The drop-in names: the reference's CPU simulator classes
// (src/hybrid_simulator.hpp, src/pic_simulator.hpp, src/flip_simulator.hpp,
// src/apic_simulator.hpp) and its CUDA ones (src/cuda_pic_simulator.hpp).
using HybridSimulator = HipHybridSimulator;
using PICSimulator = HipPICSimulator;
using FLIPSimulator = HipFLIPSimulator;
using APICSimulator = HipAPICSimulator;
using CudaPICSimulator = HipPICSimulator;
using CudaFLIPSimulator = HipFLIPSimulator;
using CudaAPICSimulator = HipAPICSimulator;

}  // namespace fluide
