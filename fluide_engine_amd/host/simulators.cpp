// simulators.cpp -- host classes of fluide/simulators.hpp over the C ABI.
#include "fluide/simulators.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <limits>
#include <map>
#include <mutex>
#include <numeric>
#include <sstream>
#include <typeinfo>

namespace fluide {

bool BoundingBox3D::IsInside(Vector3<double> p) const {
    return p.x >= _origin.x && p.x <= _origin.x + _size.x && p.y >= _origin.y && p.y <= _origin.y + _size.y &&
           p.z >= _origin.z && p.z <= _origin.z + _size.z;
}

Vector3<double> BoundingBox3D::Corner(size_t idx) const {
    return Vector3<double>(_origin.x + ((idx & 1) ? _size.x : 0.0), _origin.y + ((idx & 2) ? _size.y : 0.0),
                           _origin.z + ((idx & 4) ? _size.z : 0.0));
}

class Device {
  public:
    static std::shared_ptr<Device> Get(int device) {
        static std::mutex m;
        static std::map<int, std::weak_ptr<Device>> all;
        std::lock_guard<std::mutex> lock(m);
        std::shared_ptr<Device> d = all[device].lock();
        if (!d) {
            d = std::shared_ptr<Device>(new Device(device));
            all[device] = d;
        }
        return d;
    }
    ~Device() {
        if (ctx) fe_ctx_destroy(ctx);
    }
    fe_ctx* Context() {
        if (!ctx) {
            const int rc = fe_ctx_create(id, &ctx);
            if (rc != FE_OK) throw Error(std::string("fe_ctx_create: ") + fe_last_error(nullptr), rc);
        }
        return ctx;
    }

  private:
    explicit Device(int device) : id(device) {}
    int id;
    fe_ctx* ctx = nullptr;
};

HipHybridSimulator::HipHybridSimulator(const Vector3<size_t>& gridSize, const BoundingBox3D& domain, int method)
    : _resolution(gridSize), _domain(domain), _method(method), _device(Device::Get(0)) {}

HipHybridSimulator::~HipHybridSimulator() {
    if (_scene) fe_scene_destroy(_scene);
}

void HipHybridSimulator::Check(int rc, const char* what) const {
    if (rc == FE_OK) return;
    const char* msg = fe_last_error(_device->Context());
    throw Error(std::string(what) + ": " + (msg ? msg : "error"), rc);
}

fe_scene_desc HipHybridSimulator::Desc(int64_t capacity) const {
    fe_scene_desc d{};
    const size_t n[3] = {_resolution.x, _resolution.y, _resolution.z};
    const Vector3<double> o = _domain.GetOrigin(), L = _domain.GetSize();
    for (int a = 0; a < 3; ++a) d.n[a] = (int32_t)n[a];
    d.method = _method;
    d.origin[0] = o.x;
    d.origin[1] = o.y;
    d.origin[2] = o.z;
    d.size[0] = L.x;
    d.size[1] = L.y;
    d.size[2] = L.z;
    d.blend = _blend;
    d.max_cfl = _maxCfl;
    d.gravity[0] = _gravity.x;
    d.gravity[1] = _gravity.y;
    d.gravity[2] = _gravity.z;
    d.viscosity = _viscosity;
    d.density = _density;
    d.max_particles = capacity;
    return d;
}

static void flatten_mesh(const TriangleMesh& mesh, std::vector<double>& v, std::vector<int32_t>& t,
                         const char* who) {
    v.clear();
    t.clear();
    for (const auto& p : mesh.GetVerticies()) {
        v.push_back(p.x);
        v.push_back(p.y);
        v.push_back(p.z);
    }
    for (const auto& tr : mesh.GetTriangles())
        for (size_t q : {tr.point1Idx, tr.point2Idx, tr.point3Idx}) {
            if (q >= mesh.GetVerticies().size()) throw std::invalid_argument(std::string(who) + ": vertex index out of range");
            t.push_back((int32_t)q);
        }
}

// The scene is made on first use and then kept: more particles grow its
// storage in place, so emitters, colliders and the fluid state stay.
void HipHybridSimulator::EnsureScene(int64_t capacity) const {
    if (_scene) {
        Check(fe_scene_reserve_particles(_scene, capacity), "fe_scene_reserve_particles");
        return;
    }
    const fe_scene_desc d = Desc(std::max<int64_t>(capacity, 1));
    Check(fe_scene_create(_device->Context(), &d, &_scene), "fe_scene_create");
    PushDiffusionSolver();
    for (const auto& c : _colliders) PushCollider(*c);
    for (const auto& e : _emitters) PushEmitter(*e);
}

void HipHybridSimulator::PushEmitter(const VolumeEmitter& e) const {
    const Vector3<size_t>& r = e.Resolution();
    const int32_t res[3] = {(int32_t)r.x, (int32_t)r.y, (int32_t)r.z};
    const Vector3<double> o = e.Domain().GetOrigin(), L = e.Domain().GetSize();
    const double origin[3] = {o.x, o.y, o.z}, size[3] = {L.x, L.y, L.z};
    const double vel[3] = {e.Velocity().x, e.Velocity().y, e.Velocity().z};
    const int32_t ppb = (int32_t)e.ParticlesPerBlock();
    if (e.HasSdf())
        Check(fe_scene_add_volume_emitter_sdf(_scene, res, origin, size, e.Sdf().data(), ppb, vel, e.Seed()),
              "fe_scene_add_volume_emitter_sdf");
    else
        Check(fe_scene_add_volume_emitter(_scene, res, origin, size, e.Vertices().data(),
                                          (int64_t)e.Vertices().size() / 3, e.Triangles().data(),
                                          (int64_t)e.Triangles().size() / 3, ppb, vel, e.Seed()),
              "fe_scene_add_volume_emitter");
}

void HipHybridSimulator::AddEmitter(std::shared_ptr<Emitter> emitter) {
    if (!emitter) throw std::invalid_argument("AddEmitter: null emitter");
    auto ve = std::dynamic_pointer_cast<VolumeEmitter>(emitter);
    if (!ve) throw Unsupported("AddEmitter: VolumeEmitter is the emitter type of the GPU backend");
    if (_scene) PushEmitter(*ve);
    _emitters.push_back(std::move(ve));
}

void HipHybridSimulator::InitializeFromTriangleMesh(const TriangleMesh& mesh, uint64_t seed) {
    std::vector<double> v;
    std::vector<int32_t> t;
    flatten_mesh(mesh, v, t, "InitializeFromTriangleMesh");
    EnsureScene(0);
    Check(fe_scene_initialize_from_mesh(_scene, v.data(), (int64_t)v.size() / 3, t.data(), (int64_t)t.size() / 3,
                                        (int32_t)_particlesPerBlock, seed),
          "fe_scene_initialize_from_mesh");
}

VolumeEmitter::VolumeEmitter(const TriangleMesh& mesh, Vector3<size_t> resolution, BoundingBox3D domain,
                             const size_t& particlesPerBlock, const Vector3<double>& velocity,
                             const Vector3<double>& velocityVariance)
    : _resolution(resolution), _domain(domain), _particlesPerBlock(particlesPerBlock), _velocity(velocity),
      _velocityVariance(velocityVariance) {
    if (resolution.x > INT32_MAX || resolution.y > INT32_MAX || resolution.z > INT32_MAX)
        throw std::invalid_argument("VolumeEmitter: resolution beyond int32");
    if (particlesPerBlock > INT32_MAX) throw std::invalid_argument("VolumeEmitter: particlesPerBlock beyond int32");
    InitializeFromTriangleMesh(mesh);
}

void VolumeEmitter::InitializeFromTriangleMesh(const TriangleMesh& mesh) {
    flatten_mesh(mesh, _vertices, _triangles, "VolumeEmitter");
    _sdf.clear();
}

void VolumeEmitter::InitializeFromSdf(const ScalarGrid3D& sdf) {
    // src/emmiter/volume_emmiter.cpp:50-56: the emitter grid's values copied index by index
    const size_t nx = _resolution.x, ny = _resolution.y, nz = _resolution.z;
    if (sdf.size.x < nx || sdf.size.y < ny || sdf.size.z < nz)
        throw std::invalid_argument("VolumeEmitter::InitializeFromSdf: the grid is smaller than the emitter's");
    _sdf.assign(nx * ny * nz, 0.0);
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i) _sdf[i + nx * (j + ny * k)] = sdf(i, j, k);
}

void HipHybridSimulator::PushCollider(const TriangleMeshCollider& c) const {
    const Vector3<size_t>& n = c.GetGridSize();
    const int32_t size[3] = {(int32_t)n.x, (int32_t)n.y, (int32_t)n.z};
    const double o[3] = {c.GetGridOrigin().x, c.GetGridOrigin().y, c.GetGridOrigin().z};
    const double h[3] = {c.GetGridSpacing().x, c.GetGridSpacing().y, c.GetGridSpacing().z};
    Check(fe_scene_add_mesh_collider(_scene, size, o, h, c.Vertices().data(), (int64_t)c.Vertices().size() / 3,
                                     c.Triangles().data(), (int64_t)c.Triangles().size() / 3),
          "fe_scene_add_mesh_collider");
}

void HipHybridSimulator::AddCollider(std::shared_ptr<Collider> collider) {
    if (!collider) throw std::invalid_argument("AddCollider: null collider");
    auto mc = std::dynamic_pointer_cast<TriangleMeshCollider>(collider);
    if (!mc) throw Unsupported("AddCollider: TriangleMeshCollider is the collider type of the GPU backend");
    if (_scene) PushCollider(*mc);  // a scene created later receives every collider
    _colliders.push_back(std::move(mc));
}

void HipHybridSimulator::SetColliders(std::shared_ptr<ColliderCollection> colliders) {
    if (!colliders) throw std::invalid_argument("SetColliders: null collection");
    std::vector<std::shared_ptr<TriangleMeshCollider>> list;
    for (const auto& c : colliders->GetColliders()) {
        auto mc = std::dynamic_pointer_cast<TriangleMeshCollider>(c);
        if (!mc) throw Unsupported("SetColliders: TriangleMeshCollider is the collider type of the GPU backend");
        list.push_back(std::move(mc));
    }
    if (_scene) {
        Check(fe_scene_clear_colliders(_scene), "fe_scene_clear_colliders");
        for (const auto& c : list) PushCollider(*c);
    }
    _colliders = std::move(list);
}

void ColliderCollection::AddCollider(std::shared_ptr<Collider> collider) {
    if (!collider) throw std::invalid_argument("ColliderCollection::AddCollider: null collider");
    _colliders.push_back(std::move(collider));
}

Vector3<double> PointField::Sample(const Vector3<double>& position) const {
    // src/forces/point_field.cpp: strength * (1 / max(|x - o|, 1)) * (o - x) / |o - x|
    const double dx = _origin.x - position.x, dy = _origin.y - position.y, dz = _origin.z - position.z;
    const double len = std::sqrt(dx * dx + dy * dy + dz * dz);
    const double dist = len >= 1.0 ? len : 1.0;
    const double f = _strength * (1.0 / dist);
    return Vector3<double>(f * (dx / len), f * (dy / len), f * (dz / len));
}

JacobiIterationSolver::JacobiIterationSolver(size_t maxNumberOfIterations, size_t residualCheckInterval,
                                             double tolerance) {
    if (maxNumberOfIterations > 1000000 || residualCheckInterval > INT32_MAX)
        throw std::invalid_argument("JacobiIterationSolver: iteration counts out of range");
    _s.kind = FE_SOLVER_JACOBI;
    _s.maxIterations = (int)maxNumberOfIterations;
    _s.checkInterval = (int)residualCheckInterval;
    _s.tolerance = tolerance;
    _s.relative = false;
    _s.warmStart = true;
}

ConjugateGradientSolver::ConjugateGradientSolver(size_t maxNumberOfIterations, double tolerance) {
    if (maxNumberOfIterations > 1000000) throw std::invalid_argument("ConjugateGradientSolver: too many iterations");
    _s.kind = FE_SOLVER_CG;
    _s.maxIterations = (int)maxNumberOfIterations;
    _s.checkInterval = 0;
    _s.tolerance = tolerance;
    _s.relative = false;
    _s.warmStart = true;
}

MultigridPCGSolver::MultigridPCGSolver(size_t maxNumberOfIterations, double tolerance, bool relative)
    : ConjugateGradientSolver(maxNumberOfIterations, tolerance) {
    _s.kind = FE_SOLVER_MGPCG;
    _s.relative = relative;
}

SinglePhasePressureSolver::SinglePhasePressureSolver()
    : _solver(std::make_shared<CudaJacobiIterationSolver>(1000, 5, 1e-14)) {}

void SinglePhasePressureSolver::SetLinearSystemSolver(const std::shared_ptr<LinearSystemSolver>& solver) {
    if (!solver) throw std::invalid_argument("SetLinearSystemSolver: null solver");
    _solver = solver;
}

void HipHybridSimulator::SetPressureSolver(std::shared_ptr<PressureSolver> pressureSolver) {
    auto sp = std::dynamic_pointer_cast<SinglePhasePressureSolver>(pressureSolver);
    if (!sp) throw Unsupported("SetPressureSolver: SinglePhasePressureSolver is the pressure solver of the GPU backend");
    SetPressureSolver(sp->GetLinearSystemSolver()->Settings());
}

void HipHybridSimulator::SetDiffusionSolver(std::shared_ptr<DiffusionSolver> diffusionSolver) {
    auto be = std::dynamic_pointer_cast<BackwardEulerDiffusionSolver>(diffusionSolver);
    if (!be)
        throw Unsupported("SetDiffusionSolver: BackwardEulerDiffusionSolver is the diffusion solver of the GPU backend");
    SetDiffusionSolver(be->Settings());
}

TriangleMeshCollider::TriangleMeshCollider(const Vector3<size_t>& gridSize, const Vector3<double>& gridOrigin,
                                           const Vector3<double>& gridSpacing, const TriangleMesh& mesh)
    : _size(gridSize), _origin(gridOrigin), _spacing(gridSpacing) {
    if (gridSize.x > INT32_MAX || gridSize.y > INT32_MAX || gridSize.z > INT32_MAX)
        throw std::invalid_argument("TriangleMeshCollider: grid size beyond int32");
    flatten_mesh(mesh, _vertices, _triangles, "TriangleMeshCollider");
}

void HipHybridSimulator::PushDiffusionSolver() const {
    if (!_scene) return;
    const fe_solver_desc sd{FE_SOLVER_JACOBI, _diffusion.maxIterations, _diffusion.checkInterval, 0, 0, 0,
                            _diffusion.tolerance};
    Check(fe_scene_set_diffusion_solver(_scene, &sd), "fe_scene_set_diffusion_solver");
}

void HipHybridSimulator::PushParams() const {
    if (!_scene) return;
    const double g[3] = {_gravity.x, _gravity.y, _gravity.z};
    Check(fe_scene_set_params(_scene, _blend, _maxCfl, g, _viscosity, _density), "fe_scene_set_params");
}

// PhysicsAnimation::OnUpdate / AdvanceTimeStep (src/physics_animation.cpp:45-92).
void HipHybridSimulator::Update(const Frame& frame) {
    if (frame.GetIndex() <= _frame.GetIndex()) return;
    EnsureScene(0);
    const unsigned long frames = frame.GetIndex() - _frame.GetIndex();
    const fe_solver_desc sd{_solver.kind,       _solver.maxIterations, _solver.checkInterval, _solver.relative ? 1 : 0,
                            _solver.warmStart ? 1 : 0, 0,              _solver.tolerance};
    for (unsigned long f = 0; f < frames; ++f) {
        _time = _frame.GetTimeInSeconds();
        int32_t sub = 0;
        Check(fe_advance_frame(_scene, frame.GetTimeIntervalInSeconds(), &sd, &sub), "fe_advance_frame");
        _substeps = (unsigned int)sub;
        _time += frame.GetTimeIntervalInSeconds();
    }
    _frame = frame;
}

void HipHybridSimulator::AdvanceSingleFrame() {
    Frame f = _frame;
    Update(++f);
}

double HipHybridSimulator::Cfl(double dt) const {
    EnsureScene(0);
    double c = 0.0;
    Check(fe_cfl(_scene, dt, &c), "fe_cfl");
    return c;
}

// PICSimulator keeps a force list and applies each to every face
// (src/forces/external_force.cpp:4-30); a uniform field sums into gravity.
void HipHybridSimulator::AddExternalForce(std::shared_ptr<ExternalForce> force) {
    if (!force) return;
    const Vector3<double> s = force->Sample(_domain.GetOrigin());
    for (size_t c = 1; c < 8; ++c) {
        const Vector3<double> t = force->Sample(_domain.Corner(c));
        if (t.x != s.x || t.y != s.y || t.z != s.z)
            throw Unsupported("only uniform (DirectionalField) forces are on the GPU backend");
    }
    _gravity = Vector3<double>(_gravity.x + s.x, _gravity.y + s.y, _gravity.z + s.z);
    PushParams();
}

Vector3<double> HipHybridSimulator::GetGridSpacing() const {
    const Vector3<double> L = _domain.GetSize();
    return Vector3<double>(L.x / _resolution.x, L.y / _resolution.y, L.z / _resolution.z);  // src/hybrid_simulator.cpp:7
}

const ScalarGrid3D& HipHybridSimulator::GetFluidSdf() const {
    EnsureScene(0);
    const Vector3<double> h = GetGridSpacing(), o = _domain.GetOrigin();
    _sdf.size = _resolution;
    _sdf.spacing = h;
    _sdf.origin = Vector3<double>(o.x + h.x / 2.0, o.y + h.y / 2.0, o.z + h.z / 2.0);  // src/hybrid_simulator.cpp:15
    _sdf.data.resize(_resolution.x * _resolution.y * _resolution.z);
    Check(fe_grid_download(_scene, FE_GRID_SDF, _sdf.data.data(), _sdf.data.size() * sizeof(double)),
          "fe_grid_download");
    return _sdf;
}

size_t HipHybridSimulator::GetNumberOfParticles() const {
    if (!_scene) return 0;
    int64_t n = 0;
    Check(fe_particles_count(_scene, &n), "fe_particles_count");
    return (size_t)n;
}

void HipHybridSimulator::SetMaxClf(double maxCfl) {
    _maxCfl = maxCfl;
    PushParams();
}

void HipHybridSimulator::SetViscosity(double viscosity) {
    if (!(viscosity >= 0.0)) throw std::invalid_argument("viscosity must be >= 0");
    _viscosity = viscosity;
    PushParams();
}

void HipHybridSimulator::SetDensity(double density) {
    _density = density;
    PushParams();
}

void HipHybridSimulator::SetBlend(double blend) {
    _blend = blend;
    PushParams();
}

void HipHybridSimulator::SetPressureSolver(const PressureSolverSettings& settings) { _solver = settings; }

void HipHybridSimulator::SetDiffusionSolver(const DiffusionSolverSettings& settings) {
    if (settings.maxIterations < 0 || settings.checkInterval < 0 || !(settings.tolerance >= 0.0))
        throw std::invalid_argument("diffusion solver: maxIterations, checkInterval and tolerance must be >= 0");
    _diffusion = settings;
    PushDiffusionSolver();
}

std::vector<fe_solve_stats> HipHybridSimulator::LastDiffusionStats() const {
    std::vector<fe_solve_stats> st(3);
    if (_scene) Check(fe_last_diffusion_stats(_scene, st.data()), "fe_last_diffusion_stats");
    return st;
}

void HipHybridSimulator::InitializeParticles(const std::vector<fe_seed_shape>& shapes, uint64_t seed) {
    const fe_scene_desc d = Desc(0);
    int64_t n = 0;
    Check(fe_seed_particles(&d, shapes.data(), (int32_t)shapes.size(), seed, nullptr, 0, &n), "fe_seed_particles");
    std::vector<float> x((size_t)n * 3), v((size_t)n * 3, 0.0f), C;
    Check(fe_seed_particles(&d, shapes.data(), (int32_t)shapes.size(), seed, x.data(), n, &n), "fe_seed_particles");
    if (_method == FE_APIC) C.assign((size_t)n * 9, 0.0f);
    SetParticles(x.data(), v.data(), C.empty() ? nullptr : C.data(), n);
}

void HipHybridSimulator::SetParticles(const float* x, const float* v, const float* C, int64_t count) {
    EnsureScene(count);
    Check(fe_particles_upload(_scene, x, v, C, count), "fe_particles_upload");
}

void HipHybridSimulator::GetParticles(std::vector<float>* x, std::vector<float>* v, std::vector<float>* C) const {
    const size_t n = GetNumberOfParticles();
    if (x) x->resize(3 * n);
    if (v) v->resize(3 * n);
    if (C) C->resize(_method == FE_APIC ? 9 * n : 0);
    if (n == 0) return;
    Check(fe_particles_download(_scene, x ? x->data() : nullptr, v ? v->data() : nullptr,
                                (C && _method == FE_APIC) ? C->data() : nullptr, (int64_t)n),
          "fe_particles_download");
}

size_t HipHybridSimulator::GridBytes(int field) const {
    EnsureScene(0);
    size_t b = 0;
    Check(fe_grid_bytes(_scene, field, &b), "fe_grid_bytes");
    return b;
}

void HipHybridSimulator::GetGrid(int field, void* dst, size_t bytes) const {
    EnsureScene(0);
    Check(fe_grid_download(_scene, field, dst, bytes), "fe_grid_download");
}

void HipHybridSimulator::SetGrid(int field, const void* src, size_t bytes) {
    EnsureScene(0);
    Check(fe_grid_upload(_scene, field, src, bytes), "fe_grid_upload");
}

fe_solve_stats HipHybridSimulator::LastSolveStats() const {
    fe_solve_stats st{};
    if (_scene) Check(fe_last_solve_stats(_scene, &st), "fe_last_solve_stats");
    return st;
}

}  // namespace fluide

// ---------------------------------------------------------------------------
// Surface, OBJ files and the frame loop
// ---------------------------------------------------------------------------
namespace fluide {

void HipHybridSimulator::GetSurface(TriangleMesh& mesh) const {
    EnsureScene(0);
    int64_t nv = 0, nt = 0;
    Check(fe_surface_build(_scene, &nv, &nt), "fe_surface_build");
    std::vector<double> v((size_t)nv * 3), n((size_t)nv * 3);
    std::vector<int32_t> t((size_t)nt * 3);
    Check(fe_surface_download(_scene, v.data(), n.data(), t.data()), "fe_surface_download");
    const size_t base = mesh.GetVerticies().size();
    for (int64_t i = 0; i < nv; ++i) {
        mesh.AddVertex(Vector3<double>(v[3 * i], v[3 * i + 1], v[3 * i + 2]));
        mesh.AddNormal(Vector3<double>(n[3 * i], n[3 * i + 1], n[3 * i + 2]));
    }
    for (int64_t i = 0; i < nt; ++i)
        mesh.AddTriangle(Triangle3D_t(base + t[3 * i], base + t[3 * i + 1], base + t[3 * i + 2], base + t[3 * i]));
}

void OBJManager::Save(const std::string& path, const TriangleMesh& mesh) const {
    std::ofstream f(path.c_str());
    if (!f) throw std::runtime_error("cannot open " + path);
    f << "mtllib test.mtl\n";
    f << "o " << mesh.GetObjectName() << "\n" << std::endl;
    for (const auto& v : mesh.GetVerticies())
        f << "v " << std::to_string(v.x) << " " << std::to_string(v.y) << " " << std::to_string(v.z) << "\n";
    f << std::endl;
    for (const auto& n : mesh.GetNormals())
        f << "vn " << std::to_string(n.x) << " " << std::to_string(n.y) << " " << std::to_string(n.z) << "\n";
    f << std::endl;
    f << "s off\n" << std::endl;
    for (const auto& t : mesh.GetTriangles()) {
        f << "f " << t.point1Idx + 1 << "//" << t.normalIdx + 1 << " " << t.point2Idx + 1 << "//" << t.normalIdx + 1
          << " " << t.point3Idx + 1 << "//" << t.normalIdx + 1 << "\n";
    }
    f << std::endl;
    if (!f) throw std::runtime_error("write error on " + path);
}

void OBJManager::Load(const std::string& path, TriangleMesh& mesh) const {
    std::ifstream f(path.c_str());
    if (!f) throw std::runtime_error("cannot open " + path);
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream in(line);
        std::string tag;
        if (!(in >> tag)) continue;
        if (tag == "o") {
            std::string name;
            if (in >> name) mesh.SetObjectName(name);
        } else if (tag == "v" || tag == "vn") {
            double x, y, z;
            if (in >> x >> y >> z) (tag == "v" ? mesh.GetVerticies() : mesh.GetNormals()).emplace_back(x, y, z);
        } else if (tag == "f") {
            size_t p[3], nidx = 0;
            std::string tok;
            int k = 0;
            while (k < 3 && in >> tok) {
                const size_t slash = tok.find('/');
                p[k++] = std::stoul(tok.substr(0, slash)) - 1;
                const size_t last = tok.rfind('/');
                if (last != std::string::npos && last + 1 < tok.size()) nidx = std::stoul(tok.substr(last + 1)) - 1;
            }
            if (k == 3) mesh.AddTriangle(Triangle3D_t(p[0], p[1], p[2], nidx));
        }
    }
}

static double median_of(std::vector<double> v) {
    if (v.empty()) return std::numeric_limits<double>::quiet_NaN();
    std::sort(v.begin(), v.end());
    const size_t n = v.size();
    return n % 2 ? v[n / 2] : 0.5 * (v[n / 2 - 1] + v[n / 2]);
}

static double mean_of(const std::vector<double>& v) {
    if (v.empty()) return std::numeric_limits<double>::quiet_NaN();
    return std::accumulate(v.begin(), v.end(), 0.0) / (double)v.size();
}

// SimulationStats::PrintStats (src/simulation_stats.cpp:19-33), with its
// labels (the reference prints the mean under "Median CFL" and the median
// under "Mean CFL"; kept).
void SimulationStats::PrintStats() const {
    std::cout << "================ SIMULATION STATS ================\n";
    std::cout << "Simulator type: " << simulatorType << "\n";
    std::cout << "Grid resolution: " << gridSize.x << " x " << gridSize.y << " x " << gridSize.z << "\n";
    std::cout << "Number of particles: " << numberOfParticles << "\n";
    std::cout << "Number of Iterations: " << numberOfIterations << "\n";
    std::cout << "Time interval: " << timeIntervalInSeconds << " [s]\n";
    std::cout << "Median CFL: " << mean_of(cflPerIteration) << "\n";
    std::cout << "Mean CFL: " << median_of(cflPerIteration) << "\n";
    std::cout << "Median frame: " << median_of(iterationTimeInSeconds) << " [s]\n";
    std::cout << "Mean frame: " << mean_of(iterationTimeInSeconds) << " [s]\n";
    std::cout << "Simulation time: " << simulationTimeInSeconds << " [s]\n";
    std::cout << "==================================================\n";
}

void SimulationRunner::RunSimulation(const std::shared_ptr<HipHybridSimulator>& simulator,
                                     double timeIntervalInSeconds, size_t numberOfIterations,
                                     const std::string& simulationOutputName, const std::string& cacheFolderPath) {
    if (!simulator) throw std::invalid_argument("simulator is null");
    OBJManager obj;
    TriangleMesh surface;
    simulator->SetCurrentFrame(Frame(timeIntervalInSeconds));
    std::filesystem::create_directories(cacheFolderPath + "/simulated_frames");
    std::ofstream log(cacheFolderPath + "/log.txt");
    std::streambuf* saved = std::cout.rdbuf(log.rdbuf());
    try {
        for (size_t i = 0; i < numberOfIterations; ++i) {
            std::cout << "Iteration: " << i << "\n";
            const auto t0 = std::chrono::steady_clock::now();
            simulator->AdvanceSingleFrame();
            simulator->GetSurface(surface);
            obj.Save(cacheFolderPath + "/simulated_frames/" + simulationOutputName + "_" + std::to_string(i) + ".obj",
                     surface);
            const auto t1 = std::chrono::steady_clock::now();
            const double sec = std::chrono::duration<double>(t1 - t0).count();
            _stats.simulationTimeInSeconds += sec;
            _stats.iterationTimeInSeconds.push_back(sec);
            _stats.cflPerIteration.push_back(simulator->Cfl(timeIntervalInSeconds));
            surface.Clear();
            std::cout << std::endl;
        }
        _stats.numberOfIterations = numberOfIterations;
        _stats.timeIntervalInSeconds = timeIntervalInSeconds;
        _stats.simulatorType = typeid(*simulator).name();
        _stats.numberOfParticles = simulator->GetNumberOfParticles();
        _stats.gridSize = simulator->GetResolution();
        _stats.PrintStats();
    } catch (...) {
        std::cout.rdbuf(saved);
        throw;
    }
    std::cout.rdbuf(saved);
}

}  // namespace fluide
