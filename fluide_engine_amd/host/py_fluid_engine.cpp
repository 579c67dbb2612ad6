// py_fluid_engine.cpp -- module PyFluidEngine with the class and method names
// of the reference's pybind11 wrapper (src/python_wrapper/main.cpp:20-33,
// vector3_wrapper.cpp, bounding_box_3d_wrapper.cpp, frame_wrapper.cpp,
// exteranl_forces_wrapper.cpp, simulators_wrapper.cu:36-85), backed by the
// gfx950 simulators.  Added: particle/grid transfer as numpy arrays, analytic
// scene initialisation and the pressure-solver selection.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "fluide/simulators.hpp"

namespace py = pybind11;
using namespace fluide;

namespace {

using FArr = py::array_t<float, py::array::c_style | py::array::forcecast>;

fe_seed_shape box(std::vector<double> lo, std::vector<double> hi) {
    if (lo.size() != 3 || hi.size() != 3) throw py::value_error("box corners need 3 coordinates");
    fe_seed_shape s{};
    s.kind = 0;
    for (int a = 0; a < 3; ++a) {
        s.lo[a] = lo[a];
        s.hi[a] = hi[a];
    }
    return s;
}

fe_seed_shape sphere(std::vector<double> centre, double radius) {
    if (centre.size() != 3) throw py::value_error("sphere centre needs 3 coordinates");
    fe_seed_shape s{};
    s.kind = 1;
    for (int a = 0; a < 3; ++a) s.lo[a] = centre[a];
    s.hi[0] = radius;
    return s;
}

void set_particles(HipHybridSimulator& sim, FArr x, FArr v, py::object C) {
    if (x.ndim() != 2 || x.shape(1) != 3) throw py::value_error("x must be (n, 3)");
    if (v.ndim() != 2 || v.shape(1) != 3 || v.shape(0) != x.shape(0)) throw py::value_error("v must be (n, 3)");
    FArr c;
    const float* cp = nullptr;
    if (!C.is_none()) {
        c = C.cast<FArr>();
        if (c.ndim() != 2 || c.shape(1) != 9 || c.shape(0) != x.shape(0)) throw py::value_error("C must be (n, 9)");
        cp = c.data();
    }
    sim.SetParticles(x.data(), v.data(), cp, (int64_t)x.shape(0));
}

py::tuple get_particles(const HipHybridSimulator& sim) {
    std::vector<float> x, v, C;
    sim.GetParticles(&x, &v, &C);
    const py::ssize_t n = (py::ssize_t)(x.size() / 3);
    FArr ax({n, (py::ssize_t)3}), av({n, (py::ssize_t)3});
    std::copy(x.begin(), x.end(), ax.mutable_data());
    std::copy(v.begin(), v.end(), av.mutable_data());
    if (sim.Method() != FE_APIC) return py::make_tuple(ax, av, py::none());
    FArr ac({n, (py::ssize_t)9});
    std::copy(C.begin(), C.end(), ac.mutable_data());
    return py::make_tuple(ax, av, ac);
}

py::array_t<double> vec3_array(const std::vector<Vector3<double>>& v) {
    py::array_t<double> a({(py::ssize_t)v.size(), (py::ssize_t)3});
    auto* d = a.mutable_data();
    for (const auto& p : v) {
        *d++ = p.x;
        *d++ = p.y;
        *d++ = p.z;
    }
    return a;
}

// Grid field as a numpy array shaped (nz, ny, nx) of the field's own dims.
py::array get_grid(const HipHybridSimulator& sim, int field) {
    const Vector3<size_t> n = sim.GetResolution();
    size_t d[3] = {n.x, n.y, n.z};
    if (field >= FE_GRID_U && field <= FE_GRID_W) d[field] += 1;
    if (field >= FE_GRID_MU && field <= FE_GRID_MW) d[field - FE_GRID_MU] += 1;
    if (field >= FE_GRID_SU && field <= FE_GRID_SW) d[field - FE_GRID_SU] += 1;
    if (field >= FE_GRID_DMU && field <= FE_GRID_DMW) d[field - FE_GRID_DMU] += 1;
    const std::vector<py::ssize_t> shape = {(py::ssize_t)d[2], (py::ssize_t)d[1], (py::ssize_t)d[0]};
    py::array out;
    if (field == FE_GRID_SDF || field == FE_GRID_PRESSURE || field == FE_GRID_COLLIDER_SDF)
        out = py::array_t<double>(shape);
    else if ((field >= FE_GRID_MU && field <= FE_GRID_MW) || field == FE_GRID_FLAGS ||
             (field >= FE_GRID_DMU && field <= FE_GRID_DMW))
        out = py::array_t<uint8_t>(shape);
    else out = py::array_t<float>(shape);
    sim.GetGrid(field, out.mutable_data(), (size_t)out.nbytes());
    return out;
}

py::dict stats_dict(const fe_solve_stats& s) {
    py::dict d;
    d["iterations"] = s.iterations;
    d["converged"] = s.converged != 0;
    d["rel_residual"] = s.rel_residual;
    d["abs_residual"] = s.abs_residual;
    d["ms"] = s.ms;
    return d;
}

py::dict stats(const HipHybridSimulator& sim) { return stats_dict(sim.LastSolveStats()); }

py::list diffusion_stats(const HipHybridSimulator& sim) {
    py::list out;
    for (const fe_solve_stats& s : sim.LastDiffusionStats()) out.append(stats_dict(s));
    return out;
}

void set_diffusion_solver(HipHybridSimulator& sim, int max_iterations, int check_interval, double tolerance) {
    DiffusionSolverSettings s;
    s.maxIterations = max_iterations;
    s.checkInterval = check_interval;
    s.tolerance = tolerance;
    sim.SetDiffusionSolver(s);
}

void set_solver(HipHybridSimulator& sim, const std::string& kind, int max_iterations, double tolerance, bool relative,
                bool warm_start, int check_interval) {
    PressureSolverSettings s;
    if (kind == "jacobi") s.kind = FE_SOLVER_JACOBI;
    else if (kind == "cg") s.kind = FE_SOLVER_CG;
    else if (kind == "pcg") s.kind = FE_SOLVER_PCG_JACOBI;
    else if (kind == "mgpcg") s.kind = FE_SOLVER_MGPCG;
    else throw py::value_error("solver kind must be one of jacobi, cg, pcg, mgpcg");
    s.maxIterations = max_iterations;
    s.tolerance = tolerance;
    s.relative = relative;
    s.warmStart = warm_start;
    s.checkInterval = check_interval;
    sim.SetPressureSolver(s);
}

template <typename Sim, typename Base>
py::class_<Sim, std::shared_ptr<Sim>, Base> add_sim(py::module& m, const char* name) {
    auto c = py::class_<Sim, std::shared_ptr<Sim>, Base>(m, name)
                 .def(py::init<Vector3<size_t>, BoundingBox3D>())
                 .def("AdvanceSingleFrame", &Sim::AdvanceSingleFrame)
                 .def("InitializeFromTriangleMesh", &Sim::InitializeFromTriangleMesh, py::arg("mesh"),
                      py::arg("seed") = 0)
                 .def("AddExternalForce", &Sim::AddExternalForce)
                 .def("AddCollider", &Sim::AddCollider)
                 .def("AddEmitter", &Sim::AddEmitter)
                 .def("SetCurrentFrame", &Sim::SetCurrentFrame)
                 .def("SetViscosity", &Sim::SetViscosity)
                 .def("SetMaxClf", &Sim::SetMaxClf);
    return c;
}

}  // namespace

PYBIND11_MODULE(PyFluidEngine, m) {
    m.doc() = "Fluide_Engine hybrid particle-grid simulators on the gfx950 (MI355X) backend";
    py::register_exception<Unsupported>(m, "Unsupported", PyExc_NotImplementedError);
    py::register_exception<Error>(m, "FluideError", PyExc_RuntimeError);

    py::class_<Frame>(m, "Frame")
        .def(py::init<double>())
        .def(py::init<>())
        .def("GetIndex", &Frame::GetIndex)
        .def("GetTimeIntervalInSeconds", &Frame::GetTimeIntervalInSeconds)
        .def("GetTimeInSeconds", &Frame::GetTimeInSeconds)
        .def("Advance", py::overload_cast<>(&Frame::Advance));

    py::class_<Vector3<double>>(m, "Vector3D")
        .def(py::init<double, double, double>())
        .def_readwrite("x", &Vector3<double>::x)
        .def_readwrite("y", &Vector3<double>::y)
        .def_readwrite("z", &Vector3<double>::z);
    py::class_<Vector3<size_t>>(m, "Vector3S")
        .def(py::init<size_t, size_t, size_t>())
        .def_readwrite("x", &Vector3<size_t>::x)
        .def_readwrite("y", &Vector3<size_t>::y)
        .def_readwrite("z", &Vector3<size_t>::z);

    py::class_<BoundingBox3D>(m, "BoundingBox3D")
        .def(py::init<Vector3<double>, Vector3<double>>())
        .def("GetOrigin", &BoundingBox3D::GetOrigin)
        .def("GetSize", &BoundingBox3D::GetSize);

    py::class_<ExternalForce, std::shared_ptr<ExternalForce>>(m, "ExternalForce");
    py::class_<DirectionalField, std::shared_ptr<DirectionalField>, ExternalForce>(m, "DirectionalField")
        .def(py::init<Vector3<double>>());
    py::class_<PointField, std::shared_ptr<PointField>, ExternalForce>(m, "PointField")
        .def(py::init<Vector3<double>, double, double>());

    // src/python_wrapper/triangle_mesh_wrapper.cpp, obj_manager_wrapper.cpp,
    // simulator_runner_wrapper.cpp; the mesh also hands its arrays to numpy
    py::class_<TriangleMesh>(m, "TriangleMesh")
        .def(py::init<>())
        .def("Clear", &TriangleMesh::Clear)
        .def("GetObjectName", &TriangleMesh::GetObjectName)
        .def("SetObjectName", &TriangleMesh::SetObjectName)
        .def("GetVerticies", [](const TriangleMesh& t) { return vec3_array(t.GetVerticies()); })
        .def("GetNormals", [](const TriangleMesh& t) { return vec3_array(t.GetNormals()); })
        .def("GetTriangles", [](const TriangleMesh& t) {
            py::array_t<int64_t> a({(py::ssize_t)t.GetTriangles().size(), (py::ssize_t)4});
            auto* d = a.mutable_data();
            for (const auto& tr : t.GetTriangles()) {
                *d++ = (int64_t)tr.point1Idx;
                *d++ = (int64_t)tr.point2Idx;
                *d++ = (int64_t)tr.point3Idx;
                *d++ = (int64_t)tr.normalIdx;
            }
            return a;
        });
    // src/python_wrapper/emitters_wrapper.cpp:15-23; InitializeFromSdf takes the
    // emitter grid's values as a (nz, ny, nx) array
    py::class_<Emitter, std::shared_ptr<Emitter>>(m, "Emitter");
    py::class_<VolumeEmitter, std::shared_ptr<VolumeEmitter>, Emitter>(m, "VolumeEmitter")
        .def(py::init<const TriangleMesh&, Vector3<size_t>, BoundingBox3D, const size_t&, const Vector3<double>&,
                      const Vector3<double>&>())
        .def("InitializeFromTriangleMesh", &VolumeEmitter::InitializeFromTriangleMesh)
        .def("InitializeFromSdf",
             [](VolumeEmitter& e, py::array_t<double, py::array::c_style | py::array::forcecast> a) {
                 if (a.ndim() != 3) throw py::value_error("InitializeFromSdf: a (nz, ny, nx) array");
                 ScalarGrid3D g;
                 g.size = Vector3<size_t>((size_t)a.shape(2), (size_t)a.shape(1), (size_t)a.shape(0));
                 g.data.assign(a.data(), a.data() + a.size());
                 e.InitializeFromSdf(g);
             })
        .def("SetSeed", &VolumeEmitter::SetSeed);
    // src/python_wrapper/colliders_wrapper.cpp:14-17
    py::class_<Collider, std::shared_ptr<Collider>>(m, "Collider");
    py::class_<TriangleMeshCollider, std::shared_ptr<TriangleMeshCollider>, Collider>(m, "TriangleMeshCollider")
        .def(py::init<const Vector3<size_t>&, const Vector3<double>&, const Vector3<double>&, const TriangleMesh&>());
    py::class_<OBJManager>(m, "OBJManager")
        .def(py::init<>())
        .def("Save", &OBJManager::Save)
        .def("Load", &OBJManager::Load);
    py::class_<SimulationStats>(m, "SimulationStats")
        .def(py::init<>())
        .def("PrintStats", &SimulationStats::PrintStats)
        .def("Clear", &SimulationStats::Clear)
        .def_readwrite("simulatorType", &SimulationStats::simulatorType)
        .def_readwrite("numberOfIterations", &SimulationStats::numberOfIterations)
        .def_readwrite("gridSize", &SimulationStats::gridSize)
        .def_readwrite("numberOfParticles", &SimulationStats::numberOfParticles)
        .def_readwrite("timeIntervalInSeconds", &SimulationStats::timeIntervalInSeconds)
        .def_readwrite("simulationTimeInSeconds", &SimulationStats::simulationTimeInSeconds)
        .def_readwrite("iterationTimeInSeconds", &SimulationStats::iterationTimeInSeconds)
        .def_readwrite("cflPerIteration", &SimulationStats::cflPerIteration);

    py::class_<HipHybridSimulator, std::shared_ptr<HipHybridSimulator>>(m, "HybridSimulator")
        .def("GetSurface", &HipHybridSimulator::GetSurface)
        .def("MaxCfl", &HipHybridSimulator::MaxCfl)
        .def("GetViscosity", &HipHybridSimulator::GetViscosity)
        .def("GetDensity", &HipHybridSimulator::GetDensity)
        .def("GetParticlesPerBlock", &HipHybridSimulator::GetParticlesPerBlock)
        .def("GetOrigin", &HipHybridSimulator::GetOrigin)
        .def("GetResolution", &HipHybridSimulator::GetResolution)
        .def("GetGridSpacing", &HipHybridSimulator::GetGridSpacing)
        .def("GetNumberOfParticles", &HipHybridSimulator::GetNumberOfParticles)
        .def("GetCurrentFrame", &HipHybridSimulator::GetCurrentFrame)
        .def("GetCurrentTimeInSeconds", &HipHybridSimulator::GetCurrentTimeInSeconds)
        .def("GetNumberOfSubTimeSteps", &HipHybridSimulator::GetNumberOfSubTimeSteps)
        .def("Cfl", &HipHybridSimulator::Cfl)
        .def("SetDensity", &HipHybridSimulator::SetDensity)
        .def("SetParticlesPerBlock", &HipHybridSimulator::SetParticlesPerBlock)
        .def("GetFluidSdf",
             [](const HipHybridSimulator& s) {
                 const ScalarGrid3D& g = s.GetFluidSdf();
                 py::array_t<double> a({(py::ssize_t)g.size.z, (py::ssize_t)g.size.y, (py::ssize_t)g.size.x});
                 std::copy(g.data.begin(), g.data.end(), a.mutable_data());
                 return a;
             })
        // GPU-backend additions
        .def("InitializeParticles", &HipHybridSimulator::InitializeParticles, py::arg("shapes"), py::arg("seed") = 0)
        .def("SetParticles", &set_particles, py::arg("x"), py::arg("v"), py::arg("C") = py::none())
        .def("GetParticles", &get_particles)
        .def("GetGrid", &get_grid)
        .def("LastSolveStats", &stats)
        .def("SetPressureSolver", &set_solver, py::arg("kind"), py::arg("max_iterations") = 1000,
             py::arg("tolerance") = 1e-14, py::arg("relative") = false, py::arg("warm_start") = true,
             py::arg("check_interval") = 5)
        .def("SetDiffusionSolver", &set_diffusion_solver, py::arg("max_iterations") = 1000,
             py::arg("check_interval") = 5, py::arg("tolerance") = 1e-14)
        .def("LastDiffusionStats", &diffusion_stats);

    py::class_<fe_seed_shape>(m, "SeedShape");
    m.def("Box", &box, py::arg("lo"), py::arg("hi"), "Particle-seeding box [lo, hi)");
    m.def("Sphere", &sphere, py::arg("centre"), py::arg("radius"), "Particle-seeding sphere");

    // src/python_wrapper/simulators_wrapper.cu:47-84: APIC and FLIP derive from PIC
    add_sim<HipPICSimulator, HipHybridSimulator>(m, "PICSimulator");
    add_sim<HipAPICSimulator, HipPICSimulator>(m, "APICSimulator");
    add_sim<HipFLIPSimulator, HipPICSimulator>(m, "FLIPSimulator")
        .def("SetBlendingFactor", &HipFLIPSimulator::SetBlendingFactor)
        .def("GetBlendingFactor", &HipFLIPSimulator::GetBlendingFactor);
    py::class_<SimulationRunner>(m, "SimulationRunner")
        .def(py::init<>())
        .def("RunSimulation", &SimulationRunner::RunSimulation)
        .def("GetStats", &SimulationRunner::GetStats);

    m.attr("CudaPICSimulator") = m.attr("PICSimulator");
    m.attr("CudaFLIPSimulator") = m.attr("FLIPSimulator");
    m.attr("CudaAPICSimulator") = m.attr("APICSimulator");

    m.attr("GRID_U") = (int)FE_GRID_U;
    m.attr("GRID_V") = (int)FE_GRID_V;
    m.attr("GRID_W") = (int)FE_GRID_W;
    m.attr("GRID_SDF") = (int)FE_GRID_SDF;
    m.attr("GRID_FLAGS") = (int)FE_GRID_FLAGS;
    m.attr("GRID_PRESSURE") = (int)FE_GRID_PRESSURE;
    m.attr("GRID_DMU") = (int)FE_GRID_DMU;
    m.attr("GRID_DMV") = (int)FE_GRID_DMV;
    m.attr("GRID_DMW") = (int)FE_GRID_DMW;
    m.attr("GRID_COLLIDER_SDF") = (int)FE_GRID_COLLIDER_SDF;
}
