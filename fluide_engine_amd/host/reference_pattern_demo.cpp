// reference_pattern_demo.cpp -- the reference's own driver pattern
// (src/main.cpp:16-58: RunSimulation over a shared_ptr<PICSimulator>, an
// emitter loaded through OBJManager, gravity, viscosity, maxCfl 5,
// SimulationRunner writing one OBJ per frame) written against the reference's
// header paths, which host/compat maps onto the gfx950 backend.  It adds the
// reference's solver and collider setters (src/pic_simulator.hpp:41-43):
// SinglePhasePressureSolver with a CudaConjugateGradientSolver,
// BackwardEulerDiffusionSolver and a ColliderCollection with one
// TriangleMeshCollider.
//   reference_pattern_demo <pic|flip|apic> <resolution> <frames> <output folder>
// prints one JSON line per run.
#include "3d/collider_collection.hpp"
#include "3d/obj_manager.hpp"
#include "3d/triangle_mesh_collider.hpp"
#include "apic_simulator.hpp"
#include "cuda_pic_simulator.hpp"
#include "emmiter/volume_emmiter.hpp"
#include "flip_simulator.hpp"
#include "fluid_solvers/backward_euler_diffusion_solver.hpp"
#include "fluid_solvers/single_phase_pressure_solver.hpp"
#include "forces/directional_field.hpp"
#include "linear_system/cuda_conjugate_gradient_solver.hpp"
#include "pic_simulator.hpp"
#include "simulation_runner.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <iostream>
#include <memory>
#include <string>

// An axis-aligned box [lo, hi] as 12 outward-facing triangles.
static TriangleMesh BoxMesh(Vector3<double> lo, Vector3<double> hi, const std::string& name) {
    TriangleMesh m;
    m.SetObjectName(name);
    for (int c = 0; c < 8; ++c)
        m.AddVertex(Vector3<double>(c & 1 ? hi.x : lo.x, c & 2 ? hi.y : lo.y, c & 4 ? hi.z : lo.z));
    const int f[12][3] = {{0, 2, 1}, {1, 2, 3}, {4, 5, 6}, {5, 7, 6}, {0, 1, 4}, {1, 5, 4},
                          {2, 6, 3}, {3, 6, 7}, {0, 4, 2}, {2, 4, 6}, {1, 3, 5}, {3, 7, 5}};
    m.AddNormal(Vector3<double>(0, 0, 1));
    for (const auto& t : f) m.AddTriangle(Triangle3D(t[0], t[1], t[2], 0));
    return m;
}

static void RunSimulation(std::shared_ptr<PICSimulator> simulator, size_t numberOfIterations,
                          double timeIntervalInSeconds, std::string prefix, Vector3<size_t> size, BoundingBox3D domain,
                          const std::string& folder) {
    // the emitter goes through an OBJ file, as the reference loads it
    OBJManager objLoader;
    std::filesystem::create_directories(folder);
    objLoader.Save(folder + "/emitter.obj", BoxMesh(Vector3<double>(2.5, 0.5, 2.5), Vector3<double>(3.5, 1.5, 3.5), "emitter"));
    TriangleMesh emitterObject;
    objLoader.Load(folder + "/emitter.obj", emitterObject);

    simulator->AddExternalForce(std::make_shared<DirectionalField>(Vector3<double>(0, -9.81, 0)));
    simulator->AddEmitter(std::make_shared<VolumeEmitter>(emitterObject, size, domain, 12,
                                                          Vector3<double>(0.0, 6.0, 0.0), 0.0));
    simulator->SetViscosity(0.01);
    simulator->SetMaxClf(5);

    auto pressure = std::make_shared<SinglePhasePressureSolver>();
    pressure->SetLinearSystemSolver(std::make_shared<CudaConjugateGradientSolver>(2000, 1e-8));
    simulator->SetPressureSolver(pressure);
    simulator->SetDiffusionSolver(std::make_shared<BackwardEulerDiffusionSolver>());
    const Vector3<double> h(domain.GetSize().x / size.x, domain.GetSize().y / size.y, domain.GetSize().z / size.z);
    auto colliders = std::make_shared<ColliderCollection>(size, h, domain.GetOrigin());
    colliders->AddCollider(std::make_shared<TriangleMeshCollider>(
        size, domain.GetOrigin(), h, BoxMesh(Vector3<double>(1.0, 0.0, 1.0), Vector3<double>(2.0, 1.0, 2.0), "block")));
    simulator->SetColliders(colliders);

    SimulationRunner runner;
    runner.RunSimulation(simulator, timeIntervalInSeconds, numberOfIterations, prefix, folder + "/" + prefix + "/");
    std::cerr << "Simulation ended!\n";
    const SimulationStats st = runner.GetStats();
    std::printf("{\"simulator\": \"%s\", \"frames\": %zu, \"particles\": %zu, \"colliders\": %zu, \"max_cfl\": %g, "
                "\"viscosity\": %g}\n",
                prefix.c_str(), st.numberOfIterations, simulator->GetNumberOfParticles(),
                simulator->NumberOfColliders(), simulator->MaxCfl(), simulator->GetViscosity());
    std::fflush(stdout);
}

int main(int argc, char** argv) {
    const char* which = argc > 1 ? argv[1] : "flip";
    const size_t resolution = argc > 2 ? (size_t)std::atoi(argv[2]) : 30;
    const size_t numberOfIterations = argc > 3 ? (size_t)std::atoi(argv[3]) : 5;
    const std::string out = argc > 4 ? argv[4] : "simOutputs";
    const double timeIntervalInSeconds = 1.0 / 60.0;
    try {
        const Vector3<size_t> size(resolution, resolution, resolution);
        BoundingBox3D domain(Vector3<double>(0, 0, 0), Vector3<double>(6, 6, 6));
        std::shared_ptr<PICSimulator> simulator;
        if (!std::strcmp(which, "pic")) simulator = std::make_shared<PICSimulator>(size, domain);
        else if (!std::strcmp(which, "apic")) simulator = std::make_shared<APICSimulator>(size, domain);
        else simulator = std::make_shared<FLIPSimulator>(size, domain);
        RunSimulation(simulator, numberOfIterations, timeIntervalInSeconds, std::string(which) + "_out", size, domain,
                      out);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "reference_pattern_demo: %s\n", e.what());
        return 1;
    }
}
