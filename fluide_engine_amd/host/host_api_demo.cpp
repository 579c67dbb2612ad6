// host_api_demo.cpp -- the reference's driver pattern (src/main.cpp:52-58,
// src/simulation_runner.cpp:17-55) on the GPU simulators: build a dam-break
// scene, add gravity, advance frames, print one JSON line.
//   host_api_demo <pic|flip|apic> <n> <frames> <frame_dt> <solver: jacobi|mgpcg>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "fluide/simulators.hpp"

using namespace fluide;

int main(int argc, char** argv) {
    const char* method = argc > 1 ? argv[1] : "flip";
    const size_t n = argc > 2 ? (size_t)std::atoi(argv[2]) : 32;
    const int frames = argc > 3 ? std::atoi(argv[3]) : 5;
    const double dt = argc > 4 ? std::atof(argv[4]) : 1.0 / 240.0;
    const char* solver = argc > 5 ? argv[5] : "jacobi";
    try {
        const Vector3<size_t> size(n, n, n);
        const BoundingBox3D domain(Vector3<double>(0, 0, 0), Vector3<double>(1, 1, 1));
        std::shared_ptr<HipHybridSimulator> sim;
        if (!std::strcmp(method, "pic")) sim = std::make_shared<CudaPICSimulator>(size, domain);
        else if (!std::strcmp(method, "apic")) sim = std::make_shared<CudaAPICSimulator>(size, domain);
        else sim = std::make_shared<CudaFLIPSimulator>(size, domain);
        sim->AddExternalForce(std::make_shared<DirectionalField>(Vector3<double>(0, -9.81, 0)));
        PressureSolverSettings ps;
        if (!std::strcmp(solver, "mgpcg")) {
            ps.kind = FE_SOLVER_MGPCG;
            ps.maxIterations = 2000;
            ps.checkInterval = 0;
            ps.tolerance = 1e-6;
            ps.relative = true;
        } else {
            ps.maxIterations = 200;
            ps.tolerance = 1e-14;
        }
        sim->SetPressureSolver(ps);
        fe_seed_shape dam{};
        dam.kind = 0;
        const double lo[3] = {-1, -1, -1}, hi[3] = {0.4, 0.6, 2.0};
        for (int a = 0; a < 3; ++a) {
            dam.lo[a] = lo[a];
            dam.hi[a] = hi[a];
        }
        sim->InitializeParticles({dam}, 0);
        sim->SetCurrentFrame(Frame(dt));
        unsigned total = 0;
        for (int f = 0; f < frames; ++f) {
            sim->AdvanceSingleFrame();
            total += sim->GetNumberOfSubTimeSteps();
        }
        std::vector<float> x, v;
        sim->GetParticles(&x, &v, nullptr);
        double c[3] = {0, 0, 0};
        const size_t np = x.size() / 3;
        for (size_t p = 0; p < np; ++p)
            for (int a = 0; a < 3; ++a) c[a] += x[3 * p + a];
        for (int a = 0; a < 3; ++a) c[a] /= (double)(np ? np : 1);
        const fe_solve_stats st = sim->LastSolveStats();
        std::printf(
            "{\"method\": \"%s\", \"n\": %zu, \"particles\": %zu, \"frames\": %d, \"substeps\": %u, \"time\": %.9g, "
            "\"centroid\": [%.9g, %.9g, %.9g], \"last_iterations\": %d, \"last_rel_residual\": %.6g}\n",
            method, n, sim->GetNumberOfParticles(), frames, total, sim->GetCurrentTimeInSeconds(), c[0], c[1], c[2],
            st.iterations, st.rel_residual);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "host_api_demo: %s\n", e.what());
        return 1;
    }
}
