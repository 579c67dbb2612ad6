// global_names.hpp -- the reference's class names in the global namespace, as
// the reference declares them, for code that includes the reference's own
// header paths (the headers under host/compat/ include this file).
#pragma once

#include "fluide/simulators.hpp"

using fluide::APICSimulator;
using fluide::BackwardEulerDiffusionSolver;
using fluide::BoundingBox3D;
using fluide::Collider;
using fluide::ColliderCollection;
using fluide::ConjugateGradientSolver;
using fluide::CudaAPICSimulator;
using fluide::CudaConjugateGradientSolver;
using fluide::CudaFLIPSimulator;
using fluide::CudaJacobiIterationSolver;
using fluide::CudaPICSimulator;
using fluide::DiffusionSolver;
using fluide::DirectionalField;
using fluide::Emitter;
using fluide::ExternalForce;
using fluide::FLIPSimulator;
using fluide::ForwardEulerDiffusionSolver;
using fluide::Frame;
using fluide::HybridSimulator;
using fluide::JacobiIterationSolver;
using fluide::LinearSystemSolver;
using fluide::OBJManager;
using fluide::PICSimulator;
using fluide::PointField;
using fluide::PressureSolver;
using fluide::ScalarGrid3D;
using fluide::SimulationRunner;
using fluide::SimulationStats;
using fluide::SinglePhasePressureSolver;
using fluide::Triangle3D;
using fluide::Triangle3D_t;
using fluide::TriangleMesh;
using fluide::TriangleMeshCollider;
using fluide::Vector3;
using fluide::VolumeEmitter;
