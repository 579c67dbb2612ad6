"""The drop-in surface of the C++ and Python host layers (CPU only): the
reference's class names and hierarchy (src/main.cpp:16-58,
src/pic_simulator.hpp:29-43, src/flip_simulator.hpp, src/apic_simulator.hpp,
src/python_wrapper/simulators_wrapper.cu:36-85, colliders_wrapper.cpp,
emitters_wrapper.cpp, exteranl_forces_wrapper.cpp), and the reference's own
main.cpp building unchanged against the backend through its header paths
(fluide_engine_amd/host/compat)."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "fluide_engine_amd", "lib")
HOST = os.path.join(ROOT, "fluide_engine_amd", "host")
REF_MAIN = "/root/reference/src/main.cpp"


@pytest.fixture(scope="module")
def P():
    if LIB not in sys.path:
        sys.path.insert(0, LIB)
    import PyFluidEngine

    return PyFluidEngine


@pytest.mark.skipif(not os.path.exists(REF_MAIN), reason="the reference source tree is not on this machine")
def test_reference_main_builds_unchanged(tmp_path):
    # a copy, so that its quoted includes resolve to the compat headers rather
    # than to the reference's own files next to it
    src = tmp_path / "main.cpp"
    shutil.copyfile(REF_MAIN, src)
    exe = tmp_path / "main"
    cmd = ["g++", "-std=c++17", "-O0", "-Wall", "-I", os.path.join(HOST, "compat"), "-I", HOST, "-I",
           os.path.join(ROOT, "include"), str(src), "-o", str(exe), "-L", LIB, "-lfluide_host", "-lfluide_hip",
           "-Wl,-rpath," + LIB]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert exe.exists()


def test_every_compat_header_compiles_alone(tmp_path):
    comp = os.path.join(HOST, "compat")
    headers = sorted(os.path.relpath(os.path.join(d, f), comp) for d, _, fs in os.walk(comp) for f in fs)
    assert "pic_simulator.hpp" in headers and "linear_system/cuda_conjugate_gradient_solver.hpp" in headers
    src = tmp_path / "all.cpp"
    src.write_text("".join(f'#include "{h}"\n' for h in headers) +
                   "int main() { std::shared_ptr<PICSimulator> p; std::shared_ptr<HybridSimulator> h = p; "
                   "std::shared_ptr<FLIPSimulator> f; p = f; std::shared_ptr<APICSimulator> a; p = a; "
                   "std::shared_ptr<Collider> c = std::shared_ptr<TriangleMeshCollider>(); "
                   "std::shared_ptr<Emitter> e = std::shared_ptr<VolumeEmitter>(); "
                   "std::shared_ptr<PressureSolver> ps = std::make_shared<SinglePhasePressureSolver>(); "
                   "std::shared_ptr<DiffusionSolver> ds = std::make_shared<BackwardEulerDiffusionSolver>(); "
                   "(void)h; (void)c; (void)e; (void)ps; (void)ds; return 0; }\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", comp, "-I", HOST, "-I",
                        os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]


def test_python_class_hierarchy(P):
    assert issubclass(P.FLIPSimulator, P.PICSimulator) and issubclass(P.APICSimulator, P.PICSimulator)
    assert issubclass(P.PICSimulator, P.HybridSimulator)
    assert issubclass(P.TriangleMeshCollider, P.Collider)
    assert issubclass(P.VolumeEmitter, P.Emitter)
    assert issubclass(P.PointField, P.ExternalForce) and issubclass(P.DirectionalField, P.ExternalForce)


def test_non_uniform_force_is_rejected(P):
    sim = P.FLIPSimulator(P.Vector3S(8, 8, 8), P.BoundingBox3D(P.Vector3D(0, 0, 0), P.Vector3D(1, 1, 1)))
    with pytest.raises(NotImplementedError):
        sim.AddExternalForce(P.PointField(P.Vector3D(0.5, 0.5, 0.5), 2.0, 1.0))
    sim.AddExternalForce(P.PointField(P.Vector3D(0.5, 0.5, 0.5), 0.0, 1.0))  # zero strength: uniform
