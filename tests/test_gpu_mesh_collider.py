"""GPU parity of mesh colliders through the C ABI against the oracle:
Mesh2SDF::Build (src/3d/mesh_2_sdf.cpp:14-95) on the device, AddCollider of a
TriangleMeshCollider folded into the scene's collider SDF
(src/3d/collider_collection.cpp:33-50), and a substep with such a collider.

Bars: the mesh SDF and the scene's collider SDF bit-exact (FP64 in the
reference's operation order on both sides; the inside test's angle sum only
feeds a threshold that closed meshes clear by far away from their surface);
the substep at the collider substep bars (flags and fluid SDF bit-exact,
velocities within 2e-4 and positions within 2e-6 of their magnitudes).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from fluide_engine_amd import capi  # noqa: E402
from oracle import oracle as orc  # noqa: E402
from meshes import box, icosphere  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


MESHES = {
    "box": lambda: box((0.31, 0.22, 0.27), (0.68, 0.57, 0.74)),
    "icosphere": lambda: icosphere((0.52, 0.41, 0.47), 0.26, levels=2),
    "two": lambda: _two(),
}


def _two():
    v1, t1 = box((0.12, 0.1, 0.15), (0.4, 0.35, 0.5))
    v2, t2 = icosphere((0.7, 0.62, 0.55), 0.2, levels=1)
    return np.vstack([v1, v2]), np.vstack([t1, t2 + len(v1)])


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("grid", [((20, 18, 16), (0.0, 0.0, 0.0), (0.05, 0.055, 0.0625)),
                                  ((23, 21, 25), (-0.1, -0.05, 0.02), (0.05, 0.05, 0.05))])
def test_mesh_sdf_bit_exact(ctx, name, grid):
    size, origin, h = grid
    print(f"[gpu] mesh sdf {name} size={size}", flush=True)
    v, t = MESHES[name]()
    got = capi.mesh_sdf(ctx, size, origin, h, v, t)
    want = orc.mesh_to_sdf(size, origin, h, v, t)
    bad = got != want
    assert not bad.any(), (int(bad.sum()), float(np.abs(got - want).max()),
                           int(((got < 0) != (want < 0)).sum()))
    assert (got < 0).sum() > 0


def test_mesh_sdf_errors(ctx):
    print("[gpu] mesh sdf errors", flush=True)
    v, t = box((0.2, 0.2, 0.2), (0.5, 0.5, 0.5))
    with pytest.raises(capi.FluideError):
        capi.mesh_sdf(ctx, (8, 8, 8), (0, 0, 0), (0.1, 0.1, 0.1), v, t + 100)
    with pytest.raises(capi.FluideError):
        capi.mesh_sdf(ctx, (8, 0, 8), (0, 0, 0), (0.1, 0.1, 0.1), v, t)
    with pytest.raises(capi.FluideError):
        capi.mesh_sdf(ctx, (8, 8, 8), (0, 0, 0), (0.0, 0.1, 0.1), v, t)
    # no triangles: the fill value everywhere, nothing inside
    empty = capi.mesh_sdf(ctx, (5, 4, 3), (0, 0, 0), (0.1, 0.2, 0.1), np.zeros((0, 3)), np.zeros((0, 3), np.int32))
    assert np.all(empty == 0.2 * 12)


def test_scene_mesh_colliders_running_minimum(ctx):
    print("[gpu] scene mesh colliders", flush=True)
    n = (16, 14, 12)
    desc = capi.scene_desc(n, method=capi.FLIP, max_particles=64)
    g = capi.Scene(ctx, desc)
    o = orc.Scene(n)
    h = tuple(1.0 / q for q in n)
    v1, t1 = icosphere((0.3, 0.35, 0.5), 0.2, levels=2)
    v2, t2 = box((0.55, 0.1, 0.2), (0.9, 0.5, 0.7))
    g.add_mesh_collider(n, (0.0, 0.0, 0.0), h, v1, t1)
    o.add_mesh_collider(n, (0.0, 0.0, 0.0), h, v1, t1)
    assert np.array_equal(g.get(capi.GRID_COLLIDER_SDF), o.get(orc.COLLIDER_SDF))
    m, o2, h2 = (23, 19, 17), (-0.05, -0.02, 0.0), (0.05, 0.06, 0.065)
    g.add_mesh_collider(m, o2, h2, v2, t2)
    o.add_mesh_collider(m, o2, h2, v2, t2)
    assert np.array_equal(g.get(capi.GRID_COLLIDER_SDF), o.get(orc.COLLIDER_SDF))
    with pytest.raises(capi.FluideError):
        g.add_mesh_collider(n, (0.0, 0.0, 0.0), h, v1, t1 - 1000)
    g.close()


def test_substep_with_mesh_collider(ctx):
    print("[gpu] substep with mesh collider", flush=True)
    n = (24, 24, 24)
    desc = capi.scene_desc(n, method=capi.FLIP, blend=0.05, max_particles=8 * 24 ** 3 + 64)
    x = capi.seed_particles(desc, [capi.box((-1, -1, -1), (0.5, 0.6, 2.0))], seed=4)
    rng = np.random.default_rng(14)
    x = np.clip(x + rng.uniform(-0.004, 0.004, x.shape).astype(np.float32), 0.0, 1.0).astype(np.float32)
    v = (rng.normal(size=x.shape) * 0.3).astype(np.float32)
    g = capi.Scene(ctx, desc)
    g.upload_particles(x, v)
    o = orc.Scene(n, method=orc.FLIP, blend=0.05, max_cfl=3.0, gravity=tuple(desc.gravity))
    o.set_particles(x.astype(np.float64), v.astype(np.float64))
    h = tuple(1.0 / q for q in n)
    vm, tm = box((0.58, -0.1, 0.3), (0.8, 0.3, 0.72))  # a block on the floor, beside the water column
    g.add_mesh_collider(n, (0.0, 0.0, 0.0), h, vm, tm)
    o.add_mesh_collider(n, (0.0, 0.0, 0.0), h, vm, tm)
    dt = 1.0 / 240
    for step in range(2):
        g.substep(dt, capi.solver(capi.SOLVER_JACOBI, 200, 5, 1e-14))
        o.substep(dt, orc.solver(orc.JACOBI, 200, 5, 1e-14))
        flags = g.get(capi.GRID_FLAGS)
        # the second substep starts from FP32 positions on the GPU and FP64
        # ones in the oracle: a cell whose nearest particle sits within ~1e-7
        # of the SDF radius may flip
        assert int((flags != o.get(orc.FLAGS)).sum()) <= (0 if step == 0 else 2), step
        assert (flags == capi.BOUNDRY).sum() > 20
    gx, gv, _ = g.download_particles()
    ox, ov, _ = o.particles()
    scale = max(1.0, float(np.abs(ov).max()))
    assert float(np.abs(gv - ov).max()) <= 2e-4 * scale
    assert float(np.abs(gx - ox).max()) <= 2e-6 * max(1.0, float(np.abs(ox).max()))
    g.close()
