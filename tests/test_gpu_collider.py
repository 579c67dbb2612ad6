"""GPU parity of the collider path through the C ABI against the oracle on
the same inputs: the collider branches of BlockedBoundryConditionSolver::
ConstrainVelocity (src/fluid_solvers/blocked_boundry_condition_solver.cpp:18-291),
BOUNDRY cell and face markers, ExtrapolateIntoCollider on the fluid SDF
(src/pic_simulator.cpp:497-519), the collision response of moved particles,
the collider terms of the diffusion system and of GetSurface.

Bars:
  * markers (cells, diffusion faces), the blocked/inside decisions and the
    fluid SDF with the reference's bucket search -- bit-exact (FP64 samples
    in the reference's expressions on both sides);
  * the boundary condition's velocities -- FP32 storage on the GPU against
    FP64 in the oracle: within 1e-5 of the field's magnitude;
  * particle collision -- positions within 1e-6, velocities within 1e-5;
  * a substep with a collider -- flags bit-exact, velocities within 2e-4 and
    positions within 2e-6 of their magnitudes (the inviscid substep bar);
  * the surface -- as the surface parity test (1e-12).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from fluide_engine_amd import capi  # noqa: E402
from oracle import oracle as orc  # noqa: E402

COMPS = ((capi.GRID_U, orc.U), (capi.GRID_V, orc.V), (capi.GRID_W, orc.W))
DMARK = (capi.GRID_DMU, capi.GRID_DMV, capi.GRID_DMW)


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def _nodes(n, size=(1.0, 1.0, 1.0)):
    """Collider SDF nodes origin + h*(i,j,k), arrays in (z, y, x) order."""
    h = [size[a] / n[a] for a in range(3)]
    ax = [np.arange(n[a]) * h[a] for a in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return X, Y, Z


def _sphere(n, c, r, size=(1.0, 1.0, 1.0)):
    X, Y, Z = _nodes(n, size)
    return np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) - r


def _pair(ctx, n, csdf, max_cfl=3.0, viscosity=0.0, max_particles=64, size=(1.0, 1.0, 1.0)):
    desc = capi.scene_desc(n, method=capi.FLIP, size=size, blend=0.05, max_cfl=max_cfl, viscosity=viscosity,
                           max_particles=max_particles)
    g = capi.Scene(ctx, desc)
    o = orc.Scene(n, size=size, method=orc.FLIP, blend=0.05, max_cfl=max_cfl, gravity=tuple(desc.gravity))
    if viscosity > 0:
        o.set_viscosity(viscosity)
    if csdf is not None:
        g.set_collider_sdf(csdf)
        o.set(orc.COLLIDER_SDF, csdf)
    return g, o


def _close(got, want, rtol, what):
    scale = max(1.0, float(np.abs(want).max()))
    err = float(np.abs(got.astype(np.float64) - want).max())
    assert err <= rtol * scale, f"{what}: {err:.3e} (scale {scale:.3e})"


def _random_velocity(g, o, seed):
    rng = np.random.default_rng(seed)
    for gf, of in COMPS:
        f = rng.normal(size=g.shape(gf)).astype(np.float32)
        g.set(gf, f)
        o.set(of, f.astype(np.float64))


@pytest.mark.parametrize("case", ["sphere", "plane", "two_spheres"])
def test_bc_stage_parity(ctx, case):
    print(f"[gpu] collider bc stage {case}", flush=True)
    n = {"sphere": (20, 18, 22), "plane": (16, 16, 16), "two_spheres": (24, 20, 16)}[case]
    if case == "sphere":
        csdf = _sphere(n, (0.5, 0.4, 0.55), 0.23)
    elif case == "plane":
        X, Y, Z = _nodes(n)
        csdf = 0.61 - (0.8 * X + 0.6 * Z)  # collider where 0.8x + 0.6z > 0.61
    else:
        csdf = np.minimum(_sphere(n, (0.3, 0.3, 0.5), 0.18), _sphere(n, (0.72, 0.6, 0.4), 0.2))
    g, o = _pair(ctx, n, csdf)
    _random_velocity(g, o, seed=sum(n))
    g.run_stage(capi.STAGE_BC, 0.0)
    o.apply_bc()
    for gf, of in COMPS:
        _close(g.get(gf), o.get(of), 1e-5, f"component {gf}")
    # collider faces and walls: the zeros are exact zeros on both sides
    for gf, of in COMPS:
        assert np.array_equal(g.get(gf) == 0, o.get(of) == 0), f"zero pattern {gf}"
    g.close()


def test_bc_without_collider_after_clear(ctx):
    print("[gpu] collider cleared", flush=True)
    n = (12, 10, 14)
    g, o = _pair(ctx, n, _sphere(n, (0.5, 0.5, 0.5), 0.3))
    g.clear_colliders()
    o2 = orc.Scene(n, method=orc.FLIP, blend=0.05, max_cfl=3.0)
    _random_velocity(g, o2, seed=3)
    g.run_stage(capi.STAGE_BC, 0.0)
    o2.apply_bc()
    for gf, of in COMPS:
        assert np.array_equal(g.get(gf).astype(np.float64), o2.get(of))
    with pytest.raises(capi.FluideError):
        g.get(capi.GRID_COLLIDER_SDF)
    g.close()


def test_markers_with_collider(ctx):
    print("[gpu] collider markers", flush=True)
    n = (18, 14, 16)
    csdf = _sphere(n, (0.45, 0.3, 0.5), 0.27)
    g, o = _pair(ctx, n, csdf)
    rng = np.random.default_rng(5)
    sdf = rng.uniform(-0.5, 1.0, (n[2], n[1], n[0]))
    g.set(capi.GRID_SDF, sdf)
    o.set(orc.SDF, sdf)
    g.run_stage(capi.STAGE_MARKERS, 0.0)
    o.build_markers()
    flags = g.get(capi.GRID_FLAGS)
    assert np.array_equal(flags, o.get(orc.FLAGS))
    assert (flags == capi.BOUNDRY).sum() > 50
    assert np.array_equal(g.get(capi.GRID_COLLIDER_SDF), csdf)
    g.close()


def _particles(n, boxes, seed, jitter=0.004):
    desc = capi.scene_desc(n, method=capi.FLIP, max_particles=8 * int(np.prod(n)) + 64)
    x = capi.seed_particles(desc, [capi.box(lo, hi) for lo, hi in boxes], seed=seed)
    rng = np.random.default_rng(seed + 100)
    x = np.clip(x + rng.uniform(-jitter, jitter, x.shape).astype(np.float32), 0.0, 1.0).astype(np.float32)
    v = (rng.normal(size=x.shape) * 0.3).astype(np.float32)
    return x, v


@pytest.mark.parametrize("collider", [False, True])
def test_sdf_bucket_search_bit_exact(ctx, collider):
    """With the reference's bucket search every SDF value, positive ones
    included, is the oracle's; with a collider the values inside it are the
    extrapolated ones, and the markers follow."""
    print(f"[gpu] sdf bucket search collider={collider}", flush=True)
    n = (20, 20, 20)
    x, v = _particles(n, [((-1, -1, -1), (0.45, 0.55, 2.0)), ((0.6, 0.1, 0.2), (0.9, 0.3, 0.7))], seed=7)
    csdf = _sphere(n, (0.4, 0.2, 0.5), 0.2) if collider else None
    g, o = _pair(ctx, n, csdf, max_particles=x.shape[0] + 64)
    if not collider:
        g.set_sdf_search(capi.SDF_SEARCH_BUCKETS)
    g.upload_particles(x, v)
    o.set_particles(x.astype(np.float64), v.astype(np.float64))
    g.run_stage(capi.STAGE_SDF, 0.0)
    o.build_sdf()
    o.build_markers()
    assert np.array_equal(g.get(capi.GRID_SDF), o.get(orc.SDF))
    assert np.array_equal(g.get(capi.GRID_FLAGS), o.get(orc.FLAGS))
    g.close()


def test_move_collision_parity(ctx):
    print("[gpu] collider particle collision", flush=True)
    n = (16, 16, 16)
    csdf = _sphere(n, (0.5, 0.35, 0.5), 0.22)
    x, v = _particles(n, [((0.2, 0.45, 0.2), (0.8, 0.75, 0.8))], seed=11, jitter=0.0)
    g, o = _pair(ctx, n, csdf, max_particles=x.shape[0] + 64)
    g.upload_particles(x, v)
    o.set_particles(x.astype(np.float64), v.astype(np.float64))
    # a uniform downward grid velocity carries the particles into the sphere
    for gf, of in COMPS:
        f = np.full(g.shape(gf), -4.0 if gf == capi.GRID_V else 0.5, np.float32)
        g.set(gf, f)
        o.set(of, f.astype(np.float64))
    dt = 1.0 / 30
    g.run_stage(capi.STAGE_MOVE, dt)
    o.move_particles(dt)
    gx, gv, _ = g.download_particles()
    ox, ov, _ = o.particles()
    _close(gx, ox, 1e-6, "positions")
    _close(gv, ov, 1e-5, "velocities")
    moved = np.abs(ov - v).max(axis=1) > 0
    assert moved.sum() > 10  # some particles hit the sphere
    g.close()


@pytest.mark.parametrize("viscosity", [0.0, 0.02])
def test_substep_with_collider_parity(ctx, viscosity):
    print(f"[gpu] substep with collider viscosity={viscosity}", flush=True)
    n = (24, 24, 24)
    x, v = _particles(n, [((-1, -1, -1), (0.45, 0.6, 2.0))], seed=2)
    csdf = _sphere(n, (0.62, 0.25, 0.5), 0.18)
    g, o = _pair(ctx, n, csdf, viscosity=viscosity, max_particles=x.shape[0] + 64)
    g.upload_particles(x, v)
    o.set_particles(x.astype(np.float64), v.astype(np.float64))
    dt = 1.0 / 240
    g.substep(dt, capi.solver(capi.SOLVER_JACOBI, 200, 5, 1e-14))
    o.substep(dt, orc.solver(orc.JACOBI, 200, 5, 1e-14))
    flags = g.get(capi.GRID_FLAGS)
    assert np.array_equal(flags, o.get(orc.FLAGS))
    assert (flags == capi.BOUNDRY).sum() > 20
    assert np.array_equal(g.get(capi.GRID_SDF), o.get(orc.SDF))
    if viscosity > 0:
        for c in range(3):
            assert np.array_equal(g.get(DMARK[c]), o.diffusion_markers(c))
    gx, gv, _ = g.download_particles()
    ox, ov, _ = o.particles()
    _close(gv, ov, 2e-4, "particle velocities")
    _close(gx, ox, 2e-6, "particle positions")
    for gf, of in COMPS:
        _close(g.get(gf), o.get(of), 2e-4, f"grid {gf}")
    g.close()


def test_surface_with_collider(ctx):
    """GetSurface replaces the fluid SDF by minus the collider SDF where both
    are negative (src/hybrid_simulator.cpp:29-37)."""
    print("[gpu] surface with collider", flush=True)
    from test_gpu_surface import _compare
    n = (20, 18, 16)
    h = [1.0 / n[a] for a in range(3)]
    ax = [(np.arange(n[a]) + 0.5) * h[a] for a in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    sdf = Y - 0.45  # a pool below y = 0.45
    csdf = _sphere(n, (0.5, 0.3, 0.45), 0.2)
    g, o = _pair(ctx, n, csdf)
    g.set(capi.GRID_SDF, sdf)
    o.set(orc.SDF, sdf)
    gm, om = g.surface(), o.surface()
    _compare(gm, om)
    # the collider changes the mesh: without it the pool is one plane
    g.clear_colliders()
    plain = g.surface()
    assert plain[2].shape[0] != gm[2].shape[0]
    g.close()
