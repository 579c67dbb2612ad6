"""GPU parity of the fluid surface (MarchingCubesSolver::BuildSurface,
src/3d/marching_cubes_solver.cpp) through the C ABI against the oracle on
the same SDF, in both triangulation modes (the reference's case tables, and
the closed-surface face rule); on a simulated dam break the GPU builds its
surface from its own SDF (the reference's 8-bucket values) and the oracle its
own SDF from the same particles.  Bars: equal vertex and triangle counts; the vertex set equal
to 1e-12 (the cube vertices use the reference's FP64 expressions on both
sides; cap crossing points may differ by rounding); the triangles equal as
oriented position triples; normals equal to 1e-12 at matched vertices."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from fluide_engine_amd import capi  # noqa: E402
from oracle import oracle as orc  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def _canon_tris(v, t):
    P = v[t]  # (T, 3, 3)
    out = []
    for tri in P:
        k = min(range(3), key=lambda q: tuple(tri[q]))
        out.append(np.concatenate([tri[k], tri[(k + 1) % 3], tri[(k + 2) % 3]]))
    out = np.array(out)
    return out[np.lexsort(out.T[::-1])]


def _sorted_rows(a):
    return a[np.lexsort(a.T[::-1])]


def _compare(gm, om):
    gv, gn, gt = gm
    ov, on, ot = om
    assert gv.shape == ov.shape and gt.shape == ot.shape, (gv.shape, ov.shape, gt.shape, ot.shape)
    assert np.abs(_sorted_rows(np.hstack([gv, gn])) - _sorted_rows(np.hstack([ov, on]))).max() < 1e-12
    assert np.abs(_canon_tris(gv, gt) - _canon_tris(ov, ot)).max() < 1e-12


def _scene_pair(ctx, n, sdf):
    desc = capi.scene_desc(n, method=capi.FLIP, max_particles=16)
    g = capi.Scene(ctx, desc)
    g.set(capi.GRID_SDF, sdf)
    o = orc.Scene(n)
    o.set(orc.SDF, sdf)
    return g, o


MODES = [capi.SURFACE_REFERENCE, capi.SURFACE_CLOSED]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("case", ["sphere", "random", "slab"])
def test_surface_parity(ctx, case, mode):
    print(f"[gpu] surface {case} mode {mode}", flush=True)
    n = {"sphere": (24, 22, 26), "random": (13, 9, 11), "slab": (10, 12, 8)}[case]
    h = [1.0 / n[a] for a in range(3)]
    ax = [(np.arange(n[a]) + 0.5) * h[a] for a in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    if case == "sphere":
        sdf = np.sqrt((X - 0.5) ** 2 + (Y - 0.47) ** 2 + (Z - 0.52) ** 2) - 0.3
    elif case == "random":
        sdf = np.random.default_rng(8).uniform(-1, 1, (n[2], n[1], n[0]))
    else:
        sdf = Y - 0.31
    g, o = _scene_pair(ctx, n, sdf)
    g.set_surface_mode(mode)
    o.set_surface_mode(mode)
    _compare(g.surface(), o.surface())
    # a second build on the same scene reuses the buffers
    _compare(g.surface(), o.surface())
    g.close()


@pytest.mark.parametrize("mode", MODES)
def test_surface_of_dam_break(ctx, mode):
    """GetSurface after a frame (src/simulation_runner.cpp:35-36): the fluid
    SDF of the last substep's start, with the reference's positive values.
    The oracle builds its SDF (the reference's hash-grid search) from the
    particles the GPU held at that substep's start."""
    print(f"[gpu] surface of a dam break, mode {mode}", flush=True)
    n = 32
    desc = capi.scene_desc((n, n, n), method=capi.FLIP, blend=0.03, max_particles=8 * n ** 3)
    x = capi.seed_particles(desc, [capi.box((-1, -1, -1), (0.4, 0.6, 2.0))], seed=0)
    g = capi.Scene(ctx, desc)
    g.upload_particles(x, np.zeros_like(x))
    solver = capi.solver(capi.SOLVER_MGPCG, 500, 0, 1e-6, relative=True)
    for _ in range(5):
        g.advance_frame(1.0 / 240, solver)
    px, pv, _ = g.download_particles()  # the start of the last substep
    g.substep(1.0 / 240, solver)
    g.set_surface_mode(mode)
    gm = g.surface()
    o = orc.Scene((n, n, n))
    o.set_surface_mode(mode)
    o.set_particles(px.astype(np.float64), pv.astype(np.float64))
    o.build_sdf()
    assert np.array_equal(g.get(capi.GRID_SDF), o.get(orc.SDF))  # bucket search: every value the reference's
    _compare(gm, o.surface())
    assert len(gm[2]) > 1000
    g.close()
