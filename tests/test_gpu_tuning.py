"""Kernel variants and their autotuner (no reference counterpart): P2G as
2x2x2 face tiles or one thread per face, G2P/advection from the face arrays
or from their quad layout.  The variants of a stage must give bit-identical
results (so autotuning never changes a simulation), and the autotuner must
leave the particles untouched."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from fluide_engine_amd import capi  # noqa: E402

NAMES = {capi.PIC: "PIC", capi.FLIP: "FLIP", capi.APIC: "APIC"}
GRIDS = [capi.GRID_U, capi.GRID_V, capi.GRID_W, capi.GRID_MU, capi.GRID_MV, capi.GRID_MW]


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def _scene(ctx, shape, method, seed):
    desc = capi.scene_desc(shape, method=method, blend=0.03, max_particles=8 * int(np.prod(shape)) + 64)
    x = capi.seed_particles(desc, [capi.box((-1, -1, -1), (0.45, 0.6, 2.0))], seed=seed)
    rng = np.random.default_rng(seed)
    extra = np.array([[1.0, 0.3, 0.3], [0.999, 0.999, 0.999], [0.0, 0.0, 0.0], [0.5, 1.0, 0.01]], np.float32)
    x = np.concatenate([x, extra])
    v = (rng.normal(size=x.shape) * 0.5).astype(np.float32)
    C = (rng.normal(size=(x.shape[0], 9)) * 0.5).astype(np.float32) if method == capi.APIC else None
    s = capi.Scene(ctx, desc)
    s.upload_particles(x, v, C)
    return s


@pytest.mark.parametrize("method", [capi.PIC, capi.FLIP, capi.APIC])
@pytest.mark.parametrize("shape", [(24, 24, 24), (17, 13, 11)])
def test_p2g_variants_bit_identical(ctx, method, shape):
    print(f"[gpu] p2g variants {NAMES[method]} {shape}", flush=True)
    s = _scene(ctx, shape, method, seed=3)
    out = {}
    for variant in (capi.P2G_TILES, capi.P2G_FACES, capi.P2G_SCATTER):
        s.set_tuning(variant, capi.G2P_QUADS)
        s.run_stage(capi.STAGE_P2G)
        fields = GRIDS + ([capi.GRID_SU, capi.GRID_SV, capi.GRID_SW] if method == capi.FLIP else [])
        out[variant] = [s.get(f) for f in fields]
    for other in (capi.P2G_FACES, capi.P2G_SCATTER):
        for a, b in zip(out[capi.P2G_TILES], out[other]):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    assert any(np.any(a != 0) for a in out[capi.P2G_TILES][:3])
    s.close()


@pytest.mark.parametrize("method", [capi.PIC, capi.FLIP, capi.APIC])
def test_g2p_variants_bit_identical(ctx, method):
    print(f"[gpu] g2p variants {NAMES[method]}", flush=True)
    shape = (20, 17, 13)
    res = {}
    for variant in (capi.G2P_QUADS, capi.G2P_PAIRS):
        s = _scene(ctx, shape, method, seed=5)
        s.set_tuning(capi.P2G_TILES, variant)
        s.run_stage(capi.STAGE_P2G)
        rng = np.random.default_rng(6)
        for f in (capi.GRID_U, capi.GRID_V, capi.GRID_W):
            s.set(f, (rng.normal(size=s.shape(f)) * 0.5).astype(np.float32))
        s.run_stage(capi.STAGE_G2P)
        s.run_stage(capi.STAGE_MOVE, dt=1.0 / 240)
        res[variant] = s.download_particles()
        s.close()
    n = 3 if method == capi.APIC else 2  # x, v (and C for APIC)
    for a, b in zip(res[capi.G2P_QUADS][:n], res[capi.G2P_PAIRS][:n]):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_autotune_keeps_particles_and_results(ctx):
    """The autotuner reports both variants' times, picks one of each, leaves
    the particles as they were, and a substep after it equals a substep with
    the default variants bit for bit."""
    print("[gpu] autotune", flush=True)
    shape = (32, 32, 32)
    dt = 1.0 / 480
    solver = capi.solver(capi.SOLVER_MGPCG, 200, 0, 1e-6, relative=True, warm_start=True)
    a = _scene(ctx, shape, capi.FLIP, seed=8)
    b = _scene(ctx, shape, capi.FLIP, seed=8)
    for s in (a, b):
        s.substep(dt, solver)
    before = a.download_particles()
    (p2g, g2p), times = a.autotune(reps=3)
    assert p2g in (capi.P2G_TILES, capi.P2G_FACES, capi.P2G_SCATTER) and g2p in (capi.G2P_QUADS, capi.G2P_PAIRS)
    assert all(t > 0 for t in times["p2g_ms"] + times["g2p_ms"])
    assert a.tuning() == (p2g, g2p)
    after = a.download_particles()
    for x, y in zip(before[:2], after[:2]):  # x, v
        assert np.array_equal(x, y)
    for s in (a, b):
        s.substep(dt, solver)
    for x, y in zip(a.download_particles()[:2], b.download_particles()[:2]):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
    with pytest.raises(capi.FluideError):
        a.set_tuning(7, 0)
    a.close()
    b.close()
