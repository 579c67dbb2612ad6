"""Batches of independent scenes (fe_batch): advancing scenes together on
several host threads gives every scene bit for bit the state it reaches
advanced alone; invalid batches are rejected."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from fluide_engine_amd import capi  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def _scenes(ctx, specs):
    out = []
    for n, lo, hi in specs:
        desc = capi.scene_desc((n, n, n), method=capi.FLIP, blend=0.03, max_particles=8 * n ** 3)
        x = capi.seed_particles(desc, [capi.box(lo, hi)], seed=n)
        sc = capi.Scene(ctx, desc)
        sc.upload_particles(x, np.zeros_like(x))
        out.append(sc)
    return out


def test_batch_matches_scenes_alone(ctx):
    print("[gpu] batch of 6 scenes vs alone", flush=True)
    specs = capi.batch_scene_specs(count=6, seed=3, sizes=(16, 24, 32))
    solver = capi.solver(capi.SOLVER_MGPCG, 500, 0, 1e-6, relative=True)
    alone = _scenes(ctx, specs)
    batched = _scenes(ctx, specs)
    frames, dt = 3, 1.0 / 240
    for sc in alone:
        for _ in range(frames):
            sc.advance_frame(dt, solver)
    b = capi.Batch(ctx, batched, workers=3)
    for _ in range(frames):
        sub, ms = b.advance_frame(dt, solver)
        assert all(s >= 1 for s in sub) and all(m > 0 for m in ms)
    for a, g in zip(alone, batched):
        ax, av, _ = a.download_particles()
        gx, gv, _ = g.download_particles()
        assert np.array_equal(ax, gx) and np.array_equal(av, gv)
    b.close()
    for sc in alone + batched:
        sc.close()


def test_batch_rejects_bad_lists(ctx):
    print("[gpu] batch validation", flush=True)
    scs = _scenes(ctx, [(8, (0.1, 0.1, 0.1), (0.5, 0.5, 0.5))])
    with pytest.raises(capi.FluideError):
        capi.Batch(ctx, [scs[0], scs[0]])
    with pytest.raises(capi.FluideError):
        capi.Batch(ctx, scs, workers=0)
    other = capi.Context(0)
    with pytest.raises(capi.FluideError):
        capi.Batch(other, scs)
    other.close()
    scs[0].close()
