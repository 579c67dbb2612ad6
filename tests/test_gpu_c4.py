"""The batched-scenes workload at its specification: 96 FLIP dam breaks on
cubic grids n ~ U{24, 32, 48, 64, 96} (seed 1), one fluid box each with
extents ~ U[0.2, 0.7], 2x2x2 particles per fluid cell, multigrid-PCG to 1e-6,
frame dt = 1/480.  Bars: every scene advanced in the batch, kernel by kernel
or as captured HIP graphs, ends bit for bit where it ends advanced alone; three sampled scenes (the first of 24^3, 48^3
and 96^3) agree with the FP64 oracle over the substep in which their falling
box first needs a pressure solve (the oracle re-seeded from the GPU's
particles and pressure just before it), with the 128^3 substep bars (markers
bit-exact, velocities within 2e-6 of the largest oracle velocity, positions
within 5e-7)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from fluide_engine_amd import capi  # noqa: E402
from oracle import oracle as orc  # noqa: E402

DT = 1.0 / 480.0
THREADS = max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    print(f"\n[gpu] device {c.device_name}", flush=True)
    yield c
    c.close()


def _scene(ctx, n, lo, hi):
    desc = capi.scene_desc((n, n, n), method=capi.FLIP, blend=0.03, max_particles=8 * n ** 3)
    x = capi.seed_particles(desc, [capi.box(lo, hi)], seed=n)
    desc.max_particles = max(1, x.shape[0])
    sc = capi.Scene(ctx, desc)
    sc.upload_particles(x, np.zeros_like(x))
    return sc, x


SOLVER = capi.solver(capi.SOLVER_MGPCG, 2000, 0, 1e-6, relative=True, warm_start=True)
FRAMES = 4


@pytest.fixture(scope="module")
def alone(ctx):
    """Every scene of the specification advanced alone, kernel by kernel."""
    specs = capi.batch_scene_specs()
    assert len(specs) == 96
    print(f"[gpu] C4: 96 scenes alone, {FRAMES} frames", flush=True)
    out = []
    for n, lo, hi in specs:
        sc, _ = _scene(ctx, n, lo, hi)
        for _ in range(FRAMES):
            sc.advance_frame(DT, SOLVER)
        out.append(sc.download_particles()[:2])
        sc.close()
    return out


@pytest.mark.parametrize("graphs", [False, True])
def test_c4_batch_bit_identical_to_alone(ctx, alone, graphs):
    """Batched on 4 host threads, launched kernel by kernel or (graphs) with
    every qualifying substep as captured HIP graphs, captured on the worker
    threads while the other scenes run."""
    specs = capi.batch_scene_specs()
    print(f"[gpu] C4: the same 96 scenes batched (4 workers), graphs {graphs}", flush=True)
    scenes = [_scene(ctx, n, lo, hi)[0] for n, lo, hi in specs]
    for sc in scenes:
        sc.set_graph(graphs)
    b = capi.Batch(ctx, scenes, workers=4)
    for _ in range(FRAMES):
        sub, _ = b.advance_frame(DT, SOLVER)
        assert all(s >= 1 for s in sub)
    for i, (sc, (ax, av)) in enumerate(zip(scenes, alone)):
        gx, gv, _ = sc.download_particles()
        assert np.array_equal(gx, ax) and np.array_equal(gv, av), f"scene {i} ({specs[i][0]}^3)"
    if graphs:
        st = [sc.graph_stats() for sc in scenes]
        print(f"[gpu]   graph substeps {sum(a for a, _, _ in st)}, captures {sum(c for _, c, _ in st)}, "
              f"resumed {sum(r for _, _, r in st)}", flush=True)
        assert all(a >= FRAMES - 1 for a, _, _ in st)  # every substep after the first
    b.close()
    for sc in scenes:
        sc.close()


def test_c4_sampled_scenes_against_oracle(ctx):
    specs = capi.batch_scene_specs()
    picks = []
    for want in (24, 48, 96):
        picks.append(next(i for i, (n, _, _) in enumerate(specs) if n == want))
    for i in picks:
        n, lo, hi = specs[i]
        print(f"[gpu] C4 scene {i}: {n}^3 box {np.round(lo, 3)} - {np.round(hi, 3)}", flush=True)
        sc, x = _scene(ctx, n, lo, hi)
        solver = capi.solver(capi.SOLVER_MGPCG, 200, 0, 1e-6, relative=True, warm_start=True)
        # free fall needs no pressure (a uniform velocity has no divergence):
        # advance until the substep that does, keeping the state before it
        for step in range(400):
            px, pv, _ = sc.download_particles()
            pp = sc.get(capi.GRID_PRESSURE)
            sg = sc.substep(DT, solver)
            if sg.iterations > 0:
                break
        assert sg.iterations > 0 and sg.converged == 1, step
        gflags = sc.get(capi.GRID_FLAGS)
        gmarks = [sc.get(m) for m in (capi.GRID_MU, capi.GRID_MV, capi.GRID_MW)]
        gx, gv, _ = sc.download_particles()
        sc.close()
        o = orc.Scene((n, n, n), method=orc.FLIP, blend=0.03, threads=THREADS)
        o.set_fast_dots()
        o.set_particles(px.astype(np.float64), pv.astype(np.float64))
        o.set(orc.PRESSURE, pp)
        o.substep(DT, orc.solver(orc.CG, 5000, 0, 1e-8, relative=True, warm_start=True))
        assert np.array_equal(gflags, o.get(orc.FLAGS))
        for gm, om in zip(gmarks, (orc.MU, orc.MV, orc.MW)):
            assert np.array_equal(gm, o.get(om).astype(np.uint8))
        ox, ov, _ = o.particles()
        ev = float(np.abs(gv - ov).max() / np.abs(ov).max())
        ex = float(np.abs(gx - ox).max())
        print(f"[gpu]   {x.shape[0]} particles, substep {step}: |dv|/|v|max {ev:.2e}, |dx|max {ex:.2e}, "
              f"{sg.iterations} it", flush=True)
        assert ev <= 2e-6 and ex <= 5e-7
