"""The two kernel variants of the multigrid V-cycle (FE_MG_FUSED: one descent
and one ascent kernel per level over LDS tiles with recomputed halos;
FE_MG_SPLIT: one kernel per sweep, residual, restriction and prolongation)
apply the same preconditioner, value for value; only the r.z partials are
summed over different blocks, so the CG scalars may differ in their last FP64
bits.  Bars: the same iteration count (or one apart), both to the FP64-checked
tolerance, pressures within 1e-5 of each other relative to the largest."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from fluide_engine_amd import capi  # noqa: E402

GRID = [capi.GRID_U, capi.GRID_V, capi.GRID_W]


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    print(f"\n[gpu] device {c.device_name}", flush=True)
    yield c
    c.close()


def _system(ctx, shape, seed, noise=0.03):
    desc = capi.scene_desc(shape, method=capi.FLIP, max_particles=16)
    g = capi.Scene(ctx, desc)
    rng = np.random.default_rng(seed)
    K, J, I = np.meshgrid(*[(np.arange(m) + 0.5) / m for m in shape[::-1]], indexing="ij")
    hmax = 1.0 / min(shape)
    sdf = np.where((I < 0.45) & (J < 0.55), -0.5 * hmax, 0.5 * hmax)
    sdf[rng.uniform(size=sdf.shape) < noise] *= -1
    g.set(capi.GRID_SDF, sdf)
    vel = [rng.normal(size=g.shape(f)).astype(np.float32) for f in GRID]
    return g, vel


@pytest.mark.parametrize("shape", [(32, 32, 32), (40, 24, 17), (6, 5, 4), (64, 64, 64), (70, 45, 33), (17, 80, 9)])
def test_fused_and_split_vcycles_agree(ctx, shape):
    print(f"[gpu] mg variants {shape}", flush=True)
    g, vel = _system(ctx, shape, 21)
    out = {}
    for variant in (capi.MG_SPLIT, capi.MG_FUSED):
        g.set_mg_variant(variant)
        for f, a in zip(GRID, vel):
            g.set(f, a)
        st = g.run_stage(capi.STAGE_PRESSURE,
                         desc=capi.solver(capi.SOLVER_MGPCG, 200, 0, 1e-6, relative=True, warm_start=False))
        assert st.converged == 1 and st.rel_residual <= 1e-6, (variant, st.iterations, st.rel_residual)
        out[variant] = (st.iterations, g.get(capi.GRID_PRESSURE), [g.get(f) for f in GRID])
    (it_s, p_s, u_s), (it_f, p_f, u_f) = out[capi.MG_SPLIT], out[capi.MG_FUSED]
    print(f"[gpu]   iterations split {it_s} fused {it_f}", flush=True)
    assert abs(it_s - it_f) <= 1
    scale = float(np.abs(p_s).max())
    assert float(np.abs(p_f - p_s).max()) <= 1e-5 * scale
    for a, b in zip(u_s, u_f):
        assert float(np.abs(a.astype(np.float64) - b).max()) <= 1e-5 * max(1.0, float(np.abs(a).max()))
    g.close()


@pytest.mark.parametrize("n", [64, 128])
def test_fused_vcycle_dam_break_substeps(ctx, n):
    """Dam-break substeps (most level-0 tiles of the air hold no FLUID cell and
    are skipped) from zero and warm-started: the fused V-cycle takes the split
    one's iteration counts (or one apart) and both leave the same markers."""
    print(f"[gpu] mg variants dam break {n}^3", flush=True)
    desc = capi.scene_desc((n, n, n), method=capi.FLIP, blend=0.03, max_particles=8 * n ** 3)
    x = capi.seed_particles(desc, [capi.box((-1, -1, -1), (0.4, 0.6, 2.0))], seed=0)
    res = {}
    for variant in (capi.MG_SPLIT, capi.MG_FUSED):
        g = capi.Scene(ctx, desc)
        g.set_mg_variant(variant)
        g.upload_particles(x, np.zeros_like(x))
        its = []
        for step in range(3):
            st = g.substep(1.0 / 480, capi.solver(capi.SOLVER_MGPCG, 200, 0, 1e-6, relative=True,
                                                  warm_start=step > 0))
            assert st.converged == 1 and st.rel_residual <= 1e-6
            its.append(st.iterations)
        res[variant] = (its, g.get(capi.GRID_FLAGS), g.download_particles()[1])
        g.close()
    (its_s, fl_s, v_s), (its_f, fl_f, v_f) = res[capi.MG_SPLIT], res[capi.MG_FUSED]
    print(f"[gpu]   iterations split {its_s} fused {its_f}", flush=True)
    assert all(abs(a - b) <= 1 for a, b in zip(its_s, its_f))
    assert np.array_equal(fl_s, fl_f)
    assert float(np.abs(v_s - v_f).max()) <= 1e-5 * max(1e-3, float(np.abs(v_s).max()))
