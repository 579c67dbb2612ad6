"""GPU parity of BackwardEulerDiffusionSolver (viscosity), through the C ABI,
against the FP64 oracle on the same inputs.

Bars:
  * face markers of the diffusion solve -- bit-exact (FP64 SDF samples on both
    sides);
  * Jacobi sweep counts -- equal (same stopping rule, FP64 residual);
  * velocities after the solve -- the GPU's FP64 sweeps follow the oracle's
    operation order, so the FP32 result is the rounded FP64 value: within
    1e-6 relative of the field's magnitude (one FP32 rounding is 6e-8);
  * a full substep with viscosity -- markers bit-exact, velocities within the
    substep bar of the inviscid parity test (2e-4 relative).
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


def _pair(ctx, n, size, nu, seed):
    desc = capi.scene_desc(n, method=capi.FLIP, size=size, gravity=(0.0, 0.0, 0.0), viscosity=nu, max_particles=64)
    g = capi.Scene(ctx, desc)
    o = orc.Scene(n, size=size, method=orc.FLIP, blend=desc.blend, max_cfl=3.0, gravity=(0.0, 0.0, 0.0))
    o.set_viscosity(nu)
    rng = np.random.default_rng(seed)
    sdf = rng.uniform(-1.0, 0.7, (n[2], n[1], n[0]))
    g.set(capi.GRID_SDF, sdf)
    o.set(orc.SDF, sdf)
    for gf, of in COMPS:
        f = rng.normal(size=g.shape(gf)).astype(np.float32)
        g.set(gf, f)
        o.set(of, f.astype(np.float64))
    return g, o


def _compare(g, o, gst, ost, rtol=1e-6):
    for c, (gf, of) in enumerate(COMPS):
        assert np.array_equal(g.get(DMARK[c]), o.diffusion_markers(c)), f"markers {c}"
        assert gst[c].iterations == ost[c].iterations, (c, gst[c].iterations, ost[c].iterations)
        want = o.get(of)
        got = g.get(gf).astype(np.float64)
        scale = max(1.0, float(np.abs(want).max()))
        err = float(np.abs(got - want).max())
        assert err <= rtol * scale, f"component {c}: {err:.3e}"


@pytest.mark.parametrize("n,size,nu", [((20, 16, 12), (1.0, 0.9, 0.7), 0.01), ((17, 9, 23), (1.0, 1.0, 1.0), 0.002),
                                       ((8, 8, 8), (1.0, 1.0, 1.0), 0.3)])
def test_diffusion_stage_parity(ctx, n, size, nu):
    print(f"[gpu] diffusion stage n={n} nu={nu}", flush=True)
    g, o = _pair(ctx, n, size, nu, seed=sum(n))
    dt = 1.0 / 240
    g.run_stage(capi.STAGE_DIFFUSION, dt)
    ost = o.diffusion(dt)
    gst = g.last_diffusion_stats()
    _compare(g, o, gst, ost)
    assert all(s.converged for s in gst)
    g.close()


def test_diffusion_iteration_cap(ctx):
    """A viscosity high enough that Jacobi stops at the 1000-sweep cap
    (CudaJacobiIterationSolver(1000, 5, 1e-14)) on both sides."""
    print("[gpu] diffusion at the iteration cap", flush=True)
    n = (12, 10, 8)
    g, o = _pair(ctx, n, (1.0, 1.0, 1.0), 2.0, seed=4)
    dt = 1.0 / 60
    g.run_stage(capi.STAGE_DIFFUSION, dt)
    ost = o.diffusion(dt)
    gst = g.last_diffusion_stats()
    assert all(s.iterations == 1000 and not s.converged for s in gst)
    _compare(g, o, gst, ost)
    g.close()


def test_diffusion_solver_setting(ctx):
    print("[gpu] diffusion solver setting", flush=True)
    n = (10, 10, 10)
    g, o = _pair(ctx, n, (1.0, 1.0, 1.0), 0.05, seed=9)
    g.set_diffusion_solver(capi.solver(capi.SOLVER_JACOBI, 37, 4, 1e-30))
    o.set_diffusion_solver(orc.solver(orc.JACOBI, 37, 4, 1e-30))
    dt = 1.0 / 240
    g.run_stage(capi.STAGE_DIFFUSION, dt)
    ost = o.diffusion(dt)
    gst = g.last_diffusion_stats()
    assert all(s.iterations == 37 for s in gst)
    _compare(g, o, gst, ost)
    with pytest.raises(capi.FluideError):
        g.set_diffusion_solver(capi.solver(capi.SOLVER_CG, 100, 5, 1e-10))
    g.close()


def test_substep_with_viscosity_parity(ctx):
    print("[gpu] substep with viscosity", flush=True)
    n = 24
    desc = capi.scene_desc((n, n, n), method=capi.FLIP, blend=0.03, max_particles=8 * n ** 3 + 64, viscosity=0.02)
    x = capi.seed_particles(desc, [capi.box((-1, -1, -1), (0.4, 0.6, 2.0))], seed=2)
    rng = np.random.default_rng(12)
    x = np.clip(x + rng.uniform(-0.004, 0.004, x.shape).astype(np.float32), 0.0, 1.0).astype(np.float32)
    v = (rng.normal(size=x.shape) * 0.3).astype(np.float32)
    g = capi.Scene(ctx, desc)
    g.upload_particles(x, v)
    o = orc.Scene((n, n, n), method=orc.FLIP, blend=0.03, max_cfl=3.0, gravity=tuple(desc.gravity))
    o.set_viscosity(0.02)
    o.set_particles(x.astype(np.float64), v.astype(np.float64))
    dt = 1.0 / 240
    g.substep(dt, capi.solver(capi.SOLVER_JACOBI, 200, 5, 1e-14))
    o.substep(dt, orc.solver(orc.JACOBI, 200, 5, 1e-14))
    assert np.array_equal(g.get(capi.GRID_FLAGS), o.get(orc.FLAGS))
    for c in range(3):
        assert np.array_equal(g.get(DMARK[c]), o.diffusion_markers(c))
    gx, gv, _ = g.download_particles()
    ox, ov, _ = o.particles()
    scale = max(1.0, float(np.abs(ov).max()))
    assert float(np.abs(gv - ov).max()) <= 2e-4 * scale
    assert float(np.abs(gx - ox).max()) <= 2e-6 * max(1.0, float(np.abs(ox).max()))
    gst = g.last_diffusion_stats()
    assert all(s.iterations > 5 for s in gst)
    g.close()
