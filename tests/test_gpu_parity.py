"""GPU parity: every stage of the substep on gfx950 (through the C ABI) against
the FP64 oracle on the same seeded inputs.

Bars (stated per test):
  * integer results -- particle cell indices, P2G face markers, cell markers --
    bit-exact;
  * the fluid SDF where negative (the value that decides the markers) --
    bit-exact, FP64 on both sides;
  * FP32 grid/particle values -- relative tolerance against the FP64 oracle,
    scaled by the field's magnitude;
  * pressure -- both solvers reach the stated relative residual, checked in FP64.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from fluide_engine_amd import capi  # noqa: E402
from oracle import oracle as orc  # noqa: E402

ORC_METHOD = {capi.PIC: orc.PIC, capi.FLIP: orc.FLIP, capi.APIC: orc.APIC}
NAMES = {capi.PIC: "PIC", capi.FLIP: "FLIP", capi.APIC: "APIC"}


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    print(f"\n[gpu] device {c.device_name}", flush=True)
    yield c
    c.close()


def dam_particles(n, method, seed=0, vel_scale=0.0, jitter=0.0, cmat=False):
    desc = capi.scene_desc((n, n, n), method=method, blend=0.03, max_particles=8 * n ** 3 + 64)
    x = capi.seed_particles(desc, [capi.box((-1, -1, -1), (0.4, 0.6, 2.0))], seed=seed)
    rng = np.random.default_rng(seed + 7)
    if jitter:
        x = np.clip(x + rng.uniform(-jitter, jitter, x.shape).astype(np.float32), 0.0, 1.0).astype(np.float32)
    v = (rng.normal(size=x.shape) * vel_scale).astype(np.float32)
    C = (rng.normal(size=(x.shape[0], 9)) * vel_scale).astype(np.float32) if cmat else None
    return desc, x, v, C


def pair(ctx, desc, x, v, C=None):
    g = capi.Scene(ctx, desc)
    g.upload_particles(x, v, C)
    o = orc.Scene(tuple(desc.n), origin=tuple(desc.origin), size=tuple(desc.size),
                  method=ORC_METHOD[desc.method], blend=desc.blend, max_cfl=desc.max_cfl,
                  gravity=tuple(desc.gravity))
    o.set_particles(x.astype(np.float64), v.astype(np.float64), None if C is None else C.astype(np.float64))
    return g, o


def rel_close(a, b, rtol, what):
    scale = max(1.0, float(np.abs(b).max()))
    err = float(np.abs(a.astype(np.float64) - b).max())
    assert err <= rtol * scale, f"{what}: max abs err {err:.3e} > {rtol:.1e} * {scale:.3e}"
    return err


GRID = [(capi.GRID_U, orc.U), (capi.GRID_V, orc.V), (capi.GRID_W, orc.W)]
MARK = [(capi.GRID_MU, orc.MU), (capi.GRID_MV, orc.MV), (capi.GRID_MW, orc.MW)]


# ---------------------------------------------------------------------------
# particle -> cell indices
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [16, 24, 32, 48])
def test_particle_cells_bit_exact(ctx, n):
    print(f"[gpu] cells n={n}", flush=True)
    desc = capi.scene_desc((n, n, n), method=capi.FLIP, max_particles=200000)
    rng = np.random.default_rng(n)
    h = 1.0 / n
    x = rng.uniform(0.0, 1.0, (60000, 3)).astype(np.float32)
    # particles exactly on (FP32-rounded) cell boundaries, at the walls and outside the domain
    k = rng.integers(0, n + 1, (20000, 3))
    x = np.concatenate([x, (k * h).astype(np.float32),
                        np.nextafter((k * h).astype(np.float32), np.float32(-1)),
                        np.array([[0, 0, 0], [1, 1, 1], [-0.1, 0.5, 1.2], [1.5, -2, 0.5]], np.float32)])
    g = capi.Scene(ctx, desc)
    g.upload_particles(x, np.zeros_like(x))
    cells = g.particle_cells()
    idx = np.clip(np.floor((x.astype(np.float64) - 0.0) / (1.0 / n)), 0, n - 1).astype(np.int64)
    ref = idx[:, 0] + n * (idx[:, 1] + n * idx[:, 2])
    assert np.array_equal(cells.astype(np.int64), ref)
    g.close()


# ---------------------------------------------------------------------------
# P2G
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("method", [capi.PIC, capi.FLIP, capi.APIC])
@pytest.mark.parametrize("n", [16, 24])
def test_p2g_parity(ctx, method, n):
    print(f"[gpu] p2g {NAMES[method]} n={n}", flush=True)
    desc, x, v, C = dam_particles(n, method, seed=n, vel_scale=1.0, jitter=0.02, cmat=method == capi.APIC)
    # add particles on the walls and in the last cells (clamped stencils, Q1/Q2)
    extra = np.array([[1.0, 0.3, 0.3], [0.999, 0.999, 0.999], [0.0, 0.0, 0.0], [0.5, 1.0, 0.01]], np.float32)
    x = np.concatenate([x, extra])
    v = np.concatenate([v, np.ones((4, 3), np.float32)])
    if C is not None:
        C = np.concatenate([C, np.ones((4, 9), np.float32)])
    g, o = pair(ctx, desc, x, v, C)
    g.run_stage(capi.STAGE_P2G)
    o.p2g()
    for (gf, of), (gm, om) in zip(GRID, MARK):
        assert np.array_equal(g.get(gm), o.get(om).astype(np.uint8)), f"markers {gf}"
        rel_close(g.get(gf), o.get(of), 2e-5, f"face velocity {gf}")
    if method == capi.FLIP:
        for gs, os_ in ((capi.GRID_SU, orc.RU), (capi.GRID_SV, orc.RV), (capi.GRID_SW, orc.RW)):
            rel_close(g.get(gs), o.get(os_), 2e-5, "FLIP snapshot")
    g.close()


# ---------------------------------------------------------------------------
# SDF + cell markers
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [16, 24, 32])
def test_sdf_markers_bit_exact(ctx, n):
    print(f"[gpu] sdf n={n}", flush=True)
    desc, x, v, _ = dam_particles(n, capi.FLIP, seed=3 * n, jitter=0.6 / n)
    # sparse spray outside the block exercises isolated particles
    rng = np.random.default_rng(n)
    x = np.concatenate([x, rng.uniform(0, 1, (500, 3)).astype(np.float32)])
    g, o = pair(ctx, desc, x, np.zeros_like(x))
    g.run_stage(capi.STAGE_SDF)
    o.build_sdf()
    o.build_markers()
    flags_g, flags_o = g.get(capi.GRID_FLAGS), o.get(orc.FLAGS)
    assert np.array_equal(flags_g, flags_o)
    sg, so = g.get(capi.GRID_SDF), o.get(orc.SDF)
    neg = so < 0
    assert neg.sum() > 0
    assert np.array_equal(sg[neg], so[neg])  # FP64 on both sides, same operations
    assert np.all(sg[~neg] >= 0)
    g.close()


# ---------------------------------------------------------------------------
# extrapolation, walls, gravity
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(20, 20, 20), (13, 11, 7)])
def test_extrapolate_bc_gravity_parity(ctx, shape):
    """(13, 11, 7): face arrays of 1078, 1092 and 1144 entries, so the v and
    w validity arrays start off 4-byte boundaries and rows split the 4-face
    groups the sweep reads as words."""
    print(f"[gpu] extrapolate/bc/gravity {shape}", flush=True)
    desc = capi.scene_desc(shape, method=capi.FLIP, max_particles=16)
    g = capi.Scene(ctx, desc)
    o = orc.Scene(shape, method=orc.FLIP)
    rng = np.random.default_rng(5)
    for (gf, of), (gm, om) in zip(GRID, MARK):
        shp = g.shape(gf)
        vals = rng.normal(size=shp).astype(np.float32)
        mk = (rng.uniform(size=shp) < 0.15).astype(np.uint8)
        vals[mk == 0] = 0.0
        g.set(gf, vals); o.set(of, vals.astype(np.float64))
        g.set(gm, mk); o.set(om, mk.astype(np.int32))
    g.run_stage(capi.STAGE_EXTRAPOLATE)
    o.extrapolate_to_air()
    for gf, of in GRID:
        rel_close(g.get(gf), o.get(of), 1e-6, f"extrapolated {gf}")
    g.run_stage(capi.STAGE_BC)
    o.apply_bc()
    g.run_stage(capi.STAGE_GRAVITY, dt=1.0 / 240)
    o.apply_gravity(1.0 / 240)
    for gf, of in GRID:
        rel_close(g.get(gf), o.get(of), 1e-6, f"bc+gravity {gf}")
    assert np.all(g.get(capi.GRID_U)[:, :, 0] == 0.0)
    g.close()


# ---------------------------------------------------------------------------
# pressure projection
# ---------------------------------------------------------------------------
def _pressure_pair(ctx, n, seed, noise=0.03):
    shape = (n, n, n) if np.isscalar(n) else tuple(n)  # (nx, ny, nz)
    desc = capi.scene_desc(shape, method=capi.FLIP, max_particles=16)
    g = capi.Scene(ctx, desc)
    o = orc.Scene(shape, method=orc.FLIP)
    rng = np.random.default_rng(seed)
    # fluid: a dam block plus noise; the SDF decides the markers on both sides
    K, J, I = np.meshgrid(*[(np.arange(m) + 0.5) / m for m in shape[::-1]], indexing="ij")
    hmax = 1.0 / min(shape)
    sdf = np.where((I < 0.45) & (J < 0.55), -0.5 * hmax, 0.5 * hmax)
    sdf[rng.uniform(size=sdf.shape) < noise] *= -1
    g.set(capi.GRID_SDF, sdf); o.set(orc.SDF, sdf)
    for gf, of in GRID:
        vals = rng.normal(size=g.shape(gf)).astype(np.float32)
        g.set(gf, vals); o.set(of, vals.astype(np.float64))
    return g, o


def test_pressure_jacobi_parity(ctx):
    print("[gpu] pressure jacobi", flush=True)
    g, o = _pressure_pair(ctx, 24, 11)
    sg = g.run_stage(capi.STAGE_PRESSURE, desc=capi.solver(capi.SOLVER_JACOBI, 200, 5, 1e-14))
    so = o.pressure(orc.solver(orc.JACOBI, 200, 5, 1e-14))
    assert sg.iterations == 200 == so.iterations
    assert np.array_equal(g.get(capi.GRID_FLAGS), o.get(orc.FLAGS))
    rel_close(g.get(capi.GRID_PRESSURE), o.get(orc.PRESSURE), 2e-5, "jacobi pressure")
    for gf, of in GRID:
        rel_close(g.get(gf), o.get(of), 2e-5, f"projected {gf}")
    assert abs(sg.rel_residual - so.rel_residual) <= 1e-3 * so.rel_residual + 1e-6
    g.close()


ORC_SOLVER = {capi.SOLVER_CG: orc.CG, capi.SOLVER_PCG_JACOBI: orc.PCG_JACOBI,
              capi.SOLVER_MGPCG: orc.CG}  # the oracle's CG gives the converged answer for MGPCG


@pytest.mark.parametrize("shape", [32, (40, 24, 17), (6, 5, 4)])
@pytest.mark.parametrize("kind", [capi.SOLVER_CG, capi.SOLVER_PCG_JACOBI, capi.SOLVER_MGPCG])
def test_pressure_cg_reaches_tolerance(ctx, kind, shape):
    print(f"[gpu] pressure cg kind={kind} shape={shape}", flush=True)
    g, o = _pressure_pair(ctx, shape, 13)
    sg = g.run_stage(capi.STAGE_PRESSURE, desc=capi.solver(kind, 2000, 0, 1e-6, relative=True, warm_start=False))
    so = o.pressure(orc.solver(ORC_SOLVER[kind], 2000, 0, 1e-10, relative=True, warm_start=False))
    assert sg.converged == 1
    assert sg.rel_residual <= 1e-6 * 1.05, sg.rel_residual  # FP64 residual of the FP64 solution
    # the 1e-6 solution against the oracle's 1e-10 one
    rel_close(g.get(capi.GRID_PRESSURE), o.get(orc.PRESSURE), 1e-4, "cg pressure")
    for gf, of in GRID:
        rel_close(g.get(gf), o.get(of), 1e-4, f"projected {gf}")
    if kind == capi.SOLVER_MGPCG:
        # about 15-20 on these noisy systems (CPU model of the same V-cycle)
        assert sg.iterations <= 40, sg.iterations
    g.close()


@pytest.mark.parametrize("n", [48, 64])
@pytest.mark.parametrize("kind", [capi.SOLVER_MGPCG, capi.SOLVER_CG])
def test_first_substep_solve_converges(ctx, kind, n):
    """The first substep of a fresh dam break (zero start): the solve reaches
    the FP64-checked tolerance with the solver chunks that hold no FLUID cell
    skipped and exact zeros left outside FLUID.  At 48^3 and 64^3 one
    multigrid level (24^3, 32^3) lies between the finest level and the
    one-workgroup tail; 128^3 has two."""
    print(f"[gpu] first-substep solve {n}^3 kind {kind}", flush=True)
    desc = capi.scene_desc((n, n, n), method=capi.FLIP, blend=0.03, max_particles=8 * n ** 3)
    x = capi.seed_particles(desc, [capi.box((-1, -1, -1), (0.4, 0.6, 2.0))], seed=0)
    g = capi.Scene(ctx, desc)
    g.upload_particles(x, np.zeros_like(x))
    cap = 40 if kind == capi.SOLVER_MGPCG else 2000
    st = g.substep(1.0 / 480, capi.solver(kind, cap, 0, 1e-8, relative=True, warm_start=False))
    assert st.converged == 1 and st.rel_residual <= 1e-8, (st.iterations, st.rel_residual)
    p = g.get(capi.GRID_PRESSURE)
    assert np.all(p[g.get(capi.GRID_FLAGS) != capi.FLUID] == 0.0)  # identity rows: exactly 0
    g.close()


def test_mgpcg_iterations_and_warm_start(ctx):
    """MGPCG against diagonal PCG on the same 64^3 system: far fewer
    iterations, the same answer; a warm-started second solve of the solved
    system stops at once."""
    print("[gpu] mgpcg 64^3", flush=True)
    g, _ = _pressure_pair(ctx, 64, 17, noise=0.0)
    base = [g.get(f) for f, _ in GRID]
    sj = g.run_stage(capi.STAGE_PRESSURE, desc=capi.solver(capi.SOLVER_PCG_JACOBI, 4000, 0, 1e-7, relative=True,
                                                           warm_start=False))
    pj = g.get(capi.GRID_PRESSURE)
    for (f, _), b in zip(GRID, base):
        g.set(f, b)
    sm = g.run_stage(capi.STAGE_PRESSURE, desc=capi.solver(capi.SOLVER_MGPCG, 4000, 0, 1e-7, relative=True,
                                                           warm_start=False))
    pm = g.get(capi.GRID_PRESSURE)
    assert sj.converged == 1 and sm.converged == 1
    assert sm.rel_residual <= 1.05e-7 and sj.rel_residual <= 1.05e-7
    assert sm.iterations * 8 < sj.iterations, (sm.iterations, sj.iterations)
    rel_close(pm, pj, 2e-5, "MGPCG vs PCG pressure")
    for (f, _), b in zip(GRID, base):
        g.set(f, b)
    sw = g.run_stage(capi.STAGE_PRESSURE, desc=capi.solver(capi.SOLVER_MGPCG, 4000, 0, 1e-6, relative=True,
                                                           warm_start=True))
    assert sw.converged == 1 and sw.iterations <= 2, sw.iterations
    g.close()


def test_reference_single_phase_known_answer(ctx):
    # test/single_phase_pressuer_solver_test.cpp:61-133 on the GPU path
    print("[gpu] single-phase known answer", flush=True)
    desc = capi.scene_desc((4, 4, 4), method=capi.PIC, size=(4.0, 4.0, 4.0), max_particles=16)
    g = capi.Scene(ctx, desc)
    g.set(capi.GRID_SDF, np.full((4, 4, 4), -1.0))
    vv = np.zeros((4, 5, 4), np.float32)
    vv[:, 1:3, :] = 1.0
    g.set(capi.GRID_V, vv)
    g.run_stage(capi.STAGE_PRESSURE, desc=capi.solver(capi.SOLVER_JACOBI, 1000, 5, 1e-14))
    u, v, w = g.get(capi.GRID_U), g.get(capi.GRID_V), g.get(capi.GRID_W)
    assert np.all(np.abs(u[:3, :3, :4]) < 1e-5)
    assert np.all(np.abs(v[:3, :4, :3]) < 1e-5)
    assert np.all(np.abs(w[:4, :3, :3]) < 1e-5)
    p = g.get(capi.GRID_PRESSURE).astype(np.float64)
    assert np.allclose(p[:3, 1:3, :3] - p[:3, 0:2, :3], -1.0, atol=1e-5)
    g.close()


# ---------------------------------------------------------------------------
# G2P + advection
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("method", [capi.PIC, capi.FLIP, capi.APIC])
def test_g2p_move_parity(ctx, method):
    print(f"[gpu] g2p/move {NAMES[method]}", flush=True)
    n = 24
    desc, x, v, C = dam_particles(n, method, seed=4, vel_scale=0.5, jitter=0.01, cmat=method == capi.APIC)
    g, o = pair(ctx, desc, x, v, C)
    rng = np.random.default_rng(9)
    for gf, of in GRID:
        vals = (rng.normal(size=g.shape(gf)) * 0.5).astype(np.float32)
        g.set(gf, vals); o.set(of, vals.astype(np.float64))
    if method == capi.FLIP:
        for gs, os_ in ((capi.GRID_SU, orc.RU), (capi.GRID_SV, orc.RV), (capi.GRID_SW, orc.RW)):
            vals = (rng.normal(size=g.shape(gs)) * 0.5).astype(np.float32)
            g.set(gs, vals); o.set(os_, vals.astype(np.float64))
    g.run_stage(capi.STAGE_G2P)
    o.g2p()
    gx, gv, gC = g.download_particles()
    ox, ov, oC = o.particles()
    rel_close(gv, ov, 2e-5, "G2P velocity")
    if method == capi.APIC:
        rel_close(gC, oC, 2e-4, "APIC C")
    dt = 1.0 / 240
    g.run_stage(capi.STAGE_MOVE, dt=dt)
    o.move_particles(dt)
    gx, gv, _ = g.download_particles()
    ox, ov, _ = o.particles()
    rel_close(gx, ox, 1e-6, "advected position")
    rel_close(gv, ov, 2e-5, "velocity after wall clamp")
    # cell indices of the advected particles: bit-exact given the GPU's own positions
    cells = g.particle_cells()
    idx = np.clip(np.floor(gx.astype(np.float64) * n), 0, n - 1).astype(np.int64)
    assert np.array_equal(cells, idx[:, 0] + n * (idx[:, 1] + n * idx[:, 2]))
    g.close()


@pytest.mark.parametrize("method", [capi.PIC, capi.FLIP])
def test_g2p_move_parity_odd_shape(ctx, method):
    """G2P and advection on a (17, 13, 11) grid: odd face dims on every axis
    (the quad layout's clamped last column and row are sampled by particles
    next to the walls)."""
    print(f"[gpu] g2p/move odd shape {NAMES[method]}", flush=True)
    shape = (17, 13, 11)
    rng = np.random.default_rng(12)
    x = rng.uniform(0.0, 1.0, (3000, 3)).astype(np.float32)
    x[:8] = np.array([[0, 0, 0], [1, 1, 1], [1, 0.5, 0.5], [0.5, 1, 0.5], [0.5, 0.5, 1],
                      [0.999, 0.001, 0.999], [0.03, 0.97, 0.5], [0.97, 0.03, 0.02]], np.float32)
    v = (rng.normal(size=x.shape) * 0.5).astype(np.float32)
    desc = capi.scene_desc(shape, method=method, blend=0.03, max_particles=4096)
    g, o = pair(ctx, desc, x, v)
    for gf, of in GRID:
        vals = (rng.normal(size=g.shape(gf)) * 0.5).astype(np.float32)
        g.set(gf, vals); o.set(of, vals.astype(np.float64))
    if method == capi.FLIP:
        for gs, os_ in ((capi.GRID_SU, orc.RU), (capi.GRID_SV, orc.RV), (capi.GRID_SW, orc.RW)):
            vals = (rng.normal(size=g.shape(gs)) * 0.5).astype(np.float32)
            g.set(gs, vals); o.set(os_, vals.astype(np.float64))
    g.run_stage(capi.STAGE_G2P)
    o.g2p()
    _, gv, _ = g.download_particles()
    _, ov, _ = o.particles()
    rel_close(gv, ov, 2e-5, "G2P velocity")
    dt = 1.0 / 240
    g.run_stage(capi.STAGE_MOVE, dt=dt)
    o.move_particles(dt)
    gx, gv, _ = g.download_particles()
    ox, ov, _ = o.particles()
    rel_close(gx, ox, 1e-6, "advected position")
    rel_close(gv, ov, 2e-5, "velocity after wall clamp")
    g.close()


# ---------------------------------------------------------------------------
# full substeps
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("method", [capi.PIC, capi.FLIP, capi.APIC])
def test_substep_parity(ctx, method):
    print(f"[gpu] substep {NAMES[method]}", flush=True)
    n = 32
    desc, x, v, C = dam_particles(n, method, seed=1, vel_scale=0.3, jitter=0.005, cmat=method == capi.APIC)
    g, o = pair(ctx, desc, x, v, C)
    dt = 1.0 / 240
    sg = g.substep(dt, capi.solver(capi.SOLVER_JACOBI, 200, 5, 1e-14))
    o.substep(dt, orc.solver(orc.JACOBI, 200, 5, 1e-14))
    # markers are decided before any FP32/FP64 drift: bit-exact
    assert np.array_equal(g.get(capi.GRID_FLAGS), o.get(orc.FLAGS))
    for (gm, om) in MARK:
        assert np.array_equal(g.get(gm), o.get(om).astype(np.uint8))
    gx, gv, _ = g.download_particles()
    ox, ov, _ = o.particles()
    rel_close(gv, ov, 2e-4, "substep velocity")
    rel_close(gx, ox, 2e-6, "substep position")
    assert sg.iterations == 200
    g.close()


def test_dam_break_pic_32_run_drift(ctx):
    """32^3 PIC dam break, dt = 1/240, Jacobi 200 sweeps, 50 substeps.  The
    FP32 GPU state and the FP64 oracle diverge slowly through chaotic motion;
    the bar is on integral quantities: fluid volume (FLUID cell count) and the
    particle centroid."""
    print("[gpu] 32^3 PIC 50 substeps", flush=True)
    n = 32
    desc, x, v, _ = dam_particles(n, capi.PIC, seed=0)
    g, o = pair(ctx, desc, x, v)
    solv_g = capi.solver(capi.SOLVER_JACOBI, 200, 5, 1e-14)
    solv_o = orc.solver(orc.JACOBI, 200, 5, 1e-14)
    fluid0 = None
    for step in range(50):
        g.substep(1.0 / 240, solv_g)
        o.substep(1.0 / 240, solv_o)
        fg = int((g.get(capi.GRID_FLAGS) == capi.FLUID).sum())
        fo = int((o.get(orc.FLAGS) == orc.FLUID).sum())
        if fluid0 is None:
            fluid0 = fo
        assert abs(fg - fo) <= 0.01 * fo + 4, (step, fg, fo)
    gx, _, _ = g.download_particles()
    ox, _, _ = o.particles()
    assert np.abs(gx.mean(0) - ox.mean(0)).max() < 2e-3
    print(f"[gpu] fluid cells t0={fluid0} t50: gpu={fg} oracle={fo}", flush=True)
    g.close()
