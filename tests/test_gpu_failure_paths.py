"""Failure paths (the reference has none: no error reporting, no checks):
non-finite particle data and face velocities are reported as FE_E_NONFINITE,
a pressure solve stopped by its iteration cap is reported in the stats
(converged = 0), not raised, and the singular Neumann system of an all-fluid
closed box (the SolveSinglePhase setup of
test/single_phase_pressuer_solver_test.cpp:61-133, at a larger size) is solved
by CG and multigrid-PCG, its right-hand side being consistent."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from fluide_engine_amd import capi  # noqa: E402
from pressure_residual import host_relative_residual  # noqa: E402

GRID = [capi.GRID_U, capi.GRID_V, capi.GRID_W]


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    print(f"\n[gpu] device {c.device_name}", flush=True)
    yield c
    c.close()


def _dam(ctx, n=24):
    desc = capi.scene_desc((n, n, n), method=capi.FLIP, blend=0.03, max_particles=8 * n ** 3)
    x = capi.seed_particles(desc, [capi.box((-1, -1, -1), (0.4, 0.6, 2.0))], seed=0)
    return desc, x


@pytest.mark.parametrize("what", ["position", "velocity", "C"])
@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_non_finite_particles_rejected_at_upload(ctx, what, bad):
    print(f"[gpu] non-finite particle {what} {bad}", flush=True)
    desc, x = _dam(ctx)
    desc.method = capi.APIC
    g = capi.Scene(ctx, desc)
    v = np.zeros_like(x)
    Cm = np.zeros((x.shape[0], 9), np.float32)
    arr = {"position": x, "velocity": v, "C": Cm}[what].copy()
    arr.flat[1234] = bad
    args = {"position": (arr, v, Cm), "velocity": (x, arr, Cm), "C": (x, v, arr)}[what]
    with pytest.raises(capi.FluideError) as e:
        g.upload_particles(*args)
    assert e.value.code == capi.E_NONFINITE and "non-finite" in str(e.value)
    # the scene is unchanged and usable
    g.upload_particles(x, v, Cm)
    assert g.count() == x.shape[0]
    g.close()


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_non_finite_face_velocity_stops_the_frame(ctx, bad):
    print(f"[gpu] non-finite face velocity {bad}", flush=True)
    desc, x = _dam(ctx)
    g = capi.Scene(ctx, desc)
    g.upload_particles(x, np.zeros_like(x))
    sv = capi.solver(capi.SOLVER_MGPCG, 100, 0, 1e-6, relative=True)
    g.advance_frame(1.0 / 480, sv)
    cfl_ok = g.cfl(1.0 / 480)
    assert np.isfinite(cfl_ok)
    u = g.get(capi.GRID_U)
    u[5, 6, 7] = bad
    g.set(capi.GRID_U, u)
    with pytest.raises(capi.FluideError) as e:
        g.cfl(1.0 / 480)
    assert e.value.code == capi.E_NONFINITE
    with pytest.raises(capi.FluideError) as e:
        g.advance_frame(1.0 / 480, sv)
    assert e.value.code == capi.E_NONFINITE
    # a finite state again: the next frame runs
    u[5, 6, 7] = 0.0
    g.set(capi.GRID_U, u)
    assert g.advance_frame(1.0 / 480, sv) >= 1
    g.close()


@pytest.mark.parametrize("kind,cap", [(capi.SOLVER_CG, 5), (capi.SOLVER_MGPCG, 2), (capi.SOLVER_PCG_JACOBI, 3),
                                      (capi.SOLVER_JACOBI, 7)])
def test_iteration_cap_reported_not_raised(ctx, kind, cap):
    print(f"[gpu] iteration cap kind {kind} cap {cap}", flush=True)
    desc, x = _dam(ctx)
    g = capi.Scene(ctx, desc)
    g.upload_particles(x, np.zeros_like(x))
    rel = kind != capi.SOLVER_JACOBI
    st = g.substep(1.0 / 480, capi.solver(kind, cap, 5, 1e-12, relative=rel, warm_start=False))
    assert st.converged == 0
    assert st.iterations == cap
    assert st.rel_residual > 1e-12 and np.isfinite(st.rel_residual)
    assert g.last_stats().converged == 0
    g.close()


@pytest.mark.parametrize("kind", [capi.SOLVER_CG, capi.SOLVER_MGPCG])
def test_singular_closed_box_all_fluid(ctx, kind):
    """Every cell FLUID, walls on all sides: A is singular (constants in its
    null space) and b = div u sums to zero over the box when the wall faces are
    0, so the system is consistent and CG converges in the range of A."""
    print(f"[gpu] singular closed box kind {kind}", flush=True)
    n = 24
    desc = capi.scene_desc((n, n, n), method=capi.FLIP, max_particles=16)
    g = capi.Scene(ctx, desc)
    g.set(capi.GRID_SDF, np.full((n, n, n), -1.0))
    rng = np.random.default_rng(4)
    vel = []
    for f, ax in zip(GRID, (2, 1, 0)):
        a = rng.normal(size=g.shape(f)).astype(np.float32)
        sl = [slice(None)] * 3
        sl[ax] = 0
        a[tuple(sl)] = 0.0
        sl[ax] = -1
        a[tuple(sl)] = 0.0  # no flux through the walls
        g.set(f, a)
        vel.append(a)
    st = g.run_stage(capi.STAGE_PRESSURE, desc=capi.solver(kind, 2000, 0, 1e-6, relative=True, warm_start=False))
    flags = g.get(capi.GRID_FLAGS)
    assert np.all(flags == capi.FLUID)
    p = g.get(capi.GRID_PRESSURE)
    rel, b = host_relative_residual(flags, *vel, p, (1.0 / n,) * 3)
    print(f"[gpu]   {st.iterations} iterations, rel {st.rel_residual:.3e}, host {rel:.3e}, "
          f"sum b / |b|_1 {abs(b.sum()) / np.abs(b).sum():.2e}", flush=True)
    assert st.converged == 1 and st.rel_residual <= 1e-6
    assert rel <= 1.1e-6
    # the projected field is divergence-free to the same relative level
    u, v, w = (g.get(f).astype(np.float64) for f in GRID)
    div = (u[:, :, 1:] - u[:, :, :-1] + v[:, 1:, :] - v[:, :-1, :] + w[1:] - w[:-1]) * n
    assert np.linalg.norm(div) <= 2e-6 * np.linalg.norm(b)
    g.close()
