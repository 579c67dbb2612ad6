"""Substeps as captured HIP graphs (fe_scene_set_graph) against the same
substeps launched kernel by kernel: two scenes seeded alike, one in graph
mode, advanced side by side.  Bars: after every substep the particles, the
face velocities and the FP64 pressure are bit-identical, and so are the
solver's iteration count, convergence flag and FP64 residual.  Every graph is
replayed several times with eager work (downloads, the eager scene's own
substeps) and synchronizations between the replays.  The runs cover PIC, FLIP
and APIC, the three CG-family solvers, a tolerance change in mid-run (the
solve then needs more iterations than were captured, so the host adds them
between the two graphs), CFL-limited frames with substeps of different dt,
and scenes the graph path does not take (viscosity: every substep eager)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from fluide_engine_amd import capi  # noqa: E402

DT = 1.0 / 480.0


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    print(f"\n[gpu] device {c.device_name}", flush=True)
    yield c
    c.close()


def _pair(ctx, n, method, lo=(-1, -1, -1), hi=(0.4, 0.6, 2.0), v0=None, viscosity=0.0):
    out = []
    for graph in (False, True):
        desc = capi.scene_desc((n, n, n), method=method, blend=0.03, max_particles=8 * n ** 3, viscosity=viscosity)
        x = capi.seed_particles(desc, [capi.box(lo, hi)], seed=7)
        v = np.zeros_like(x) if v0 is None else np.tile(np.asarray(v0, np.float32), (x.shape[0], 1))
        sc = capi.Scene(ctx, desc)
        sc.upload_particles(x, v)
        sc.set_graph(graph)
        out.append(sc)
    return out


def _same(e, g, tag, apic=False):
    ex, ev, eC = e.download_particles()
    gx, gv, gC = g.download_particles()
    assert np.array_equal(ex, gx) and np.array_equal(ev, gv), tag
    if apic:
        assert np.array_equal(eC, gC), tag
    for f in (capi.GRID_U, capi.GRID_V, capi.GRID_W, capi.GRID_PRESSURE, capi.GRID_FLAGS):
        assert np.array_equal(e.get(f), g.get(f)), (tag, f)


def _same_stats(se, sg, tag):
    assert (se.iterations, se.converged) == (sg.iterations, sg.converged), (tag, se.iterations, sg.iterations)
    assert se.rel_residual == sg.rel_residual and se.abs_residual == sg.abs_residual, tag


SOLVERS = {"mgpcg": capi.SOLVER_MGPCG, "cg": capi.SOLVER_CG, "pcg": capi.SOLVER_PCG_JACOBI}


@pytest.mark.parametrize("method", [capi.PIC, capi.FLIP, capi.APIC])
@pytest.mark.parametrize("solver", ["mgpcg", "cg", "pcg"])
def test_graph_substeps_bit_identical(ctx, method, solver):
    print(f"[gpu] graph substeps method {method} solver {solver}", flush=True)
    e, g = _pair(ctx, 32, method)
    kind = SOLVERS[solver]
    cap = 400 if kind != capi.SOLVER_MGPCG else 100
    for step in range(12):
        # a looser tolerance first, then a tighter one: the recaptured graph
        # holds the loose solves' iteration count + 2, fewer than the tight
        # solve needs, so the host adds iterations between the graphs
        tol = 1e-3 if step < 5 else 1e-7
        sd = capi.solver(kind, cap, 0, tol, relative=True, warm_start=True)
        se = e.substep(DT, sd)
        sg = g.substep(DT, sd)
        _same_stats(se, sg, step)
        _same(e, g, step, apic=method == capi.APIC)
    n_graph, captures, resumed = g.graph_stats()
    print(f"[gpu]   graph substeps {n_graph}, captures {captures}, resumed {resumed}", flush=True)
    assert e.graph_stats() == (0, 0, 0)
    assert n_graph == 11  # the first substep runs eagerly (its bins come from no advection)
    assert captures >= 4  # one pair per particle-buffer parity, again after the tolerance change
    assert resumed >= 1
    e.close()
    g.close()


def test_graph_frames_with_cfl_substeps(ctx):
    """A dam thrown sideways: frames split into CFL substeps of different dt,
    which reach the captured kernels through device memory."""
    print("[gpu] graph frames with CFL substeps", flush=True)
    e, g = _pair(ctx, 32, capi.FLIP, v0=(12.0, 0.0, 0.0))
    sd = capi.solver(capi.SOLVER_MGPCG, 200, 0, 1e-6, relative=True, warm_start=True)
    counts = []
    for frame in range(4):
        ne = e.advance_frame(1.0 / 120, sd)
        ng = g.advance_frame(1.0 / 120, sd)
        assert ne == ng
        counts.append(ng)
        _same(e, g, frame)
    n_graph, captures, _ = g.graph_stats()
    print(f"[gpu]   substeps per frame {counts}, graph substeps {n_graph}, captures {captures}", flush=True)
    assert max(counts) >= 2
    assert n_graph == sum(counts) - 1
    e.close()
    g.close()


def test_graph_off_for_viscous_scenes_and_after_disable(ctx):
    print("[gpu] graph mode on a viscous scene, then switched off", flush=True)
    e, g = _pair(ctx, 24, capi.FLIP, viscosity=0.05)
    sd = capi.solver(capi.SOLVER_MGPCG, 200, 0, 1e-6, relative=True, warm_start=True)
    for step in range(3):
        _same_stats(e.substep(DT, sd), g.substep(DT, sd), step)
        _same(e, g, step)
    assert g.graph_stats()[0] == 0  # viscosity: every substep eager
    e.close()
    g.close()
    e, g = _pair(ctx, 24, capi.FLIP)
    for step in range(6):
        if step == 3:
            g.set_graph(False)
        _same_stats(e.substep(DT, sd), g.substep(DT, sd), step)
        _same(e, g, step)
    assert g.graph_stats()[0] == 2  # substeps 1 and 2
    e.close()
    g.close()
