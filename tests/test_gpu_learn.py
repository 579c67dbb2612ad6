"""GPU checks of the learned pressure warm start (no reference counterpart)
against the numpy restatement in oracle/net_oracle.py:

  * one forward + backward pass through the C ABI on random parameters and
    inputs: the loss within 1e-2 relative, the prediction within 2e-2 of its
    scale and every layer's weight and bias gradient within 3e-2 relative
    (L2) of the restatement, which rounds to BF16 at the same points (the
    summation orders differ, so a value near a BF16 rounding boundary may land
    one ulp, 2^-8 relative, apart);
  * Adam on the device lowers the training loss (the relative A-norm error
    after the best scaling) on solver samples;
  * on a held-out system the trained guess keeps a fraction of the zero
    guess's A-norm error, and CG and multigrid-PCG from it reach the same
    tolerance and pressure in no more iterations.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from fluide_engine_amd import capi  # noqa: E402
from oracle import net_oracle as no  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("loss_kind", ["mse", "energy"])
@pytest.mark.parametrize("shape", [(16, 16, 16), (24, 16, 32)])
def test_forward_backward_matches_restatement(ctx, shape, loss_kind):
    print(f"[gpu] net forward/backward {shape} {loss_kind}", flush=True)
    nz, ny, nx = shape
    rng = np.random.default_rng(sum(shape))
    flags = rng.choice([0, 1, 1, 1, 2], size=shape).astype(np.uint8)
    flags[0] = 0
    ih2 = (float(nx * nx), float(ny * ny), float(nz * nz))
    if loss_kind == "mse":
        b = (rng.normal(size=shape) * (flags == 1)).astype(np.float32)
        target = rng.normal(size=shape).astype(np.float32)
    else:  # b = A x for a solution x on the FLUID cells
        target = np.where(flags == 1, rng.normal(size=shape), 0).astype(np.float32)
        b = no.apply_A(target, flags, ih2).astype(np.float32)
    params = no.random_params(7)
    net = capi.Net(ctx, seed=1)
    net.set_loss(capi.NET_LOSS_MSE if loss_kind == "mse" else capi.NET_LOSS_ENERGY)
    flat = capi.net_flatten(params)
    assert flat.size == capi.Net.param_count()
    net.set_params(flat)
    assert np.array_equal(net.params(), flat)
    loss, g, pred = net.loss_and_grads(b, flags, target)
    ref = no.Net(params)
    X0 = ref.features(b, flags)
    rloss, rg, acts = ref.loss_and_grads(X0, target, flags == 1, loss_kind=loss_kind, b=b, flags=flags, ih2=ih2)
    assert abs(loss - rloss) <= 1e-2 * abs(rloss), (loss, rloss)
    scale = float(np.abs(acts["out"]).max())
    assert float(np.abs(pred - acts["out"]).max()) <= 2e-2 * scale
    gd = capi.net_unflatten(g)
    for name, _, _ in capi.NET_LAYERS + [("head", 0, 0)]:
        assert _rel(gd[name][0], rg[name][0]) < 3e-2, (name, "W", _rel(gd[name][0], rg[name][0]))
        assert _rel(np.atleast_1d(gd[name][1]), np.atleast_1d(rg[name][1])) < 3e-2, (name, "b")
    net.close()


def _dam_scene(ctx, n, box_hi, seed=0):
    desc = capi.scene_desc((n, n, n), method=capi.FLIP, blend=0.03, max_particles=8 * n ** 3)
    x = capi.seed_particles(desc, [capi.box((-1, -1, -1), box_hi)], seed=seed)
    s = capi.Scene(ctx, desc)
    s.upload_particles(x, np.zeros_like(x))
    return s


def test_training_lowers_loss_and_warm_start_gains(ctx):
    """Trained on 16 systems of four dam breaks, the network's guess for a
    held-out system (a fifth dam) has a fraction of the zero guess's A-norm
    error -- the quantity CG reduces -- and neither CG nor multigrid-PCG needs
    more iterations from it; both reach the same pressure."""
    print("[gpu] net training and warm start", flush=True)
    n = 32
    dt = 1.0 / 240
    solver = capi.solver(capi.SOLVER_MGPCG, 500, 0, 1e-6, relative=True, warm_start=True)
    net = capi.Net(ctx, seed=3)
    boxes = [(0.4, 0.6, 2.0), (0.6, 0.4, 2.0), (2.0, 0.3, 0.5), (0.5, 0.7, 0.6)]
    for k, hi in enumerate(boxes):
        s = _dam_scene(ctx, n, hi, seed=k)
        for step in range(6):
            s.substep(dt, solver)
            if step >= 2:
                net.add_sample(s)
        s.close()
    assert net.sample_count() == 16
    losses = net.train(600, learning_rate=2e-3, seed=5)
    assert np.all(np.isfinite(losses))
    assert losses[-30:].mean() < 0.2 * losses[:30].mean(), (losses[:30].mean(), losses[-30:].mean())

    s = _dam_scene(ctx, n, (0.45, 0.55, 1.5), seed=9)
    for _ in range(4):
        s.substep(dt, solver)
    s.set_warm_start_net(net)
    stages = [capi.STAGE_P2G, capi.STAGE_SDF, capi.STAGE_EXTRAPOLATE, capi.STAGE_BC, capi.STAGE_GRAVITY, capi.STAGE_BC]
    for st in stages:
        s.run_stage(st, dt)
    u0 = [s.get(f) for f in (capi.GRID_U, capi.GRID_V, capi.GRID_W)]

    def pressure(kind, mode, iters=2000):
        for f, a in zip((capi.GRID_U, capi.GRID_V, capi.GRID_W), u0):
            s.set(f, a)
        st = s.run_stage(capi.STAGE_PRESSURE, dt, capi.solver(kind, iters, 0, 1e-6, relative=True, warm_start=mode))
        return st, s.get(capi.GRID_PRESSURE)

    st0, p_star = pressure(capi.SOLVER_CG, capi.WARM_START_ZERO)
    stl, p_l = pressure(capi.SOLVER_CG, capi.WARM_START_LEARNED)
    _, x0 = pressure(capi.SOLVER_CG, capi.WARM_START_LEARNED, iters=0)  # the guess itself
    mg0, _ = pressure(capi.SOLVER_MGPCG, capi.WARM_START_ZERO)
    mgl, _ = pressure(capi.SOLVER_MGPCG, capi.WARM_START_LEARNED)
    flags = s.get(capi.GRID_FLAGS)
    ih2 = (float(n * n),) * 3
    e = x0 - p_star
    rel_energy = float(np.sum(e * no.apply_A(e, flags, ih2)) / np.sum(p_star * no.apply_A(p_star, flags, ih2)))
    print(f"[gpu] held-out: relative A-norm error^2 of the guess {rel_energy:.4f} (zero guess 1); CG iterations "
          f"{st0.iterations} -> {stl.iterations}, multigrid-PCG {mg0.iterations} -> {mgl.iterations}", flush=True)
    assert rel_energy < 0.25
    assert st0.rel_residual <= 1e-6 and stl.rel_residual <= 1e-6
    assert np.abs(p_l - p_star).max() <= 1e-4 * np.abs(p_star).max()
    assert stl.iterations <= st0.iterations and mgl.iterations <= mg0.iterations
    assert s.warm_start_scale() != 0
    s.close()
    net.close()
