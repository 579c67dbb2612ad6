"""Pins the oracle's collider path (the collider branches of
BlockedBoundryConditionSolver::ConstrainVelocity,
src/fluid_solvers/blocked_boundry_condition_solver.cpp:18-291, BOUNDRY cell
markers, and the SDF collision response of moved particles) with a planar
collider, whose SDF is linear so that every sample and gradient is exact:

  * no collider: the boundary condition is the wall zeroing only;
  * inside the collider every face carries the tangential part of the
    sampled velocity: for a collider face with normal -x, u is 0 and v, w
    keep what the extrapolation (depth 3) carried into the blocked faces;
  * faces between a collider node and a free node are zero;
  * cells outside the fluid whose centre samples the collider negative are
    BOUNDRY;
  * a particle that ends inside the collider lands on its plane and loses
    its inward normal velocity, keeping the tangential one.
"""
import numpy as np

from oracle import oracle as orc


def _nodes(n):
    h = 1.0 / n
    x = np.arange(n) * h  # the collider SDF is node centred at the origin
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    return X, Y, Z, h


def _scene(n):
    return orc.Scene((n, n, n), method=orc.FLIP, blend=0.05, max_cfl=3.0, gravity=(0.0, 0.0, 0.0))


def test_no_collider_is_walls_only():
    n = 8
    s = _scene(n)
    rng = np.random.default_rng(1)
    f = [rng.normal(size=s.shape(w)) for w in (orc.U, orc.V, orc.W)]
    for w, a in zip((orc.U, orc.V, orc.W), f):
        s.set(w, a)
    s.apply_bc()
    u, v, w = s.get(orc.U), s.get(orc.V), s.get(orc.W)
    assert np.all(u[:, :, 0] == 0) and np.all(u[:, :, -1] == 0)
    assert np.all(v[:, 0, :] == 0) and np.all(v[:, -1, :] == 0)
    assert np.all(w[0] == 0) and np.all(w[-1] == 0)
    assert np.array_equal(u[:, :, 1:-1], f[0][:, :, 1:-1])
    assert np.array_equal(v[:, 1:-1, :], f[1][:, 1:-1, :])


def test_planar_collider_tangential_and_boundary_faces():
    n = 12
    X, Y, Z, h = _nodes(n)
    x0 = 0.63
    s = _scene(n)
    s.set(orc.COLLIDER_SDF, x0 - X)  # collider where x > x0, outward normal -x
    a, b, c = 0.7, -0.4, 0.25
    for w, val in zip((orc.U, orc.V, orc.W), (a, b, c)):
        s.set(w, np.full(s.shape(w), val))
    s.apply_bc()
    u, v, w = s.get(orc.U), s.get(orc.V), s.get(orc.W)
    xu = np.arange(n + 1) * h          # u faces at x = i h
    xc = (np.arange(n) + 0.5) * h      # v, w faces at cell-centre x
    inside_u = xu > x0
    inside_c = xc > x0
    # inside the collider: tangential part of the (uniform) velocity; blocked
    # faces hold what the extrapolation carried into them (3 layers), deeper
    # ones stay 0
    reach = inside_c & (xc < x0 + 3 * h)
    assert np.all(u[:, 1:-1, :][:, :, inside_u] == 0)
    assert np.allclose(v[:, 1:-1, :][:, :, reach], b, atol=1e-12)
    assert np.allclose(w[1:-1, :, :][:, :, reach], c, atol=1e-12)
    assert np.all(v[:, 1:-1, :][:, :, xc > x0 + 4 * h] == 0)
    # outside the collider and away from walls: untouched
    out_u = xu < x0 - h
    assert np.allclose(u[1:-1, 1:-1, :][:, :, out_u][:, :, 1:], a)
    # faces between a collider node and a free node are zero
    solid = (x0 - np.arange(n) * h) < 0
    for i in range(n - 1):
        if solid[i] != solid[i + 1]:
            assert np.all(u[:, :, i + 1] == 0)


def test_boundry_markers_from_collider():
    n = 10
    X, Y, Z, h = _nodes(n)
    s = _scene(n)
    s.set(orc.COLLIDER_SDF, Y - 0.35)  # collider below y = 0.35
    s.set(orc.SDF, np.full((n, n, n), 1.0))  # no fluid
    s.build_markers()
    flags = s.get(orc.FLAGS)
    yc = (np.arange(n) + 0.5) * h
    want = np.where(yc < 0.35, orc.BOUNDRY, orc.AIR)
    assert np.array_equal(flags, np.broadcast_to(want[None, :, None], flags.shape))


def test_particle_collision_on_a_plane():
    n = 16
    X, Y, Z, h = _nodes(n)
    s = _scene(n)
    s.set(orc.COLLIDER_SDF, 0.3 - Y)  # collider above y = 0.3
    for w in (orc.U, orc.V, orc.W):
        s.set(w, np.zeros(s.shape(w)))
    x = np.array([[0.5, 0.33, 0.5], [0.4, 0.2, 0.6]])
    v = np.array([[0.2, 1.5, -0.1], [0.3, -0.5, 0.0]])
    s.set_particles(x, v)
    s.move_particles(1.0 / 240)
    xo, vo, _ = s.particles()
    assert abs(xo[0, 1] - 0.3) < 1e-12 and np.allclose(xo[0, [0, 2]], x[0, [0, 2]])
    assert np.allclose(vo[0], [0.2, 0.0, -0.1])  # inward (+y) component removed
    assert np.allclose(xo[1], x[1]) and np.allclose(vo[1], v[1])  # outside: untouched
