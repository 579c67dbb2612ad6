"""Pins the oracle's BackwardEulerDiffusionSolver restatement
(src/fluid_solvers/backward_euler_diffusion_solver.cpp) without retyping it:

  * face markers against the separately pinned ScalarGrid3D sample
    (orc_sample_array) at face positions computed here from the data origins
    of test/face_centered_grid3d_test.cpp:48-54;
  * a Neumann cosine mode, the closed-form eigenvector of the all-FLUID
    system, per axis on an anisotropic grid (catches a wrong c_a, a swapped
    axis or a dropped boundary term);
  * conservation: without BOUNDRY faces every row and column of the system
    sums to 1, so the solve preserves the sum of each component;
  * the maximum principle with BOUNDRY faces (a convex combination);
  * a dense direct solve of the implicit heat step written as fluxes
    (Dirichlet at BOUNDRY faces, no flux at AIR faces and domain edges).
"""
import numpy as np
import pytest

from oracle import oracle as orc


def _scene(n, size=(1.0, 1.0, 1.0)):
    return orc.Scene(n, size=size, method=orc.FLIP, blend=0.05, max_cfl=3.0, gravity=(0.0, 0.0, 0.0))


def _face_shape(n, c):
    nx, ny, nz = n
    return [(nz, ny, nx + 1), (nz, ny + 1, nx), (nz + 1, ny, nx)][c]


def _face_pos(n, h, c, origin=(0.0, 0.0, 0.0)):
    """GetX/Y/ZPos: data origin (o + h*(0 or 0.5 per axis)) + h * index, as (nz, ny, nx, 3)."""
    shp = _face_shape(n, c)
    k, j, i = np.meshgrid(*[np.arange(s) for s in shp], indexing="ij")
    idx = [i, j, k]
    return np.stack([(origin[a] + h[a] * (0.0 if a == c else 0.5)) + h[a] * idx[a] for a in range(3)], axis=-1)


def _set_sdf(s, sdf):
    s.set(orc.SDF, sdf)


def test_markers_match_sampled_sdf():
    n = (7, 6, 5)
    size = (1.0, 0.9, 0.7)
    h = tuple(size[a] / n[a] for a in range(3))
    rng = np.random.default_rng(3)
    s = _scene(n, size)
    sdf = rng.uniform(-1, 1, (n[2], n[1], n[0]))
    _set_sdf(s, sdf)
    s.set_viscosity(0.01)
    s.diffusion(1.0 / 240)
    so = tuple(0.5 * h[a] for a in range(3))
    for c in range(3):
        pos = _face_pos(n, h, c)
        want = np.empty(pos.shape[:3], np.uint8)
        for idx in np.ndindex(*pos.shape[:3]):
            v = orc.sample_array(sdf, so, h, pos[idx])
            want[idx] = orc.FLUID if v < 0 else orc.AIR  # the collider SDF is kMaxDistance
        got = s.diffusion_markers(c)
        assert np.array_equal(got, want), c
        assert 0 < int((got == orc.FLUID).sum()) < got.size


@pytest.mark.parametrize("c,axis,m", [(0, 0, 1), (0, 1, 2), (1, 0, 3), (1, 1, 1), (2, 2, 2), (2, 0, 1)])
def test_cosine_mode_closed_form(c, axis, m):
    """All faces FLUID: each axis is a Neumann chain of N faces, whose mode
    cos(pi m (i + 1/2) / N) has Laplacian eigenvalue 2 - 2 cos(pi m / N), so
    the implicit step divides it by 1 + c_axis (2 - 2 cos(pi m / N))."""
    n = (10, 8, 6)
    size = (1.0, 1.2, 0.9)
    h = tuple(size[a] / n[a] for a in range(3))
    dt, nu = 1.0 / 100, 0.02
    s = _scene(n, size)
    _set_sdf(s, -np.ones((n[2], n[1], n[0])))
    s.set_viscosity(nu)
    s.set_diffusion_solver(orc.solver(orc.JACOBI, 20000, 5, 1e-13))
    shp = _face_shape(n, c)
    N = shp[2 - axis]
    i = np.arange(N)
    mode = np.cos(np.pi * m * (i + 0.5) / N)
    bshape = [1, 1, 1]
    bshape[2 - axis] = N
    field = np.broadcast_to(mode.reshape(bshape), shp).copy()
    what = (orc.U, orc.V, orc.W)[c]
    for cc in range(3):  # the other components at zero
        s.set((orc.U, orc.V, orc.W)[cc], np.zeros(_face_shape(n, cc)))
    s.set(what, field)
    st = s.diffusion(dt)
    cax = dt * nu / (h[axis] * h[axis])
    want = field / (1.0 + cax * (2.0 - 2.0 * np.cos(np.pi * m / N)))
    got = s.get(what)
    assert np.abs(got - want).max() < 1e-12
    assert st[c].iterations < 20000


def test_conservation_without_boundary_faces():
    n = (9, 7, 6)
    rng = np.random.default_rng(5)
    s = _scene(n)
    _set_sdf(s, rng.uniform(-1, 0.6, (n[2], n[1], n[0])))
    s.set_viscosity(0.05)
    s.set_diffusion_solver(orc.solver(orc.JACOBI, 20000, 5, 1e-14))
    before = []
    for c, what in enumerate((orc.U, orc.V, orc.W)):
        f = rng.normal(size=_face_shape(n, c))
        s.set(what, f)
        before.append(f)
    s.diffusion(1.0 / 120)
    for c, what in enumerate((orc.U, orc.V, orc.W)):
        got = s.get(what)
        assert abs(got.sum() - before[c].sum()) < 1e-10 * np.abs(before[c]).sum()
        assert np.abs(got - before[c]).max() > 1e-3  # something diffused


def _dense_flux_solve(u, mark, cax):
    """Implicit heat step as fluxes: for a FLUID face, x - sum over in-range
    neighbours of c_a * (y_nb - x) = u, with y_nb = x_nb for FLUID, u_nb fixed
    for BOUNDRY, and no flux for AIR; every other face keeps u."""
    shp = u.shape
    nf = u.size
    A = np.zeros((nf, nf))
    rhs = u.reshape(-1).copy()
    lin = np.arange(nf).reshape(shp)
    for idx in np.ndindex(*shp):
        r = lin[idx]
        A[r, r] = 1.0
        if mark[idx] != orc.FLUID:
            continue
        for ax in range(3):  # array axis 2 = x, 1 = y, 0 = z
            arr_ax = 2 - ax
            for d in (-1, 1):
                nb = list(idx)
                nb[arr_ax] += d
                if not 0 <= nb[arr_ax] < shp[arr_ax]:
                    continue
                nb = tuple(nb)
                if mark[nb] == orc.AIR:
                    continue
                A[r, r] += cax[ax]
                if mark[nb] == orc.FLUID:
                    A[r, lin[nb]] -= cax[ax]
                else:
                    rhs[r] += cax[ax] * u[nb]
    return np.linalg.solve(A, rhs).reshape(shp)


def test_dense_solve_with_boundary_faces_and_max_principle():
    n = (5, 4, 4)
    size = (1.0, 0.8, 0.8)
    h = tuple(size[a] / n[a] for a in range(3))
    dt, nu = 1.0 / 60, 0.03
    rng = np.random.default_rng(11)
    s = _scene(n, size)
    _set_sdf(s, rng.uniform(-1, 0.5, (n[2], n[1], n[0])))
    # collider SDF is node centred at the origin: negative in the x > 0.6 slab
    xs = np.arange(n[0]) * h[0]
    csdf = np.broadcast_to((0.6 - xs)[None, None, :], (n[2], n[1], n[0])).copy()
    s.set(orc.COLLIDER_SDF, csdf)
    s.set_viscosity(nu)
    s.set_diffusion_solver(orc.solver(orc.JACOBI, 50000, 5, 1e-14))
    fields = []
    for c, what in enumerate((orc.U, orc.V, orc.W)):
        f = rng.normal(size=_face_shape(n, c))
        s.set(what, f)
        fields.append(f)
    s.diffusion(dt)
    cax = tuple(dt * nu / (h[a] * h[a]) for a in range(3))
    seen_boundary = False
    for c, what in enumerate((orc.U, orc.V, orc.W)):
        mark = s.diffusion_markers(c)
        seen_boundary |= bool((mark == orc.BOUNDRY).any())
        want = _dense_flux_solve(fields[c], mark, cax)
        got = s.get(what)
        assert np.abs(got - want).max() < 1e-11, c
        assert got.min() >= fields[c].min() - 1e-12 and got.max() <= fields[c].max() + 1e-12
    assert seen_boundary


def test_constant_field_is_fixed_point():
    n = (6, 6, 6)
    rng = np.random.default_rng(2)
    s = _scene(n)
    _set_sdf(s, rng.uniform(-1, 1, (n[2], n[1], n[0])))
    s.set_viscosity(0.5)
    for c, what in enumerate((orc.U, orc.V, orc.W)):
        s.set(what, np.full(_face_shape(n, c), 0.75 + c))
    st = s.diffusion(1.0 / 30)
    for c, what in enumerate((orc.U, orc.V, orc.W)):
        assert np.abs(s.get(what) - (0.75 + c)).max() < 1e-14
        assert st[c].iterations == 6  # first check (after sweep 5) finds the residual below 1e-14


def test_substep_runs_diffusion_only_with_viscosity():
    desc_n = (12, 12, 12)
    x = np.array([[0.3, 0.3, 0.5], [0.35, 0.32, 0.5], [0.4, 0.3, 0.55]] * 4, dtype=np.float64)
    x += np.random.default_rng(0).uniform(0, 0.05, x.shape)
    v = np.random.default_rng(1).normal(size=x.shape)
    out = []
    for nu in (0.0, 0.0, 0.2):
        s = _scene(desc_n)
        s.set_particles(x, v)
        s.set_viscosity(nu)
        s.substep(1.0 / 240, orc.solver(orc.JACOBI, 200, 5, 1e-14))
        out.append(s.particles()[1])
    assert np.array_equal(out[0], out[1])
    assert np.abs(out[2] - out[0]).max() > 1e-6
