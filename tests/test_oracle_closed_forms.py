"""Pin the FP64 oracle where the reference ships no test: closed forms,
invariants, brute force on tiny inputs and dense direct solves.  Each check is
chosen so that a plausible restatement mistake (a wrong origin offset, a
transposed index, a dropped term, a sign flip) fails it.  CPU only.
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc

rng = np.random.default_rng(1234)


def hat_weight_1d(xhat, node):
    """Independent form of the 1-D linear weight: the hat function max(0, 1-|x-node|)."""
    return max(0.0, 1.0 - abs(xhat - node))


def test_barycentric_clamp_rules():
    # src/common/math_utils.hpp:48-73: below range -> (0, 0); above -> (high-1, 1)
    assert orc.barycentric(-0.3, 0, 5) == (0, 0.0)
    assert orc.barycentric(7.2, 0, 5) == (4, 1.0)
    assert orc.barycentric(4.0, 0, 5) == (4, 0.0)
    assert orc.barycentric(4.5, 0, 5) == (4, 0.5)
    assert orc.barycentric(5.0, 0, 5) == (4, 1.0)
    i, f = orc.barycentric(2.25, 0, 5)
    assert (i, f) == (2, 0.25)
    assert orc.barycentric(3.7, 0, 0) == (0, 0.0)


def test_weights_match_hat_functions_interior():
    size, h = (8, 8, 8), (0.125, 0.125, 0.125)
    origin = (0.0, 0.0625, 0.0625)
    for _ in range(200):
        x = rng.uniform(0.2, 0.8, 3)
        idx, w = orc.coords_weights(size, origin, h, x)
        assert w.sum() == pytest.approx(1.0, abs=1e-15)
        xhat = [(x[a] - origin[a]) / h[a] for a in range(3)]
        for q in range(8):
            ref = 1.0
            for a in range(3):
                ref *= hat_weight_1d(xhat[a], idx[q, a])
            assert w[q] == pytest.approx(ref, abs=1e-14)
        # node order (000,100,010,110,001,101,011,111)
        assert list(idx[1] - idx[0]) == [1, 0, 0]
        assert list(idx[2] - idx[0]) == [0, 1, 0]
        assert list(idx[4] - idx[0]) == [0, 0, 1]


def test_grad_weights_are_derivatives_of_weights():
    size, h = (8, 8, 8), (0.125, 0.25, 0.5)
    origin = (0.0, 0.1, 0.2)
    eps = 1e-7
    for _ in range(50):
        x = np.array(origin) + np.array(h) * rng.uniform(0.5, 5.5, 3)  # inside the unclamped range
        idx, g = orc.coords_grad_weights(size, origin, h, x)
        _, w0 = orc.coords_weights(size, origin, h, x)
        for a in range(3):
            xp = x.copy()
            xp[a] += eps
            idx2, w1 = orc.coords_weights(size, origin, h, xp)
            if not np.array_equal(idx, idx2):
                continue
            fd = (w1 - w0) / eps
            assert np.allclose(g[:, a], fd, atol=1e-5)


def _box_scene(n, method, blend=0.05):
    return orc.Scene((n, n, n), method=method, blend=blend)


def test_p2g_uniform_velocity_and_markers():
    for method in (orc.PIC, orc.FLIP, orc.APIC):
        s = _box_scene(8, method)
        x = rng.uniform(0.1, 0.9, (300, 3))
        v = np.tile([0.7, -1.25, 2.5], (300, 1))
        s.set_particles(x, v)
        s.p2g()
        for comp, mk, val in ((orc.U, orc.MU, 0.7), (orc.V, orc.MV, -1.25), (orc.W, orc.MW, 2.5)):
            a, m = s.get(comp), s.get(mk)
            marked = a[m == 1]
            # a marked face is the weighted average (= val) unless every weight on it
            # was zero (Q2: zero-weight nodes of a clamped stencil are still marked)
            is_val = np.abs(marked - val) < 1e-13
            assert np.all(is_val | (marked == 0.0))
            assert is_val.mean() > 0.9
            assert np.all(a[m == 0] == 0.0)


def _brute_p2g(x, v, n, comp, dims_for_clamp, h):
    """Weighted average by brute force with hat weights (valid when no clamp applies)."""
    nx, ny, nz = n
    shape = {0: (nz, ny, nx + 1), 1: (nz, ny + 1, nx), 2: (nz + 1, ny, nx)}[comp]
    off = [0.5, 0.5, 0.5]
    off[comp] = 0.0
    num = np.zeros(shape)
    den = np.zeros(shape)
    for p in range(x.shape[0]):
        xhat = [x[p, a] / h - off[a] for a in range(3)]
        base = [int(math.floor(xhat[a])) for a in range(3)]
        for dk in (0, 1):
            for dj in (0, 1):
                for di in (0, 1):
                    I, J, K = base[0] + di, base[1] + dj, base[2] + dk
                    w = hat_weight_1d(xhat[0], I) * hat_weight_1d(xhat[1], J) * hat_weight_1d(xhat[2], K)
                    num[K, J, I] += v[p, comp] * w
                    den[K, J, I] += w
    out = np.zeros(shape)
    nz_ = den > 0
    out[nz_] = num[nz_] / den[nz_]
    return out


def test_p2g_brute_force_interior():
    n = 6
    h = 1.0 / n
    x = rng.uniform(0.26, 0.74, (60, 3))  # stencils stay clear of every clamp
    v = rng.normal(size=(60, 3))
    for method in (orc.PIC, orc.FLIP):
        s = _box_scene(n, method)
        s.set_particles(x, v)
        s.p2g()
        for comp, what in ((0, orc.U), (1, orc.V), (2, orc.W)):
            ref = _brute_p2g(x, v, (n, n, n), comp, None, h)
            assert np.allclose(s.get(what), ref, atol=1e-13)


def test_pic_p2g_never_writes_last_face_but_flip_does():
    # Q1: PIC clamps the normal-axis stencil at n-2 (cell dims), FLIP at n-1.
    n = 4
    x = np.array([[0.99, 0.5, 0.5], [0.5, 0.99, 0.5], [0.5, 0.5, 0.99]])
    v = np.ones((3, 3))
    pic, flip = _box_scene(n, orc.PIC), _box_scene(n, orc.FLIP)
    for s in (pic, flip):
        s.set_particles(x, v)
        s.p2g()
    assert np.all(pic.get(orc.MU)[:, :, n] == 0)
    assert np.all(pic.get(orc.MV)[:, n, :] == 0)
    assert np.all(pic.get(orc.MW)[n, :, :] == 0)
    assert flip.get(orc.MU)[:, :, n].any()
    # Q2: PIC marks the zero-weight node of a clamped stencil (particle in the last cell)
    assert pic.get(orc.MU)[2, 2, n - 2] == 1


def test_apic_reproduces_affine_field():
    # APIC P2G with C = grad(v) reproduces an affine field on faces whose stencil
    # particles never hit the transverse clamp (src/apic_simulator.cpp:54-63).
    n = 8
    h = 1.0 / n
    A = rng.normal(size=(3, 3))
    c = rng.normal(size=3)
    x = rng.uniform(0.25, 0.75, (4000, 3))
    v = x @ A.T + c
    Cm = np.repeat(A.reshape(1, 9), x.shape[0], axis=0)
    s = _box_scene(n, orc.APIC)
    s.set_particles(x, v, Cm)
    s.p2g()
    for comp, what, mk in ((0, orc.U, orc.MU), (1, orc.V, orc.MV), (2, orc.W, orc.MW)):
        a, m = s.get(what), s.get(mk)
        off = np.array([0.5, 0.5, 0.5])
        off[comp] = 0.0
        kk, jj, ii = np.nonzero(m)
        pos = np.stack([(ii + off[0]) * h, (jj + off[1]) * h, (kk + off[2]) * h], axis=1)
        expect = pos @ A[comp] + c[comp]
        assert np.allclose(a[kk, jj, ii], expect, atol=1e-12)


def test_flip_snapshot_and_g2p_blend():
    n = 8
    x = rng.uniform(0.2, 0.8, (500, 3))
    v = rng.normal(size=(500, 3))
    for blend in (0.0, 0.03, 1.0):
        s = _box_scene(n, orc.FLIP, blend=blend)
        s.set_particles(x, v)
        s.p2g()
        assert np.array_equal(s.get(orc.RU), s.get(orc.U))  # Q4 snapshot
        pic = _box_scene(n, orc.PIC)
        pic.set(orc.U, s.get(orc.U)); pic.set(orc.V, s.get(orc.V)); pic.set(orc.W, s.get(orc.W))
        pic.set_particles(x, v)
        pic.g2p()
        _, vpic, _ = pic.particles()
        s.g2p()  # grid unchanged -> delta = 0 -> v_flip = v_p
        _, vout, _ = s.particles()
        assert np.allclose(vout, (1 - blend) * v + blend * vpic, atol=1e-13)


def test_g2p_reproduces_linear_field():
    # trilinear interpolation is exact for fields linear in x, y, z
    n = 8
    h = 1.0 / n
    s = _box_scene(n, orc.PIC)
    A = rng.normal(size=(3, 3))
    c = rng.normal(size=3)
    for comp, what in ((0, orc.U), (1, orc.V), (2, orc.W)):
        shp = s.shape(what)
        off = np.array([0.5, 0.5, 0.5])
        off[comp] = 0.0
        K, J, I = np.meshgrid(*[np.arange(d) for d in shp], indexing="ij")
        pos = np.stack([(I + off[0]) * h, (J + off[1]) * h, (K + off[2]) * h], axis=-1)
        s.set(what, pos @ A[comp] + c[comp])
    x = rng.uniform(0.1, 0.9, (200, 3))
    s.set_particles(x, np.zeros_like(x))
    s.g2p()
    _, v, _ = s.particles()
    assert np.allclose(v, x @ A.T + c, atol=1e-12)


def test_move_particles_constant_field_and_clamp():
    n = 8
    s = _box_scene(n, orc.PIC)
    vel = np.array([0.3, -0.2, 0.1])
    for comp, what in ((0, orc.U), (1, orc.V), (2, orc.W)):
        s.set(what, np.full(s.shape(what), vel[comp]))
    x = rng.uniform(0.2, 0.8, (100, 3))
    x[0] = [0.001, 0.5, 0.999]
    v = np.tile([5.0, 6.0, 7.0], (100, 1))
    s.set_particles(x, v)
    dt = 0.05
    s.move_particles(dt)
    xo, vo, _ = s.particles()
    assert np.allclose(xo[1:], x[1:] + dt * vel, atol=1e-14)
    assert np.all(vo[1:] == v[1:])
    # particle 0 crosses z = 1: clamped to the wall, that component zeroed (src/pic_simulator.cpp:255-295)
    assert xo[0, 2] == 1.0 and vo[0, 2] == 0.0 and vo[0, 0] == 5.0


def test_sdf_sign_matches_brute_force():
    for n in (8, 12):
        s = _box_scene(n, orc.FLIP)
        x = rng.uniform(0.0, 1.0, (int(2.5 * n ** 3 * 0.3), 3))
        s.set_particles(x, np.zeros_like(x))
        s.build_sdf()
        s.build_markers()
        h = 1.0 / n
        r = 1.2 * h / math.sqrt(2.0)
        c = (np.arange(n) + 0.5) * h
        K, J, I = np.meshgrid(c, c, c, indexing="ij")
        centres = np.stack([I, J, K], axis=-1).reshape(-1, 3)
        d = np.sqrt(((centres[:, None, :] - x[None, :, :]) ** 2).sum(-1)).min(axis=1).reshape(n, n, n)
        fluid = d < r
        flags = s.get(orc.FLAGS)
        assert np.array_equal(flags == orc.FLUID, fluid)
        sdf = s.get(orc.SDF)
        neg = sdf < 0
        assert np.allclose(sdf[neg], d[neg] - r, atol=1e-15)


def test_build_system_by_hand():
    # 3x1x1 line: FLUID FLUID AIR, h = 0.5 -> 1/h^2 = 4
    s = orc.Scene((3, 1, 1), size=(1.5, 0.5, 0.5), method=orc.PIC)
    s.set(orc.SDF, np.array([-1.0, -1.0, 1.0]).reshape(1, 1, 3))
    u = np.array([0.0, 1.0, 3.0, 0.0]).reshape(1, 1, 4)
    s.set(orc.U, u)
    s.build_markers()
    s.build_system()
    b = s.get(orc.RHS).ravel()
    assert list(s.get(orc.FLAGS).ravel()) == [orc.FLUID, orc.FLUID, orc.AIR]
    assert b[0] == pytest.approx((1.0 - 0.0) / 0.5) and b[1] == pytest.approx((3.0 - 1.0) / 0.5) and b[2] == 0.0


def test_cg_and_jacobi_match_dense_solve():
    # a random fluid/air mask with at least one air cell per connected fluid region
    n = 5
    solved = 0
    for trial in range(3):
        mask = rng.uniform(size=(n, n, n)) < 0.7
        mask[0, 0, 0] = False
        s = orc.Scene((n, n, n), method=orc.PIC)
        s.set(orc.SDF, np.where(mask, -1.0, 1.0))
        for comp, what in ((0, orc.U), (1, orc.V), (2, orc.W)):
            s.set(what, rng.normal(size=s.shape(what)))
        s.build_markers()
        s.build_system()
        flags = s.get(orc.FLAGS)
        b = s.get(orc.RHS)
        # independent dense assembly from the flags (5-point rule of BuildSystem)
        N = n ** 3
        A = np.zeros((N, N))
        idx = lambda i, j, k: i + n * (j + n * k)
        h2 = float(n * n)
        for k in range(n):
            for j in range(n):
                for i in range(n):
                    r = idx(i, j, k)
                    if flags[k, j, i] != orc.FLUID:
                        A[r, r] = 1.0
                        continue
                    for (di, dj, dk) in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                        ii, jj, kk = i + di, j + dj, k + dk
                        if not (0 <= ii < n and 0 <= jj < n and 0 <= kk < n):
                            continue
                        f = flags[kk, jj, ii]
                        if f == orc.BOUNDRY:
                            continue
                        A[r, r] += h2
                        if f == orc.FLUID:
                            A[r, idx(ii, jj, kk)] -= h2
        # singular rows (fluid islands with no air) are removed by the random mask check
        try:
            xstar = np.linalg.solve(A, b.reshape(-1))
        except np.linalg.LinAlgError:
            continue
        assert np.allclose(A, A.T)
        st = s.pressure(orc.solver(orc.CG, 2000, 0, 1e-13, relative=True, warm_start=False))
        p = s.get(orc.PRESSURE).reshape(-1)
        assert np.allclose(p, xstar, atol=1e-8 * max(1.0, np.abs(xstar).max()))
        assert st.rel_residual < 1e-12
        s2 = orc.Scene((n, n, n), method=orc.PIC)
        s2.set(orc.SDF, np.where(mask, -1.0, 1.0))
        for comp, what in ((0, orc.U), (1, orc.V), (2, orc.W)):
            s2.set(what, s.get(what))
        st2 = s2.pressure(orc.solver(orc.PCG_JACOBI, 2000, 0, 1e-13, relative=True, warm_start=False))
        assert st2.rel_residual < 1e-12
        solved += 1
    assert solved >= 1


def test_apply_pressure_only_plus_faces():
    # Q6: a FLUID cell at i=1 next to AIR at i=0: face i=1 (its -side) is not corrected
    s = orc.Scene((3, 1, 1), size=(3.0, 1.0, 1.0), method=orc.PIC)
    s.set(orc.SDF, np.array([1.0, -1.0, 1.0]).reshape(1, 1, 3))
    s.set(orc.U, np.zeros((1, 1, 4)))
    s.build_markers()
    s.build_system()
    s.set(orc.PRESSURE, np.array([0.0, 2.0, 0.0]).reshape(1, 1, 3))
    s.apply_pressure()
    u = s.get(orc.U).ravel()
    assert u[1] == 0.0  # -side face of the FLUID cell: untouched
    assert u[2] == pytest.approx(0.0 - 2.0)  # +side: (p2 - p1)/h


def test_walls_and_gravity_closed_form():
    s = _box_scene(4, orc.FLIP)
    for what in (orc.U, orc.V, orc.W):
        s.set(what, np.ones(s.shape(what)))
    s.apply_gravity(0.1)
    s.apply_bc()
    u, v, w = s.get(orc.U), s.get(orc.V), s.get(orc.W)
    assert np.all(u[:, :, 0] == 0) and np.all(u[:, :, -1] == 0) and np.all(u[:, :, 1:-1] == 1.0)
    assert np.all(v[:, 0, :] == 0) and np.all(v[:, -1, :] == 0)
    assert np.allclose(v[:, 1:-1, :], 1.0 + 0.1 * -9.81)
    assert np.all(w[0] == 0) and np.all(w[-1] == 0)


def test_cfl_closed_form():
    s = _box_scene(4, orc.FLIP)
    s.set(orc.U, np.full(s.shape(orc.U), -2.0))
    dt = 0.01
    assert s.cfl(dt) == pytest.approx(2.0 * dt / 0.25)
    s.set(orc.U, np.zeros(s.shape(orc.U)))
    s.set(orc.V, np.full(s.shape(orc.V), 1.0))
    # Q13: the y component adds dt*(-9.8)
    assert s.cfl(dt) == pytest.approx((1.0 - 9.8 * dt) * dt / 0.25)


# --- G2P and advection on multilinear fields --------------------------------
# A field F(x) = k + A.x + B*x*y*z is multilinear, so trilinear interpolation
# reproduces it exactly inside every cell and the gradient weights give its
# exact gradient.  With the clamps of GetBarycentric the sample and gradient
# are those of F at a clamped point q, which gives closed forms for G2P (PIC,
# FLIP delta, APIC C) and for the midpoint rule of MoveParticles.  Each check
# is shown to fail on a planted restatement mistake (Scene.plant_mistake).

def _multilinear(n, k, A, B):
    """Node values of F_c = k[c] + A[c].x + B[c]*x*y*z on the three face grids."""
    h = 1.0 / n
    out = []
    for comp in range(3):
        dims = [n, n, n]
        dims[comp] += 1
        off = np.array([0.5, 0.5, 0.5])
        off[comp] = 0.0
        K, J, I = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
        X, Y, Z = (I + off[0]) * h, (J + off[1]) * h, (K + off[2]) * h
        out.append(k[comp] + A[comp, 0] * X + A[comp, 1] * Y + A[comp, 2] * Z + B[comp] * X * Y * Z)
    return out


def _clamped_point(n, comp, p, dims):
    """The point whose value GetBarycentric samples: each axis clamped to the node range."""
    h = 1.0 / n
    o = np.array([0.5, 0.5, 0.5]) * h
    o[comp] = 0.0
    return o + h * np.clip((p - o) / h, 0.0, np.array(dims, dtype=np.float64) - 1.0)


def _field_value(k, A, B, comp, q):
    return k[comp] + q @ A[comp] + B[comp] * q[:, 0] * q[:, 1] * q[:, 2]


def _field_gradient(A, B, comp, q):
    g = np.tile(A[comp], (q.shape[0], 1))
    g[:, 0] += B[comp] * q[:, 1] * q[:, 2]
    g[:, 1] += B[comp] * q[:, 0] * q[:, 2]
    g[:, 2] += B[comp] * q[:, 0] * q[:, 1]
    return g


def _set_fields(s, what3, fields):
    for what, f in zip(what3, fields):
        s.set(what, f)


def _apic_c_expected(n, A, B, x):
    """src/apic_simulator.cpp:147-175: row c of C is the gradient of component c,
    taken with the cell-count stencil (GetSize(), not the face dims) at the
    position clamped to [o + h/2, o + L - h/2] on the two transverse axes."""
    h = 1.0 / n
    C = np.empty((x.shape[0], 3, 3))
    for comp in range(3):
        pc = x.copy()
        for ax in range(3):
            if ax != comp:
                pc[:, ax] = np.clip(pc[:, ax], h / 2, 1.0 - h / 2)
        q = _clamped_point(n, comp, pc, (n, n, n))
        C[:, comp, :] = _field_gradient(A, B, comp, q)
    return C.reshape(-1, 9)


def _apic_g2p(n, k, A, B, x, plant=orc.PLANT_NONE):
    s = _box_scene(n, orc.APIC)
    s.plant_mistake(plant)
    _set_fields(s, (orc.U, orc.V, orc.W), _multilinear(n, k, A, B))
    s.set_particles(x, np.zeros_like(x), np.zeros((x.shape[0], 9)))
    s.g2p()
    _, v, Cm = s.particles()
    return v, Cm


def test_apic_g2p_affine_matrix_is_field_gradient():
    n = 8
    lr = np.random.default_rng(77)
    A = lr.normal(size=(3, 3))  # not symmetric: a transposed C cannot pass
    B = lr.normal(size=3)
    k = lr.normal(size=3)
    x = lr.uniform(0.0, 1.0, (3000, 3))  # includes the clamped rims next to every wall
    v, Cm = _apic_g2p(n, k, A, B, x)
    # velocity: FaceCenteredGrid3D::Sample with the face dims, unclamped position
    for comp in range(3):
        dims = [n, n, n]
        dims[comp] += 1
        q = _clamped_point(n, comp, x, dims)
        assert np.allclose(v[:, comp], _field_value(k, A, B, comp, q), atol=1e-12)
    expect = _apic_c_expected(n, A, B, x)
    assert np.allclose(Cm, expect, atol=1e-11)
    # interior particles see the plain field gradient
    inner = np.all((x > 1.0 / n) & (x < 1.0 - 1.0 / n), axis=1)
    for comp in range(3):
        assert np.allclose(Cm[inner].reshape(-1, 3, 3)[:, comp, :], _field_gradient(A, B, comp, x[inner]),
                           atol=1e-11)


@pytest.mark.parametrize("plant", [orc.PLANT_APIC_TRANSPOSE, orc.PLANT_APIC_FACE_DIMS])
def test_apic_g2p_check_fails_on_planted_mistake(plant):
    n = 8
    lr = np.random.default_rng(77)
    A = lr.normal(size=(3, 3))
    B = lr.normal(size=3)
    k = lr.normal(size=3)
    x = lr.uniform(0.0, 1.0, (3000, 3))
    _, Cm = _apic_g2p(n, k, A, B, x, plant)
    assert not np.allclose(Cm, _apic_c_expected(n, A, B, x), atol=1e-11)


def _flip_g2p(n, blend, snap, new, x, v, plant=orc.PLANT_NONE):
    s = _box_scene(n, orc.FLIP, blend=blend)
    s.plant_mistake(plant)
    _set_fields(s, (orc.RU, orc.RV, orc.RW), snap)
    _set_fields(s, (orc.U, orc.V, orc.W), new)
    s.set_particles(x, v)
    s.g2p()
    return s.particles()[1]


def _flip_expected(n, blend, f_snap, f_new, x, v):
    """src/flip_simulator.cpp:121-161: v_flip = v_p + Sample(u_new - u_snapshot);
    v = Lerp(v_flip, Sample(u_new), blend) = (1 - blend) v_flip + blend Sample(u_new)."""
    out = np.empty_like(v)
    for comp in range(3):
        dims = [n, n, n]
        dims[comp] += 1
        q = _clamped_point(n, comp, x, dims)
        new = _field_value(*f_new, comp, q)
        flip = v[:, comp] + new - _field_value(*f_snap, comp, q)
        out[:, comp] = flip if blend == 0.0 else (1 - blend) * flip + blend * new
    return out


def _flip_fields(seed):
    lr = np.random.default_rng(seed)
    f_snap = (lr.normal(size=3), lr.normal(size=(3, 3)), lr.normal(size=3))
    f_new = (lr.normal(size=3), lr.normal(size=(3, 3)), lr.normal(size=3))
    x = lr.uniform(0.0, 1.0, (1500, 3))
    v = lr.normal(size=(1500, 3))
    return f_snap, f_new, x, v


@pytest.mark.parametrize("blend", [0.0, 0.03, 0.5])
def test_flip_g2p_delta_on_changed_grid(blend):
    n = 8
    f_snap, f_new, x, v = _flip_fields(5)
    out = _flip_g2p(n, blend, _multilinear(n, *f_snap), _multilinear(n, *f_new), x, v)
    assert np.allclose(out, _flip_expected(n, blend, f_snap, f_new, x, v), atol=1e-12)


def test_flip_g2p_check_fails_on_planted_delta_sign():
    n = 8
    f_snap, f_new, x, v = _flip_fields(5)
    out = _flip_g2p(n, 0.03, _multilinear(n, *f_snap), _multilinear(n, *f_new), x, v, orc.PLANT_FLIP_DELTA_SIGN)
    assert not np.allclose(out, _flip_expected(n, 0.03, f_snap, f_new, x, v), atol=1e-12)


def _advect(n, A, c, x, dt, max_cfl, plant=orc.PLANT_NONE):
    s = orc.Scene((n, n, n), method=orc.PIC, max_cfl=max_cfl)
    s.plant_mistake(plant)
    _set_fields(s, (orc.U, orc.V, orc.W), _multilinear(n, c, A, np.zeros(3)))
    v = np.ones_like(x)
    s.set_particles(x, v)
    s.move_particles(dt)
    xo, vo, _ = s.particles()
    assert np.array_equal(vo, v)  # no wall reached: velocities untouched
    return xo


def _midpoint_closed_form(A, c, x, dt, m):
    """m = max(maxCfl, 1) midpoint steps of d = dt/m on u(x) = A x + c
    (src/pic_simulator.cpp:239-268): x <- (I + dA + d^2 A^2/2) x + (d I + d^2 A/2) c."""
    d = dt / m
    eye = np.eye(3)
    M = eye + d * A + 0.5 * d * d * (A @ A)
    t = (d * eye + 0.5 * d * d * A) @ c
    for _ in range(m):
        x = x @ M.T + t
    return x


@pytest.mark.parametrize("max_cfl,m", [(3.0, 3), (5.0, 5), (2.5, 2), (0.5, 1)])
def test_move_particles_midpoint_on_linear_field(max_cfl, m):
    n = 8
    lr = np.random.default_rng(11)
    A = lr.normal(size=(3, 3))
    c = lr.normal(size=3)
    x = lr.uniform(0.3, 0.7, (400, 3))
    dt = 0.05
    assert np.abs(x @ A.T + c).max() < 3.0  # with dt = 0.05 every path stays inside [h/2, 1 - h/2]
    xo = _advect(n, A, c, x, dt, max_cfl)
    assert np.allclose(xo, _midpoint_closed_form(A, c, x, dt, m), atol=1e-12)


def test_move_particles_check_fails_on_planted_euler():
    n = 8
    lr = np.random.default_rng(11)
    A = lr.normal(size=(3, 3))
    c = lr.normal(size=3)
    x = lr.uniform(0.3, 0.7, (400, 3))
    xo = _advect(n, A, c, x, 0.05, 3.0, orc.PLANT_EULER_ADVECTION)
    assert not np.allclose(xo, _midpoint_closed_form(A, c, x, 0.05, 3), atol=1e-12)
