"""Pin the FP64 oracle against the reference's own known-answer tests.

Each test restates the inputs and expected values of one gtest case in the
reference tree (cited per test) and runs them through the oracle's restated
primitive.  CPU only.
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc


def test_math_utils_lerp_family():
    # test/math_utils_test.cpp:5-31 -- Lerp/Bilerp/Trilerp known answers
    assert orc.trilerp([0, 2, 0, 2, 0, 2, 0, 2], 0.5, 0.5, 0.5) == 1.0
    # Bilerp(0,2,0,2,.5,.5) == trilerp with equal z-layers
    assert orc.trilerp([0, 2, 0, 2, 0, 2, 0, 2], 0.5, 0.5, 0.0) == 1.0
    # Lerp(0,2,.5) = 1 is the x-only case
    assert orc.trilerp([0, 2, 0, 0, 0, 0, 0, 0], 0.5, 0.0, 0.0) == 1.0


def test_blas_known_answers():
    # test/blas_test.cpp:5-46
    v1 = np.full((2, 2, 2), 2.0)
    v2 = np.full((2, 2, 2), 3.0)
    assert orc.blas_dot(v1, v2) == 48.0
    Ac, Ar, Au, Af = (np.full((2, 2, 2), c) for c in (2.0, 3.0, 5.0, 4.0))
    r = orc.residual_rows(Ac, Ar, Au, Af, v2, v1)
    assert np.all(r == -25.0)
    assert math.sqrt(orc.blas_dot(v1, v1)) == math.sqrt(32)


def test_face_grid_data_origin_and_sample():
    # test/face_centered_grid3d_test.cpp:48-54 -- x data origin (0, .5, .5) at h = 1
    s = orc.Scene((2, 2, 2), size=(2.0, 2.0, 2.0))
    u = s.get(orc.U)
    assert u.shape == (2, 2, 3)
    # test/face_centered_grid3d_test.cpp:65-81: 3^3 grid, h = .5, Sample(.25,.25,.25) = (2, 3, 4.5)
    h = (0.5, 0.5, 0.5)
    x = np.ones((3, 3, 4))  # (nz, ny, nx+1)
    y = np.full((3, 4, 3), 2.0)
    z = np.full((4, 3, 3), 3.0)
    x[0, 0, 1] = 3.0
    y[0, 1, 0] = 4.0
    z[1, 0, 0] = 6.0
    p = (0.25, 0.25, 0.25)
    xo, yo, zo = (0.0, 0.25, 0.25), (0.25, 0.0, 0.25), (0.25, 0.25, 0.0)
    assert orc.sample_array(x, xo, h, p) == 2.0
    assert orc.sample_array(y, yo, h, p) == 3.0
    assert orc.sample_array(z, zo, h, p) == 4.5


def test_scalar_grid_sample():
    # test/scalar_grid3d_test.cpp:41-72
    g = np.zeros((3, 3, 3))  # (z, y, x)
    for (i, j, k) in [(0, 1, 0), (1, 1, 0), (0, 1, 1), (1, 1, 1), (2, 1, 1), (1, 1, 2), (2, 1, 2), (0, 1, 2), (2, 1, 0)]:
        g[k, j, i] = 1.0
    for (i, j, k) in [(0, 2, 0), (1, 2, 0), (0, 2, 1), (1, 2, 1), (2, 2, 1), (1, 2, 2), (2, 2, 2), (0, 2, 2), (2, 2, 0)]:
        g[k, j, i] = 4.0
    o, h = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    for y, expected in [(0, 0), (0.5, 0.5), (0.75, 0.75), (1, 1), (1.25, 1.75), (1.5, 2.5), (2, 4)]:
        assert orc.sample_array(g, o, h, (1.0, y, 1.0)) == expected
    # test/scalar_grid3d_test.cpp:74-86 -- origin .5, node (1,1,1) = 4
    g2 = np.ones((3, 3, 3))
    g2[1, 1, 1] = 4.0
    o2 = (0.5, 0.5, 0.5)
    for (i, j, k), e in [((0, 0, 0), 1), ((1, 0, 0), 1), ((0, 1, 0), 1), ((1, 1, 0), 1), ((0, 0, 1), 1), ((1, 1, 1), 4)]:
        assert orc.sample_array(g2, o2, h, (0.5 + i, 0.5 + j, 0.5 + k)) == e


def test_hash_grid_searcher_known_answers():
    # test/hash_grid_particle_searcher_test.cpp:5-28 (HasNearbyPoint with 7^3 buckets of width 2)
    pts = [(-0.1, -0.1, -0.1), (-0.1, 1, -0.5), (0.5, 0.7, 0.5), (1, 1, 1), (1.1, 1.1, 1.1), (1.5, 0.5, 1),
           (2, 1, 1), (2, 2, 1), (2, 2, 2), (5, 5, 5), (5, 5, 5.001), (3.1, 2.1, 5)]
    assert orc.hash_grid_query(pts, 7, 2.0, (3.5, 3.5, 3.5), 1.0) == []
    assert orc.hash_grid_query(pts, 7, 2.0, (5.5, 5.5, 5.5), 1.0) != []
    assert orc.hash_grid_query(pts, 7, 2.0, (1, 1, 1), 1.0) != []
    assert orc.hash_grid_query(pts, 7, 2.0, (1, 2, 3), 2.0) != []
    # test/hash_grid_particle_searcher_test.cpp:54-76 (ForEachNearbyPoint -> {0, 2})
    pts2 = [(0, 1, 3), (2, 5, 4), (-1, 3, 0)]
    assert orc.hash_grid_query(pts2, 4, 2.0 * math.sqrt(10), (0, 0, 0), math.sqrt(10.0)) == [0, 2]


def test_single_phase_pressure_solve():
    # test/single_phase_pressuer_solver_test.cpp:61-133: 4^3 all-fluid box, v = 1 on
    # interior y faces, default Jacobi (1000 sweeps, check every 5, |r| < 1e-14).
    s = orc.Scene((4, 4, 4), size=(4.0, 4.0, 4.0), method=orc.PIC)
    s.set(orc.SDF, np.full((4, 4, 4), -1.0))
    v = np.zeros((4, 5, 4))
    v[:, 1:3, :] = 1.0  # j in {1, 2}  (the test sets j != 0 && j != 3 for j < 4)
    v[:, 3, :] = 0.0
    s.set(orc.V, v)
    st = s.pressure(orc.solver(orc.JACOBI, 1000, 5, 1e-14))
    assert st.iterations <= 1000
    u, vv, w = s.get(orc.U), s.get(orc.V), s.get(orc.W)
    assert np.all(np.abs(u[:3, :3, :4]) < 1e-6)
    assert np.all(np.abs(vv[:3, :4, :3]) < 1e-6)
    assert np.all(np.abs(w[:4, :3, :3]) < 1e-6)
    p = s.get(orc.PRESSURE)
    d = p[:3, 1:3, :3] - p[:3, 0:2, :3]
    assert np.allclose(d, -1.0, atol=1e-6)
    flags = s.get(orc.FLAGS)
    assert np.all(flags == orc.FLUID)


def test_extrapolate_reference_case_by_hand():
    # The input of test/pic_simulator_test.cpp:78-123 (print-only there); expected
    # values computed by hand from the neighbour-mean rule of math_utils.hpp:186-253.
    vel = np.zeros((3, 3, 3))  # (z, y, x)
    valid = np.zeros((3, 3, 3), dtype=np.int32)
    for (i, j, k), val in [((0, 0, 0), 2.0), ((1, 1, 0), 8.0), ((1, 0, 0), 4.0)]:
        vel[k, j, i] = val
        valid[k, j, i] = 1
    out = orc.extrapolate_to_region(vel, valid, 1)
    expect = {(0, 1, 0): 5.0, (2, 0, 0): 4.0, (0, 0, 1): 2.0, (1, 0, 1): 4.0, (1, 1, 1): 8.0, (2, 1, 0): 8.0,
              (1, 2, 0): 8.0, (0, 2, 0): 0.0, (0, 0, 0): 2.0, (1, 1, 0): 8.0, (1, 0, 0): 4.0}
    for (i, j, k), e in expect.items():
        assert out[k, j, i] == e, (i, j, k)
    # second iteration grows from the new valid set: (0,2,0) sees (0,1,0)=5 and (1,2,0)=8
    out2 = orc.extrapolate_to_region(vel, valid, 2)
    assert out2[0, 2, 0] == pytest.approx((5.0 + 8.0) / 2)
