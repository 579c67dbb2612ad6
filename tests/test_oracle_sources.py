"""Pins the oracle's scene sources: PICSimulator::InitializeFromTriangleMesh +
InitializeParticles (src/pic_simulator.cpp:33-39,521-548) and
VolumeEmitter::Emitt (src/emmiter/volume_emmiter.cpp:58-85).

  * the generator that stands in for std::rand is SplitMix64 (published test
    vector: 0xE220A8397B1DCDAF is the first output from state 0), checked
    against a plain-Python restatement and for uniformity;
  * InitializeFromTriangleMesh: the fluid SDF is Mesh2SDF of the mesh on the
    cell-centred grid; a cell gets particlesPerBlock particles iff that SDF
    sampled at its corner is negative, each inside its cell, at rest;
  * Emitt: nothing in the first substep (the collider SDF still holds its
    initial 0), afterwards particlesPerBlock particles with the start
    velocity in every emitter node with emitterSdf <= 0, fluidSdf >= 0 and
    colliderSdf > 0; a collider over the emitter stops it.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from meshes import box

M64 = (1 << 64) - 1


def splitmix64(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def uniform01_py(seed, call, n, axis):
    r = splitmix64(splitmix64(splitmix64(seed) ^ call) ^ (3 * n + axis))
    return (r >> 11) * 2.0 ** -53


def test_generator():
    assert splitmix64(0) == 0xE220A8397B1DCDAF  # first output of SplitMix64 from state 0
    for seed, call, n, axis in [(0, 0, 0, 0), (7, 3, 12345, 2), (2 ** 63 + 5, 99, 10 ** 9, 1)]:
        assert orc.uniform01(seed, call, n, axis) == uniform01_py(seed, call, n, axis)
    u = np.array([orc.uniform01(11, 0, n, a) for n in range(4000) for a in range(3)])
    assert np.all((u >= 0) & (u < 1))
    hist, _ = np.histogram(u, bins=10, range=(0, 1))
    chi2 = float(((hist - u.size / 10) ** 2 / (u.size / 10)).sum())
    assert chi2 < 30  # 9 degrees of freedom: p < 1e-3 beyond ~27.9


def test_initialize_from_mesh():
    n = (12, 10, 14)
    s = orc.Scene(n)
    h = tuple(1.0 / q for q in n)
    lo, hi = (0.23, 0.31, 0.18), (0.61, 0.72, 0.55)
    v, t = box(lo, hi)
    s.initialize_from_mesh(v, t, 8, seed=5)
    so = tuple(0.5 * q for q in h)
    assert np.array_equal(s.get(orc.SDF), orc.mesh_to_sdf(n, so, h, v, t))
    sdf = s.get(orc.SDF)
    picked = []
    for i in range(n[0]):
        for j in range(n[1]):
            for k in range(n[2]):
                if orc.sample_array(sdf, so, h, (i * h[0], j * h[1], k * h[2])) < 0:
                    picked.append((i, j, k))
    x, vel, _ = s.particles()
    assert x.shape[0] == 8 * len(picked) > 0
    assert np.all(vel == 0)
    cells = np.repeat(np.array(picked), 8, axis=0)
    lo_c = cells * np.array(h)
    assert np.all((x >= lo_c) & (x <= lo_c + np.array(h)))
    # the picked corners lie inside the box, up to the interpolation band of one cell
    corners = np.array(picked) * np.array(h)
    assert np.all((corners > np.array(lo) - np.array(h)) & (corners < np.array(hi) + np.array(h)))
    # appended, not replaced
    s.initialize_from_mesh(v, t, 2, seed=6)
    assert s.particles()[0].shape[0] == 10 * len(picked)


def test_emitter_first_substep_and_conditions():
    n = (10, 10, 10)
    h = (0.1, 0.1, 0.1)
    s = orc.Scene(n, gravity=(0.0, 0.0, 0.0))
    v, t = box((0.22, 0.61, 0.33), (0.48, 0.87, 0.57))
    s.add_volume_emitter(n, (0.0, 0.0, 0.0), h, v, t, 4, (0.0, 6.0, 0.0), seed=1)
    s.substep(1.0 / 240, orc.solver(orc.JACOBI, 10, 5, 1e-14))
    assert s.particles()[0].shape[0] == 0  # Emitt ran before BuildCollider: colliderSdf == 0
    s.emit()
    x, vel, _ = s.particles()
    esdf = orc.mesh_to_sdf(n, (0.0, 0.0, 0.0), h, v, t)
    fsdf = s.get(orc.SDF)
    want = [(i, j, k) for i in range(10) for j in range(10) for k in range(10)
            if esdf[k, j, i] <= 0 and fsdf[k, j, i] >= 0]
    assert x.shape[0] == 4 * len(want) > 0
    assert np.all(vel == np.array([0.0, 6.0, 0.0]))
    cells = np.repeat(np.array(want), 4, axis=0) * 0.1
    assert np.all((x >= cells) & (x <= cells + 0.1))
    # positions: h * u + pos with the generator keyed by (seed, call 1, particle, axis)
    assert x[5, 2] == 0.1 * orc.uniform01(1, 1, 5, 2) + cells[5, 2]
    with pytest.raises(ValueError):
        s.add_volume_emitter((9, 10, 10), (0.0, 0.0, 0.0), h, v, t, 4, (0.0, 0.0, 0.0), seed=1)


def test_emitter_blocked_by_collider():
    n = (10, 10, 10)
    h = (0.1, 0.1, 0.1)
    s = orc.Scene(n, gravity=(0.0, 0.0, 0.0))
    v, t = box((0.22, 0.61, 0.33), (0.48, 0.87, 0.57))
    s.add_volume_emitter(n, (0.0, 0.0, 0.0), h, v, t, 4, (0.0, 0.0, 0.0), seed=1)
    s.add_mesh_collider(n, (0.0, 0.0, 0.0), h, *box((0.1, 0.5, 0.2), (0.6, 0.95, 0.7)))
    s.substep(1.0 / 240, orc.solver(orc.JACOBI, 10, 5, 1e-14))
    s.emit()
    assert s.particles()[0].shape[0] == 0
