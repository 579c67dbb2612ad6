"""CPU-side checks of the C ABI library: it loads, exports every entry point
include/fluide_hip.h declares, and its host-only seeding is deterministic with
the exact particle counts of the benchmark scenes.  No GPU compute is issued."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from fluide_engine_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "fluide_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fe_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = capi.load()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(capi.EXPORTED)
    assert b"gfx950" in L.fe_version()


def test_seeding_counts_match_benchmark_scenes():
    dam = [capi.box((-1, -1, -1), (0.4, 0.6, 2.0))]
    x32 = capi.seed_particles(capi.scene_desc((32, 32, 32)), dam, seed=0)
    assert x32.shape == (63232, 3)  # 13*19*32 cells x 8
    d128 = capi.scene_desc((128, 128, 128))
    x128 = capi.seed_particles(d128, dam, seed=0)
    assert x128.shape == (4021248, 3)  # 51*77*128 cells x 8
    assert x128.min() >= 0.0 and x128.max() <= 1.0
    assert (x128[:, 0] <= 0.4062500).all() and (x128[:, 1] <= 0.6015625).all()


def test_seeding_is_deterministic_and_jittered_per_subcell():
    d = capi.scene_desc((8, 8, 8))
    shapes = [capi.box((0.0, 0.0, 0.0), (0.5, 0.5, 0.5))]
    a = capi.seed_particles(d, shapes, seed=3)
    b = capi.seed_particles(d, shapes, seed=3)
    c = capi.seed_particles(d, shapes, seed=4)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    h = 1.0 / 8
    # 8 particles per cell, one in each half-cell octant
    cells = np.floor(a / h).astype(int)
    sub = np.floor((a - cells * h) / (h / 2)).astype(int)
    key = cells[:, 0] + 8 * (cells[:, 1] + 8 * cells[:, 2])
    for k in np.unique(key)[:20]:
        s = sub[key == k]
        assert len(s) == 8
        assert len({tuple(t) for t in s}) == 8


def test_sphere_seeding_and_invalid_args():
    d = capi.scene_desc((32, 32, 32))
    x = capi.seed_particles(d, [capi.sphere((0.5, 0.5, 0.5), 0.25)], seed=0)
    r = np.linalg.norm(x - 0.5, axis=1)
    assert r.max() < 0.25 + np.sqrt(3) / 32
    L = capi.load()
    cnt = C.c_int64()
    bad = capi.scene_desc((0, 4, 4))
    assert L.fe_seed_particles(C.byref(bad), None, 0, 0, None, 0, C.byref(cnt)) != 0


def test_context_without_device_reports_error():
    pytest.importorskip("torch")
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(capi.FluideError, match="no HIP device|gfx950"):
        capi.Context(0)
