"""fe_mesh_sdf_noalloc (no device allocation: its buffers live in the device's
static scratch array) against fe_mesh_sdf through the C ABI: bit-identical on
two calls in a row, the second with a smaller mesh and grid (the scratch keeps
the first call's contents), plus its size and argument errors."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from fluide_engine_amd import capi  # noqa: E402
from meshes import box, icosphere  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    print(f"\n[gpu] device {c.device_name}", flush=True)
    yield c
    c.close()


def test_noalloc_matches_allocating_entry_point(ctx):
    calls = [
        # 40^3 grid, an icosphere of 320 triangles
        ((40, 40, 40), (0.0, 0.0, 0.0), (1.0 / 39,) * 3, icosphere((0.45, 0.5, 0.55), 0.3, levels=2)),
        # then a smaller anisotropic grid and a box of 12 triangles
        ((17, 12, 9), (-0.1, 0.05, 0.0), (0.07, 0.09, 0.11), box((0.2, 0.25, 0.1), (0.8, 0.7, 0.6))),
    ]
    for size, origin, h, (v, t) in calls:
        print(f"[gpu] mesh SDF no-alloc: grid {size}, {len(t)} triangles", flush=True)
        ref = capi.mesh_sdf(ctx, size, origin, h, v, t)
        got = capi.mesh_sdf_noalloc(ctx, size, origin, h, v, t)
        assert got.shape == ref.shape
        assert np.array_equal(got, ref)
        assert (got < 0).any() and (got > 0).any()


def test_noalloc_limits_and_errors(ctx):
    v, t = box((0.2, 0.2, 0.2), (0.8, 0.8, 0.8))
    with pytest.raises(capi.FluideError, match="at most"):
        capi.mesh_sdf_noalloc(ctx, (129, 128, 128), (0, 0, 0), (0.01, 0.01, 0.01), v, t)  # > 2^21 nodes
    big = np.zeros((65537, 3), np.int32)
    with pytest.raises(capi.FluideError, match="at most"):
        capi.mesh_sdf_noalloc(ctx, (8, 8, 8), (0, 0, 0), (0.1, 0.1, 0.1), v, big)
    with pytest.raises(capi.FluideError, match="out of range"):  # detected on the device, no access out of bounds
        capi.mesh_sdf_noalloc(ctx, (8, 8, 8), (0, 0, 0), (0.1, 0.1, 0.1), v, t + 100)
    # the scratch is usable after an error
    ref = capi.mesh_sdf(ctx, (8, 8, 8), (0, 0, 0), (0.125, 0.125, 0.125), v, t)
    assert np.array_equal(capi.mesh_sdf_noalloc(ctx, (8, 8, 8), (0, 0, 0), (0.125, 0.125, 0.125), v, t), ref)
