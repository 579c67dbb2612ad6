"""The host FP64 residual used by the GPU pressure tests, checked against the
oracle's own BuildSystem and CG on small systems (CPU only)."""
import numpy as np

from oracle import oracle as orc
from pressure_residual import host_relative_residual


def test_host_residual_matches_oracle_system():
    """The numpy residual above against the oracle's own BuildSystem + CG on a
    small noisy system (CPU part of this file's checks)."""
    n = 12
    rng = np.random.default_rng(3)
    o = orc.Scene((n, n, n), method=orc.FLIP)
    K, J, I = np.meshgrid(*[(np.arange(n) + 0.5) / n] * 3, indexing="ij")
    sdf = np.where((I < 0.5) & (J < 0.6), -0.5 / n, 0.5 / n)
    sdf[rng.uniform(size=sdf.shape) < 0.05] *= -1
    o.set(orc.SDF, sdf)
    for what in (orc.U, orc.V, orc.W):
        o.set(what, rng.normal(size=o.shape(what)))
    u, v, w = o.get(orc.U), o.get(orc.V), o.get(orc.W)
    st = o.pressure(orc.solver(orc.CG, 400, 0, 1e-5, relative=True, warm_start=False))
    rel, b = host_relative_residual(o.get(orc.FLAGS), u, v, w, o.get(orc.PRESSURE), (1.0 / n,) * 3)
    assert np.array_equal(b, o.get(orc.RHS))
    assert abs(rel - st.rel_residual) <= 1e-9, (rel, st.rel_residual)
    # a wrong pressure is seen: the zero vector, and one with a flipped sign
    p = o.get(orc.PRESSURE)
    assert host_relative_residual(o.get(orc.FLAGS), u, v, w, np.zeros_like(p), (1.0 / n,) * 3)[0] == 1.0
    assert host_relative_residual(o.get(orc.FLAGS), u, v, w, -p, (1.0 / n,) * 3)[0] > 1.5
