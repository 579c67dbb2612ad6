"""Host FP64 residual of the pressure system, written out from the reference's
BuildSystem rule; used by the tests to check the GPU's pressure independently
of its own residual kernel."""
import numpy as np

AIR, FLUID, BOUNDRY = 0, 1, 2


def host_relative_residual(flags, u, v, w, p, h):
    """||b - A p||_2 / ||b||_2 in FP64 with arrays indexed [k, j, i].

    FLUID rows: b = divergence of the face velocities (DivergenceAtCallCenter,
    src/grid_systems/face_centered_grid3d.cpp:157-166); the centre counts
    1/h^2 for every neighbour inside the grid that is not BOUNDRY, and every
    FLUID neighbour enters with -1/h^2.  Other rows are the identity with
    b = 0 (src/fluid_solvers/single_phase_pressure_solver.cpp:99-164).
    """
    fluid = flags == FLUID
    open_ = flags != BOUNDRY
    u, v, w, p = (a.astype(np.float64) for a in (u, v, w, p))
    div = ((u[:, :, 1:] - u[:, :, :-1]) / h[0] + (v[:, 1:, :] - v[:, :-1, :]) / h[1]
           + (w[1:, :, :] - w[:-1, :, :]) / h[2])
    b = np.where(fluid, div, 0.0)
    centre = np.zeros_like(p)
    off = np.zeros_like(p)
    for axis, hh in ((2, h[0]), (1, h[1]), (0, h[2])):
        ih2 = 1.0 / (hh * hh)
        n = p.shape[axis]
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[axis] = slice(0, n - 1)
        hi[axis] = slice(1, n)
        lo, hi = tuple(lo), tuple(hi)
        # + neighbour of the cells [0, n-1), - neighbour of the cells [1, n)
        centre[lo] += ih2 * open_[hi]
        centre[hi] += ih2 * open_[lo]
        off[lo] += ih2 * np.where(fluid[hi], p[hi], 0.0)
        off[hi] += ih2 * np.where(fluid[lo], p[lo], 0.0)
    ap = np.where(fluid, centre * p - off, p)
    r = b - ap
    return float(np.linalg.norm(r) / np.linalg.norm(b)), b
