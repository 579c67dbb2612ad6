"""TEST INFRASTRUCTURE: ctypes view of the FP64 CPU oracle (oracle/fe_oracle.cpp).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module; the product library never does.  Every scene method mirrors one stage
of the reference's PICSimulator / FLIPSimulator / APICSimulator (the citations
live next to each function in fe_oracle.cpp).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libfe_oracle.so")

PIC, FLIP, APIC = 0, 1, 2
AIR, FLUID, BOUNDRY = 0, 1, 2
U, V, W, MU, MV, MW, RU, RV, RW, SDF, FLAGS, PRESSURE, RHS, COLLIDER_SDF = range(14)
JACOBI, CG, PCG_JACOBI = 0, 1, 2
SURFACE_REFERENCE, SURFACE_CLOSED = 0, 1
# deliberate restatement mistakes for the pinning tests (fe_oracle.h ORC_PLANT_*)
PLANT_NONE, PLANT_APIC_TRANSPOSE, PLANT_FLIP_DELTA_SIGN, PLANT_EULER_ADVECTION, PLANT_APIC_FACE_DIMS = range(5)


class SolverDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("max_iterations", C.c_int32),
        ("check_interval", C.c_int32),
        ("relative", C.c_int32),
        ("warm_start", C.c_int32),
        ("pad", C.c_int32),
        ("tolerance", C.c_double),
    ]


class SolveStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("pad", C.c_int32),
        ("rel_residual", C.c_double),
        ("abs_residual", C.c_double),
        ("ms", C.c_double),
    ]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [ip, dp, dp, C.c_int, C.c_double, C.c_double, dp]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_threads.argtypes = [C.c_void_p, C.c_int]
        L.orc_plant_mistake.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_fast_dots.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_surface_mode.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_particles.argtypes = [C.c_void_p, dp, dp, dp, C.c_int64]
        L.orc_get_particles.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc_num_particles.argtypes = [C.c_void_p]
        L.orc_num_particles.restype = C.c_int64
        L.orc_grid_count.argtypes = [C.c_void_p, C.c_int]
        L.orc_grid_count.restype = C.c_int64
        L.orc_get_grid.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.orc_set_grid.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        for name in ("orc_p2g", "orc_build_sdf", "orc_build_markers", "orc_extrapolate_to_air", "orc_apply_bc",
                     "orc_build_system", "orc_apply_pressure", "orc_g2p"):
            getattr(L, name).argtypes = [C.c_void_p]
        L.orc_apply_gravity.argtypes = [C.c_void_p, C.c_double]
        L.orc_move_particles.argtypes = [C.c_void_p, C.c_double]
        L.orc_pressure.argtypes = [C.c_void_p, C.POINTER(SolverDesc), C.POINTER(SolveStats)]
        L.orc_substep.argtypes = [C.c_void_p, C.c_double, C.POINTER(SolverDesc), C.POINTER(SolveStats)]
        L.orc_cfl.argtypes = [C.c_void_p, C.c_double]
        L.orc_cfl.restype = C.c_double
        L.orc_num_substeps.argtypes = [C.c_void_p, C.c_double]
        L.orc_num_substeps.restype = C.c_uint
        L.orc_advance_frame.argtypes = [C.c_void_p, C.c_double, C.POINTER(SolverDesc)]
        L.orc_advance_frame.restype = C.c_int
        L.orc_last_stats.argtypes = [C.c_void_p, C.POINTER(SolveStats)]
        L.orc_barycentric.argtypes = [C.c_double, C.c_int, C.c_int, ip, dp]
        L.orc_coords_weights.argtypes = [ip, dp, dp, dp, ip, dp]
        L.orc_coords_grad_weights.argtypes = [ip, dp, dp, dp, ip, dp]
        L.orc_trilerp.argtypes = [dp, C.c_double, C.c_double, C.c_double]
        L.orc_trilerp.restype = C.c_double
        L.orc_sample_array.argtypes = [dp, ip, dp, dp, dp]
        L.orc_sample_array.restype = C.c_double
        L.orc_extrapolate_to_region.argtypes = [dp, ip, ip, C.c_int, dp]
        L.orc_hash_grid_query.argtypes = [dp, C.c_int64, C.c_int, C.c_double, dp, C.c_double,
                                          C.POINTER(C.c_int64), C.c_int]
        L.orc_hash_grid_query.restype = C.c_int
        L.orc_blas_dot.argtypes = [dp, dp, ip]
        L.orc_blas_dot.restype = C.c_double
        L.orc_solve_rows.argtypes = [dp, dp, dp, dp, dp, dp, ip, C.POINTER(SolverDesc), C.POINTER(SolveStats)]
        L.orc_residual_rows.argtypes = [dp, dp, dp, dp, dp, dp, ip, dp]
        L.orc_spmv_rows.argtypes = [dp, dp, dp, dp, dp, ip, dp]
        L.orc_set_viscosity.argtypes = [C.c_void_p, C.c_double]
        L.orc_set_diffusion_solver.argtypes = [C.c_void_p, C.POINTER(SolverDesc)]
        L.orc_diffusion.argtypes = [C.c_void_p, C.c_double, C.POINTER(SolveStats)]
        L.orc_diffusion_markers.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.orc_surface.argtypes = [C.c_void_p, C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.POINTER(C.c_double)),
                                  C.POINTER(C.POINTER(C.c_int32)), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.orc_free.argtypes = [C.c_void_p]
        i32p = C.POINTER(C.c_int32)
        L.orc_mesh_to_sdf.argtypes = [ip, dp, dp, dp, C.c_int64, i32p, C.c_int64, dp]
        L.orc_add_mesh_collider.argtypes = [C.c_void_p, ip, dp, dp, dp, C.c_int64, i32p, C.c_int64]
        L.orc_closest_on_triangle.argtypes = [dp, dp, dp, dp, dp]
        L.orc_mesh_inside.argtypes = [dp, C.c_int64, i32p, C.c_int64, dp]
        L.orc_add_volume_emitter.argtypes = [C.c_void_p, ip, dp, dp, dp, C.c_int64, i32p, C.c_int64, C.c_int, dp,
                                             C.c_uint64]
        L.orc_initialize_from_mesh.argtypes = [C.c_void_p, dp, C.c_int64, i32p, C.c_int64, C.c_int, C.c_uint64]
        L.orc_uniform01.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int]
        L.orc_uniform01.restype = C.c_double
        L.orc_emit.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def solver(kind=JACOBI, max_iterations=1000, check_interval=5, tolerance=1e-14, relative=False, warm_start=True):
    return SolverDesc(kind, max_iterations, check_interval, int(relative), int(warm_start), 0, tolerance)


class Scene:
    """One oracle scene; arrays are FP64, x-fastest, shapes returned as (nz, ny, nx)."""

    def __init__(self, n, origin=(0.0, 0.0, 0.0), size=(1.0, 1.0, 1.0), method=FLIP, blend=0.05,
                 max_cfl=3.0, gravity=(0.0, -9.81, 0.0), threads=1):
        L = lib()
        self.n = tuple(int(x) for x in n)
        nn = (C.c_int * 3)(*self.n)
        o = (C.c_double * 3)(*origin)
        s = (C.c_double * 3)(*size)
        g = (C.c_double * 3)(*gravity)
        self.h = tuple(size[a] / self.n[a] for a in range(3))
        self.origin, self.size, self.method = tuple(origin), tuple(size), method
        self._s = L.orc_create(nn, o, s, method, blend, max_cfl, g)
        L.orc_set_threads(self._s, threads)

    def __del__(self):
        if getattr(self, "_s", None):
            lib().orc_destroy(self._s)
            self._s = None

    def shape(self, what):
        nx, ny, nz = self.n
        if what in (U, MU, RU):
            return (nz, ny, nx + 1)
        if what in (V, MV, RV):
            return (nz, ny + 1, nx)
        if what in (W, MW, RW):
            return (nz + 1, ny, nx)
        return (nz, ny, nx)

    def set_threads(self, t):
        lib().orc_set_threads(self._s, t)

    def set_fast_dots(self, on=True):
        """Solver dot products in storage order over fixed blocks (tests on large grids)."""
        lib().orc_set_fast_dots(self._s, 1 if on else 0)

    def set_surface_mode(self, mode):
        """SURFACE_REFERENCE (the reference's case tables, the default) or SURFACE_CLOSED."""
        lib().orc_set_surface_mode(self._s, mode)

    def plant_mistake(self, which):
        """Pinning tests only: switch on one deliberate restatement mistake (PLANT_*)."""
        lib().orc_plant_mistake(self._s, which)

    def set_particles(self, x, v, Cmat=None):
        x = _f64(x).reshape(-1, 3)
        v = _f64(v).reshape(-1, 3)
        cp = None
        if Cmat is not None:
            Cmat = _f64(Cmat).reshape(-1, 9)
            cp = _dp(Cmat)
        lib().orc_set_particles(self._s, _dp(x), _dp(v), cp, x.shape[0])

    def particles(self):
        n = lib().orc_num_particles(self._s)
        x = np.empty((n, 3))
        v = np.empty((n, 3))
        Cm = np.empty((n, 9))
        lib().orc_get_particles(self._s, _dp(x), _dp(v), _dp(Cm))
        return x, v, Cm

    def get(self, what):
        shp = self.shape(what)
        if what in (MU, MV, MW):
            out = np.empty(shp, dtype=np.int32)
        elif what == FLAGS:
            out = np.empty(shp, dtype=np.uint8)
        else:
            out = np.empty(shp, dtype=np.float64)
        assert lib().orc_get_grid(self._s, what, out.ctypes.data_as(C.c_void_p)) == 0
        return out

    def set(self, what, arr):
        shp = self.shape(what)
        if what in (MU, MV, MW):
            a = np.ascontiguousarray(arr, dtype=np.int32).reshape(shp)
        elif what == FLAGS:
            a = np.ascontiguousarray(arr, dtype=np.uint8).reshape(shp)
        else:
            a = np.ascontiguousarray(arr, dtype=np.float64).reshape(shp)
        assert lib().orc_set_grid(self._s, what, a.ctypes.data_as(C.c_void_p)) == 0

    # stages (each restates one reference function; see fe_oracle.cpp)
    def p2g(self):
        lib().orc_p2g(self._s)

    def build_sdf(self):
        lib().orc_build_sdf(self._s)

    def build_markers(self):
        lib().orc_build_markers(self._s)

    def extrapolate_to_air(self):
        lib().orc_extrapolate_to_air(self._s)

    def apply_bc(self):
        lib().orc_apply_bc(self._s)

    def apply_gravity(self, dt):
        lib().orc_apply_gravity(self._s, dt)

    def build_system(self):
        lib().orc_build_system(self._s)

    def apply_pressure(self):
        lib().orc_apply_pressure(self._s)

    def g2p(self):
        lib().orc_g2p(self._s)

    def move_particles(self, dt):
        lib().orc_move_particles(self._s, dt)

    def pressure(self, desc):
        st = SolveStats()
        lib().orc_pressure(self._s, C.byref(desc), C.byref(st))
        return st

    def substep(self, dt, desc):
        st = SolveStats()
        lib().orc_substep(self._s, dt, C.byref(desc), C.byref(st))
        return st

    def set_viscosity(self, nu):
        lib().orc_set_viscosity(self._s, nu)

    def set_diffusion_solver(self, desc):
        lib().orc_set_diffusion_solver(self._s, C.byref(desc))

    def diffusion(self, dt):
        """BackwardEulerDiffusionSolver::Solve on u, v, w in place; stats per component."""
        st = (SolveStats * 3)()
        lib().orc_diffusion(self._s, dt, st)
        return list(st)

    def diffusion_markers(self, c):
        out = np.empty(self.shape((U, V, W)[c]), dtype=np.uint8)
        lib().orc_diffusion_markers(self._s, c, out.ctypes.data_as(C.c_void_p))
        return out

    def add_mesh_collider(self, size, origin, h, vertices, triangles):
        """AddCollider(TriangleMeshCollider(size, origin, h, mesh)); size is (nx, ny, nz)."""
        v = _f64(np.asarray(vertices).reshape(-1, 3))
        t = np.ascontiguousarray(np.asarray(triangles).reshape(-1, 3), dtype=np.int32)
        sz = (C.c_int * 3)(*[int(q) for q in size])
        o = (C.c_double * 3)(*origin)
        hh = (C.c_double * 3)(*h)
        rc = lib().orc_add_mesh_collider(self._s, sz, o, hh, _dp(v), v.shape[0],
                                         t.ctypes.data_as(C.POINTER(C.c_int32)), t.shape[0])
        if rc != 0:
            raise ValueError("triangle vertex index out of range")

    def add_volume_emitter(self, size, origin, h, vertices, triangles, particles_per_block, velocity, seed):
        """AddEmitter(VolumeEmitter(mesh, resolution, domain, particlesPerBlock, velocity, 0)); the emitter
        grid (size, origin, h) must have the scene's size."""
        v = _f64(np.asarray(vertices).reshape(-1, 3))
        t = np.ascontiguousarray(np.asarray(triangles).reshape(-1, 3), dtype=np.int32)
        rc = lib().orc_add_volume_emitter(self._s, (C.c_int * 3)(*[int(q) for q in size]), (C.c_double * 3)(*origin),
                                          (C.c_double * 3)(*h), _dp(v), v.shape[0],
                                          t.ctypes.data_as(C.POINTER(C.c_int32)), t.shape[0],
                                          int(particles_per_block), (C.c_double * 3)(*velocity), int(seed))
        if rc != 0:
            raise ValueError("bad emitter (vertex index out of range or grid size differs from the scene's)")

    def emit(self):
        lib().orc_emit(self._s)

    def initialize_from_mesh(self, vertices, triangles, particles_per_block, seed):
        """InitializeFromTriangleMesh: Mesh2SDF into the fluid SDF, then InitializeParticles (appends)."""
        v = _f64(np.asarray(vertices).reshape(-1, 3))
        t = np.ascontiguousarray(np.asarray(triangles).reshape(-1, 3), dtype=np.int32)
        rc = lib().orc_initialize_from_mesh(self._s, _dp(v), v.shape[0], t.ctypes.data_as(C.POINTER(C.c_int32)),
                                            t.shape[0], int(particles_per_block), int(seed))
        if rc != 0:
            raise ValueError("triangle vertex index out of range")

    def surface(self):
        """MarchingCubesSolver::BuildSurface: (vertices (V,3), normals (V,3), triangles (T,3))."""
        L = lib()
        vp, np_, tp = C.POINTER(C.c_double)(), C.POINTER(C.c_double)(), C.POINTER(C.c_int32)()
        nv, nt = C.c_int64(), C.c_int64()
        L.orc_surface(self._s, C.byref(vp), C.byref(np_), C.byref(tp), C.byref(nv), C.byref(nt))
        try:
            v = np.ctypeslib.as_array(vp, shape=(max(nv.value, 1) * 3,))[: nv.value * 3].reshape(-1, 3).copy()
            n = np.ctypeslib.as_array(np_, shape=(max(nv.value, 1) * 3,))[: nv.value * 3].reshape(-1, 3).copy()
            t = np.ctypeslib.as_array(tp, shape=(max(nt.value, 1) * 3,))[: nt.value * 3].reshape(-1, 3).copy()
        finally:
            for ptr in (vp, np_, tp):
                L.orc_free(C.cast(ptr, C.c_void_p))
        return v, n, t

    def cfl(self, dt):
        return lib().orc_cfl(self._s, dt)

    def advance_frame(self, frame_dt, desc):
        return lib().orc_advance_frame(self._s, frame_dt, C.byref(desc))


# ---- primitives -----------------------------------------------------------

def barycentric(x, i_low, i_high):
    i = C.c_int()
    f = C.c_double()
    lib().orc_barycentric(x, i_low, i_high, C.byref(i), C.byref(f))
    return i.value, f.value


def coords_weights(size, origin, h, x):
    idx = np.zeros(24, dtype=np.int32)
    w = np.zeros(8)
    lib().orc_coords_weights(_ip(np.array(size, dtype=np.int32)), _dp(_f64(origin)), _dp(_f64(h)), _dp(_f64(x)),
                             _ip(idx), _dp(w))
    return idx.reshape(8, 3), w


def coords_grad_weights(size, origin, h, x):
    idx = np.zeros(24, dtype=np.int32)
    g = np.zeros(24)
    lib().orc_coords_grad_weights(_ip(np.array(size, dtype=np.int32)), _dp(_f64(origin)), _dp(_f64(h)),
                                  _dp(_f64(x)), _ip(idx), _dp(g))
    return idx.reshape(8, 3), g.reshape(8, 3)


def trilerp(vals, fx, fy, fz):
    return lib().orc_trilerp(_dp(_f64(vals)), fx, fy, fz)


def sample_array(data_zyx, origin, h, p):
    d = _f64(data_zyx)
    dims = np.array([d.shape[2], d.shape[1], d.shape[0]], dtype=np.int32)
    return lib().orc_sample_array(_dp(d), _ip(dims), _dp(_f64(origin)), _dp(_f64(h)), _dp(_f64(p)))


def extrapolate_to_region(inp_zyx, valid_zyx, iterations):
    d = _f64(inp_zyx)
    va = np.ascontiguousarray(valid_zyx, dtype=np.int32)
    dims = np.array([d.shape[2], d.shape[1], d.shape[0]], dtype=np.int32)
    out = np.empty_like(d)
    lib().orc_extrapolate_to_region(_dp(d), _ip(va), _ip(dims), iterations, _dp(out))
    return out


def hash_grid_query(points, buckets, spacing, q, radius):
    pts = _f64(points).reshape(-1, 3)
    out = np.zeros(max(1, pts.shape[0]), dtype=np.int64)
    cnt = lib().orc_hash_grid_query(_dp(pts), pts.shape[0], buckets, spacing, _dp(_f64(q)), radius,
                                    out.ctypes.data_as(C.POINTER(C.c_int64)), out.shape[0])
    return sorted(out[:cnt].tolist())


def blas_dot(a, b):
    a = _f64(a)
    b = _f64(b)
    dims = np.array([a.shape[2], a.shape[1], a.shape[0]], dtype=np.int32)
    return lib().orc_blas_dot(_dp(a), _dp(b), _ip(dims))


def solve_rows(Ac, Ar, Au, Af, b, x0, desc):
    arrs = [_f64(a) for a in (Ac, Ar, Au, Af, b)]
    x = _f64(x0).copy()
    dims = np.array([x.shape[2], x.shape[1], x.shape[0]], dtype=np.int32)
    st = SolveStats()
    lib().orc_solve_rows(*[_dp(a) for a in arrs], _dp(x), _ip(dims), C.byref(desc), C.byref(st))
    return x, st


def residual_rows(Ac, Ar, Au, Af, b, x):
    arrs = [_f64(a) for a in (Ac, Ar, Au, Af, b, x)]
    dims = np.array([arrs[0].shape[2], arrs[0].shape[1], arrs[0].shape[0]], dtype=np.int32)
    out = np.empty_like(arrs[0])
    lib().orc_residual_rows(*[_dp(a) for a in arrs], _ip(dims), _dp(out))
    return out


def spmv_rows(Ac, Ar, Au, Af, v):
    arrs = [_f64(a) for a in (Ac, Ar, Au, Af, v)]
    dims = np.array([arrs[0].shape[2], arrs[0].shape[1], arrs[0].shape[0]], dtype=np.int32)
    out = np.empty_like(arrs[0])
    lib().orc_spmv_rows(*[_dp(a) for a in arrs], _ip(dims), _dp(out))
    return out


def mesh_to_sdf(size, origin, h, vertices, triangles):
    """Mesh2SDF::Build on a node grid of size (nx, ny, nz): array (nz, ny, nx)."""
    v = _f64(np.asarray(vertices).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(triangles).reshape(-1, 3), dtype=np.int32)
    out = np.empty((size[2], size[1], size[0]), np.float64)
    sz = (C.c_int * 3)(*[int(q) for q in size])
    rc = lib().orc_mesh_to_sdf(sz, (C.c_double * 3)(*origin), (C.c_double * 3)(*h), _dp(v), v.shape[0],
                               t.ctypes.data_as(C.POINTER(C.c_int32)), t.shape[0], _dp(out))
    if rc != 0:
        raise ValueError("triangle vertex index out of range")
    return out


def closest_on_triangle(x, p0, p1, p2):
    """Collisions::ClossestPointOnTriangle."""
    out = np.empty(3)
    args = [_f64(np.asarray(a, dtype=np.float64)) for a in (x, p0, p1, p2)]
    lib().orc_closest_on_triangle(*[_dp(a) for a in args], _dp(out))
    return out


def mesh_inside(vertices, triangles, x):
    """TriangleMesh::IsInside."""
    v = _f64(np.asarray(vertices).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(triangles).reshape(-1, 3), dtype=np.int32)
    xx = _f64(np.asarray(x, dtype=np.float64))
    return bool(lib().orc_mesh_inside(_dp(v), v.shape[0], t.ctypes.data_as(C.POINTER(C.c_int32)), t.shape[0],
                                      _dp(xx)))


def uniform01(seed, call, n, axis):
    """The generator standing in for std::rand() / RAND_MAX in particle seeding."""
    return lib().orc_uniform01(int(seed), int(call), int(n), int(axis))
