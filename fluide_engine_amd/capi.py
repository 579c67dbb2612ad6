"""ctypes binding of the C ABI in include/fluide_hip.h (libfluide_hip.so).

The library is the only compute path: if it is missing or no gfx950 device is
present, every entry point raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfluide_hip.so")

PIC, FLIP, APIC = 0, 1, 2
AIR, FLUID, BOUNDRY = 0, 1, 2
E_INVALID, E_HIP, E_NOMEM, E_UNSUPPORTED, E_NO_DEVICE, E_NONFINITE = -1, -2, -3, -4, -5, -6
SOLVER_JACOBI, SOLVER_CG, SOLVER_PCG_JACOBI, SOLVER_MGPCG = 0, 1, 2, 3
SDF_SEARCH_CELLS, SDF_SEARCH_BUCKETS = 0, 1
WARM_START_ZERO, WARM_START_PREVIOUS, WARM_START_LEARNED = 0, 1, 2
NET_SCALE_RESIDUAL, NET_SCALE_ENERGY = 0, 1
NET_LOSS_MSE, NET_LOSS_ENERGY = 0, 1
P2G_TILES, P2G_FACES, P2G_SCATTER = 0, 1, 2
G2P_QUADS, G2P_PAIRS = 0, 1
MG_FUSED, MG_SPLIT = 0, 1
SURFACE_REFERENCE, SURFACE_CLOSED = 0, 1
(GRID_U, GRID_V, GRID_W, GRID_MU, GRID_MV, GRID_MW, GRID_SU, GRID_SV, GRID_SW, GRID_SDF, GRID_FLAGS,
 GRID_PRESSURE, GRID_RHS, GRID_DMU, GRID_DMV, GRID_DMW, GRID_COLLIDER_SDF) = range(17)
(STAGE_SORT, STAGE_P2G, STAGE_SDF, STAGE_EXTRAPOLATE, STAGE_BC, STAGE_GRAVITY, STAGE_PRESSURE, STAGE_G2P,
 STAGE_MOVE, STAGE_MARKERS, STAGE_DIFFUSION, STAGE_EMIT) = range(12)
STAGE_NAMES = ["sort", "p2g", "sdf", "extrapolate", "bc", "gravity", "pressure", "g2p_move", "move", "markers",
               "diffusion", "emit"]

EXPORTED = [
    "fe_ctx_create", "fe_ctx_destroy", "fe_last_error", "fe_ctx_device_name", "fe_scene_create", "fe_scene_destroy",
    "fe_scene_sync", "fe_particles_upload", "fe_particles_download", "fe_particles_count", "fe_substep",
    "fe_advance_frame", "fe_cfl", "fe_run_stage", "fe_grid_download", "fe_grid_upload", "fe_grid_bytes",
    "fe_particle_cells", "fe_seed_particles", "fe_stage_times", "fe_version", "fe_kernel_timer",
    "fe_kernel_timer_read", "fe_last_solve_stats", "fe_copy_probe", "fe_scene_set_params",
    "fe_scene_set_diffusion_solver", "fe_last_diffusion_stats", "fe_batch_create", "fe_batch_destroy",
    "fe_batch_advance_frame", "fe_surface_build", "fe_surface_download", "fe_scene_clear_colliders",
    "fe_scene_set_sdf_search", "fe_scene_add_mesh_collider", "fe_mesh_sdf", "fe_scene_initialize_from_mesh",
    "fe_scene_add_volume_emitter", "fe_scene_clear_emitters", "fe_scene_reserve_particles",
    "fe_scene_particle_capacity", "fe_scene_add_volume_emitter_sdf", "fe_net_create", "fe_net_destroy",
    "fe_net_param_count", "fe_net_get_params", "fe_net_set_params", "fe_net_add_sample", "fe_net_sample_count",
    "fe_net_clear_samples", "fe_net_train", "fe_net_loss_and_grads", "fe_scene_set_warm_start_net",
    "fe_scene_warm_start_scale", "fe_net_set_scale_mode", "fe_net_set_loss", "fe_scene_get_tuning",
    "fe_scene_set_tuning", "fe_scene_autotune", "fe_mesh_sdf_noalloc", "fe_scene_set_surface_mode",
    "fe_scene_set_graph", "fe_scene_graph_stats",
]
(KERNEL_NONE, KERNEL_CG_SPMV, KERNEL_CG_UPDATE, KERNEL_P2G, KERNEL_G2P_MOVE, KERNEL_SDF, KERNEL_JACOBI,
 KERNEL_MG_SMOOTH, KERNEL_MG_DOWN, KERNEL_MG_UP) = range(10)


class SolverDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("max_iterations", C.c_int32), ("check_interval", C.c_int32),
                ("relative", C.c_int32), ("warm_start", C.c_int32), ("reserved", C.c_int32),
                ("tolerance", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("p2g", C.c_int32), ("g2p", C.c_int32), ("mg", C.c_int32)]


class TuningTimes(C.Structure):
    _fields_ = [("p2g_ms", C.c_double * 3), ("g2p_ms", C.c_double * 2)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("rel_residual", C.c_double),
                ("abs_residual", C.c_double), ("ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SceneDesc(C.Structure):
    _fields_ = [("n", C.c_int32 * 3), ("method", C.c_int32), ("origin", C.c_double * 3), ("size", C.c_double * 3),
                ("blend", C.c_double), ("max_cfl", C.c_double), ("gravity", C.c_double * 3),
                ("viscosity", C.c_double), ("density", C.c_double), ("max_particles", C.c_int64)]


class SeedShape(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("lo", C.c_double * 3), ("hi", C.c_double * 3)]


class FluideError(RuntimeError):
    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


_lib = None


def load():
    """Load libfluide_hip.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FluideError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "(make -C fluide_engine_amd/csrc); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.fe_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.fe_ctx_destroy.argtypes = [vp]
    L.fe_last_error.argtypes = [vp]
    L.fe_last_error.restype = C.c_char_p
    L.fe_ctx_device_name.argtypes = [vp, C.c_char_p, C.c_size_t]
    L.fe_scene_create.argtypes = [vp, C.POINTER(SceneDesc), C.POINTER(vp)]
    L.fe_scene_destroy.argtypes = [vp]
    L.fe_scene_sync.argtypes = [vp]
    fp = C.POINTER(C.c_float)
    L.fe_particles_upload.argtypes = [vp, fp, fp, fp, C.c_int64]
    L.fe_particles_download.argtypes = [vp, fp, fp, fp, C.c_int64]
    L.fe_particles_count.argtypes = [vp, C.POINTER(C.c_int64)]
    L.fe_substep.argtypes = [vp, C.c_double, C.POINTER(SolverDesc), C.POINTER(SolveStats)]
    L.fe_advance_frame.argtypes = [vp, C.c_double, C.POINTER(SolverDesc), C.POINTER(C.c_int32)]
    L.fe_cfl.argtypes = [vp, C.c_double, C.POINTER(C.c_double)]
    L.fe_run_stage.argtypes = [vp, C.c_int32, C.c_double, C.POINTER(SolverDesc), C.POINTER(SolveStats)]
    L.fe_grid_download.argtypes = [vp, C.c_int32, vp, C.c_size_t]
    L.fe_grid_upload.argtypes = [vp, C.c_int32, vp, C.c_size_t]
    L.fe_grid_bytes.argtypes = [vp, C.c_int32, C.POINTER(C.c_size_t)]
    L.fe_particle_cells.argtypes = [vp, C.POINTER(C.c_int32), C.c_int64]
    L.fe_seed_particles.argtypes = [C.POINTER(SceneDesc), C.POINTER(SeedShape), C.c_int32, C.c_uint64, fp,
                                    C.c_int64, C.POINTER(C.c_int64)]
    L.fe_stage_times.argtypes = [vp, C.POINTER(C.c_double), C.c_int32]
    L.fe_kernel_timer.argtypes = [vp, C.c_int32]
    L.fe_last_solve_stats.argtypes = [vp, C.POINTER(SolveStats)]
    L.fe_kernel_timer_read.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.c_int32]
    L.fe_copy_probe.argtypes = [vp, C.c_int64, C.POINTER(C.c_double)]
    L.fe_scene_set_params.argtypes = [vp, C.c_double, C.c_double, C.POINTER(C.c_double), C.c_double, C.c_double]
    L.fe_scene_set_diffusion_solver.argtypes = [vp, C.POINTER(SolverDesc)]
    L.fe_last_diffusion_stats.argtypes = [vp, C.POINTER(SolveStats)]
    L.fe_batch_create.argtypes = [vp, C.POINTER(vp), C.c_int32, C.c_int32, C.POINTER(vp)]
    L.fe_batch_destroy.argtypes = [vp]
    L.fe_surface_build.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.fe_scene_clear_colliders.argtypes = [vp]
    L.fe_scene_set_sdf_search.argtypes = [vp, C.c_int32]
    L.fe_scene_set_graph.argtypes = [vp, C.c_int32]
    L.fe_scene_graph_stats.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    i32p = C.POINTER(C.c_int32)
    dp = C.POINTER(C.c_double)
    L.fe_scene_add_mesh_collider.argtypes = [vp, i32p, dp, dp, dp, C.c_int64, i32p, C.c_int64]
    L.fe_mesh_sdf.argtypes = [vp, i32p, dp, dp, dp, C.c_int64, i32p, C.c_int64, dp]
    L.fe_mesh_sdf_noalloc.argtypes = [vp, i32p, dp, dp, dp, C.c_int64, i32p, C.c_int64, dp]
    L.fe_scene_initialize_from_mesh.argtypes = [vp, dp, C.c_int64, i32p, C.c_int64, C.c_int32, C.c_uint64]
    L.fe_scene_add_volume_emitter.argtypes = [vp, i32p, dp, dp, dp, C.c_int64, i32p, C.c_int64, C.c_int32, dp,
                                              C.c_uint64]
    L.fe_scene_clear_emitters.argtypes = [vp]
    L.fe_scene_add_volume_emitter_sdf.argtypes = [vp, i32p, dp, dp, dp, C.c_int32, dp, C.c_uint64]
    fp = C.POINTER(C.c_float)
    L.fe_net_create.argtypes = [vp, C.c_uint64, C.POINTER(vp)]
    L.fe_net_destroy.argtypes = [vp]
    L.fe_net_param_count.argtypes = [C.POINTER(C.c_int64)]
    L.fe_net_get_params.argtypes = [vp, fp, C.c_int64]
    L.fe_net_set_params.argtypes = [vp, fp, C.c_int64]
    L.fe_net_add_sample.argtypes = [vp, vp]
    L.fe_net_sample_count.argtypes = [vp, C.POINTER(C.c_int64)]
    L.fe_net_clear_samples.argtypes = [vp]
    L.fe_net_train.argtypes = [vp, C.c_int32, C.c_double, C.c_uint64, fp]
    L.fe_net_loss_and_grads.argtypes = [vp, i32p, dp, fp, C.POINTER(C.c_uint8), fp, dp, fp, fp]
    L.fe_net_set_loss.argtypes = [vp, C.c_int32]
    L.fe_scene_set_warm_start_net.argtypes = [vp, vp]
    L.fe_scene_warm_start_scale.argtypes = [vp, dp]
    L.fe_net_set_scale_mode.argtypes = [vp, C.c_int32]
    L.fe_scene_reserve_particles.argtypes = [vp, C.c_int64]
    L.fe_scene_particle_capacity.argtypes = [vp, C.POINTER(C.c_int64)]
    L.fe_surface_download.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int32)]
    L.fe_batch_advance_frame.argtypes = [vp, C.c_double, C.POINTER(SolverDesc), C.POINTER(C.c_int32),
                                         C.POINTER(C.c_double)]
    L.fe_version.restype = C.c_char_p
    for name in EXPORTED:
        getattr(L, name).restype = C.c_char_p if name in ("fe_last_error", "fe_version") else C.c_int
    _lib = L
    return L


def solver(kind=SOLVER_JACOBI, max_iterations=1000, check_interval=5, tolerance=1e-14, relative=False,
           warm_start=True):
    """Solver descriptor; the defaults are SinglePhasePressureSolver's
    CudaJacobiIterationSolver(1000, 5, 1e-14) (single_phase_pressure_solver.cpp:11)."""
    return SolverDesc(kind, max_iterations, check_interval, int(relative), int(warm_start), 0, tolerance)


def scene_desc(n, method=FLIP, origin=(0.0, 0.0, 0.0), size=(1.0, 1.0, 1.0), blend=0.05, max_cfl=3.0,
               gravity=(0.0, -9.81, 0.0), max_particles=0, viscosity=0.0, density=1.0):
    d = SceneDesc()
    d.n[:] = list(n)
    d.method = method
    d.origin[:] = list(origin)
    d.size[:] = list(size)
    d.blend = blend
    d.max_cfl = max_cfl
    d.gravity[:] = list(gravity)
    d.viscosity = viscosity
    d.density = density
    d.max_particles = max_particles
    return d


def box(lo, hi):
    s = SeedShape()
    s.kind = 0
    s.lo[:] = list(lo)
    s.hi[:] = list(hi)
    return s


def sphere(centre, radius):
    s = SeedShape()
    s.kind = 1
    s.lo[:] = list(centre)
    s.hi[:] = [radius, 0.0, 0.0]
    return s


def batch_scene_specs(count=96, seed=1, sizes=(24, 32, 48, 64, 96), extent=(0.2, 0.7)):
    """The batched-scenes workload: `count` cubic grids with n drawn uniformly
    from `sizes`, each with one axis-aligned fluid box of extents drawn from
    U[extent] per axis, placed uniformly inside the unit cube; returns a list
    of (n, lo, hi)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.choice(sizes))
        ext = rng.uniform(extent[0], extent[1], 3)
        lo = rng.uniform(0.0, 1.0 - ext)
        out.append((n, tuple(float(a) for a in lo), tuple(float(a) for a in lo + ext)))
    return out


def seed_particles(desc, shapes, seed=0):
    """Deterministic 2x2x2-per-cell seeding (host only; see fe_seed_particles)."""
    L = load()
    arr = (SeedShape * max(1, len(shapes)))(*shapes)
    cnt = C.c_int64()
    rc = L.fe_seed_particles(C.byref(desc), arr, len(shapes), seed, None, 0, C.byref(cnt))
    if rc != 0:
        raise FluideError(f"fe_seed_particles failed ({rc})")
    x = np.empty((cnt.value, 3), dtype=np.float32)
    rc = L.fe_seed_particles(C.byref(desc), arr, len(shapes), seed,
                             x.ctypes.data_as(C.POINTER(C.c_float)), cnt.value, C.byref(cnt))
    if rc != 0:
        raise FluideError(f"fe_seed_particles failed ({rc})")
    return x


def _fptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def _mesh_args(size, origin, spacing, vertices, triangles):
    """ctypes arguments (size, origin, spacing, vertices, nv, triangles, nt) of a
    mesh on a node grid; the arrays are kept alive on the returned tuple."""
    v = np.ascontiguousarray(np.asarray(vertices, dtype=np.float64).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(triangles, dtype=np.int32).reshape(-1, 3))
    sz = (C.c_int32 * 3)(*[int(q) for q in size])
    o = (C.c_double * 3)(*[float(q) for q in origin])
    h = (C.c_double * 3)(*[float(q) for q in spacing])
    args = _MeshArgs((sz, o, h, v.ctypes.data_as(C.POINTER(C.c_double)), v.shape[0],
                      t.ctypes.data_as(C.POINTER(C.c_int32)), t.shape[0]))
    args.keep = (v, t)
    return args


class _MeshArgs(tuple):
    pass


def mesh_sdf(ctx, size, origin, spacing, vertices, triangles):
    """Mesh2SDF::Build on the device: the SDF of a triangle mesh on a node grid
    of size (nx, ny, nz), returned as (nz, ny, nx) float64."""
    out = np.empty((int(size[2]), int(size[1]), int(size[0])), np.float64)
    args = _mesh_args(size, origin, spacing, vertices, triangles)
    ctx.check(ctx.L.fe_mesh_sdf(ctx._h, *args, out.ctypes.data_as(C.POINTER(C.c_double))), "fe_mesh_sdf")
    return out


def mesh_sdf_noalloc(ctx, size, origin, spacing, vertices, triangles):
    """mesh_sdf without device allocation (the device's static scratch array;
    at most 65536 vertices, 65536 triangles and 2^21 nodes per call)."""
    out = np.empty((int(size[2]), int(size[1]), int(size[0])), np.float64)
    args = _mesh_args(size, origin, spacing, vertices, triangles)
    ctx.check(ctx.L.fe_mesh_sdf_noalloc(ctx._h, *args, out.ctypes.data_as(C.POINTER(C.c_double))),
              "fe_mesh_sdf_noalloc")
    return out


class Context:
    def __init__(self, device=-1):
        L = load()
        h = C.c_void_p()
        rc = L.fe_ctx_create(device, C.byref(h))
        if rc != 0:
            raise FluideError(f"fe_ctx_create failed ({rc}): {L.fe_last_error(None).decode()}")
        self._h = h
        self.L = L

    @property
    def device_name(self):
        buf = C.create_string_buffer(64)
        self.L.fe_ctx_device_name(self._h, buf, 64)
        return buf.value.decode()

    def check(self, rc, what):
        if rc != 0:
            raise FluideError(f"{what} failed ({rc}): {self.L.fe_last_error(self._h).decode()}", rc)

    def copy_probe(self, nbytes):
        """One plain device copy of nbytes (the byte-counter calibration kernel); returns ms."""
        ms = C.c_double()
        self.check(self.L.fe_copy_probe(self._h, int(nbytes), C.byref(ms)), "fe_copy_probe")
        return ms.value

    def close(self):
        if getattr(self, "_h", None):
            self.L.fe_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Scene:
    """One fe_scene: the device state of a HybridSimulator."""

    def __init__(self, ctx: Context, desc: SceneDesc):
        self.ctx, self.desc, self.L = ctx, desc, ctx.L
        h = C.c_void_p()
        ctx.check(self.L.fe_scene_create(ctx._h, C.byref(desc), C.byref(h)), "fe_scene_create")
        self._h = h
        self.n = tuple(desc.n)

    def close(self):
        if getattr(self, "_h", None):
            self.L.fe_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def shape(self, field):
        nx, ny, nz = self.n
        if field in (GRID_U, GRID_MU, GRID_SU, GRID_DMU):
            return (nz, ny, nx + 1)
        if field in (GRID_V, GRID_MV, GRID_SV, GRID_DMV):
            return (nz, ny + 1, nx)
        if field in (GRID_W, GRID_MW, GRID_SW, GRID_DMW):
            return (nz + 1, ny, nx)
        return (nz, ny, nx)

    @staticmethod
    def dtype(field):
        if field in (GRID_MU, GRID_MV, GRID_MW, GRID_FLAGS, GRID_DMU, GRID_DMV, GRID_DMW):
            return np.uint8
        if field in (GRID_SDF, GRID_PRESSURE, GRID_COLLIDER_SDF):
            return np.float64
        return np.float32

    def upload_particles(self, x, v, Cm=None):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
        v = np.ascontiguousarray(v, dtype=np.float32).reshape(-1, 3)
        cm = None if Cm is None else np.ascontiguousarray(Cm, dtype=np.float32).reshape(-1, 9)
        self.ctx.check(self.L.fe_particles_upload(self._h, _fptr(x), _fptr(v), _fptr(cm), x.shape[0]),
                       "fe_particles_upload")

    def count(self):
        c = C.c_int64()
        self.ctx.check(self.L.fe_particles_count(self._h, C.byref(c)), "fe_particles_count")
        return c.value

    def download_particles(self):
        n = self.count()
        x = np.empty((n, 3), np.float32)
        v = np.empty((n, 3), np.float32)
        cm = np.empty((n, 9), np.float32)
        self.ctx.check(self.L.fe_particles_download(self._h, _fptr(x), _fptr(v), _fptr(cm), n),
                       "fe_particles_download")
        return x, v, cm

    def particle_cells(self):
        n = self.count()
        out = np.empty(n, np.int32)
        self.ctx.check(self.L.fe_particle_cells(self._h, out.ctypes.data_as(C.POINTER(C.c_int32)), n),
                       "fe_particle_cells")
        return out

    def get(self, field):
        out = np.empty(self.shape(field), dtype=self.dtype(field))
        self.ctx.check(self.L.fe_grid_download(self._h, field, out.ctypes.data_as(C.c_void_p), out.nbytes),
                       "fe_grid_download")
        return out

    def set(self, field, arr):
        a = np.ascontiguousarray(arr, dtype=self.dtype(field)).reshape(self.shape(field))
        self.ctx.check(self.L.fe_grid_upload(self._h, field, a.ctypes.data_as(C.c_void_p), a.nbytes),
                       "fe_grid_upload")

    def substep(self, dt, desc=None):
        st = SolveStats()
        desc = desc or solver()
        self.ctx.check(self.L.fe_substep(self._h, dt, C.byref(desc), C.byref(st)), "fe_substep")
        return st

    def advance_frame(self, frame_dt, desc=None):
        n = C.c_int32()
        desc = desc or solver()
        self.ctx.check(self.L.fe_advance_frame(self._h, frame_dt, C.byref(desc), C.byref(n)), "fe_advance_frame")
        return n.value

    def cfl(self, dt):
        out = C.c_double()
        self.ctx.check(self.L.fe_cfl(self._h, dt, C.byref(out)), "fe_cfl")
        return out.value

    def run_stage(self, stage, dt=0.0, desc=None):
        st = SolveStats()
        # the diffusion stage takes the scene's own diffusion solver when desc is None
        d = None if (desc is None and stage == STAGE_DIFFUSION) else C.byref(desc or solver())
        self.ctx.check(self.L.fe_run_stage(self._h, stage, dt, d, C.byref(st)),
                       f"fe_run_stage({STAGE_NAMES[stage] if 0 <= stage < len(STAGE_NAMES) else stage})")
        return st

    def stage_times(self):
        out = (C.c_double * 16)()
        self.ctx.check(self.L.fe_stage_times(self._h, out, 16), "fe_stage_times")
        return {STAGE_NAMES[i]: out[i] for i in range(len(STAGE_NAMES)) if out[i] > 0}

    def sync(self):
        self.ctx.check(self.L.fe_scene_sync(self._h), "fe_scene_sync")

    def last_stats(self):
        st = SolveStats()
        self.ctx.check(self.L.fe_last_solve_stats(self._h, C.byref(st)), "fe_last_solve_stats")
        return st

    def surface(self):
        """Fluid surface of the current SDF: (vertices (V,3), normals (V,3), triangles (T,3) int32)."""
        nv, nt = C.c_int64(), C.c_int64()
        self.ctx.check(self.L.fe_surface_build(self._h, C.byref(nv), C.byref(nt)), "fe_surface_build")
        v = np.empty((nv.value, 3), np.float64)
        n = np.empty((nv.value, 3), np.float64)
        t = np.empty((nt.value, 3), np.int32)
        dp = C.POINTER(C.c_double)
        self.ctx.check(self.L.fe_surface_download(self._h, v.ctypes.data_as(dp), n.ctypes.data_as(dp),
                                                  t.ctypes.data_as(C.POINTER(C.c_int32))), "fe_surface_download")
        return v, n, t

    def set_collider_sdf(self, phi):
        """Upload the collider SDF (nz, ny, nx FP64 at the nodes origin + h*(i,j,k))."""
        self.set(GRID_COLLIDER_SDF, np.ascontiguousarray(phi, dtype=np.float64))

    def set_surface_mode(self, mode):
        """SURFACE_REFERENCE (the reference's case tables, the default) or SURFACE_CLOSED."""
        self.ctx.check(self.L.fe_scene_set_surface_mode(self._h, int(mode)), "fe_scene_set_surface_mode")

    def set_sdf_search(self, mode):
        self.ctx.check(self.L.fe_scene_set_sdf_search(self._h, mode), "fe_scene_set_sdf_search")

    def set_graph(self, enable):
        """Run qualifying substeps as captured HIP graphs (bit-identical results)."""
        self.ctx.check(self.L.fe_scene_set_graph(self._h, 1 if enable else 0), "fe_scene_set_graph")

    def graph_stats(self):
        """(graph substeps run, captures made, substeps that needed host iterations)."""
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self.ctx.check(self.L.fe_scene_graph_stats(self._h, C.byref(a), C.byref(b), C.byref(c)),
                       "fe_scene_graph_stats")
        return a.value, b.value, c.value

    def add_mesh_collider(self, size, origin, spacing, vertices, triangles):
        """AddCollider(TriangleMeshCollider(size, origin, spacing, mesh)); size is (nx, ny, nz)."""
        args = _mesh_args(size, origin, spacing, vertices, triangles)
        self.ctx.check(self.L.fe_scene_add_mesh_collider(self._h, *args), "fe_scene_add_mesh_collider")

    def initialize_from_mesh(self, vertices, triangles, particles_per_block, seed=0):
        """InitializeFromTriangleMesh: Mesh2SDF into the fluid SDF, then particles appended in its cells."""
        a = _mesh_args((1, 1, 1), (0, 0, 0), (1, 1, 1), vertices, triangles)
        self.ctx.check(self.L.fe_scene_initialize_from_mesh(self._h, a[3], a[4], a[5], a[6], int(particles_per_block),
                                                            int(seed)), "fe_scene_initialize_from_mesh")

    def add_volume_emitter(self, resolution, domain_origin, domain_size, vertices, triangles, particles_per_block,
                           velocity, seed=0):
        """AddEmitter(VolumeEmitter(mesh, resolution, domain, particlesPerBlock, velocity, variance))."""
        a = _mesh_args(resolution, domain_origin, domain_size, vertices, triangles)
        vel = (C.c_double * 3)(*[float(q) for q in velocity])
        self.ctx.check(self.L.fe_scene_add_volume_emitter(self._h, *a, int(particles_per_block), vel, int(seed)),
                       "fe_scene_add_volume_emitter")

    def reserve_particles(self, capacity):
        self.ctx.check(self.L.fe_scene_reserve_particles(self._h, int(capacity)), "fe_scene_reserve_particles")

    def capacity(self):
        c = C.c_int64()
        self.ctx.check(self.L.fe_scene_particle_capacity(self._h, C.byref(c)), "fe_scene_particle_capacity")
        return c.value

    def set_warm_start_net(self, net):
        self.ctx.check(self.L.fe_scene_set_warm_start_net(self._h, net._h if net is not None else None),
                       "fe_scene_set_warm_start_net")

    def warm_start_scale(self):
        a = C.c_double()
        self.ctx.check(self.L.fe_scene_warm_start_scale(self._h, C.byref(a)), "fe_scene_warm_start_scale")
        return a.value

    def tuning(self):
        """(p2g variant, g2p variant) in use."""
        t = Tuning()
        self.ctx.check(self.L.fe_scene_get_tuning(self._h, C.byref(t)), "fe_scene_get_tuning")
        return t.p2g, t.g2p

    def set_tuning(self, p2g, g2p, mg=None):
        t = Tuning()
        self.ctx.check(self.L.fe_scene_get_tuning(self._h, C.byref(t)), "fe_scene_get_tuning")
        t.p2g, t.g2p = p2g, g2p
        if mg is not None:
            t.mg = mg
        self.ctx.check(self.L.fe_scene_set_tuning(self._h, C.byref(t)), "fe_scene_set_tuning")

    def mg_variant(self):
        t = Tuning()
        self.ctx.check(self.L.fe_scene_get_tuning(self._h, C.byref(t)), "fe_scene_get_tuning")
        return t.mg

    def set_mg_variant(self, mg):
        p2g, g2p = self.tuning()
        self.set_tuning(p2g, g2p, mg)

    def autotune(self, reps=5):
        """Times the kernel variants on the current particles and keeps the
        fastest; returns ((p2g, g2p), {"p2g_ms": [tiles, faces, scatter], "g2p_ms": [quads, pairs]})."""
        t, tm = Tuning(), TuningTimes()
        self.ctx.check(self.L.fe_scene_autotune(self._h, int(reps), C.byref(t), C.byref(tm)), "fe_scene_autotune")
        return (t.p2g, t.g2p), {"p2g_ms": list(tm.p2g_ms), "g2p_ms": list(tm.g2p_ms)}

    def clear_emitters(self):
        self.ctx.check(self.L.fe_scene_clear_emitters(self._h), "fe_scene_clear_emitters")

    def clear_colliders(self):
        self.ctx.check(self.L.fe_scene_clear_colliders(self._h), "fe_scene_clear_colliders")

    def set_diffusion_solver(self, desc):
        self.ctx.check(self.L.fe_scene_set_diffusion_solver(self._h, C.byref(desc)), "fe_scene_set_diffusion_solver")

    def last_diffusion_stats(self):
        st = (SolveStats * 3)()
        self.ctx.check(self.L.fe_last_diffusion_stats(self._h, st), "fe_last_diffusion_stats")
        return list(st)

    def kernel_timer(self, kernel_id):
        """Time every launch of one kernel with HIP events on the scene's stream."""
        self.ctx.check(self.L.fe_kernel_timer(self._h, kernel_id), "fe_kernel_timer")

    def kernel_timer_read(self, reset=False):
        ms = C.c_double()
        n = C.c_int64()
        self.ctx.check(self.L.fe_kernel_timer_read(self._h, C.byref(ms), C.byref(n), int(reset)),
                       "fe_kernel_timer_read")
        return ms.value, n.value


class Batch:
    """fe_batch: independent scenes of one context advanced together by host
    worker threads, each scene on its own stream."""

    def __init__(self, ctx: Context, scenes, workers=4):
        self.ctx, self.L = ctx, ctx.L
        self.scenes = list(scenes)
        arr = (C.c_void_p * len(self.scenes))(*[sc._h for sc in self.scenes])
        h = C.c_void_p()
        ctx.check(self.L.fe_batch_create(ctx._h, arr, len(self.scenes), workers, C.byref(h)), "fe_batch_create")
        self._h = h

    def advance_frame(self, frame_dt, desc=None):
        """One frame of every scene; returns (substeps, wall ms) per scene."""
        n = len(self.scenes)
        sub = (C.c_int32 * n)()
        ms = (C.c_double * n)()
        desc = desc or solver()
        self.ctx.check(self.L.fe_batch_advance_frame(self._h, frame_dt, C.byref(desc), sub, ms),
                       "fe_batch_advance_frame")
        return list(sub), list(ms)

    def close(self):
        if getattr(self, "_h", None):
            self.L.fe_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Net:
    """The learned pressure warm start (fe_net_*)."""

    def __init__(self, ctx, seed=0):
        self.ctx, self.L = ctx, ctx.L
        h = C.c_void_p()
        ctx.check(self.L.fe_net_create(ctx._h, int(seed), C.byref(h)), "fe_net_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self.L.fe_net_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @staticmethod
    def param_count():
        c = C.c_int64()
        load().fe_net_param_count(C.byref(c))
        return c.value

    def params(self):
        out = np.empty(self.param_count(), np.float32)
        self.ctx.check(self.L.fe_net_get_params(self._h, _fptr(out), out.size), "fe_net_get_params")
        return out

    def set_params(self, flat):
        a = np.ascontiguousarray(flat, dtype=np.float32)
        self.ctx.check(self.L.fe_net_set_params(self._h, _fptr(a), a.size), "fe_net_set_params")

    def set_scale_mode(self, mode):
        self.ctx.check(self.L.fe_net_set_scale_mode(self._h, int(mode)), "fe_net_set_scale_mode")

    def add_sample(self, scene):
        self.ctx.check(self.L.fe_net_add_sample(self._h, scene._h), "fe_net_add_sample")

    def sample_count(self):
        c = C.c_int64()
        self.ctx.check(self.L.fe_net_sample_count(self._h, C.byref(c)), "fe_net_sample_count")
        return c.value

    def clear_samples(self):
        self.ctx.check(self.L.fe_net_clear_samples(self._h), "fe_net_clear_samples")

    def train(self, steps, learning_rate=1e-3, seed=0):
        losses = np.empty(int(steps), np.float32)
        self.ctx.check(self.L.fe_net_train(self._h, int(steps), float(learning_rate), int(seed), _fptr(losses)),
                       "fe_net_train")
        return losses

    def set_loss(self, mode):
        self.ctx.check(self.L.fe_net_set_loss(self._h, int(mode)), "fe_net_set_loss")

    def loss_and_grads(self, b, flags, target, h=None):
        """b, flags, target as (nz, ny, nx) arrays, h the spacing per axis (x, y, z)
        or None for 1/n; returns (loss, flat gradients, prediction)."""
        b = np.ascontiguousarray(b, dtype=np.float32)
        f = np.ascontiguousarray(flags, dtype=np.uint8)
        t = np.ascontiguousarray(target, dtype=np.float32)
        n = (C.c_int32 * 3)(b.shape[2], b.shape[1], b.shape[0])
        loss = C.c_double()
        g = np.empty(self.param_count(), np.float32)
        pred = np.empty(b.shape, np.float32)
        hh = None if h is None else (C.c_double * 3)(*[float(q) for q in h])
        self.ctx.check(self.L.fe_net_loss_and_grads(self._h, n, hh, _fptr(b), f.ctypes.data_as(C.POINTER(C.c_uint8)),
                                                    _fptr(t), C.byref(loss), _fptr(g), _fptr(pred)),
                       "fe_net_loss_and_grads")
        return loss.value, g, pred


NET_LAYERS = [("enc0", 16, 16), ("enc1", 16, 32), ("enc2", 32, 48), ("enc3", 48, 64),
              ("dec2", 112, 48), ("dec1", 80, 32), ("dec0", 48, 16)]


def net_unflatten(flat):
    """Flat parameters -> {layer: (W (27, ci, co), b (co,)), 'head': (hw (16,), hb)}."""
    out, o = {}, 0
    for name, ci, co in NET_LAYERS:
        w = flat[o:o + 27 * ci * co].reshape(27, ci, co)
        o += 27 * ci * co
        out[name] = (w, flat[o:o + co])
        o += co
    out["head"] = (flat[o:o + 16], flat[o + 16])
    return out


def net_flatten(params):
    parts = []
    for name, ci, co in NET_LAYERS:
        w, b = params[name]
        parts += [np.asarray(w, np.float32).reshape(-1), np.asarray(b, np.float32).reshape(-1)]
    hw, hb = params["head"]
    parts += [np.asarray(hw, np.float32).reshape(-1), np.asarray([hb], np.float32)]
    return np.concatenate(parts)
