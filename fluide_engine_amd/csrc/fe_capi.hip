// fe_capi.hip -- the C ABI of include/fluide_hip.h: scene lifetime, state
// transfer and the substep orchestration of PICSimulator::OnAdvanceTimeStep
// (src/pic_simulator.cpp:111-237) over the gfx950 kernels.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "fe_learn.hpp"
#include "fe_stencil.hpp"

using namespace fe;

namespace {

thread_local std::string g_err_no_ctx;

int set_err(fe_ctx* ctx, int code, const std::string& msg) {
    if (ctx) {
        std::lock_guard<std::mutex> lock(ctx->err_mu);
        ctx->err = msg;
    } else
        g_err_no_ctx = msg;
    return code;
}

int hip_fail(fe_ctx* ctx, hipError_t e, const char* what) {
    return set_err(ctx, FE_E_HIP, std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")");
}

#define HIPCHK(ctx, call)                                   \
    do {                                                    \
        hipError_t _e = (call);                             \
        if (_e != hipSuccess) return hip_fail(ctx, _e, #call); \
    } while (0)

// Device arrays carry 64 bytes of slack past their end, so kernels may read
// whole aligned words that straddle an array's last element.
template <typename T>
hipError_t dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc((void**)p, count * sizeof(T) + 64);
}

void build_geom(fe_scene* s) {
    Geom& g = s->g;
    const fe_scene_desc& d = s->desc;
    g.method = d.method;
    for (int a = 0; a < 3; ++a) {
        g.n[a] = d.n[a];
        g.od[a] = d.origin[a];
        g.Ld[a] = d.size[a];
        g.hd[a] = d.size[a] / d.n[a];  // src/hybrid_simulator.cpp:7
        g.o[a] = (float)g.od[a];
        g.h[a] = (float)g.hd[a];
        g.ihd[a] = 1.0 / g.hd[a];
        g.ih[a] = (float)g.ihd[a];
        g.L[a] = (float)g.Ld[a];
    }
    // FaceCenteredGrid3D::CalculateDataOrigins (face_centered_grid3d.cpp:324-332)
    for (int c = 0; c < 3; ++c)
        for (int a = 0; a < 3; ++a) {
            g.cod[c][a] = g.od[a] + g.hd[a] * (a == c ? 0.0 : 0.5);
            g.co[c][a] = (float)g.cod[c][a];
        }
}

size_t field_bytes(const fe_scene* s, int32_t field) {
    const Geom& g = s->g;
    switch (field) {
        case FE_GRID_U: case FE_GRID_SU: return face_count(g, 0) * sizeof(float);
        case FE_GRID_V: case FE_GRID_SV: return face_count(g, 1) * sizeof(float);
        case FE_GRID_W: case FE_GRID_SW: return face_count(g, 2) * sizeof(float);
        case FE_GRID_MU: return face_count(g, 0);
        case FE_GRID_MV: return face_count(g, 1);
        case FE_GRID_MW: return face_count(g, 2);
        case FE_GRID_SDF: case FE_GRID_PRESSURE: return cell_count(g) * sizeof(double);
        case FE_GRID_FLAGS: return cell_count(g);
        case FE_GRID_RHS: return cell_count(g) * sizeof(float);
        case FE_GRID_DMU: return face_count(g, 0);
        case FE_GRID_DMV: return face_count(g, 1);
        case FE_GRID_DMW: return face_count(g, 2);
        case FE_GRID_COLLIDER_SDF: return cell_count(g) * sizeof(double);
        default: return 0;
    }
}

void* field_ptr(fe_scene* s, int32_t field) {
    switch (field) {
        case FE_GRID_U: return s->u[0];
        case FE_GRID_V: return s->u[1];
        case FE_GRID_W: return s->u[2];
        case FE_GRID_SU: return s->snap[0];
        case FE_GRID_SV: return s->snap[1];
        case FE_GRID_SW: return s->snap[2];
        case FE_GRID_MU: return s->mark[0];
        case FE_GRID_MV: return s->mark[1];
        case FE_GRID_MW: return s->mark[2];
        case FE_GRID_SDF: return s->sdf;
        case FE_GRID_FLAGS: return s->flags;
        case FE_GRID_PRESSURE: return s->p;
        case FE_GRID_RHS: return s->b;
        case FE_GRID_DMU: return s->dmark[0];
        case FE_GRID_DMV: return s->dmark[1];
        case FE_GRID_DMW: return s->dmark[2];
        case FE_GRID_COLLIDER_SDF: return s->csdf;
        default: return nullptr;
    }
}

bool diffusion_field(int32_t field) { return field == FE_GRID_DMU || field == FE_GRID_DMV || field == FE_GRID_DMW; }

int check_viscosity(fe_ctx* ctx, double viscosity) {
    if (!(viscosity >= 0.0) || std::isinf(viscosity)) return set_err(ctx, FE_E_INVALID, "viscosity must be finite and >= 0");
    return FE_OK;
}

// BackwardEulerDiffusionSolver's linear solver: Jacobi only (its default,
// src/fluid_solvers/backward_euler_diffusion_solver.cpp:10).
int check_diffusion_solver(fe_ctx* ctx, const fe_solver_desc* d) {
    if (!d) return set_err(ctx, FE_E_INVALID, "solver descriptor is NULL");
    if (d->kind != FE_SOLVER_JACOBI)
        return set_err(ctx, FE_E_UNSUPPORTED, "the diffusion solve runs the Jacobi solver only");
    if (d->max_iterations < 0 || d->max_iterations > 1000000)
        return set_err(ctx, FE_E_INVALID, "max_iterations out of range");
    if (d->check_interval < 0) return set_err(ctx, FE_E_INVALID, "check_interval must be >= 0");
    if (!(d->tolerance >= 0.0)) return set_err(ctx, FE_E_INVALID, "tolerance must be >= 0");
    return FE_OK;
}

int check_solver(fe_ctx* ctx, const fe_solver_desc* d) {
    if (!d) return set_err(ctx, FE_E_INVALID, "solver descriptor is NULL");
    if (d->kind < FE_SOLVER_JACOBI || d->kind > FE_SOLVER_MGPCG)
        return set_err(ctx, FE_E_UNSUPPORTED, "unknown solver kind " + std::to_string(d->kind));
    if (d->max_iterations < 0 || d->max_iterations > 1000000)
        return set_err(ctx, FE_E_INVALID, "max_iterations out of range");
    if (!(d->tolerance >= 0.0)) return set_err(ctx, FE_E_INVALID, "tolerance must be >= 0");
    if (d->warm_start < FE_WARM_START_ZERO || d->warm_start > FE_WARM_START_LEARNED)
        return set_err(ctx, FE_E_INVALID, "unknown warm_start " + std::to_string(d->warm_start));
    return FE_OK;
}

void ev(fe_scene* s, int i) { (void)hipEventRecord(s->ev[i], s->stream); }

// Stage sequence markers for fe_stage_times.
struct StageClock {
    fe_scene* s;
    int used[kMaxStages] = {0};
    int order[kMaxStages];
    int n = 0;
    explicit StageClock(fe_scene* sc) : s(sc) {}
    void begin(int stage) {
        order[n] = stage;
        ev(s, n);
        ++n;
    }
    void end() { ev(s, n); }
    void collect() {
        for (int i = 0; i < kMaxStages; ++i) s->stage_ms[i] = 0.0;
        for (int i = 0; i < n; ++i) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, s->ev[i], s->ev[i + 1]) == hipSuccess) s->stage_ms[order[i]] += ms;
        }
    }
};

int ensure_sorted(fe_scene* s) {
    if (!s->sorted) HIPCHK(s->ctx, launch_sort(s, false));
    return FE_OK;
}

// The pressure solve's stats from the host mirror of its scalars and the
// final FP64 residual sums.
void solve_stats(fe_scene* s, const fe_solver_desc* d, double rr, double bb, float ms, fe_solve_stats* st) {
    const SolverScalars& h = *s->sc_host;
    fe_solve_stats& ls = s->last_stats;
    ls.iterations = d->kind == FE_SOLVER_JACOBI ? (h.done ? h.it : d->max_iterations) : h.it;
    ls.abs_residual = std::sqrt(rr);
    ls.rel_residual = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
    ls.ms = ms;
    ls.converged = h.converged;
    if (st) *st = ls;
}

// Pressure stage: SinglePhasePressureSolver::Solve (single_phase_pressure_solver.cpp:21-28)
// minus the markers (built with the SDF), plus the FP64 residual for the stats.
int pressure(fe_scene* s, const fe_solver_desc* d, bool walls_after, fe_solve_stats* st, int ev_start) {
    if (d->warm_start == FE_WARM_START_LEARNED && !s->net)
        return set_err(s->ctx, FE_E_INVALID, "FE_WARM_START_LEARNED needs a network (fe_scene_set_warm_start_net)");
    HIPCHK(s->ctx, launch_pressure_solve(s, d));
    HIPCHK(s->ctx, launch_final_residual(s));
    HIPCHK(s->ctx, launch_apply_pressure(s, walls_after));
    hipEvent_t stop = s->ev[kMaxStages];
    HIPCHK(s->ctx, hipEventRecord(stop, s->stream));
    double rr = 0.0, bb = 0.0;
    HIPCHK(s->ctx, read_final_residual(s, &rr, &bb));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, s->ev[ev_start], stop);
    solve_stats(s, d, rr, bb, ms, st);
    return FE_OK;
}

int emit(fe_scene* s) {
    bool over = false;
    hipError_t e = launch_emit(s, &over);
    if (over)
        return set_err(s->ctx, FE_E_INVALID, "emitted particles exceed 2^31 - 1 particles");
    HIPCHK(s->ctx, e);
    return FE_OK;
}

// The substep up to the pressure solve: OnBeginAdvanceTimeStep
// (src/pic_simulator.cpp:156-196), ApplyBoundryCondition, ComputeExternalForces
// (:203-212) and ComputeDiffusion (:214-222).  Launches only.
int substep_head(fe_scene* s, double dt, StageClock* clk) {
    fe_ctx* ctx = s->ctx;
    // particles binned by cell (keys come fused from the previous advection)
    if (clk) clk->begin(FE_STAGE_SORT);
    if (!s->sorted) HIPCHK(ctx, launch_sort(s, s->keys_ready));
    s->keys_ready = false;
    if (clk) clk->begin(FE_STAGE_P2G);
    HIPCHK(ctx, launch_p2g(s));
    if (clk) clk->begin(FE_STAGE_SDF);
    HIPCHK(ctx, launch_sdf(s));
    HIPCHK(ctx, launch_markers(s));  // BuildMarkers reads only the SDF: done once here
    // ApplyBoundryCondition is the wall zeroing without a collider (fused into
    // the gravity and pressure-gradient kernels) and the full
    // ConstrainVelocity with one
    const bool col = s->has_collider;
    if (clk) clk->begin(FE_STAGE_EXTRAPOLATE);
    HIPCHK(ctx, launch_extrapolate(s));
    HIPCHK(ctx, launch_bc(s));
    if (clk) clk->begin(FE_STAGE_GRAVITY);
    HIPCHK(ctx, launch_gravity(s, dt, !col));
    if (col) HIPCHK(ctx, launch_bc(s));
    // ComputeDiffusion when viscosity > 0, then ApplyBoundryCondition
    if (s->desc.viscosity > 0.0) {
        if (clk) clk->begin(FE_STAGE_DIFFUSION);
        HIPCHK(ctx, launch_diffusion(s, dt, &s->diff_desc));
        HIPCHK(ctx, launch_bc(s));
    }
    return FE_OK;
}

// The substep after the pressure solve and its application:
// ComputeAdvection (src/pic_simulator.cpp:231-237).  Launches only.
int substep_tail(fe_scene* s, double dt, StageClock* clk) {
    fe_ctx* ctx = s->ctx;
    const bool col = s->has_collider;
    if (col) HIPCHK(ctx, launch_bc(s));
    if (clk) clk->begin(FE_STAGE_EXTRAPOLATE);
    HIPCHK(ctx, launch_extrapolate(s));
    HIPCHK(ctx, launch_bc(s));
    if (clk) clk->begin(FE_STAGE_G2P);
    HIPCHK(ctx, launch_g2p_move(s, dt, true, true, !col));
    HIPCHK(ctx, launch_collide(s));  // moved particles inside the collider (keys recomputed by the sort)
    s->keys_ready = !col;
    return FE_OK;
}

// ---------------------------------------------------------------------------
// Substeps as HIP graphs.  A scene in graph mode runs each qualifying substep
// as two captured graphs: (a) the stages up to the pressure solve and the
// solve's first iterations, ending with a copy of the solver state to the
// host; (b) the final residual, the pressure application and the advection,
// ending with a copy of the residual sums.  Between them the host adds
// iterations if the solve has not stopped (the same kernels, launched as in
// an eager substep) and moves the solution into s->p if the iteration count
// left it in the partner buffer.  dt reaches the captured kernels through
// device memory, written by a one-thread kernel before (a).  Each graph reads
// the buffers its capture saw, and the particle ping-pong alternates with
// every sort, so a scene keeps one pair per parity of s->cur.  The number of
// captured iterations is the previous solve's count + 2 (at least 4); a pair
// is captured again when anything it baked in changes (the key below), after
// two substeps in a row needed host iterations, or after eight that stopped
// more than four iterations early.
// ---------------------------------------------------------------------------
__global__ void k_set_f64(double* __restrict__ a, double v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *a = v;
}

std::mutex g_capture_mu;  // one capture at a time in the process

struct GraphKey {
    uint64_t gen;
    fe_solver_desc desc;
    double* p;
    int64_t np, cap;
    float* px0;
    fe_tuning tune;
};

bool graph_eligible(const fe_scene* s, const fe_solver_desc* d) {
    return s->graph_mode && s->emitters.empty() && !(s->desc.viscosity > 0.0) && !s->has_collider &&
           d->kind != FE_SOLVER_JACOBI && d->warm_start != FE_WARM_START_LEARNED && s->keys_ready && !s->sorted &&
           s->np > 0 && s->kt_id == FE_KERNEL_NONE && (d->kind != FE_SOLVER_MGPCG || !s->mg.empty());
}

bool same_solver(const fe_solver_desc& a, const fe_solver_desc& b) {
    return a.kind == b.kind && a.max_iterations == b.max_iterations && a.check_interval == b.check_interval &&
           a.relative == b.relative && a.warm_start == b.warm_start && a.tolerance == b.tolerance;
}

void graph_release(fe::SubstepGraph& G) {
    if (G.a) (void)hipGraphExecDestroy(G.a);
    if (G.b) (void)hipGraphExecDestroy(G.b);
    G = fe::SubstepGraph{};
}

bool graph_key_matches(const fe::SubstepGraph& G, const fe_scene* s, const fe_solver_desc* d) {
    return G.valid && G.gen == s->graph_gen && same_solver(G.desc, *d) && G.p == s->p &&
           G.np == s->np && G.cap == s->cap && G.px0 == s->P[0].x && G.tune.p2g == s->tune.p2g &&
           G.tune.g2p == s->tune.g2p && G.tune.mg == s->tune.mg;
}

// Captures what `body` launches on the scene's stream into an executable graph.
template <typename F>
int capture(fe_scene* s, hipGraphExec_t* out, F body) {
    fe_ctx* ctx = s->ctx;
    std::lock_guard<std::mutex> lock(g_capture_mu);
    HIPCHK(ctx, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    s->capturing = true;
    const int rc = body();
    s->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(s->stream, &g);
    if (rc != FE_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ee != hipSuccess) return hip_fail(ctx, ee, "hipStreamEndCapture");
    const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return hip_fail(ctx, ei, "hipGraphInstantiate");
    return FE_OK;
}

int substep_graph(fe_scene* s, double dt, const fe_solver_desc* d, fe_solve_stats* st) {
    fe_ctx* ctx = s->ctx;
    fe::SubstepGraph& G = s->graph[s->cur & 1];
    const int precond = d->kind == FE_SOLVER_PCG_JACOBI ? 1 : (d->kind == FE_SOLVER_MGPCG ? 2 : 0);
    const bool fresh = !graph_key_matches(G, s, d) || G.resume_streak >= 2 || G.idle_streak >= 8;
    ++s->substeps_run;
    hipLaunchKernelGGL(k_set_f64, dim3(1), dim3(64), 0, s->stream, s->dt_dev, dt);
    HIPCHK(ctx, hipEventRecord(s->ev[0], s->stream));
    int rc;
    if (fresh) {
        graph_release(G);
        const int prev = s->last_cg_iters[precond];
        G.iters = std::max(4, std::min(prev + 2, d->max_iterations + 1));
        G.gen = s->graph_gen;
        G.desc = *d;
        G.p = s->p;
        G.np = s->np;
        G.cap = s->cap;
        G.px0 = s->P[0].x;
        G.tune = s->tune;
        rc = capture(s, &G.a, [&]() -> int {
            const int r = substep_head(s, dt, nullptr);
            if (r != FE_OK) return r;
            HIPCHK(ctx, capture_pressure_head(s, d, G.iters, &G.run_after_a));
            return FE_OK;
        });
        if (rc != FE_OK) {
            graph_release(G);
            return rc;
        }
        // the host state the head leaves (the capture ran its launch functions)
        G.cur_after = s->cur;
        G.sdf_from_cells_after = s->sdf_from_cells;
        ++s->graph_captures;
    }
    HIPCHK(ctx, hipGraphLaunch(G.a, s->stream));
    HIPCHK(ctx, hipStreamSynchronize(s->stream));
    s->cur = G.cur_after;
    s->sorted = true;
    s->keys_ready = false;
    s->count_zero = true;  // the head's k_scatter counted the histogram down
    s->sdf_from_cells = G.sdf_from_cells_after;
    bool resumed = false;
    HIPCHK(ctx, finish_pressure_after_head(s, d, G.run_after_a, &resumed));
    const int iters = s->sc_host->it;
    if (resumed) {
        ++s->graph_resumed;
        ++G.resume_streak;
    } else {
        G.resume_streak = 0;
    }
    G.idle_streak = (!resumed && G.iters > iters + 4) ? G.idle_streak + 1 : 0;
    if (!G.b) {
        const SolverShape sh = solver_shape_n(cell_count(s->g));
        rc = capture(s, &G.b, [&]() -> int {
            HIPCHK(ctx, launch_final_residual(s));
            HIPCHK(ctx, launch_apply_pressure(s, true));
            const int r = substep_tail(s, dt, nullptr);
            if (r != FE_OK) return r;
            HIPCHK(ctx, hipMemcpyAsync(s->res_host, s->partials + 2 * sh.blocks, 2 * sizeof(double),
                                       hipMemcpyDeviceToHost, s->stream));
            return FE_OK;
        });
        if (rc != FE_OK) {
            graph_release(G);
            return rc;
        }
        G.valid = true;
    }
    HIPCHK(ctx, hipGraphLaunch(G.b, s->stream));
    HIPCHK(ctx, hipEventRecord(s->ev[1], s->stream));
    HIPCHK(ctx, hipStreamSynchronize(s->stream));
    s->sorted = false;
    s->keys_ready = true;
    s->count_zero = false;  // the advection's histogram of the next bins
    ++s->graph_substeps;
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, s->ev[0], s->ev[1]);
    for (int i = 0; i < kMaxStages; ++i) s->stage_ms[i] = 0.0;
    solve_stats(s, d, s->res_host[0], s->res_host[1], ms, st);
    kt_flush(s);
    return FE_OK;
}

int substep(fe_scene* s, double dt, const fe_solver_desc* d, fe_solve_stats* st) {
    if (graph_eligible(s, d)) return substep_graph(s, dt, d, st);
    fe_ctx* ctx = s->ctx;
    // Emitter::Emitt (OnBeginAdvanceTimeStep, src/pic_simulator.cpp:158-163),
    // before BuildCollider: the first substep sees a collider SDF of 0
    if (!s->emitters.empty()) {
        const int rc = emit(s);
        if (rc != FE_OK) return rc;
    }
    ++s->substeps_run;
    StageClock clk(s);
    int rc = substep_head(s, dt, &clk);
    if (rc != FE_OK) return rc;
    // ComputePressure (:224-229)
    const bool col = s->has_collider;
    const int pev = clk.n;
    clk.begin(FE_STAGE_PRESSURE);
    rc = pressure(s, d, !col, st, pev);
    if (rc != FE_OK) return rc;
    rc = substep_tail(s, dt, &clk);
    if (rc != FE_OK) return rc;
    clk.end();
    HIPCHK(ctx, hipStreamSynchronize(s->stream));
    clk.collect();
    kt_flush(s);
    return FE_OK;
}

// Mesh2SDF of one triangle mesh on a node grid, into a new device buffer
// *out (freed by the caller); synchronous on st.
int mesh_sdf_device(fe_ctx* ctx, hipStream_t st, const int32_t size[3], const double origin[3],
                    const double spacing[3], const double* vertices, int64_t nv, const int32_t* triangles,
                    int64_t nt, NodeGrid3& G, double** out) {
    *out = nullptr;
    if (!size || !origin || !spacing || nv < 0 || nt < 0 || (nv > 0 && !vertices) || (nt > 0 && !triangles))
        return set_err(ctx, FE_E_INVALID, "mesh collider: null pointer or negative count");
    for (int a = 0; a < 3; ++a) {
        if (size[a] < 1 || !(spacing[a] > 0.0) || !std::isfinite(spacing[a]) || !std::isfinite(origin[a]))
            return set_err(ctx, FE_E_INVALID, "mesh collider: grid sizes must be >= 1, spacings finite and > 0");
        G.n[a] = size[a];
        G.o[a] = origin[a];
        G.h[a] = spacing[a];
    }
    const int64_t nn = (int64_t)size[0] * size[1] * size[2];
    std::vector<double> tv;
    std::vector<int32_t> tb;
    double fill = 0.0;
    if (!mesh_prepare(G, vertices, nv, triangles, nt, tv, tb, fill))
        return set_err(ctx, FE_E_INVALID,
                       "mesh collider: a triangle's vertex index is out of range, or a coordinate / h.x does not "
                       "fit an int");
    double* tvd = nullptr;
    int32_t* tbd = nullptr;
    double* o = nullptr;
    auto cleanup = [&]() {
        if (tvd) (void)hipFree(tvd);
        if (tbd) (void)hipFree(tbd);
    };
    hipError_t e;
    if ((e = hipMalloc((void**)&tvd, sizeof(double) * std::max<size_t>(tv.size(), 1))) != hipSuccess ||
        (e = hipMalloc((void**)&tbd, sizeof(int32_t) * std::max<size_t>(tb.size(), 1))) != hipSuccess ||
        (e = hipMalloc((void**)&o, sizeof(double) * nn)) != hipSuccess) {
        cleanup();
        if (o) (void)hipFree(o);
        return hip_fail(ctx, e, "mesh collider allocation");
    }
    if ((nt > 0 && ((e = hipMemcpyAsync(tvd, tv.data(), sizeof(double) * tv.size(), hipMemcpyHostToDevice, st)) !=
                        hipSuccess ||
                    (e = hipMemcpyAsync(tbd, tb.data(), sizeof(int32_t) * tb.size(), hipMemcpyHostToDevice, st)) !=
                        hipSuccess)) ||
        (e = launch_mesh_sdf(st, G, tvd, tbd, nt, fill, o)) != hipSuccess || (e = hipStreamSynchronize(st)) != hipSuccess) {
        cleanup();
        (void)hipFree(o);
        return hip_fail(ctx, e, "mesh collider SDF");
    }
    cleanup();
    *out = o;
    return FE_OK;
}

}  // namespace

extern "C" {

const char* fe_version(void) { return "fluide-hip 0.1 (gfx950)"; }

// Plain streaming copy (16 B per lane and iteration), the reference kernel for
// scaling hardware byte counters.
__global__ __launch_bounds__(256) void k_copy_probe(const float4* __restrict__ a, float4* __restrict__ b, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) b[i] = a[i];
}

int fe_copy_probe(fe_ctx* ctx, int64_t bytes, double* ms) {
    if (!ctx || bytes <= 0 || bytes % 16 != 0) return set_err(ctx, FE_E_INVALID, "bytes must be a positive multiple of 16");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    float4 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc((void**)&a, bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&b, bytes);
    if (e == hipSuccess) e = hipMemset(a, 0, bytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        const int64_t n = bytes / 16;
        const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
        e = hipEventRecord(e0, nullptr);
        hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, nullptr, (const float4*)a, b, n);
        if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float t = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (ms) *ms = t;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess) return hip_fail(ctx, e, "fe_copy_probe");
    return FE_OK;
}

int fe_ctx_create(int device, fe_ctx** out) {
    if (!out) return set_err(nullptr, FE_E_INVALID, "out is NULL");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return set_err(nullptr, FE_E_NO_DEVICE, std::string("no HIP device: ") + hipGetErrorName(e));
    if (device < 0) {
        if (hipGetDevice(&device) != hipSuccess) device = 0;
    }
    if (device >= count) return set_err(nullptr, FE_E_INVALID, "device index out of range");
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) return hip_fail(nullptr, e, "hipGetDeviceProperties");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_err(nullptr, FE_E_NO_DEVICE,
                       std::string("device is ") + prop.gcnArchName + "; this library is built for gfx950 only");
    // the device address of the no-allocation entry point's scratch array
    if ((e = hipSetDevice(device)) != hipSuccess || (e = noalloc_setup(device)) != hipSuccess)
        return hip_fail(nullptr, e, "scratch set-up (hipGetSymbolAddress)");
    auto* ctx = new fe_ctx();
    ctx->device = device;
    std::snprintf(ctx->name, sizeof(ctx->name), "%s", prop.gcnArchName);
    *out = ctx;
    return FE_OK;
}

int fe_ctx_destroy(fe_ctx* ctx) {
    delete ctx;
    return FE_OK;
}

const char* fe_last_error(const fe_ctx* ctx) { return ctx ? ctx->err.c_str() : g_err_no_ctx.c_str(); }

int fe_ctx_device_name(const fe_ctx* ctx, char* buf, size_t len) {
    if (!ctx || !buf || len == 0) return FE_E_INVALID;
    std::snprintf(buf, len, "%s", ctx->name);
    return FE_OK;
}

int fe_scene_create(fe_ctx* ctx, const fe_scene_desc* d, fe_scene** out) {
    if (!ctx || !d || !out) return set_err(ctx, FE_E_INVALID, "NULL argument");
    *out = nullptr;
    for (int a = 0; a < 3; ++a) {
        if (d->n[a] < 2 || d->n[a] > 2048) return set_err(ctx, FE_E_INVALID, "grid size per axis must be in [2, 2048]");
        if (!(d->size[a] > 0.0)) return set_err(ctx, FE_E_INVALID, "domain size must be positive");
    }
    if ((int64_t)d->n[0] * d->n[1] * d->n[2] > (int64_t)1 << 30)
        return set_err(ctx, FE_E_INVALID, "grid has more than 2^30 cells");
    if (d->method < FE_PIC || d->method > FE_APIC) return set_err(ctx, FE_E_INVALID, "unknown method");
    if (check_viscosity(ctx, d->viscosity) != FE_OK) return FE_E_INVALID;
    if (d->max_particles < 0 || d->max_particles > ((int64_t)1 << 31) - 1)
        return set_err(ctx, FE_E_INVALID, "max_particles out of range");
    if (!(d->max_cfl >= 0.0) || d->max_cfl > 64) return set_err(ctx, FE_E_INVALID, "max_cfl out of range");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    auto* s = new fe_scene();
    s->ctx = ctx;
    s->desc = *d;
    build_geom(s);
    s->depth = (int)std::ceil(d->max_cfl);                               // ceil(maxCfl), src/pic_simulator.cpp:491
    s->move_substeps = (int)(unsigned)std::max(d->max_cfl, 1.0);         // src/pic_simulator.cpp:255
    const Geom& g = s->g;
    const int64_t nc = cell_count(g);
    const int64_t nf[3] = {face_count(g, 0), face_count(g, 1), face_count(g, 2)};
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t r) {
        if (r != hipSuccess && e == hipSuccess) e = r;
    };
    chk(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    for (int c = 0; c < 3; ++c) {
        chk(dalloc(&s->u[c], nf[c]));
        chk(dalloc(&s->snap[c], nf[c]));
        chk(dalloc(&s->quad[c], nf[c]));
        if (d->method == FE_FLIP) chk(dalloc(&s->squad[c], nf[c]));
        chk(dalloc(&s->mark[c], nf[c]));
    }
    chk(dalloc(&s->valid[0], nf[0] + nf[1] + nf[2]));
    chk(dalloc(&s->valid[1], nf[0] + nf[1] + nf[2]));
    chk(dalloc(&s->sdf, nc));
    chk(dalloc(&s->flags, nc));
    chk(dalloc(&s->code, nc));
    chk(dalloc(&s->p, nc));
    chk(dalloc(&s->p2, nc));
    chk(dalloc(&s->b, nc));
    chk(dalloc(&s->r, nc));
    chk(dalloc(&s->z, nc));
    chk(dalloc(&s->d[0], nc));
    chk(dalloc(&s->d[1], nc));
    chk(dalloc(&s->q, nc));
    // CG: 2 + 2 partials per block, multigrid: 1 more (r.z of the V-cycle)
    s->npartials = solver_shape_n(nc).blocks * 6;
    chk(dalloc(&s->partials, s->npartials));
    chk(dalloc(&s->chunk_fluid, solver_shape_n(nc).blocks));
    chk(dalloc(&s->sc, 2));  // two state slots (CG alternates them)
    chk(hipHostMalloc((void**)&s->sc_host, sizeof(SolverScalars), hipHostMallocDefault));
    chk(dalloc(&s->cflmax, 4));
    chk(dalloc(&s->dt_dev, 1));
    chk(hipHostMalloc((void**)&s->res_host, 2 * sizeof(double), hipHostMallocDefault));
    const int64_t cap = std::max<int64_t>(d->max_particles, 1);
    s->cap = cap;
    for (int b = 0; b < 2; ++b) {
        chk(dalloc(&s->P[b].x, 3 * cap));
        chk(dalloc(&s->P[b].v, 3 * cap));
        chk(dalloc(&s->P[b].pid, cap));
        if (d->method == FE_APIC) chk(dalloc(&s->P[b].C, 9 * cap));
    }
    chk(dalloc(&s->pos4, cap));
    chk(dalloc(&s->key, cap));
    chk(dalloc(&s->order, cap));
    chk(dalloc(&s->cell_count, nc + 1));
    chk(dalloc(&s->cell_start, nc + 1));
    chk(dalloc(&s->cursor, nc));
    chk(dalloc(&s->scan_sums, (nc + 1023) / 1024 + 1));
    chk(dalloc(&s->p2g_bound, 3 * kP2GBoundSlots));
    for (int i = 0; i <= kMaxStages; ++i) chk(hipEventCreate(&s->ev[i]));
    if (e != hipSuccess) {
        int rc = hip_fail(ctx, e, "scene allocation");
        fe_scene_destroy(s);
        return rc == FE_E_HIP && e == hipErrorOutOfMemory ? FE_E_NOMEM : rc;
    }
    // zero state (fill kernels, no memset nodes)
    for (int c = 0; c < 3; ++c) {
        chk(launch_fill_f32(s->stream, s->u[c], 0.0f, nf[c]));
        chk(launch_fill_f32(s->stream, s->snap[c], 0.0f, nf[c]));
        chk(launch_fill_u8(s->stream, s->mark[c], 0, nf[c]));
    }
    chk(hipMemsetAsync(s->p, 0, nc * sizeof(double), s->stream));
    chk(hipMemsetAsync(s->p2, 0, nc * sizeof(double), s->stream));
    chk(launch_fill_f32(s->stream, s->b, 0.0f, nc));
    chk(launch_fill_f32(s->stream, s->d[0], 0.0f, nc));
    chk(launch_fill_f32(s->stream, s->d[1], 0.0f, nc));
    chk(launch_fill_u8(s->stream, s->flags, FE_AIR, nc));
    chk(hipMemsetAsync(s->sdf, 0, nc * sizeof(double), s->stream));
    chk(hipMemsetAsync(s->sc, 0, 2 * sizeof(SolverScalars), s->stream));
    chk(hipMemsetAsync(s->cflmax, 0, 4 * sizeof(unsigned long long), s->stream));
    chk(hipStreamSynchronize(s->stream));
    if (e != hipSuccess) {
        int rc = hip_fail(ctx, e, "scene initialisation");
        fe_scene_destroy(s);
        return rc;
    }
    *out = s;
    return FE_OK;
}

int fe_scene_set_params(fe_scene* s, double blend, double max_cfl, const double gravity[3], double viscosity,
                        double density) {
    if (!s || !gravity) return FE_E_INVALID;
    fe_ctx* ctx = s->ctx;
    if (check_viscosity(ctx, viscosity) != FE_OK) return FE_E_INVALID;
    if (!(max_cfl >= 0.0) || max_cfl > 64) return set_err(ctx, FE_E_INVALID, "max_cfl out of range");
    if (!(blend >= 0.0 && blend <= 1.0)) return set_err(ctx, FE_E_INVALID, "blend must be in [0, 1]");
    s->desc.blend = blend;
    s->desc.max_cfl = max_cfl;
    for (int a = 0; a < 3; ++a) s->desc.gravity[a] = gravity[a];
    s->desc.viscosity = viscosity;
    s->desc.density = density;
    s->depth = (int)std::ceil(max_cfl);                        // ceil(maxCfl), src/pic_simulator.cpp:491
    s->move_substeps = (int)(unsigned)std::max(max_cfl, 1.0);  // src/pic_simulator.cpp:255
    ++s->graph_gen;  // captured substeps hold gravity, blend and the step counts by value
    return FE_OK;
}

int fe_scene_destroy(fe_scene* s) {
    if (!s) return FE_OK;
    if (s->ctx) (void)hipSetDevice(s->ctx->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    auto F = [](void* p) {
        if (p) (void)hipFree(p);
    };
    for (int c = 0; c < 3; ++c) {
        F(s->u[c]);
        F(s->snap[c]);
        F(s->quad[c]);
        F(s->squad[c]);
        F(s->mark[c]);
    }
    F(s->valid[0]); F(s->valid[1]);
    F(s->sdf); F(s->flags); F(s->code); F(s->p); F(s->b); F(s->r); F(s->z); F(s->p2);
    F(s->d[0]); F(s->d[1]); F(s->q); F(s->partials); F(s->chunk_fluid); F(s->sc); F(s->cflmax);
    mg_free(s);
    diffusion_free(s);
    surface_free(s);
    collider_free(s);
    sources_free(s);
    if (s->sc_host) (void)hipHostFree(s->sc_host);
    if (s->res_host) (void)hipHostFree(s->res_host);
    F(s->dt_dev);
    graph_release(s->graph[0]);
    graph_release(s->graph[1]);
    for (int b = 0; b < 2; ++b) {
        F(s->P[b].x); F(s->P[b].v); F(s->P[b].C); F(s->P[b].pid);
    }
    F(s->pos4); F(s->key); F(s->order); F(s->cell_count); F(s->cell_start); F(s->cursor); F(s->scan_sums);
    F(s->p2g_bound);
    for (int i = 0; i <= kMaxStages; ++i)
        if (s->ev[i]) (void)hipEventDestroy(s->ev[i]);
    for (auto e : s->kt_ev)
        if (e) (void)hipEventDestroy(e);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return FE_OK;
}

int fe_scene_sync(fe_scene* s) {
    if (!s) return FE_E_INVALID;
    HIPCHK(s->ctx, hipStreamSynchronize(s->stream));
    return FE_OK;
}

int fe_particles_upload(fe_scene* s, const float* x, const float* v, const float* C, int64_t count) {
    if (!s) return FE_E_INVALID;
    fe_ctx* ctx = s->ctx;
    if (count < 0 || count > s->cap)
        return set_err(ctx, FE_E_INVALID,
                       "particle count " + std::to_string(count) + " exceeds max_particles " + std::to_string(s->cap));
    if (count > 0 && (!x || !v)) return set_err(ctx, FE_E_INVALID, "x and v are required");
    for (int64_t i = 0; i < 3 * count; ++i)
        if (!std::isfinite(x[i]) || !std::isfinite(v[i]))
            return set_err(ctx, FE_E_NONFINITE, "particle " + std::to_string(i / 3) + " has a non-finite " +
                                                    (std::isfinite(x[i]) ? "velocity" : "position"));
    if (C && s->desc.method == FE_APIC)
        for (int64_t i = 0; i < 9 * count; ++i)
            if (!std::isfinite(C[i]))
                return set_err(ctx, FE_E_NONFINITE, "particle " + std::to_string(i / 9) + " has a non-finite C");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    s->np = count;
    Particles& P = s->P[s->cur];
    Particles& T = s->P[1 - s->cur];  // staging
    if (count > 0) {
        HIPCHK(ctx, hipMemcpyAsync(T.x, x, sizeof(float) * 3 * count, hipMemcpyHostToDevice, s->stream));
        HIPCHK(ctx, launch_deinterleave(s, T.x, 3, P.x));
        HIPCHK(ctx, hipMemcpyAsync(T.v, v, sizeof(float) * 3 * count, hipMemcpyHostToDevice, s->stream));
        HIPCHK(ctx, launch_deinterleave(s, T.v, 3, P.v));
        if (s->desc.method == FE_APIC) {
            if (C) {
                HIPCHK(ctx, hipMemcpyAsync(T.C, C, sizeof(float) * 9 * count, hipMemcpyHostToDevice, s->stream));
                HIPCHK(ctx, launch_deinterleave(s, T.C, 9, P.C));
            } else {
                HIPCHK(ctx, launch_fill_f32(s->stream, P.C, 0.0f, 9 * s->cap));
            }
        }
        HIPCHK(ctx, launch_iota(s, P.pid, count));
    }
    s->sorted = false;
    s->keys_ready = false;
    s->sdf_from_cells = 0;  // the bins no longer hold the particles the SDF came from
    HIPCHK(ctx, hipStreamSynchronize(s->stream));
    return FE_OK;
}

int fe_particles_download(fe_scene* s, float* x, float* v, float* C, int64_t count) {
    if (!s) return FE_E_INVALID;
    fe_ctx* ctx = s->ctx;
    if (count != s->np) return set_err(ctx, FE_E_INVALID, "count must equal the particle count");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (count == 0) return FE_OK;
    Particles& P = s->P[s->cur];
    Particles& T = s->P[1 - s->cur];
    if (x) {
        HIPCHK(ctx, launch_unpermute(s, P.x, 3, T.x));
        HIPCHK(ctx, hipMemcpyAsync(x, T.x, sizeof(float) * 3 * count, hipMemcpyDeviceToHost, s->stream));
    }
    if (v) {
        HIPCHK(ctx, launch_unpermute(s, P.v, 3, T.v));
        HIPCHK(ctx, hipMemcpyAsync(v, T.v, sizeof(float) * 3 * count, hipMemcpyDeviceToHost, s->stream));
    }
    if (C) {
        if (s->desc.method != FE_APIC) {
            std::memset(C, 0, sizeof(float) * 9 * count);
        } else {
            HIPCHK(ctx, launch_unpermute(s, P.C, 9, T.C));
            HIPCHK(ctx, hipMemcpyAsync(C, T.C, sizeof(float) * 9 * count, hipMemcpyDeviceToHost, s->stream));
        }
    }
    HIPCHK(ctx, hipStreamSynchronize(s->stream));
    return FE_OK;
}

int fe_particles_count(const fe_scene* s, int64_t* count) {
    if (!s || !count) return FE_E_INVALID;
    *count = s->np;
    return FE_OK;
}

int fe_particle_cells(fe_scene* s, int32_t* cells, int64_t count) {
    if (!s || (!cells && count > 0)) return FE_E_INVALID;
    fe_ctx* ctx = s->ctx;
    if (count != s->np) return set_err(ctx, FE_E_INVALID, "count must equal the particle count");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (count == 0) return FE_OK;
    int32_t* tmp = (int32_t*)s->P[1 - s->cur].x;  // staging (3*cap floats >= cap ints)
    HIPCHK(ctx, launch_cells_by_pid(s, tmp));
    HIPCHK(ctx, hipMemcpyAsync(cells, tmp, sizeof(int32_t) * count, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(ctx, hipStreamSynchronize(s->stream));
    s->keys_ready = false;  // the key buffer was overwritten
    return FE_OK;
}

int fe_substep(fe_scene* s, double dt, const fe_solver_desc* d, fe_solve_stats* st) {
    if (!s) return FE_E_INVALID;
    fe_ctx* ctx = s->ctx;
    int rc = check_solver(ctx, d);
    if (rc != FE_OK) return rc;
    if (!(dt >= 0.0) || !std::isfinite(dt)) return set_err(ctx, FE_E_INVALID, "dt must be finite and >= 0");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    return substep(s, dt, d, st);
}

int fe_cfl(fe_scene* s, double dt, double* out) {
    if (!s || !out) return FE_E_INVALID;
    fe_ctx* ctx = s->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, launch_cfl(s, dt));
    unsigned long long bits[2] = {0, 0};
    HIPCHK(ctx, hipMemcpyAsync(bits, s->cflmax + 2 * (s->cfl_slot & 1), sizeof(bits), hipMemcpyDeviceToHost,
                               s->stream));
    HIPCHK(ctx, hipStreamSynchronize(s->stream));
    ++s->cfl_slot;
    double maxv;
    std::memcpy(&maxv, &bits[0], sizeof(double));
    const double minH = std::min(s->g.hd[0], std::min(s->g.hd[1], s->g.hd[2]));
    *out = maxv * dt / minH;
    if (bits[1])
        return set_err(ctx, FE_E_NONFINITE, "a face velocity is not finite (NaN or Inf): the state has diverged");
    return FE_OK;
}

int fe_advance_frame(fe_scene* s, double frame_dt, const fe_solver_desc* d, int32_t* substeps_out) {
    if (!s) return FE_E_INVALID;
    fe_ctx* ctx = s->ctx;
    int rc = check_solver(ctx, d);
    if (rc != FE_OK) return rc;
    if (!(frame_dt >= 0.0) || !std::isfinite(frame_dt)) return set_err(ctx, FE_E_INVALID, "frame_dt must be >= 0");
    // PhysicsAnimation::AdvanceTimeStep (src/physics_animation.cpp:64-92)
    double remaining = frame_dt;
    int32_t count = 0;
    while (remaining > 0.0000000001) {
        double cfl = 0.0;
        if ((rc = fe_cfl(s, remaining, &cfl)) != FE_OK) return rc;
        // PICSimulator::NumberOfSubTimeSteps (src/pic_simulator.cpp:349-354)
        const double ratio = s->desc.max_cfl > 0 ? std::ceil(cfl / s->desc.max_cfl) : 1.0;
        const unsigned n = (unsigned)std::max(ratio, 1.0);
        if (n > 100000) return set_err(ctx, FE_E_INVALID, "CFL demands more than 1e5 substeps (unstable state?)");
        const double sub = remaining / (double)n;
        if ((rc = substep(s, sub, d, nullptr)) != FE_OK) return rc;
        remaining -= sub;
        ++count;
    }
    if (substeps_out) *substeps_out = count;
    return FE_OK;
}

// ---------------------------------------------------------------------------
// Batches of independent scenes.  Every scene owns a stream; host worker
// threads take scenes largest first from a shared counter and advance each by
// one frame, so the kernels of different scenes overlap on the device's
// hardware queues and the host-side waits of one scene (solver checks, CFL)
// hide behind the work of the others.
// ---------------------------------------------------------------------------
int fe_batch_create(fe_ctx* ctx, fe_scene* const* scenes, int32_t count, int32_t workers, fe_batch** out) {
    if (!ctx || !out || (count > 0 && !scenes) || count < 0) return set_err(ctx, FE_E_INVALID, "NULL argument");
    *out = nullptr;
    if (workers < 1 || workers > 64) return set_err(ctx, FE_E_INVALID, "workers must be in [1, 64]");
    auto* b = new fe_batch();
    b->ctx = ctx;
    b->workers = workers;
    for (int32_t i = 0; i < count; ++i) {
        if (!scenes[i] || scenes[i]->ctx != ctx) {
            delete b;
            return set_err(ctx, FE_E_INVALID, "batch scene " + std::to_string(i) + " is NULL or of another context");
        }
        for (int32_t j = 0; j < i; ++j)
            if (scenes[j] == scenes[i]) {
                delete b;
                return set_err(ctx, FE_E_INVALID, "scene listed twice in a batch");
            }
        b->scenes.push_back(scenes[i]);
    }
    // largest scenes first: the longest frames start early (LPT order)
    b->order.resize(b->scenes.size());
    for (size_t i = 0; i < b->order.size(); ++i) b->order[i] = (int32_t)i;
    std::stable_sort(b->order.begin(), b->order.end(), [&](int32_t x, int32_t y) {
        const fe_scene* a = b->scenes[x];
        const fe_scene* c = b->scenes[y];
        return cell_count(a->g) + a->np > cell_count(c->g) + c->np;
    });
    *out = b;
    return FE_OK;
}

int fe_batch_destroy(fe_batch* b) {
    delete b;
    return FE_OK;
}

int fe_batch_advance_frame(fe_batch* b, double frame_dt, const fe_solver_desc* d, int32_t* substeps, double* ms) {
    if (!b) return FE_E_INVALID;
    fe_ctx* ctx = b->ctx;
    int rc = check_solver(ctx, d);
    if (rc != FE_OK) return rc;
    const int n = (int)b->scenes.size();
    std::atomic<int> next{0};
    std::vector<int> rcs(n, FE_OK);
    auto work = [&]() {
        if (hipSetDevice(ctx->device) != hipSuccess) return;
        for (;;) {
            const int t = next.fetch_add(1);
            if (t >= n) return;
            const int i = b->order[t];
            const auto t0 = std::chrono::steady_clock::now();
            int32_t sub = 0;
            rcs[i] = fe_advance_frame(b->scenes[i], frame_dt, d, &sub);
            const auto t1 = std::chrono::steady_clock::now();
            if (substeps) substeps[i] = sub;
            if (ms) ms[i] = std::chrono::duration<double, std::milli>(t1 - t0).count();
        }
    };
    const int nw = std::min(b->workers, std::max(n, 1));
    std::vector<std::thread> pool;
    for (int w = 1; w < nw; ++w) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    if (next.load() < n) return set_err(ctx, FE_E_HIP, "hipSetDevice failed on a batch worker thread");
    for (int i = 0; i < n; ++i)
        if (rcs[i] != FE_OK) return rcs[i];
    return FE_OK;
}

int fe_run_stage(fe_scene* s, int32_t stage, double dt, const fe_solver_desc* d, fe_solve_stats* st) {
    if (!s) return FE_E_INVALID;
    fe_ctx* ctx = s->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int rc = FE_OK;
    StageClock clk(s);
    clk.begin(stage >= 0 && stage < kMaxStages ? stage : 0);
    switch (stage) {
        case FE_STAGE_SORT:
            HIPCHK(ctx, launch_sort(s, false));
            break;
        case FE_STAGE_P2G:
            if ((rc = ensure_sorted(s)) != FE_OK) return rc;
            HIPCHK(ctx, launch_p2g(s));
            break;
        case FE_STAGE_SDF:
            if ((rc = ensure_sorted(s)) != FE_OK) return rc;
            HIPCHK(ctx, launch_sdf(s));
            HIPCHK(ctx, launch_markers(s));
            break;
        case FE_STAGE_MARKERS:
            HIPCHK(ctx, launch_markers(s));
            break;
        case FE_STAGE_EXTRAPOLATE:
            HIPCHK(ctx, launch_extrapolate(s));
            break;
        case FE_STAGE_BC:
            HIPCHK(ctx, launch_bc(s));
            break;
        case FE_STAGE_GRAVITY:
            HIPCHK(ctx, launch_gravity(s, dt, false));
            break;
        case FE_STAGE_PRESSURE:
            if ((rc = check_solver(ctx, d)) != FE_OK) return rc;
            HIPCHK(ctx, launch_markers(s));
            return pressure(s, d, false, st, 0);
        case FE_STAGE_G2P:
            HIPCHK(ctx, launch_g2p_move(s, dt, true, false, false));
            break;
        case FE_STAGE_DIFFUSION: {
            const fe_solver_desc* dd = d ? d : &s->diff_desc;
            if ((rc = check_diffusion_solver(ctx, dd)) != FE_OK) return rc;
            HIPCHK(ctx, launch_diffusion(s, dt, dd));
            break;
        }
        case FE_STAGE_MOVE:
            HIPCHK(ctx, launch_g2p_move(s, dt, false, true, false));
            HIPCHK(ctx, launch_collide(s));
            break;
        case FE_STAGE_EMIT:
            if ((rc = emit(s)) != FE_OK) return rc;
            break;
        default:
            return set_err(ctx, FE_E_INVALID, "unknown stage " + std::to_string(stage));
    }
    s->keys_ready = false;
    clk.end();
    HIPCHK(ctx, hipStreamSynchronize(s->stream));
    clk.collect();
    kt_flush(s);
    return FE_OK;
}

int fe_last_solve_stats(const fe_scene* s, fe_solve_stats* st) {
    if (!s || !st) return FE_E_INVALID;
    *st = s->last_stats;
    return FE_OK;
}

int fe_scene_add_mesh_collider(fe_scene* s, const int32_t grid_size[3], const double grid_origin[3],
                               const double grid_spacing[3], const double* vertices, int64_t vertex_count,
                               const int32_t* triangles, int64_t triangle_count) {
    if (!s) return FE_E_INVALID;
    fe_ctx* ctx = s->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    NodeGrid3 G{};
    double* c = nullptr;
    int rc = mesh_sdf_device(ctx, s->stream, grid_size, grid_origin, grid_spacing, vertices, vertex_count, triangles,
                             triangle_count, G, &c);
    if (rc != FE_OK) return rc;
    hipError_t e = collider_merge(s, G, c);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(c);
    if (e != hipSuccess) return hip_fail(ctx, e, "collider_merge");
    return FE_OK;
}

int fe_mesh_sdf(fe_ctx* ctx, const int32_t grid_size[3], const double grid_origin[3], const double grid_spacing[3],
                const double* vertices, int64_t vertex_count, const int32_t* triangles, int64_t triangle_count,
                double* sdf) {
    if (!ctx || !sdf) return FE_E_INVALID;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = nullptr;
    HIPCHK(ctx, hipStreamCreate(&st));
    NodeGrid3 G{};
    double* c = nullptr;
    int rc = mesh_sdf_device(ctx, st, grid_size, grid_origin, grid_spacing, vertices, vertex_count, triangles,
                             triangle_count, G, &c);
    if (rc == FE_OK) {
        const size_t bytes = sizeof(double) * (size_t)G.n[0] * G.n[1] * G.n[2];
        hipError_t e = hipMemcpy(sdf, c, bytes, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(ctx, e, "fe_mesh_sdf download");
    }
    if (c) (void)hipFree(c);
    (void)hipStreamDestroy(st);
    return rc;
}

// fe_mesh_sdf with every device buffer in the device's scratch array
// (fe_noalloc.hip): no device allocation, no stream object, synchronous.
int fe_mesh_sdf_noalloc(fe_ctx* ctx, const int32_t grid_size[3], const double grid_origin[3],
                        const double grid_spacing[3], const double* vertices, int64_t vertex_count,
                        const int32_t* triangles, int64_t triangle_count, double* sdf) {
    if (!ctx || !sdf || !grid_size || !grid_origin || !grid_spacing || vertex_count < 0 || triangle_count < 0 ||
        (vertex_count > 0 && !vertices) || (triangle_count > 0 && !triangles))
        return set_err(ctx, FE_E_INVALID, "fe_mesh_sdf_noalloc: null pointer or negative count");
    NodeGrid3 G{};
    for (int a = 0; a < 3; ++a) {
        if (grid_size[a] < 1 || !(grid_spacing[a] > 0.0) || !std::isfinite(grid_spacing[a]) ||
            !std::isfinite(grid_origin[a]))
            return set_err(ctx, FE_E_INVALID,
                           "fe_mesh_sdf_noalloc: grid sizes must be >= 1, spacings finite and > 0");
        G.n[a] = grid_size[a];
        G.o[a] = grid_origin[a];
        G.h[a] = grid_spacing[a];
    }
    const int64_t nn = (int64_t)grid_size[0] * grid_size[1] * grid_size[2];
    if (nn > kNaMaxNodes || vertex_count > kNaMaxVerts || triangle_count > kNaMaxTris)
        return set_err(ctx, FE_E_INVALID,
                       "fe_mesh_sdf_noalloc: at most " + std::to_string(kNaMaxNodes) + " grid nodes, " +
                           std::to_string(kNaMaxVerts) + " vertices and " + std::to_string(kNaMaxTris) +
                           " triangles per call (fe_mesh_sdf has no such limit)");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    NoallocSlices sl{};
    HIPCHK(ctx, noalloc_slices(ctx->device, &sl));
    // Mesh2SDF::Build's fill value (largest spacing times the summed sizes), as mesh_prepare
    const double hx = G.h[0], hy = G.h[1], hz = G.h[2];
    const double hmax = (hx >= hy && hx >= hz) ? hx : ((hy >= hx && hy >= hz) ? hy : hz);
    const double fill = hmax * (double)((size_t)G.n[0] + (size_t)G.n[1] + (size_t)G.n[2]);
    const hipStream_t st = nullptr;  // the null stream: every copy below is synchronous
    if (triangle_count > 0) {
        const int32_t zero = 0;
        HIPCHK(ctx, hipMemcpy(sl.bad, &zero, sizeof(zero), hipMemcpyHostToDevice));
        if (vertex_count > 0)
            HIPCHK(ctx, hipMemcpy(sl.verts, vertices, sizeof(double) * 3 * (size_t)vertex_count,
                                  hipMemcpyHostToDevice));
        HIPCHK(ctx, hipMemcpy(sl.tris, triangles, sizeof(int32_t) * 3 * (size_t)triangle_count,
                              hipMemcpyHostToDevice));
        HIPCHK(ctx, launch_mesh_prepare(st, G, sl.verts, vertex_count, sl.tris, triangle_count, sl.tv, sl.tb, sl.bad));
        int32_t flag = 0;
        HIPCHK(ctx, hipMemcpy(&flag, sl.bad, sizeof(flag), hipMemcpyDeviceToHost));
        if (flag)
            return set_err(ctx, FE_E_INVALID,
                           "mesh collider: a triangle's vertex index is out of range, or a coordinate / h.x does not "
                           "fit an int");
    }
    HIPCHK(ctx, launch_mesh_sdf(st, G, sl.tv, sl.tb, triangle_count, fill, sl.sdf));
    HIPCHK(ctx, hipMemcpy(sdf, sl.sdf, sizeof(double) * (size_t)nn, hipMemcpyDeviceToHost));
    return FE_OK;
}

int fe_scene_initialize_from_mesh(fe_scene* s, const double* vertices, int64_t vertex_count,
                                  const int32_t* triangles, int64_t triangle_count, int32_t particles_per_block,
                                  uint64_t seed) {
    if (!s) return FE_E_INVALID;
    fe_ctx* ctx = s->ctx;
    if (particles_per_block < 0) return set_err(ctx, FE_E_INVALID, "particles_per_block must be >= 0");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const Geom& g = s->g;
    const int32_t size[3] = {g.n[0], g.n[1], g.n[2]};
    // the fluid SDF grid: origin o + h/2 (src/hybrid_simulator.cpp:13-15)
    const double so[3] = {g.od[0] + g.hd[0] / 2.0, g.od[1] + g.hd[1] / 2.0, g.od[2] + g.hd[2] / 2.0};
    NodeGrid3 G{};
    double* c = nullptr;
    int rc = mesh_sdf_device(ctx, s->stream, size, so, g.hd, vertices, vertex_count, triangles, triangle_count, G, &c);
    if (rc != FE_OK) return rc;
    hipError_t e = hipMemcpyAsync(s->sdf, c, sizeof(double) * cell_count(g), hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(c);
    if (e != hipSuccess) return hip_fail(ctx, e, "fe_scene_initialize_from_mesh");
    bool over = false;
    e = launch_initialize_from_sdf(s, particles_per_block, seed, &over);
    if (over)
        return set_err(ctx, FE_E_INVALID, "particles from the mesh exceed 2^31 - 1 particles");
    HIPCHK(ctx, e);
    HIPCHK(ctx, hipStreamSynchronize(s->stream));
    return FE_OK;
}

int fe_scene_add_volume_emitter(fe_scene* s, const int32_t resolution[3], const double domain_origin[3],
                                const double domain_size[3], const double* vertices, int64_t vertex_count,
                                const int32_t* triangles, int64_t triangle_count, int32_t particles_per_block,
                                const double velocity[3], uint64_t seed) {
    if (!s) return FE_E_INVALID;
    fe_ctx* ctx = s->ctx;
    if (!resolution || !domain_origin || !domain_size || !velocity)
        return set_err(ctx, FE_E_INVALID, "volume emitter: null argument");
    for (int a = 0; a < 3; ++a) {
        if (resolution[a] != s->g.n[a])
            return set_err(ctx, FE_E_INVALID,
                           "volume emitter: its resolution must equal the scene's (Emitt reads the fluid and collider "
                           "SDFs at the emitter's indices)");
        if (!std::isfinite(velocity[a])) return set_err(ctx, FE_E_INVALID, "volume emitter: velocity must be finite");
    }
    if (particles_per_block < 0) return set_err(ctx, FE_E_INVALID, "particles_per_block must be >= 0");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    // src/emmiter/volume_emmiter.cpp:15-19: spacing = domain size / resolution, origin = domain origin
    const double h[3] = {domain_size[0] / resolution[0], domain_size[1] / resolution[1],
                         domain_size[2] / resolution[2]};
    Emitter em;
    int rc = mesh_sdf_device(ctx, s->stream, resolution, domain_origin, h, vertices, vertex_count, triangles,
                             triangle_count, em.grid, &em.sdf);
    if (rc != FE_OK) return rc;
    em.ppb = particles_per_block;
    for (int a = 0; a < 3; ++a) em.vel[a] = velocity[a];
    em.seed = seed;
    s->emitters.push_back(em);
    return FE_OK;
}

int fe_scene_get_tuning(const fe_scene* s, fe_tuning* out) {
    if (!s || !out) return FE_E_INVALID;
    *out = s->tune;
    return FE_OK;
}

int fe_scene_set_tuning(fe_scene* s, const fe_tuning* t) {
    if (!s || !t) return FE_E_INVALID;
    if (t->p2g < FE_P2G_TILES || t->p2g > FE_P2G_SCATTER || t->g2p < FE_G2P_QUADS || t->g2p > FE_G2P_PAIRS ||
        t->mg < FE_MG_FUSED || t->mg > FE_MG_SPLIT)
        return set_err(s->ctx, FE_E_INVALID, "unknown kernel variant");
    s->tune = *t;
    ++s->graph_gen;
    return FE_OK;
}

int fe_scene_autotune(fe_scene* s, int32_t reps, fe_tuning* out, fe_tuning_times* times) {
    if (!s) return FE_E_INVALID;
    if (reps < 1 || reps > 1000) return set_err(s->ctx, FE_E_INVALID, "reps must be in [1, 1000]");
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    HIPCHK(s->ctx, autotune(s, reps, times));
    ++s->graph_gen;
    if (out) *out = s->tune;
    return FE_OK;
}

int fe_scene_reserve_particles(fe_scene* s, int64_t capacity) {
    if (!s) return FE_E_INVALID;
    if (capacity < 0 || capacity > INT32_MAX)
        return set_err(s->ctx, FE_E_INVALID, "capacity must be in [0, 2^31 - 1]");
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    HIPCHK(s->ctx, reserve_particles(s, capacity));
    ++s->graph_gen;
    return FE_OK;
}

int fe_scene_particle_capacity(const fe_scene* s, int64_t* capacity) {
    if (!s || !capacity) return FE_E_INVALID;
    *capacity = s->cap;
    return FE_OK;
}

int fe_scene_add_volume_emitter_sdf(fe_scene* s, const int32_t resolution[3], const double domain_origin[3],
                                    const double domain_size[3], const double* sdf, int32_t particles_per_block,
                                    const double velocity[3], uint64_t seed) {
    if (!s) return FE_E_INVALID;
    fe_ctx* ctx = s->ctx;
    if (!resolution || !domain_origin || !domain_size || !velocity || !sdf)
        return set_err(ctx, FE_E_INVALID, "volume emitter: null argument");
    for (int a = 0; a < 3; ++a) {
        if (resolution[a] != s->g.n[a])
            return set_err(ctx, FE_E_INVALID, "volume emitter: its resolution must equal the scene's");
        if (!std::isfinite(velocity[a])) return set_err(ctx, FE_E_INVALID, "volume emitter: velocity must be finite");
    }
    if (particles_per_block < 0) return set_err(ctx, FE_E_INVALID, "particles_per_block must be >= 0");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    Emitter em;
    for (int a = 0; a < 3; ++a) {
        em.grid.n[a] = resolution[a];
        em.grid.o[a] = domain_origin[a];
        em.grid.h[a] = domain_size[a] / resolution[a];
        em.vel[a] = velocity[a];
    }
    const size_t bytes = sizeof(double) * (size_t)cell_count(s->g);
    HIPCHK(ctx, hipMalloc((void**)&em.sdf, bytes));
    hipError_t e = hipMemcpy(em.sdf, sdf, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(em.sdf);
        return hip_fail(ctx, e, "volume emitter SDF upload");
    }
    em.ppb = particles_per_block;
    em.seed = seed;
    s->emitters.push_back(em);
    return FE_OK;
}

int fe_scene_clear_emitters(fe_scene* s) {
    if (!s) return FE_E_INVALID;
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    HIPCHK(s->ctx, hipStreamSynchronize(s->stream));
    for (Emitter& em : s->emitters)
        if (em.sdf) (void)hipFree(em.sdf);
    s->emitters.clear();
    return FE_OK;
}

int fe_net_create(fe_ctx* ctx, uint64_t seed, fe_net** out) {
    if (!ctx || !out) return FE_E_INVALID;
    *out = nullptr;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    fe_net* n = new fe_net();
    n->ctx = ctx;
    hipError_t e = net_alloc(n, seed);
    if (e != hipSuccess) {
        net_free(n);
        delete n;
        return hip_fail(ctx, e, "fe_net_create");
    }
    *out = n;
    return FE_OK;
}

int fe_net_destroy(fe_net* n) {
    if (!n) return FE_OK;
    (void)hipSetDevice(n->ctx->device);
    net_free(n);
    delete n;
    return FE_OK;
}

int fe_net_param_count(int64_t* count) {
    if (!count) return FE_E_INVALID;
    *count = net_param_count();
    return FE_OK;
}

int fe_net_get_params(fe_net* n, float* params, int64_t count) {
    if (!n || !params) return FE_E_INVALID;
    if (count != net_param_count()) return set_err(n->ctx, FE_E_INVALID, "count must be fe_net_param_count");
    HIPCHK(n->ctx, hipSetDevice(n->ctx->device));
    HIPCHK(n->ctx, hipStreamSynchronize(n->stream));
    HIPCHK(n->ctx, net_get_params(n, params));
    return FE_OK;
}

int fe_net_set_params(fe_net* n, const float* params, int64_t count) {
    if (!n || !params) return FE_E_INVALID;
    if (count != net_param_count()) return set_err(n->ctx, FE_E_INVALID, "count must be fe_net_param_count");
    HIPCHK(n->ctx, hipSetDevice(n->ctx->device));
    HIPCHK(n->ctx, hipStreamSynchronize(n->stream));
    HIPCHK(n->ctx, net_set_params(n, params));
    return FE_OK;
}

static int net_grid_ok(fe_ctx* ctx, const int n[3]) {
    for (int a = 0; a < 3; ++a)
        if (n[a] < 8 || n[a] % 8 != 0)
            return set_err(ctx, FE_E_INVALID, "the warm-start network needs every grid axis a multiple of 8");
    return FE_OK;
}

int fe_net_add_sample(fe_net* n, fe_scene* s) {
    if (!n || !s) return FE_E_INVALID;
    int rc = net_grid_ok(n->ctx, s->g.n);
    if (rc != FE_OK) return rc;
    HIPCHK(n->ctx, hipSetDevice(n->ctx->device));
    hipError_t e = net_add_sample(n, s);
    if (e == hipErrorInvalidValue)
        return set_err(n->ctx, FE_E_INVALID, "the scene has no solved pressure system with FLUID cells to sample");
    HIPCHK(n->ctx, e);
    return FE_OK;
}

int fe_net_sample_count(const fe_net* n, int64_t* count) {
    if (!n || !count) return FE_E_INVALID;
    *count = (int64_t)n->samples.size();
    return FE_OK;
}

int fe_net_clear_samples(fe_net* n) {
    if (!n) return FE_E_INVALID;
    HIPCHK(n->ctx, hipSetDevice(n->ctx->device));
    HIPCHK(n->ctx, hipStreamSynchronize(n->stream));
    for (auto& smp : n->samples) {
        (void)hipFree(smp.b);
        (void)hipFree(smp.flags);
        (void)hipFree(smp.target);
        (void)hipFree(smp.code);
    }
    n->samples.clear();
    return FE_OK;
}

int fe_net_train(fe_net* n, int32_t steps, double learning_rate, uint64_t seed, float* losses) {
    if (!n) return FE_E_INVALID;
    if (steps <= 0) return set_err(n->ctx, FE_E_INVALID, "steps must be > 0");
    if (n->samples.empty()) return set_err(n->ctx, FE_E_INVALID, "the network has no samples");
    if (!(learning_rate > 0.0) || !std::isfinite(learning_rate))
        return set_err(n->ctx, FE_E_INVALID, "learning_rate must be finite and > 0");
    HIPCHK(n->ctx, hipSetDevice(n->ctx->device));
    HIPCHK(n->ctx, net_train(n, steps, (float)learning_rate, seed));
    if (losses) HIPCHK(n->ctx, hipMemcpy(losses, n->losses, sizeof(float) * steps, hipMemcpyDeviceToHost));
    return FE_OK;
}

int fe_net_set_loss(fe_net* n, int32_t mode) {
    if (!n) return FE_E_INVALID;
    if (mode != FE_NET_LOSS_MSE && mode != FE_NET_LOSS_ENERGY)
        return set_err(n->ctx, FE_E_INVALID, "unknown loss " + std::to_string(mode));
    n->loss_mode = mode;
    return FE_OK;
}

int fe_net_loss_and_grads(fe_net* n, const int32_t dims[3], const double h[3], const float* b, const uint8_t* flags,
                          const float* target, double* loss, float* grads, float* prediction) {
    if (!n || !dims || !b || !flags || !target) return FE_E_INVALID;
    int rc = net_grid_ok(n->ctx, dims);
    if (rc != FE_OK) return rc;
    fe_ctx* ctx = n->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int64_t V = (int64_t)dims[0] * dims[1] * dims[2];
    float *db = nullptr, *dt = nullptr, *dl = nullptr;
    uint8_t* df = nullptr;
    uint16_t* dc = nullptr;
    auto cleanup = [&]() {
        if (db) (void)hipFree(db);
        if (dt) (void)hipFree(dt);
        if (dl) (void)hipFree(dl);
        if (df) (void)hipFree(df);
        if (dc) (void)hipFree(dc);
    };
    hipError_t e = hipSuccess;
    if ((e = hipMalloc((void**)&db, sizeof(float) * V)) != hipSuccess ||
        (e = hipMalloc((void**)&dt, sizeof(float) * V)) != hipSuccess ||
        (e = hipMalloc((void**)&dl, sizeof(float))) != hipSuccess || (e = hipMalloc((void**)&df, V)) != hipSuccess ||
        (e = hipMalloc((void**)&dc, sizeof(uint16_t) * V)) != hipSuccess ||
        (e = hipMemcpy(db, b, sizeof(float) * V, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(dt, target, sizeof(float) * V, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(df, flags, V, hipMemcpyHostToDevice)) != hipSuccess) {
        cleanup();
        return hip_fail(ctx, e, "fe_net_loss_and_grads upload");
    }
    double ss = 0.0;
    for (int64_t v = 0; v < V; ++v) ss += (double)b[v] * b[v];
    const double rms = std::sqrt(ss / (double)V);
    LossIn in;
    double energy = 0.0;
    for (int64_t v = 0; v < V; ++v)
        if (flags[v] == FE_FLUID) {
            ++in.nfluid;
            energy += (double)b[v] * target[v];
        }
    in.target = dt;
    in.flags = df;
    in.code = dc;
    in.b = db;
    in.energy = energy;
    for (int a = 0; a < 3; ++a) {
        const double ha = h ? h[a] : 1.0 / dims[a];
        in.ih2[a] = 1.0 / (ha * ha);
    }
    if ((e = net_ensure(n, dims)) == hipSuccess && (e = net_codes(n, dims, df, dc)) == hipSuccess &&
        (e = net_forward(n, db, df, rms > 0 ? rms : 1.0)) == hipSuccess && (e = net_backward(n, in, dl)) == hipSuccess)
        e = hipStreamSynchronize(n->stream);
    if (e == hipSuccess && loss) e = hipMemcpy(loss, n->scal, sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && prediction) e = hipMemcpy(prediction, n->out, sizeof(float) * V, hipMemcpyDeviceToHost);
    if (e == hipSuccess && grads) {
        float* g = grads;
        for (auto& L : n->L) {
            const int64_t wsz = 27LL * L.cin * L.cout;
            if (e == hipSuccess) e = hipMemcpy(g, L.gw, sizeof(float) * wsz, hipMemcpyDeviceToHost);
            g += wsz;
            if (e == hipSuccess) e = hipMemcpy(g, L.gb, sizeof(float) * L.cout, hipMemcpyDeviceToHost);
            g += L.cout;
        }
        if (e == hipSuccess) e = hipMemcpy(g, n->ghead, sizeof(float) * 17, hipMemcpyDeviceToHost);
    }
    cleanup();
    HIPCHK(ctx, e);
    return FE_OK;
}

int fe_scene_set_warm_start_net(fe_scene* s, fe_net* n) {
    if (!s) return FE_E_INVALID;
    if (n) {
        int rc = net_grid_ok(s->ctx, s->g.n);
        if (rc != FE_OK) return rc;
        if (n->ctx->device != s->ctx->device)
            return set_err(s->ctx, FE_E_INVALID, "the network and the scene are on different devices");
    }
    s->net = n;
    return FE_OK;
}

int fe_net_set_scale_mode(fe_net* n, int32_t mode) {
    if (!n) return FE_E_INVALID;
    if (mode != FE_NET_SCALE_RESIDUAL && mode != FE_NET_SCALE_ENERGY)
        return set_err(n->ctx, FE_E_INVALID, "unknown scale mode " + std::to_string(mode));
    n->scale_mode = mode;
    return FE_OK;
}

int fe_scene_warm_start_scale(const fe_scene* s, double* alpha) {
    if (!s || !alpha) return FE_E_INVALID;
    *alpha = s->net_alpha;
    return FE_OK;
}

int fe_scene_clear_colliders(fe_scene* s) {
    if (!s) return FE_E_INVALID;
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    HIPCHK(s->ctx, hipStreamSynchronize(s->stream));
    collider_free(s);
    return FE_OK;
}

int fe_scene_set_sdf_search(fe_scene* s, int32_t mode) {
    if (!s) return FE_E_INVALID;
    if (mode != FE_SDF_SEARCH_CELLS && mode != FE_SDF_SEARCH_BUCKETS)
        return set_err(s->ctx, FE_E_INVALID, "unknown SDF search mode " + std::to_string(mode));
    s->sdf_buckets = mode == FE_SDF_SEARCH_BUCKETS;
    ++s->graph_gen;
    return FE_OK;
}

int fe_scene_set_graph(fe_scene* s, int32_t enable) {
    if (!s) return FE_E_INVALID;
    if (enable != 0 && enable != 1) return set_err(s->ctx, FE_E_INVALID, "enable must be 0 or 1");
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    HIPCHK(s->ctx, hipStreamSynchronize(s->stream));
    s->graph_mode = enable;
    if (!enable) {
        graph_release(s->graph[0]);
        graph_release(s->graph[1]);
    }
    return FE_OK;
}

int fe_scene_graph_stats(const fe_scene* s, int64_t* substeps, int64_t* captures, int64_t* resumed) {
    if (!s) return FE_E_INVALID;
    if (substeps) *substeps = s->graph_substeps;
    if (captures) *captures = s->graph_captures;
    if (resumed) *resumed = s->graph_resumed;
    return FE_OK;
}

int fe_surface_build(fe_scene* s, int64_t* vertex_count, int64_t* triangle_count) {
    if (!s) return FE_E_INVALID;
    fe_ctx* ctx = s->ctx;
    const Geom& g = s->g;
    const int64_t nedges = (int64_t)(g.n[0] - 1) * g.n[1] * g.n[2] + (int64_t)g.n[0] * (g.n[1] - 1) * g.n[2] +
                           (int64_t)g.n[0] * g.n[1] * (g.n[2] - 1);
    if (nedges >= ((int64_t)1 << 31) - 1)
        return set_err(ctx, FE_E_UNSUPPORTED, "surface extraction needs fewer than 2^31 node edges");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, surface_build(s));
    const McState& m = s->mc;
    if (vertex_count) *vertex_count = m.nv3 + (int64_t)m.cap_pos.size() / 3;
    if (triangle_count) *triangle_count = m.nt3 + (int64_t)m.cap_tri.size() / 3;
    return FE_OK;
}

int fe_scene_set_surface_mode(fe_scene* s, int32_t mode) {
    if (!s) return FE_E_INVALID;
    if (mode != FE_SURFACE_REFERENCE && mode != FE_SURFACE_CLOSED)
        return set_err(s->ctx, FE_E_INVALID, "unknown surface mode " + std::to_string(mode));
    s->surface_mode = mode;
    return FE_OK;
}

int fe_surface_download(fe_scene* s, double* vertices, double* normals, int32_t* triangles) {
    if (!s) return FE_E_INVALID;
    fe_ctx* ctx = s->ctx;
    if (!s->mc.built) return set_err(ctx, FE_E_INVALID, "no surface has been built on this scene");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, surface_download(s, vertices, normals, triangles));
    return FE_OK;
}

int fe_scene_set_diffusion_solver(fe_scene* s, const fe_solver_desc* d) {
    if (!s) return FE_E_INVALID;
    int rc = check_diffusion_solver(s->ctx, d);
    if (rc != FE_OK) return rc;
    s->diff_desc = *d;
    return FE_OK;
}

int fe_last_diffusion_stats(const fe_scene* s, fe_solve_stats stats[3]) {
    if (!s || !stats) return FE_E_INVALID;
    for (int c = 0; c < 3; ++c) stats[c] = s->diff_stats[c];
    return FE_OK;
}

int fe_kernel_timer(fe_scene* s, int32_t kernel_id) {
    if (!s) return FE_E_INVALID;
    fe_ctx* ctx = s->ctx;
    if (kernel_id < FE_KERNEL_NONE || kernel_id > FE_KERNEL_MG_UP)
        return set_err(ctx, FE_E_INVALID, "unknown kernel id " + std::to_string(kernel_id));
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(s->stream));
    kt_flush(s);
    if (kernel_id != FE_KERNEL_NONE && s->kt_ev.empty()) {
        s->kt_ev.resize(2 * kTimerPairs, nullptr);
        s->kt_tag.assign(kTimerPairs, 0);
        for (auto& e : s->kt_ev) HIPCHK(ctx, hipEventCreate(&e));
    }
    s->kt_id = kernel_id;
    s->kt_ms = 0.0;
    s->kt_launches = 0;
    return FE_OK;
}

int fe_kernel_timer_read(fe_scene* s, double* total_ms, int64_t* launches, int32_t reset) {
    if (!s) return FE_E_INVALID;
    fe_ctx* ctx = s->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(s->stream));
    kt_flush(s);
    if (total_ms) *total_ms = s->kt_ms;
    if (launches) *launches = s->kt_launches;
    if (reset) {
        s->kt_ms = 0.0;
        s->kt_launches = 0;
    }
    return FE_OK;
}

int fe_grid_bytes(const fe_scene* s, int32_t field, size_t* bytes) {
    if (!s || !bytes) return FE_E_INVALID;
    *bytes = field_bytes(s, field);
    return *bytes ? FE_OK : FE_E_INVALID;
}

int fe_grid_download(fe_scene* s, int32_t field, void* dst, size_t bytes) {
    if (!s || !dst) return FE_E_INVALID;
    fe_ctx* ctx = s->ctx;
    const size_t need = field_bytes(s, field);
    if (need == 0) return set_err(ctx, FE_E_INVALID, "unknown grid field " + std::to_string(field));
    if (bytes != need) return set_err(ctx, FE_E_INVALID, "byte count mismatch: need " + std::to_string(need));
    if (!field_ptr(s, field))
        return set_err(ctx, FE_E_INVALID, field == FE_GRID_COLLIDER_SDF ? "the scene has no collider"
                                                                     : "no diffusion solve has run on this scene");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipMemcpyAsync(dst, field_ptr(s, field), need, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(ctx, hipStreamSynchronize(s->stream));
    return FE_OK;
}

int fe_grid_upload(fe_scene* s, int32_t field, const void* src, size_t bytes) {
    if (!s || !src) return FE_E_INVALID;
    fe_ctx* ctx = s->ctx;
    const size_t need = field_bytes(s, field);
    if (need == 0) return set_err(ctx, FE_E_INVALID, "unknown grid field " + std::to_string(field));
    if (bytes != need) return set_err(ctx, FE_E_INVALID, "byte count mismatch: need " + std::to_string(need));
    if (diffusion_field(field)) return set_err(ctx, FE_E_INVALID, "diffusion markers are download-only");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (field == FE_GRID_COLLIDER_SDF && !s->csdf) {  // first collider: ColliderCollection storage
        HIPCHK(ctx, hipMalloc((void**)&s->csdf, need));
        s->has_collider = true;
    }
    HIPCHK(ctx, hipMemcpyAsync(field_ptr(s, field), src, need, hipMemcpyHostToDevice, s->stream));
    HIPCHK(ctx, hipStreamSynchronize(s->stream));
    if (field == FE_GRID_SDF) s->sdf_from_cells = 0;  // the caller's SDF is the surface's
    return FE_OK;
}

int fe_stage_times(const fe_scene* s, double* ms, int32_t count) {
    if (!s || !ms) return FE_E_INVALID;
    for (int i = 0; i < count; ++i) ms[i] = i < kMaxStages ? s->stage_ms[i] : 0.0;
    return FE_OK;
}

}  // extern "C"
