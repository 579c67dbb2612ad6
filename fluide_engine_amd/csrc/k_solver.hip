// k_solver.hip -- pressure projection of SinglePhasePressureSolver
// (src/fluid_solvers/single_phase_pressure_solver.cpp) on gfx950.
//
// The reference stores A as {center,right,up,front} doubles per cell
// (src/linear_system/linear_system.hpp:9-15, 32 B/cell).  Here A is never
// stored: it is rebuilt inside every stencil from the 1-byte cell markers with
// BuildSystem's rules (:99-164) -- center = sum of 1/h^2 over in-range,
// non-BOUNDRY neighbours; off-diagonal -1/h^2 between two FLUID cells;
// non-FLUID rows are the identity.  Vectors are FP32; every dot product and
// norm accumulates in FP64 with a deterministic two-level reduction whose last
// block also updates the solver scalars on the device, so the CG loop runs
// without host round trips.
#include <algorithm>
#include <utility>

#include "fe_kernels.hpp"

namespace fe {

// ---------------------------------------------------------------------------
// Cell coordinates and the matrix-free 7-point row.
// ---------------------------------------------------------------------------
struct Cell {
    int i, j, k;
    int64_t c;
};

__device__ inline Cell cell_of(const Geom& g, int64_t c) {
    Cell r;
    r.c = c;
    r.i = (int)(c % g.n[0]);
    r.j = (int)((c / g.n[0]) % g.n[1]);
    r.k = (int)(c / ((int64_t)g.n[0] * g.n[1]));
    return r;
}

struct Row {
    float center;
    float off[6];   // coefficient of neighbour (i-1, i+1, j-1, j+1, k-1, k+1)
    int64_t nb[6];  // neighbour index (valid when off != 0)
    bool fluid;
};

__device__ inline Row make_row(const Geom& g, const uint8_t* __restrict__ flags, const Cell& x, const float ih2[3]) {
    Row r;
    const int64_t sy = g.n[0], sz = (int64_t)g.n[0] * g.n[1];
    r.nb[0] = x.c - 1; r.nb[1] = x.c + 1;
    r.nb[2] = x.c - sy; r.nb[3] = x.c + sy;
    r.nb[4] = x.c - sz; r.nb[5] = x.c + sz;
    const bool in[6] = {x.i > 0, x.i + 1 < g.n[0], x.j > 0, x.j + 1 < g.n[1], x.k > 0, x.k + 1 < g.n[2]};
    r.fluid = flags[x.c] == FE_FLUID;
    float center = 0.0f;
#pragma unroll
    for (int q = 0; q < 6; ++q) r.off[q] = 0.0f;
    if (r.fluid) {
        // BuildSystem's accumulation order: +x, -x, +y, -y, +z, -z
        const int order[6] = {1, 0, 3, 2, 5, 4};
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            const int q = order[t];
            if (!in[q]) continue;
            const uint8_t fnb = flags[r.nb[q]];
            if (fnb == FE_BOUNDRY) continue;
            center += ih2[q >> 1];
            if (fnb == FE_FLUID) r.off[q] = -ih2[q >> 1];
        }
    } else {
        center = 1.0f;
    }
    r.center = center;
    return r;
}

// (A v)(c) in the reference SpMV order (src/linear_system/blas.cpp:72-83).
template <typename F>
__device__ inline float row_apply(const Row& r, float vc, F vnb) {
    float acc = r.center * vc;
#pragma unroll
    for (int q = 0; q < 6; ++q)
        if (r.off[q] != 0.0f) acc += r.off[q] * vnb(r.nb[q]);
    return acc;
}

// ---------------------------------------------------------------------------
// Deterministic FP64 reduction of NV values per thread.  Every block writes its
// partials; the last block to arrive (ticket counter) sums them in a fixed
// order and runs `fin` on the totals.
// ---------------------------------------------------------------------------
template <int NV, typename Fin>
__device__ inline void reduce_finish(double (&v)[NV], double* partials, uint32_t* counter, Fin fin) {
    __shared__ double sm[NV][kBlock / 64];
    __shared__ int is_last;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double x = v[q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if (lane == 0) sm[q][wid] = x;
    }
    __syncthreads();
    const int64_t nblocks = gridDim.x;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double t = 0.0;
#pragma unroll
            for (int w = 0; w < kBlock / 64; ++w) t += sm[q][w];
            partials[(int64_t)blockIdx.x * NV + q] = t;
        }
        __threadfence();
        const uint32_t ticket = atomicAdd(counter, 1u);
        is_last = (ticket == (uint32_t)(nblocks - 1));
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = 0.0;
    for (int64_t b = threadIdx.x; b < nblocks; b += kBlock) {
#pragma unroll
        for (int q = 0; q < NV; ++q)
            acc[q] += __hip_atomic_load(&partials[b * NV + q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double x = acc[q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if (lane == 0) sm[q][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double t = 0.0;
#pragma unroll
            for (int w = 0; w < kBlock / 64; ++w) t += sm[q][w];
            tot[q] = t;
        }
        fin(tot);
        *counter = 0;
    }
}

__device__ inline bool stop_rule(const SolverScalars* sc, double sigma, double rr) {
    if (sc->relative) return sqrt(rr) <= sc->tol * sqrt(sc->bnorm2);
    return !(sigma > sc->tol * sc->tol);
}

// ---------------------------------------------------------------------------
// Right-hand side: DivergenceAtCallCenter (face_centered_grid3d.cpp:157-166)
// on FLUID cells, 0 elsewhere; also ||b||^2.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_rhs(Geom g, FaceSet u, const uint8_t* __restrict__ flags,
                                                float* __restrict__ b, double* partials, SolverScalars* sc,
                                                int64_t padded) {
    const int64_t blk = xcd_remap(blockIdx.x, padded);
    const int64_t c = blk * kBlock + threadIdx.x;
    double v[1] = {0.0};
    if (c < cell_count(g)) {
        const Cell x = cell_of(g, c);
        float bb = 0.0f;
        if (flags[c] == FE_FLUID) {
            const int64_t iu = (int64_t)x.i + (int64_t)(g.n[0] + 1) * ((int64_t)x.j + (int64_t)g.n[1] * x.k);
            const int64_t iv = (int64_t)x.i + (int64_t)g.n[0] * ((int64_t)x.j + (int64_t)(g.n[1] + 1) * x.k);
            const int64_t iw = c;
            const float left = u.a[0][iu], right = u.a[0][iu + 1];
            const float down = u.a[1][iv], up = u.a[1][iv + g.n[0]];
            const float back = u.a[2][iw], front = u.a[2][iw + (int64_t)g.n[0] * g.n[1]];
            bb = (right - left) / g.h[0] + (up - down) / g.h[1] + (front - back) / g.h[2];
        }
        b[c] = bb;
        v[0] = (double)bb * (double)bb;
    }
    reduce_finish<1>(v, partials, &sc->counter, [&](const double* t) { sc->bnorm2 = t[0]; });
}


// FP64 row product with the same matrix (for residuals of the FP64 solution).
template <typename F>
__device__ inline double row_apply64(const Row& r, double xc, F xnb, const double ih2[3]) {
    if (!r.fluid) return xc;
    double center = 0.0, acc = 0.0;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        if (r.off[q] != 0.0f) acc -= ih2[q >> 1] * xnb(r.nb[q]);
    }
    // center: recount in FP64 from the FP32 row (exact: sums of identical powers of 1/h^2 per axis)
    center = (double)r.center;
    return center * xc + acc;
}

// ---------------------------------------------------------------------------
// Jacobi sweep (JacobiIterationSolver::Relax, src/linear_system/jacobi_iteration_solver.cpp:39-76):
// x' = (b - sum_offdiag A x) / center, on FP32 working copies of the FP64
// pressure.  Once the convergence check has fired the sweep only copies, so
// the host's buffer rotation stays consistent.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_jacobi(Geom g, const uint8_t* __restrict__ flags,
                                                   const float* __restrict__ b, const float* __restrict__ xo,
                                                   float* __restrict__ xn, const SolverScalars* sc, float ihx2,
                                                   float ihy2, float ihz2, int64_t padded) {
    const int64_t blk = xcd_remap(blockIdx.x, padded);
    const int64_t c = blk * kBlock + threadIdx.x;
    if (c >= cell_count(g)) return;
    if (sc->done) {
        xn[c] = xo[c];
        return;
    }
    const float ih2[3] = {ihx2, ihy2, ihz2};
    const Row r = make_row(g, flags, cell_of(g, c), ih2);
    float off = 0.0f;
#pragma unroll
    for (int q = 0; q < 6; ++q)
        if (r.off[q] != 0.0f) off += r.off[q] * xo[r.nb[q]];
    xn[c] = (b[c] - off) / r.center;
}

// ||b - A x||_2 < tol check (JacobiIterationSolver::CalculateTolerance, :78-82), FP64 sums.
__global__ __launch_bounds__(kBlock) void k_jacobi_check(Geom g, const uint8_t* __restrict__ flags,
                                                         const float* __restrict__ b, const float* __restrict__ x,
                                                         double* partials, SolverScalars* sc, float ihx2,
                                                         float ihy2, float ihz2, int it, int64_t padded) {
    if (sc->done) return;  // uniform across the grid
    const int64_t blk = xcd_remap(blockIdx.x, padded);
    const int64_t c = blk * kBlock + threadIdx.x;
    double v[1] = {0.0};
    if (c < cell_count(g)) {
        const float ih2[3] = {ihx2, ihy2, ihz2};
        const Row r = make_row(g, flags, cell_of(g, c), ih2);
        const double ih2d[3] = {ihx2, ihy2, ihz2};
        const double ax = row_apply64(r, (double)x[c], [&](int64_t n) { return (double)x[n]; }, ih2d);
        const double res = (double)b[c] - ax;
        v[0] = res * res;
    }
    reduce_finish<1>(v, partials, &sc->counter, [&](const double* t) {
        if (sqrt(t[0]) < sc->tol) {
            sc->done = 1;
            sc->converged = 1;
            sc->it = it + 1;
        }
    });
}

__global__ __launch_bounds__(kBlock) void k_f64_to_f32(const double* __restrict__ a, float* __restrict__ o,
                                                       int64_t n, int zero) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c < n) o[c] = zero ? 0.0f : (float)a[c];
}

__global__ __launch_bounds__(kBlock) void k_f32_to_f64(const float* __restrict__ a, double* __restrict__ o,
                                                       int64_t n) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c < n) o[c] = (double)a[c];
}

__global__ __launch_bounds__(kBlock) void k_zero_f64(double* __restrict__ x, int64_t n) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c < n) x[c] = 0.0;
}

// ---------------------------------------------------------------------------
// Conjugate gradient (ConjugateGradientSolver::Solve, src/linear_system/conjugate_gradient_solver.cpp:14-55)
// in two fused kernels per iteration, FP32 r, z, d, q and an FP64 solution x:
//   spmv:   d = z + beta*d_old (on the fly, also at the 6 neighbours),
//           q = A d, alpha = sigma / (d.q)
//   update: x' = alpha*d + x (FP64); r = -alpha*q + r, or r = b - A x' in FP64
//           on residual replacement (every 50 iterations or after sigma grew,
//           :32-49; x' at the neighbours is formed on the fly from x and d),
//           z = M^-1 r, sigma' = r.z, beta = sigma'/sigma, stopping rule.
// With the relative rule a stop is only accepted on a freshly replaced
// residual, so the decision is taken on b - A x, not on the recurrence.
// x and d ping-pong between two buffers; z is r itself without a preconditioner.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_cg_init(Geom g, const uint8_t* __restrict__ flags,
                                                    const float* __restrict__ b, const double* __restrict__ x,
                                                    float* __restrict__ r, float* __restrict__ z,
                                                    float* __restrict__ d0, double* partials, SolverScalars* sc,
                                                    float ihx2, float ihy2, float ihz2, int64_t padded) {
    const int64_t blk = xcd_remap(blockIdx.x, padded);
    const int64_t c = blk * kBlock + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (c < cell_count(g)) {
        const float ih2[3] = {ihx2, ihy2, ihz2};
        const double ih2d[3] = {ihx2, ihy2, ihz2};
        const Row row = make_row(g, flags, cell_of(g, c), ih2);
        const double ax = row_apply64(row, x[c], [&](int64_t n) { return x[n]; }, ih2d);
        const float rc = (float)((double)b[c] - ax);
        r[c] = rc;
        float zc = rc;
        if (sc->precond) {
            zc = rc / row.center;
            z[c] = zc;
        }
        d0[c] = 0.0f;
        v[0] = (double)rc * (double)zc;
        v[1] = (double)rc * (double)rc;
    }
    reduce_finish<2>(v, partials, &sc->counter, [&](const double* t) {
        sc->sigma = t[0];
        sc->rr = t[1];
        sc->beta = 0.0;
        sc->it = 0;
        sc->trigger = 0;
        const bool stop = stop_rule(sc, t[0], t[1]);
        sc->converged = stop ? 1 : 0;
        sc->done = (stop || sc->max_it <= 0) ? 1 : 0;
    });
}

__global__ __launch_bounds__(kBlock) void k_cg_spmv(Geom g, const uint8_t* __restrict__ flags,
                                                    const float* __restrict__ z, const float* __restrict__ dold,
                                                    float* __restrict__ dnew, float* __restrict__ q,
                                                    double* partials, SolverScalars* sc, float ihx2, float ihy2,
                                                    float ihz2, int64_t padded) {
    if (sc->done) return;  // uniform
    const int64_t blk = xcd_remap(blockIdx.x, padded);
    const int64_t c = blk * kBlock + threadIdx.x;
    double v[1] = {0.0};
    if (c < cell_count(g)) {
        const float beta = (float)sc->beta;
        const float ih2[3] = {ihx2, ihy2, ihz2};
        const Row row = make_row(g, flags, cell_of(g, c), ih2);
        auto D = [&](int64_t n) { return beta * dold[n] + z[n]; };  // AXpY(beta, d, s, &d)
        const float dc = D(c);
        const float qc = row_apply(row, dc, D);
        dnew[c] = dc;
        q[c] = qc;
        v[0] = (double)dc * (double)qc;
    }
    reduce_finish<1>(v, partials, &sc->counter, [&](const double* t) {
        sc->dq = t[0];
        sc->alpha = sc->sigma / t[0];
    });
}

__global__ __launch_bounds__(kBlock) void k_cg_update(Geom g, const uint8_t* __restrict__ flags,
                                                      const float* __restrict__ b, const double* __restrict__ xo,
                                                      double* __restrict__ xn, const float* __restrict__ d,
                                                      const float* __restrict__ q, float* __restrict__ r,
                                                      float* __restrict__ z, double* partials, SolverScalars* sc,
                                                      float ihx2, float ihy2, float ihz2, int64_t padded) {
    const int64_t blk = xcd_remap(blockIdx.x, padded);
    const int64_t c = blk * kBlock + threadIdx.x;
    if (sc->done) {  // uniform: keep the rotation consistent
        if (c < cell_count(g)) xn[c] = xo[c];
        return;
    }
    const int it = sc->it;
    const bool replace = sc->trigger || (it % 50 == 0 && it > 0);
    double v[2] = {0.0, 0.0};
    if (c < cell_count(g)) {
        const double alpha = sc->alpha;
        const float ih2[3] = {ihx2, ihy2, ihz2};
        const double ih2d[3] = {ihx2, ihy2, ihz2};
        const Row row = make_row(g, flags, cell_of(g, c), ih2);
        auto X = [&](int64_t n) { return alpha * (double)d[n] + xo[n]; };
        const double xc = X(c);
        xn[c] = xc;
        float rc;
        if (replace)
            rc = (float)((double)b[c] - row_apply64(row, xc, X, ih2d));
        else
            rc = -(float)alpha * q[c] + r[c];
        r[c] = rc;
        float zc = rc;
        if (sc->precond) {
            zc = rc / row.center;
            z[c] = zc;
        }
        v[0] = (double)rc * (double)zc;
        v[1] = (double)rc * (double)rc;
    }
    reduce_finish<2>(v, partials, &sc->counter, [&](const double* t) {
        const double sold = sc->sigma;
        sc->sigma_old = sold;
        sc->sigma = t[0];
        sc->rr = t[1];
        sc->trigger = t[0] > sold ? 1 : 0;
        sc->beta = t[0] / sold;
        sc->it = it + 1;
        bool stop = stop_rule(sc, t[0], t[1]);
        if (stop && sc->relative && !replace) {  // confirm on the true residual next iteration
            stop = false;
            sc->trigger = 1;
        }
        sc->converged = stop ? 1 : 0;
        sc->done = (stop || sc->it >= sc->max_it) ? 1 : 0;
    });
}

// Final ||b - A p||_2 and ||b||_2 in FP64 on the FP64 solution.
__global__ __launch_bounds__(kBlock) void k_final_residual(Geom g, const uint8_t* __restrict__ flags,
                                                           const float* __restrict__ b, const double* __restrict__ x,
                                                           double* partials, SolverScalars* sc, double ihx2,
                                                           double ihy2, double ihz2, int64_t padded) {
    const int64_t blk = xcd_remap(blockIdx.x, padded);
    const int64_t c = blk * kBlock + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (c < cell_count(g)) {
        const Cell cc = cell_of(g, c);
        const float ones[3] = {1.0f, 1.0f, 1.0f};
        const Row row = make_row(g, flags, cc, ones);  // structure only; weights in FP64 below
        double ax;
        if (row.fluid) {
            const double ih2[3] = {ihx2, ihy2, ihz2};
            const bool in[6] = {cc.i > 0, cc.i + 1 < g.n[0], cc.j > 0, cc.j + 1 < g.n[1], cc.k > 0, cc.k + 1 < g.n[2]};
            double center = 0.0, off = 0.0;
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                if (!in[q]) continue;
                const uint8_t fnb = flags[row.nb[q]];
                if (fnb == FE_BOUNDRY) continue;
                center += ih2[q >> 1];
                if (fnb == FE_FLUID) off -= ih2[q >> 1] * x[row.nb[q]];
            }
            ax = center * x[c] + off;
        } else {
            ax = x[c];
        }
        const double res = (double)b[c] - ax;
        v[0] = res * res;
        v[1] = (double)b[c] * (double)b[c];
    }
    reduce_finish<2>(v, partials, &sc->counter, [&](const double* t) {
        sc->rr = t[0];
        sc->bnorm2 = t[1];
    });
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static inline dim3 grid1(int64_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

hipError_t launch_final_residual(fe_scene* s) {
    const Geom& g = s->g;
    const int64_t nc = cell_count(g);
    const int64_t padded = pad_blocks((nc + kBlock - 1) / kBlock);
    const double ihd[3] = {1.0 / g.hd[0], 1.0 / g.hd[1], 1.0 / g.hd[2]};
    hipLaunchKernelGGL(k_final_residual, dim3((unsigned)padded), dim3(kBlock), 0, s->stream, g, s->flags, s->b, s->p,
                       s->partials, s->sc, ihd[0] * ihd[0], ihd[1] * ihd[1], ihd[2] * ihd[2], padded);
    return hipGetLastError();
}

static hipError_t read_scalars(fe_scene* s) {
    hipError_t e = hipMemcpyAsync(s->sc_host, s->sc, sizeof(SolverScalars), hipMemcpyDeviceToHost, s->stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s->stream);
}

hipError_t launch_pressure_solve(fe_scene* s, const fe_solver_desc* d) {
    const Geom& g = s->g;
    const int64_t nc = cell_count(g);
    const int64_t padded = pad_blocks((nc + kBlock - 1) / kBlock);
    // BuildSystem's invH * invH (single_phase_pressure_solver.cpp:102-103); exact
    // in FP32 for power-of-two spacings
    const double ihd[3] = {1.0 / g.hd[0], 1.0 / g.hd[1], 1.0 / g.hd[2]};
    const float ih2[3] = {(float)(ihd[0] * ihd[0]), (float)(ihd[1] * ihd[1]), (float)(ihd[2] * ihd[2])};
    hipError_t e;

    SolverScalars init{};
    init.tol = d->tolerance;
    init.relative = d->relative;
    init.precond = d->kind == FE_SOLVER_PCG_JACOBI ? 1 : 0;
    init.max_it = d->max_iterations;
    if ((e = hipMemcpyAsync(s->sc, &init, sizeof(SolverScalars), hipMemcpyHostToDevice, s->stream)) != hipSuccess)
        return e;
    FaceSet u{{s->u[0], s->u[1], s->u[2]}};
    hipLaunchKernelGGL(k_rhs, dim3((unsigned)padded), dim3(kBlock), 0, s->stream, g, u, s->flags, s->b, s->partials,
                       s->sc, padded);
    if (!d->warm_start) hipLaunchKernelGGL(k_zero_f64, grid1(nc), dim3(kBlock), 0, s->stream, s->p, nc);

    if (d->kind == FE_SOLVER_JACOBI) {
        // The reference's CudaJacobiIterationSolver uploads the previous x (warm start).
        float* xa = s->r;
        float* xb = s->z;
        hipLaunchKernelGGL(k_f64_to_f32, grid1(nc), dim3(kBlock), 0, s->stream, s->p, xa, nc, 0);
        for (int it = 0; it < d->max_iterations; ++it) {
            kt_begin(s, FE_KERNEL_JACOBI);
            hipLaunchKernelGGL(k_jacobi, dim3((unsigned)padded), dim3(kBlock), 0, s->stream, g, s->flags, s->b, xa,
                               xb, s->sc, ih2[0], ih2[1], ih2[2], padded);
            kt_end(s, FE_KERNEL_JACOBI);
            std::swap(xa, xb);
            if (it != 0 && d->check_interval > 0 && it % d->check_interval == 0)
                hipLaunchKernelGGL(k_jacobi_check, dim3((unsigned)padded), dim3(kBlock), 0, s->stream, g, s->flags,
                                   s->b, xa, s->partials, s->sc, ih2[0], ih2[1], ih2[2], it, padded);
        }
        hipLaunchKernelGGL(k_f32_to_f64, grid1(nc), dim3(kBlock), 0, s->stream, xa, s->p, nc);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = read_scalars(s)) != hipSuccess) return e;
        kt_flush(s);
        return hipSuccess;
    }

    // CG / PCG
    hipLaunchKernelGGL(k_cg_init, dim3((unsigned)padded), dim3(kBlock), 0, s->stream, g, s->flags, s->b, s->p, s->r,
                       s->z, s->d[0], s->partials, s->sc, ih2[0], ih2[1], ih2[2], padded);
    double* xa = s->p;
    double* xb = s->p2;
    int dcur = 0;
    const int batch = 16;
    int launched = 0;
    const float* zr = init.precond ? s->z : s->r;
    if ((e = read_scalars(s)) != hipSuccess) return e;
    kt_flush(s);  // pairs of earlier kernels are complete after the sync
    while (launched < d->max_iterations && !s->sc_host->done) {
        const int nb = std::min(batch, d->max_iterations - launched);
        const int it0 = s->sc_host->it;
        for (int t = 0; t < nb; ++t) {
            kt_begin(s, FE_KERNEL_CG_SPMV);
            hipLaunchKernelGGL(k_cg_spmv, dim3((unsigned)padded), dim3(kBlock), 0, s->stream, g, s->flags, zr,
                               s->d[dcur], s->d[1 - dcur], s->q, s->partials, s->sc, ih2[0], ih2[1], ih2[2], padded);
            kt_end(s, FE_KERNEL_CG_SPMV);
            dcur = 1 - dcur;
            kt_begin(s, FE_KERNEL_CG_UPDATE);
            hipLaunchKernelGGL(k_cg_update, dim3((unsigned)padded), dim3(kBlock), 0, s->stream, g, s->flags, s->b, xa,
                               xb, s->d[dcur], s->q, s->r, s->z, s->partials, s->sc, ih2[0], ih2[1], ih2[2], padded);
            kt_end(s, FE_KERNEL_CG_UPDATE);
            std::swap(xa, xb);
        }
        launched += nb;
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = read_scalars(s)) != hipSuccess) return e;
        // only the launches that iterated count for the kernel timer (the rest exit at once)
        s->kt_used = std::min(s->kt_used, std::max(0, s->sc_host->it - it0));
        kt_flush(s);
    }
    if (xa != s->p) {
        s->p2 = s->p;
        s->p = xa;
    }
    return hipSuccess;
}

}  // namespace fe
