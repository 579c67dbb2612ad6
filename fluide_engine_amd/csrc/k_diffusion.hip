// k_diffusion.hip -- BackwardEulerDiffusionSolver (viscosity) on the device.
//
// src/fluid_solvers/backward_euler_diffusion_solver.cpp:20-242, called by
// PICSimulator::ComputeDiffusion (src/pic_simulator.cpp:214-222) when the
// viscosity is positive.  Per face component:
//   * face markers: FLUID iff the fluid SDF sampled at the face position is
//     negative, else BOUNDRY iff the collider SDF is, else AIR (BuildX/Y/ZMarkers,
//     :71-128);
//   * system (BuildMatrix :136-198, BuildVectors :200-242): a FLUID row has
//     centre 1 + c_a per in-range non-AIR neighbour and -c_a to every FLUID
//     neighbour, b = u + c_a * u(neighbour) per BOUNDRY neighbour; other rows
//     are the identity with b = u; x0 = u.  c_a = dt * viscosity / h_a^2.
//   * solver: the reference's CudaJacobiIterationSolver(1000, 5, 1e-14) (:10)
//     semantics: sweep, and after sweep it (it > 0, it % interval == 0) stop
//     when ||b - A x||_2 < tolerance.
// Everything is FP64 (x, b, the row arithmetic in the reference's operation
// order with contraction off, the residual); the face velocities stay FP32
// and are exact in FP64.  The 16-bit row code has the pressure
// solver's layout: bits 0-5 coupled (FLUID) neighbours in the order i-1, i+1,
// j-1, j+1, k-1, k+1; bits 6-11 in-range non-AIR neighbours per axis; bit 12
// the row is FLUID.
#include <algorithm>
#include <cmath>

#include "fe_kernels.hpp"
#include "fe_stencil.hpp"

namespace fe {

namespace {

struct FaceDims {
    int nx, ny, nz;
};

__device__ inline void face_ijk(const FaceDims& d, int64_t f, int& i, int& j, int& k) {
    split_index(f, d.nx, d.ny, i, j, k);
}

// Markers and row codes of component C, and x0 = u.
__global__ __launch_bounds__(kBlock) void k_diff_markers(Geom g, int C, FaceDims d, const double* __restrict__ sdf,
                                                         const double* __restrict__ csdf, uint8_t* __restrict__ mark,
                                                         int64_t padded) {
#pragma clang fp contract(off)
    const int64_t f = xcd_remap(blockIdx.x, padded) * kBlock + threadIdx.x;
    const int64_t nf = (int64_t)d.nx * d.ny * d.nz;
    if (f >= nf) return;
    int idx[3];
    face_ijk(d, f, idx[0], idx[1], idx[2]);
    double p[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) p[a] = g.cod[C][a] + g.hd[a] * (double)idx[a];  // GetX/Y/ZPos
    uint8_t m = FE_AIR;
    if (sample_sdf64(g, sdf, p) < 0)
        m = FE_FLUID;
    else if (csdf && sample_collider64(g, csdf, p) < 0)
        m = FE_BOUNDRY;
    mark[f] = m;
}

__global__ __launch_bounds__(kBlock) void k_diff_build(FaceDims d, const uint8_t* __restrict__ mark,
                                                       const float* __restrict__ u, uint16_t* __restrict__ code,
                                                       double* __restrict__ x0, double* __restrict__ b, double cx,
                                                       double cy, double cz, int64_t padded) {
#pragma clang fp contract(off)
    const int64_t f = xcd_remap(blockIdx.x, padded) * kBlock + threadIdx.x;
    const int64_t nf = (int64_t)d.nx * d.ny * d.nz;
    if (f >= nf) return;
    uint16_t cd = 0;
    double bb = (double)u[f];
    if (mark[f] == FE_FLUID) {
        int i, j, k;
        face_ijk(d, f, i, j, k);
        const int64_t sy = d.nx, sz = (int64_t)d.nx * d.ny;
        const bool in[6] = {i > 0, i + 1 < d.nx, j > 0, j + 1 < d.ny, k > 0, k + 1 < d.nz};
        const int64_t nb[6] = {f - 1, f + 1, f - sy, f + sy, f - sz, f + sz};
        cd = kFluidBit;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            if (!in[q]) continue;
            const uint8_t m = mark[nb[q]];
            if (m == FE_AIR) continue;
            cd = (uint16_t)(cd + (1u << (6 + 2 * (q >> 1))));
            if (m == FE_FLUID) cd |= (uint16_t)(1u << q);
        }
        // BuildVectors' BOUNDRY terms in its order: i+1, i-1, j+1, j-1, k+1, k-1
        const double c[3] = {cx, cy, cz};
        const int order[6] = {1, 0, 3, 2, 5, 4};
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            const int q = order[t];
            if (in[q] && mark[nb[q]] == FE_BOUNDRY) bb += c[q >> 1] * (double)u[nb[q]];
        }
    }
    code[f] = cd;
    x0[f] = (double)u[f];
    b[f] = bb;
}

// Centre 1 + c_a per non-AIR neighbour, summed in BuildMatrix's order (x, y, z).
__device__ inline double diff_center(uint16_t cd, const double (&c)[3]) {
#pragma clang fp contract(off)
    double center = 1.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int cnt = (cd >> (6 + 2 * a)) & 3;
        if (cnt > 0) center += c[a];
        if (cnt > 1) center += c[a];
    }
    return center;
}

// Sum of the off-diagonal terms A_ij x_j of a FLUID row in the reference
// Jacobi order (i-1, i+1, j-1, j+1, k-1, k+1); uncoupled terms are zero.
__device__ inline double diff_offdiag(uint16_t cd, int64_t f, int64_t sy, int64_t sz, const double (&c)[3],
                                      const double* __restrict__ x) {
#pragma clang fp contract(off)
    const int64_t nb[6] = {f - 1, f + 1, f - sy, f + sy, f - sz, f + sz};
    double r = 0.0;
#pragma unroll
    for (int q = 0; q < 6; ++q)
        if (cd & (1u << q)) r += (-c[q >> 1]) * x[nb[q]];
    return r;
}

__global__ __launch_bounds__(kBlock) void k_diff_jacobi(FaceDims d, const uint16_t* __restrict__ code,
                                                        const double* __restrict__ b, const double* __restrict__ xo,
                                                        double* __restrict__ xn, double cx, double cy, double cz,
                                                        const SolverScalars* sc, int64_t padded) {
#pragma clang fp contract(off)
    if (sc->done) return;  // converged at an earlier check: x stays in the other buffer
    const int64_t f = xcd_remap(blockIdx.x, padded) * kBlock + threadIdx.x;
    const int64_t nf = (int64_t)d.nx * d.ny * d.nz;
    if (f >= nf) return;
    const uint16_t cd = code[f];
    const double bb = b[f];
    if (!(cd & kFluidBit)) {
        xn[f] = bb;  // identity row: (b - 0) / 1
        return;
    }
    const double c[3] = {cx, cy, cz};
    const double off = diff_offdiag(cd, f, d.nx, (int64_t)d.nx * d.ny, c, xo);
    xn[f] = (bb - off) / diff_center(cd, c);
}

// Partials of ||b - A x||^2 (BLAS::Residual, src/linear_system/blas.cpp:21-37).
__global__ __launch_bounds__(kBlock) void k_diff_residual(FaceDims d, const uint16_t* __restrict__ code,
                                                          const double* __restrict__ b,
                                                          const double* __restrict__ x, double cx, double cy,
                                                          double cz, const SolverScalars* sc,
                                                          double* __restrict__ part, int64_t chunk, int64_t padded) {
#pragma clang fp contract(off)
    const int64_t nf = (int64_t)d.nx * d.ny * d.nz;
    const Span sp = block_span(nf, chunk, padded);
    const double c[3] = {cx, cy, cz};
    double v[1] = {0.0};
    if (!sc->done) {
        for (int64_t f = sp.c0 + threadIdx.x; f < sp.c1; f += kBlock) {
            const uint16_t cd = code[f];
            double r = b[f];
            if (cd & kFluidBit)
                r = (r - diff_center(cd, c) * x[f]) - diff_offdiag(cd, f, d.nx, (int64_t)d.nx * d.ny, c, x);
            else
                r = r - x[f];
            v[0] += r * r;
        }
    }
    write_partials<1>(v, part);
}

// One block: the convergence check after sweep `it` (0-based).
__global__ __launch_bounds__(kBlock) void k_diff_check(const double* __restrict__ part, int nparts, int it,
                                                       SolverScalars* sc) {
    if (sc->done) return;
    double t[1];
    sum_partials<1>(part, nparts, t);
    if (threadIdx.x == 0) {
        sc->rr = t[0];
        if (sqrt(t[0]) < sc->tol) {
            sc->done = 1;
            sc->converged = 1;
            sc->it = it + 1;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_diff_store(const double* __restrict__ x, float* __restrict__ u,
                                                       int64_t n) {
    const int64_t f = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (f < n) u[f] = (float)x[f];  // output->x/y/z(i, j, k) = _system.x(i, j, k)
}

FaceDims face_dims(const Geom& g, int c) { return {face_dim(g, c, 0), face_dim(g, c, 1), face_dim(g, c, 2)}; }

}  // namespace

hipError_t diffusion_alloc(fe_scene* s) {
    if (s->dx[0]) return hipSuccess;
    int64_t nf = 0;
    for (int c = 0; c < 3; ++c) nf = std::max<int64_t>(nf, face_count(s->g, c));
    hipError_t e;
    for (int b = 0; b < 2; ++b)
        if ((e = hipMalloc((void**)&s->dx[b], sizeof(double) * nf)) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&s->dcode, sizeof(uint16_t) * nf)) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&s->db, sizeof(double) * nf)) != hipSuccess) return e;
    for (int c = 0; c < 3; ++c)
        if ((e = hipMalloc((void**)&s->dmark[c], face_count(s->g, c))) != hipSuccess) return e;
    return hipSuccess;
}

void diffusion_free(fe_scene* s) {
    for (int b = 0; b < 2; ++b) {
        if (s->dx[b]) (void)hipFree(s->dx[b]);
        s->dx[b] = nullptr;
    }
    if (s->dcode) (void)hipFree(s->dcode);
    s->dcode = nullptr;
    if (s->db) (void)hipFree(s->db);
    s->db = nullptr;
    for (int c = 0; c < 3; ++c) {
        if (s->dmark[c]) (void)hipFree(s->dmark[c]);
        s->dmark[c] = nullptr;
    }
}

// ComputeDiffusion without its trailing ApplyBoundryCondition: the three
// component solves in order x, y, z, each a Jacobi solve on the scene's stream
// in batches of sweeps with one host read of the stopping flag per batch.
hipError_t launch_diffusion(fe_scene* s, double dt, const fe_solver_desc* d) {
    const Geom& g = s->g;
    const double nu = s->desc.viscosity;
    const double c[3] = {(dt * nu) / (g.hd[0] * g.hd[0]), (dt * nu) / (g.hd[1] * g.hd[1]),
                         (dt * nu) / (g.hd[2] * g.hd[2])};
    hipError_t e;
    if ((e = diffusion_alloc(s)) != hipSuccess) return e;
    constexpr int kBatch = 50;  // sweeps between host reads of the stopping flag
    for (int comp = 0; comp < 3; ++comp) {
        const FaceDims fd = face_dims(g, comp);
        const int64_t nf = face_count(g, comp);
        const int64_t padded = pad_blocks((nf + kBlock - 1) / kBlock);
        const SolverShape sh = solver_shape_n(nf);
        if (sh.blocks > s->npartials) return hipErrorInvalidValue;  // partial sums must fit the scene's buffer
        hipLaunchKernelGGL(k_diff_markers, dim3((unsigned)padded), dim3(kBlock), 0, s->stream, g, comp, fd,
                           (const double*)s->sdf, (const double*)(s->has_collider ? s->csdf : nullptr),
                           s->dmark[comp], padded);
        hipLaunchKernelGGL(k_diff_build, dim3((unsigned)padded), dim3(kBlock), 0, s->stream, fd,
                           (const uint8_t*)s->dmark[comp], (const float*)s->u[comp], s->dcode, s->dx[0], s->db, c[0],
                           c[1], c[2], padded);
        SolverScalars init{};
        init.tol = d->tolerance;
        init.max_it = d->max_iterations;
        init.it = d->max_iterations;
        if ((e = hipMemcpyAsync(s->sc, &init, sizeof(SolverScalars), hipMemcpyHostToDevice, s->stream)) != hipSuccess)
            return e;
        int it = 0;
        bool done = false;
        while (it < d->max_iterations && !done) {
            const int end = std::min(d->max_iterations, it + kBatch);
            for (; it < end; ++it) {
                hipLaunchKernelGGL(k_diff_jacobi, dim3((unsigned)padded), dim3(kBlock), 0, s->stream, fd,
                                   (const uint16_t*)s->dcode, (const double*)s->db, (const double*)s->dx[it & 1],
                                   s->dx[(it + 1) & 1], c[0], c[1], c[2], (const SolverScalars*)s->sc, padded);
                if (it != 0 && d->check_interval > 0 && it % d->check_interval == 0) {
                    hipLaunchKernelGGL(k_diff_residual, dim3((unsigned)sh.blocks), dim3(kBlock), 0, s->stream, fd,
                                       (const uint16_t*)s->dcode, (const double*)s->db,
                                       (const double*)s->dx[(it + 1) & 1], c[0], c[1], c[2],
                                       (const SolverScalars*)s->sc, s->partials, sh.chunk, sh.blocks);
                    hipLaunchKernelGGL(k_diff_check, dim3(1), dim3(kBlock), 0, s->stream, (const double*)s->partials,
                                       (int)sh.blocks, it, s->sc);
                }
            }
            if ((e = hipGetLastError()) != hipSuccess) return e;
            if ((e = hipMemcpyAsync(s->sc_host, s->sc, sizeof(SolverScalars), hipMemcpyDeviceToHost, s->stream)) !=
                hipSuccess)
                return e;
            if ((e = hipStreamSynchronize(s->stream)) != hipSuccess) return e;
            done = s->sc_host->done != 0;
        }
        const int executed = done ? s->sc_host->it : d->max_iterations;
        hipLaunchKernelGGL(k_diff_store, dim3((unsigned)((nf + kBlock - 1) / kBlock)), dim3(kBlock), 0, s->stream,
                           (const double*)s->dx[executed & 1], s->u[comp], nf);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        fe_solve_stats& st = s->diff_stats[comp];
        st = fe_solve_stats{};
        st.iterations = executed;
        st.converged = done ? 1 : 0;
        st.abs_residual = d->check_interval > 0 && executed > d->check_interval ? std::sqrt(s->sc_host->rr) : -1.0;
        st.rel_residual = -1.0;
        st.ms = 0.0;
    }
    return hipSuccess;
}

}  // namespace fe
