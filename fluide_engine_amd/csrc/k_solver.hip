// k_solver.hip -- pressure projection of SinglePhasePressureSolver
// (src/fluid_solvers/single_phase_pressure_solver.cpp) on gfx950.
//
// The reference stores A as {center,right,up,front} doubles per cell
// (src/linear_system/linear_system.hpp:9-15, 32 B/cell).  Here BuildSystem's
// rules (:99-164) -- center = sum of 1/h^2 over in-range, non-BOUNDRY
// neighbours; off-diagonal -1/h^2 between two FLUID cells; non-FLUID rows are
// the identity -- are packed once per solve into a 16-bit code per cell
// (which neighbours are coupled, how many neighbours count per axis), so the
// iteration kernels read 2 B/cell of matrix and never divide an index.
// r, z, d, q are FP32, the solution is FP64, and every dot product and norm
// accumulates in FP64.
//
// Reductions: each kernel runs at most 1024 blocks over contiguous cell chunks
// and leaves one FP64 partial per block and value.  The next kernel's blocks
// each sum those partials in the same fixed order, so all blocks derive
// bit-identical scalars with no atomics and no inter-block fences (the kernel
// boundary orders the writes); block 0 records the solver state.
#include <algorithm>
#include <utility>

#include "fe_stencil.hpp"

namespace fe {



// ---------------------------------------------------------------------------
// BuildSystem (single_phase_pressure_solver.cpp:99-164): the row codes and the
// right-hand side b = DivergenceAtCallCenter (face_centered_grid3d.cpp:157-166)
// on FLUID cells, 0 elsewhere; partials of b.b.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_build_system(Geom g, FaceSet u, const uint8_t* __restrict__ flags,
                                                         uint16_t* __restrict__ code, float* __restrict__ b,
                                                         double* __restrict__ pout, uint8_t* __restrict__ cfl,
                                                         int64_t chunk, int64_t padded, SolverScalars* __restrict__ sc,
                                                         SolverScalars init) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *sc = init;  // the solver's start state (read by the later kernels)
    const int64_t nc = cell_count(g);
    const Span sp = block_span(nc, chunk, padded);
    const uint32_t nx = (uint32_t)g.n[0], ny = (uint32_t)g.n[1], nz = (uint32_t)g.n[2];
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    double v[1] = {0.0};
    int fluid = 0;
    // Groups of G cells: their flags loaded together, then every FLUID cell's
    // neighbour flags and faces together (two dependent load rounds per group
    // instead of per cell); the cells are then written in order.
    const FaceSet fs = u;
    constexpr int G = 4;
    for (int64_t c0 = sp.c0 + threadIdx.x; c0 < sp.c1; c0 += (int64_t)G * kBlock) {
        uint8_t fc[G];
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const int64_t c = c0 + (int64_t)u * kBlock;
            fc[u] = c < sp.c1 ? flags[c] : (uint8_t)FE_AIR;
        }
        uint8_t fn[G][6];
        float fv[G][6];
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const int64_t c = c0 + (int64_t)u * kBlock;
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                fn[u][q] = (uint8_t)FE_BOUNDRY;
                fv[u][q] = 0.0f;
            }
            if (fc[u] != FE_FLUID) continue;
            const uint32_t cu = (uint32_t)c;  // grids hold fewer than 2^30 cells
            const uint32_t i = cu % nx, j = (cu / nx) % ny, k = cu / (nx * ny);
            const bool in[6] = {i > 0, i + 1 < nx, j > 0, j + 1 < ny, k > 0, k + 1 < nz};
            const int64_t nb[6] = {c - 1, c + 1, c - sy, c + sy, c - sz, c + sz};
#pragma unroll
            for (int q = 0; q < 6; ++q)
                if (in[q]) fn[u][q] = flags[nb[q]];
            const int64_t iu = (int64_t)i + (int64_t)(nx + 1) * ((int64_t)j + (int64_t)ny * k);
            const int64_t iv = (int64_t)i + (int64_t)nx * ((int64_t)j + (int64_t)(ny + 1) * k);
            fv[u][0] = fs.a[0][iu];
            fv[u][1] = fs.a[0][iu + 1];
            fv[u][2] = fs.a[1][iv];
            fv[u][3] = fs.a[1][iv + nx];
            fv[u][4] = fs.a[2][c];
            fv[u][5] = fs.a[2][c + sz];
        }
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const int64_t c = c0 + (int64_t)u * kBlock;
            if (c >= sp.c1) break;
            float bb = 0.0f;
            uint16_t cd = 0;
            if (fc[u] == FE_FLUID) {
                cd = kFluidBit;
                fluid = 1;
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    const uint8_t f = fn[u][q];  // FE_BOUNDRY outside the grid
                    if (f == FE_BOUNDRY) continue;
                    cd = (uint16_t)(cd + (1u << (6 + 2 * (q >> 1))));
                    if (f == FE_FLUID) cd |= (uint16_t)(1u << q);
                }
                bb = (fv[u][1] - fv[u][0]) / g.h[0] + (fv[u][3] - fv[u][2]) / g.h[1] + (fv[u][5] - fv[u][4]) / g.h[2];
            }
            code[c] = cd;
            b[c] = bb;
            v[0] += (double)bb * (double)bb;
        }
    }
    // chunks without FLUID cells hold zeros in every CG and level-0 multigrid
    // vector throughout the solve, so those kernels skip them
    fluid = __syncthreads_or(fluid);
    if (threadIdx.x == 0) cfl[sp.lb] = (uint8_t)(fluid != 0);
    write_partials<1>(v, pout);
}

// ---------------------------------------------------------------------------
// Jacobi sweep (JacobiIterationSolver::Relax, src/linear_system/jacobi_iteration_solver.cpp:39-76):
// x' = (b - sum_offdiag A x) / center, on FP32 working copies of the FP64
// pressure.  When a residual check ran before this sweep its partials are
// summed first (CalculateTolerance, :78-82: ||b - A x||_2 < tol stops the
// solve); after the stop every sweep only copies, so the buffer rotation stays
// consistent.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_jacobi(Geom g, const uint16_t* __restrict__ code,
                                                   const float* __restrict__ b, const float* __restrict__ xo,
                                                   float* __restrict__ xn, SolverScalars* sc,
                                                   const double* __restrict__ pcheck, int nparts, int check_it,
                                                   float ihx2, float ihy2, float ihz2, int64_t chunk,
                                                   int64_t padded) {
    // block 0 may set sc->done during this kernel: read it once per block so
    // every thread takes the same path through the barriers below
    __shared__ int sdone;
    if (threadIdx.x == 0) sdone = sc->done;
    __syncthreads();
    bool done = sdone != 0;
    if (pcheck && !done) {
        double t[1];
        sum_partials<1>(pcheck, nparts, t);
        if (sqrt(t[0]) < sc->tol) {
            done = true;
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                sc->done = 1;
                sc->converged = 1;
                sc->it = check_it + 1;
            }
        }
    }
    const int64_t nc = cell_count(g);
    const Span sp = block_span(nc, chunk, padded);
    const float ih2[3] = {ihx2, ihy2, ihz2};
    const Stride st = stride_of(g.n[0], g.n[1], g.n[2]);
    auto X = [&](int64_t n) { return xo[n]; };
    for (int64_t c = sp.c0 + threadIdx.x; c < sp.c1; c += kBlock) {
        if (done) {
            xn[c] = xo[c];
            continue;
        }
        const uint16_t cd = code[c];
        const float center = code_center(cd, ih2);
        // off-diagonal sum in the reference order (i-1, i+1, j-1, j+1, k-1, k+1)
        const float off = code_apply(cd, 0.0f, c, st, ih2, 0.0f, X);
        xn[c] = (b[c] - off) / center;
    }
}

// Partials of ||b - A x||^2 for the Jacobi convergence check (FP64 sums).
__global__ __launch_bounds__(kBlock) void k_jacobi_residual(Geom g, const uint16_t* __restrict__ code,
                                                            const float* __restrict__ b,
                                                            const float* __restrict__ x, const SolverScalars* sc,
                                                            double* __restrict__ pout, float ihx2, float ihy2,
                                                            float ihz2, int64_t chunk, int64_t padded) {
    const int64_t nc = cell_count(g);
    const Span sp = block_span(nc, chunk, padded);
    const double ih2d[3] = {ihx2, ihy2, ihz2};
    const Stride st = stride_of(g.n[0], g.n[1], g.n[2]);
    double v[1] = {0.0};
    if (!sc->done) {
        for (int64_t c = sp.c0 + threadIdx.x; c < sp.c1; c += kBlock) {
            const double ax = code_apply64(code[c], c, st, ih2d, (double)x[c], [&](int64_t n) { return (double)x[n]; });
            const double res = (double)b[c] - ax;
            v[0] += res * res;
        }
    }
    write_partials<1>(v, pout);
}

__global__ __launch_bounds__(kBlock) void k_f64_to_f32(const double* __restrict__ a, float* __restrict__ o,
                                                       int64_t n) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c < n) o[c] = (float)a[c];
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

// What a residual b - A x on the FP64 solution reads for cell c.
struct Stencil64In {
    uint16_t cd;
    float bc;
    double xc, nb[6];
};

__device__ inline Stencil64In stencil64_in(int64_t c, const uint16_t* __restrict__ code,
                                           const double* __restrict__ x, const float* __restrict__ b,
                                           const Stride& st) {
    Stencil64In t;
    t.cd = code[c];
    t.bc = b[c];
    t.xc = x[c];
    code_neighbours(c, st, [&](int64_t n) { return x[n]; }, t.nb);
    return t;
}

// ---------------------------------------------------------------------------
// Conjugate gradient (ConjugateGradientSolver::Solve, src/linear_system/conjugate_gradient_solver.cpp:14-55),
// two kernels per iteration:
//   spmv:   [prologue: sigma' = r.z, r.r of the last update -> beta, trigger,
//           stopping rule, next replacement flag]  d = z + beta*d_old (also at
//           the 6 neighbours, on the fly), q = A d, partials of d.q
//   update: [prologue: alpha = sigma / d.q]  x' = alpha*d + x (FP64);
//           r = -alpha*q + r, or r = b - A x' in FP64 on residual replacement
//           (every 50 iterations or after sigma grew, :32-49; x' at the
//           neighbours is formed on the fly from x and d); z = M^-1 r;
//           partials of r.z and r.r
// With the relative rule a stop is only accepted on a freshly replaced
// residual, so the decision is taken on b - A x, not on the recurrence (with
// the multigrid preconditioner the next V-cycle's first kernel tests the
// replaced residual, so a confirmed stop costs no V-cycle and no spmv).
// The FP32 recurrence drifts from b - A x by about eps32 * ||A|| * sum|alpha d|,
// which grows with the residual at the last recomputation: the residual is
// also recomputed whenever ||r|| has fallen 100x below that value ("reliable
// updates", Sleijpen and van der Vorst 1996), so the drift stays well below
// ||r|| and beta stays meaningful across recomputations.
// ---------------------------------------------------------------------------
constexpr double kReliableDrop2 = 1e-4;  // (100x in ||r||)^2
__global__ __launch_bounds__(kBlock) void k_cg_init(Geom g, const uint16_t* __restrict__ code,
                                                    const float* __restrict__ b, double* __restrict__ x,
                                                    float* __restrict__ r, float* __restrict__ z,
                                                    float* __restrict__ d0, float* __restrict__ mgx, SolverScalars* sc,
                                                    const double* __restrict__ pbb, double* __restrict__ pout,
                                                    int nparts, float ihx2, float ihy2, float ihz2, int64_t chunk,
                                                    int64_t padded) {
    const int64_t nc = cell_count(g);
    const Span sp = block_span(nc, chunk, padded);
    const Stride st = stride_of(g.n[0], g.n[1], g.n[2]);
    // the thread's first group of cells loaded before the partial sum is
    // waited for (it does not depend on it)
    constexpr int G = 4;
    const int64_t g0 = sp.c0 + threadIdx.x;
    const bool early = g0 + (int64_t)(G - 1) * kBlock < sp.c1;
    Stencil64In raw[G];
    if (early) {
#pragma unroll
        for (int u = 0; u < G; ++u) raw[u] = stencil64_in(g0 + (int64_t)u * kBlock, code, x, b, st);
    }
    double bb[1];
    sum_partials<1>(pbb, nparts, bb);
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->bnorm2 = bb[0];
    const float ih2[3] = {ihx2, ihy2, ihz2};
    const double ih2d[3] = {ihx2, ihy2, ihz2};
    const int pre = sc->precond;  // 0 none, 1 diagonal, 2 multigrid (z and r.z come later)
    double v[2] = {0.0, 0.0};
    auto use = [&](int64_t c, const Stencil64In& t) {
        // the identity rows of non-FLUID cells have the solution 0: the
        // start takes it there (FLUID rows couple only to FLUID cells), so
        // r is 0 outside FLUID from the start and stays so
        const bool fl = (t.cd & kFluidBit) != 0;
        if (!fl && t.xc != 0.0) x[c] = 0.0;  // written only where it is not 0 already
        const float rc = fl ? (float)((double)t.bc - code_apply64_nb(t.cd, ih2d, t.xc, t.nb)) : 0.0f;
        r[c] = rc;
        float zc = rc;
        if (pre == 1) {
            zc = rc / code_center(t.cd, ih2);
            z[c] = zc;
        }
        if (mgx) mgx[c] = mg_weight(t.cd, ih2) * rc;  // the V-cycle's first sweep from zero
        d0[c] = 0.0f;
        if (pre != 2) v[0] += (double)rc * (double)zc;
        v[1] += (double)rc * (double)rc;
    };
    if (early) {
#pragma unroll
        for (int u = 0; u < G; ++u) use(g0 + (int64_t)u * kBlock, raw[u]);
    }
    span_groups_from<G, Stencil64In>(
        early ? g0 + (int64_t)G * kBlock : g0, sp.c1, [&](int64_t c) { return stencil64_in(c, code, x, b, st); },
        use);
    write_partials<2>(v, pout);
}

// State slots alternate with the iteration (spmv k reads slot k&1 and writes
// slot (k+1)&1), so no block can see a half-updated state.
__global__ __launch_bounds__(kBlock) void k_cg_spmv(Geom g, const uint16_t* __restrict__ code,
                                                    const float* __restrict__ z, const float* __restrict__ dold,
                                                    float* __restrict__ dnew, float* __restrict__ q,
                                                    const SolverScalars* __restrict__ in,
                                                    SolverScalars* __restrict__ out, const double* __restrict__ pin,
                                                    const double* __restrict__ pz, int nz, double* __restrict__ pout,
                                                    int nparts, float ihx2, float ihy2, float ihz2,
                                                    const uint8_t* __restrict__ cfl, int64_t chunk, int64_t padded) {
    // The chunk's FLUID flag, then the loads of the thread's first group of
    // cells, issued before the partial sums and the stop flag are waited for
    // (they depend on neither; after a stop they go unused).
    const int64_t nc = cell_count(g);
    const Span sp = block_span(nc, chunk, padded);
    const Stride st = stride_of(g.n[0], g.n[1], g.n[2]);
    const bool body = !cfl || cfl[sp.lb] != 0;
    constexpr int G = 4;
    struct Raw {
        uint16_t cd;
        float zc, dc, zn[6], dn[6];
    };
    const int64_t g0 = sp.c0 + threadIdx.x;
    const bool pre = body && g0 + (int64_t)(G - 1) * kBlock < sp.c1;
    Raw raw[G];
    if (pre) {
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const int64_t c = g0 + (int64_t)u * kBlock;
            raw[u].cd = code[c];
            raw[u].zc = z[c];
            raw[u].dc = dold[c];
            code_neighbours(c, st, [&](int64_t n) { return z[n]; }, raw[u].zn);
            code_neighbours(c, st, [&](int64_t n) { return dold[n]; }, raw[u].dn);
        }
    }
    // ---- state update from the last update's r.z and r.r (every block alike);
    // with the multigrid preconditioner r.z comes from the V-cycle's last sweep
    double t[2];
    if (pz) {  // r.z from the V-cycle's last sweep, r.r from the update: one pass
        double u[2];
        sum_partials_two(pz, 1, 0, nz, pin, 2, 1, nparts, u);
        t[0] = u[0];
        t[1] = u[1];
    } else {
        sum_partials<2>(pin, nparts, t);
    }
    if (in->done) {  // uniform: an earlier spmv stopped the solve
        if (blockIdx.x == 0 && threadIdx.x == 0) *out = *in;
        return;
    }
    const int it_prev = in->it;
    double beta, rr_ref;
    int it, trig, done, conv, repl, confirm = 0;
    if (it_prev < 0) {  // first iteration: partials come from k_cg_init
        it = 0;
        beta = 0.0;
        trig = 0;
        rr_ref = t[1];
        conv = stop_rule(in, in->bnorm2, t[0], t[1]) ? 1 : 0;
        done = (conv || in->max_it <= 0) ? 1 : 0;
    } else {
        rr_ref = in->replace ? t[1] : in->rr_ref;
        const double sold = in->sigma;
        it = it_prev + 1;
        trig = t[0] > sold ? 1 : 0;
        beta = t[0] / sold;
        bool stop = stop_rule(in, in->bnorm2, t[0], t[1]);
        if (stop && in->relative && !in->replace) {  // confirm on the true residual
            stop = false;
            trig = 1;
            confirm = 1;
        }
        conv = stop ? 1 : 0;
        done = (stop || it >= in->max_it) ? 1 : 0;
    }
    repl = (trig || (it % 50 == 0 && it > 0) || t[1] < kReliableDrop2 * rr_ref) ? 1 : 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        SolverScalars o = *in;
        o.sigma_old = in->sigma;
        o.sigma = t[0];
        o.rr = t[1];
        o.beta = beta;
        o.it = it;
        o.trigger = trig;
        o.converged = conv;
        o.done = done;
        o.replace = repl;
        o.rr_ref = rr_ref;
        o.confirm = confirm;
        *out = o;
    }
    if (done) return;
    // ---- body
    if (!body) {  // no FLUID cell: d, q stay 0
        if (threadIdx.x == 0) pout[blockIdx.x] = 0.0;
        return;
    }
    const float ih2[3] = {ihx2, ihy2, ihz2};
    const float fb = (float)beta;
    auto D = [&](int64_t n) { return fb * dold[n] + z[n]; };  // AXpY(beta, d, s, &d)
    double v[1] = {0.0};
    struct In {
        uint16_t cd;
        float dc, nb[6];
    };
    auto use = [&](int64_t c, const In& t) {
        const float qc = code_apply_nb(t.cd, code_center(t.cd, ih2), ih2, t.dc, t.nb);
        dnew[c] = t.dc;
        q[c] = qc;
        v[0] += (double)t.dc * (double)qc;
    };
    if (pre) {  // the prefetched group: D from the raw values, as D(n) forms it
#pragma unroll
        for (int u = 0; u < G; ++u) {
            In t;
            t.cd = raw[u].cd;
            t.dc = fb * raw[u].dc + raw[u].zc;
#pragma unroll
            for (int k = 0; k < 6; ++k) t.nb[k] = fb * raw[u].dn[k] + raw[u].zn[k];
            use(g0 + (int64_t)u * kBlock, t);
        }
    }
    span_groups_from<G, In>(
        pre ? g0 + (int64_t)G * kBlock : g0, sp.c1,
        [&](int64_t c) {
            In t;
            t.cd = code[c];
            t.dc = D(c);
            code_neighbours(c, st, D, t.nb);
            return t;
        },
        use);
    write_partials<1>(v, pout);
}

__global__ __launch_bounds__(kBlock) void k_cg_update(Geom g, const uint16_t* __restrict__ code,
                                                      const float* __restrict__ b, const double* __restrict__ xo,
                                                      double* __restrict__ xn, const float* __restrict__ d,
                                                      const float* __restrict__ q, float* __restrict__ r,
                                                      float* __restrict__ z, float* __restrict__ mgx, SolverScalars* sc,
                                                      const double* __restrict__ pin, double* __restrict__ pout,
                                                      int nparts, float ihx2, float ihy2, float ihz2,
                                                      const uint8_t* __restrict__ cfl, int64_t chunk, int64_t padded) {
    const int64_t nc = cell_count(g);
    const Span sp = block_span(nc, chunk, padded);
    // The chunk's FLUID flag, then the loads of the thread's first group of
    // cells, issued before the partial sum and the solver state are waited for
    // (they depend on neither; after a stop, or when r is recomputed, they go
    // unused).
    const bool body = !cfl || cfl[sp.lb] != 0;
    constexpr int G = 4;
    struct In {
        double x;
        float d, q, r;
        uint16_t cd;
    };
    const int64_t g0 = sp.c0 + threadIdx.x;
    const bool pre = body && g0 + (int64_t)(G - 1) * kBlock < sp.c1;
    In raw[G];
    if (pre) {
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const int64_t c = g0 + (int64_t)u * kBlock;
            raw[u].x = xo[c];
            raw[u].d = d[c];
            raw[u].q = q[c];
            raw[u].r = r[c];
            raw[u].cd = code[c];
        }
    }
    double t[1];
    sum_partials<1>(pin, nparts, t);
    if (sc->done) {  // uniform: keep the x rotation consistent
        for (int64_t c = sp.c0 + threadIdx.x; c < sp.c1; c += kBlock) xn[c] = xo[c];
        return;
    }
    const double alpha = sc->sigma / t[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->dq = t[0];
        sc->alpha = alpha;
    }
    if (!body) {  // no FLUID cell: x, r, z and the first sweep stay 0
        if (threadIdx.x == 0) pout[2 * blockIdx.x] = pout[2 * blockIdx.x + 1] = 0.0;
        return;
    }
    const bool replace = sc->replace != 0;
    const int pre_kind = sc->precond;
    const float ih2[3] = {ihx2, ihy2, ihz2};
    const double ih2d[3] = {ihx2, ihy2, ihz2};
    const Stride st = stride_of(g.n[0], g.n[1], g.n[2]);
    const float fa = (float)alpha;
    auto X = [&](int64_t n) { return alpha * (double)d[n] + xo[n]; };
    double v[2] = {0.0, 0.0};
    auto finish = [&](int64_t c, uint16_t cd, float rc) {
        r[c] = rc;
        float zc = rc;
        if (pre_kind == 1) {
            zc = rc / code_center(cd, ih2);
            z[c] = zc;
        }
        if (mgx) mgx[c] = mg_weight(cd, ih2) * rc;  // the V-cycle's first sweep from zero
        if (pre_kind != 2) v[0] += (double)rc * (double)zc;
        v[1] += (double)rc * (double)rc;
    };
    if (replace) {  // r = b - A x recomputed in FP64
        for (int64_t c = sp.c0 + threadIdx.x; c < sp.c1; c += kBlock) {
            const double xc = X(c);
            xn[c] = xc;
            const uint16_t cd = code[c];
            finish(c, cd, (float)((double)b[c] - code_apply64(cd, c, st, ih2d, xc, X)));
        }
    } else {
        auto use = [&](int64_t c, const In& t) {
            xn[c] = alpha * (double)t.d + t.x;
            finish(c, t.cd, -fa * t.q + t.r);
        };
        if (pre) {
#pragma unroll
            for (int u = 0; u < G; ++u) use(g0 + (int64_t)u * kBlock, raw[u]);
        }
        span_groups_from<G, In>(
            pre ? g0 + (int64_t)G * kBlock : g0, sp.c1,
            [&](int64_t c) {
                In t;
                t.x = xo[c];
                t.d = d[c];
                t.q = q[c];
                t.r = r[c];
                t.cd = (pre_kind == 1 || mgx) ? code[c] : (uint16_t)0;  // for z or the first sweep
                return t;
            },
            use);
    }
    write_partials<2>(v, pout);
}

// Final ||b - A p||^2 and ||b||^2 in FP64 on the FP64 solution, with FP64
// coefficients 1/h^2.
__global__ __launch_bounds__(kBlock) void k_final_residual(Geom g, const uint16_t* __restrict__ code,
                                                           const float* __restrict__ b, const double* __restrict__ x,
                                                           double* __restrict__ pout, double ihx2, double ihy2,
                                                           double ihz2, int64_t chunk, int64_t padded) {
    const int64_t nc = cell_count(g);
    const Span sp = block_span(nc, chunk, padded);
    const double ih2[3] = {ihx2, ihy2, ihz2};
    const Stride st = stride_of(g.n[0], g.n[1], g.n[2]);
    double v[2] = {0.0, 0.0};
    span_groups<4, Stencil64In>(
        sp.c0, sp.c1, [&](int64_t c) { return stencil64_in(c, code, x, b, st); },
        [&](int64_t c, const Stencil64In& t) {
            const double res = (double)t.bc - code_apply64_nb(t.cd, ih2, t.xc, t.nb);
            v[0] += res * res;
            v[1] += (double)t.bc * (double)t.bc;
        });
    write_partials<2>(v, pout);
}

// One block: totals of NV-wide partials into out[0..NV).
template <int NV>
__global__ __launch_bounds__(kBlock) void k_finish(const double* __restrict__ part, int nparts, double* out) {
    double t[NV];
    sum_partials<NV>(part, nparts, t);
    if (threadIdx.x == 0)
        for (int q = 0; q < NV; ++q) out[q] = t[q];
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static inline dim3 grid1(int64_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

static SolverShape solver_shape(const Geom& g) { return solver_shape_n(cell_count(g)); }

hipError_t launch_final_residual(fe_scene* s) {
    const Geom& g = s->g;
    const SolverShape sh = solver_shape(g);
    double* part = s->partials;
    hipLaunchKernelGGL(k_final_residual, dim3((unsigned)sh.blocks), dim3(kBlock), 0, s->stream, g, s->code, s->b,
                       s->p, part, g.ihd[0] * g.ihd[0], g.ihd[1] * g.ihd[1], g.ihd[2] * g.ihd[2], sh.chunk, sh.blocks);
    // totals (||r||^2, ||b||^2) land after the partials; read_final_residual fetches them
    hipLaunchKernelGGL(k_finish<2>, dim3(1), dim3(kBlock), 0, s->stream, part, (int)sh.blocks,
                       part + 2 * sh.blocks);
    return hipGetLastError();
}

static hipError_t read_scalars(fe_scene* s, int slot) {
    hipError_t e = hipMemcpyAsync(s->sc_host, s->sc + slot, sizeof(SolverScalars), hipMemcpyDeviceToHost, s->stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s->stream);
}

hipError_t read_final_residual(fe_scene* s, double* rr, double* bb) {
    const SolverShape sh = solver_shape(s->g);
    double tot[2] = {0.0, 0.0};
    hipError_t e = hipMemcpyAsync(tot, s->partials + 2 * sh.blocks, sizeof(tot), hipMemcpyDeviceToHost, s->stream);
    if (e != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s->stream)) != hipSuccess) return e;
    *rr = tot[0];
    *bb = tot[1];
    return hipSuccess;
}

// Zeros for up to 8 buffers in one launch: buffer q takes blocks
// [first[q], first[q + 1]), four 32-bit words per thread (one 16-byte store;
// every buffer starts 256-byte aligned).
struct ZeroList {
    int n = 0;
    uint32_t* p[8] = {};
    int64_t words[8] = {};
    int64_t first[9] = {};
};

__global__ __launch_bounds__(kBlock) void k_zero_many(ZeroList L) {
    const int64_t blk = blockIdx.x;
    int q = 0;
    while (q + 1 < L.n && blk >= L.first[q + 1]) ++q;
    const int64_t w0 = ((blk - L.first[q]) * kBlock + threadIdx.x) * 4;
    uint32_t* __restrict__ p = L.p[q];
    if (w0 + 4 <= L.words[q]) {
        *reinterpret_cast<uint4*>(p + w0) = make_uint4(0u, 0u, 0u, 0u);
    } else {
        for (int64_t w = w0; w < L.words[q]; ++w) p[w] = 0u;
    }
}

static void zero_add(ZeroList& L, void* p, int64_t bytes) {
    if (!p || bytes <= 0) return;
    L.p[L.n] = static_cast<uint32_t*>(p);
    L.words[L.n] = bytes / 4;
    L.first[L.n + 1] = L.first[L.n] + (L.words[L.n] + 4 * kBlock - 1) / (4 * kBlock);
    ++L.n;
}

// The solver's start: its scalars, BuildSystem and the initial guess.
static hipError_t solve_begin(fe_scene* s, const fe_solver_desc* d, SolverScalars& init) {
    const Geom& g = s->g;
    const int64_t nc = cell_count(g);
    const SolverShape sh = solver_shape(g);
    init = SolverScalars{};
    init.tol = d->tolerance;
    init.relative = d->relative;
    init.precond = d->kind == FE_SOLVER_PCG_JACOBI ? 1 : (d->kind == FE_SOLVER_MGPCG ? 2 : 0);
    init.max_it = d->max_iterations;
    init.it = -1;
    // the scalars' start goes with BuildSystem's launch (a kernel argument
    // rather than a host copy, so a captured substep carries the values)
    FaceSet u{{s->u[0], s->u[1], s->u[2]}};
    hipLaunchKernelGGL(k_build_system, dim3((unsigned)sh.blocks), dim3(kBlock), 0, s->stream, g, u, s->flags, s->code,
                       s->b, s->partials, s->chunk_fluid, sh.chunk, sh.blocks, s->sc, init);
    if (d->warm_start == FE_WARM_START_LEARNED) {
        if (!s->net) return hipErrorInvalidValue;
        hipError_t e = net_warm_start(s->net, s);
        if (e != hipSuccess) return e;
    } else if (!d->warm_start) {
        hipLaunchKernelGGL(k_zero_f64, grid1(nc), dim3(kBlock), 0, s->stream, s->p, nc);
    }
    return hipGetLastError();
}

// Jacobi sweeps (JacobiIterationSolver::Solve), host-driven.
static hipError_t jacobi_solve(fe_scene* s, const fe_solver_desc* d) {
    const Geom& g = s->g;
    const int64_t nc = cell_count(g);
    const SolverShape sh = solver_shape(g);
    const int nparts = (int)sh.blocks;
    double* PA = s->partials;
    double* PB = s->partials + 2 * sh.blocks;
    const float ih2[3] = {(float)(g.ihd[0] * g.ihd[0]), (float)(g.ihd[1] * g.ihd[1]), (float)(g.ihd[2] * g.ihd[2])};
    hipError_t e;
    // The reference's CudaJacobiIterationSolver uploads the previous x (warm start).
    float* xa = s->r;
    float* xb = s->z;
    hipLaunchKernelGGL(k_f64_to_f32, grid1(nc), dim3(kBlock), 0, s->stream, s->p, xa, nc);
    bool pending = false;
    int pending_it = 0;
    for (int it = 0; it < d->max_iterations; ++it) {
        launch_timed(s, FE_KERNEL_JACOBI, 0, k_jacobi, dim3((unsigned)sh.blocks), dim3(kBlock), g,
                     (const uint16_t*)s->code, (const float*)s->b, (const float*)xa, xb, s->sc,
                     (const double*)(pending ? PB : nullptr), nparts, pending_it, ih2[0], ih2[1], ih2[2], sh.chunk,
                     sh.blocks);
        std::swap(xa, xb);
        pending = false;
        if (it != 0 && d->check_interval > 0 && it % d->check_interval == 0) {
            hipLaunchKernelGGL(k_jacobi_residual, dim3((unsigned)sh.blocks), dim3(kBlock), 0, s->stream, g,
                               s->code, s->b, xa, s->sc, PB, ih2[0], ih2[1], ih2[2], sh.chunk, sh.blocks);
            pending = true;
            pending_it = it;
        }
    }
    if (pending)  // a check after the last sweep: the loop ends anyway, only record convergence
        hipLaunchKernelGGL(k_finish<1>, dim3(1), dim3(kBlock), 0, s->stream, PB, nparts, PA);
    hipLaunchKernelGGL(k_f32_to_f64, grid1(nc), dim3(kBlock), 0, s->stream, xa, s->p, nc);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = read_scalars(s, 0)) != hipSuccess) return e;
    if (pending && !s->sc_host->done) {
        double tot = 0.0;
        if ((e = hipMemcpy(&tot, PA, sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess) return e;
        if (std::sqrt(tot) < d->tolerance) {
            s->sc_host->done = 1;
            s->sc_host->converged = 1;
            s->sc_host->it = pending_it + 1;
        }
    }
    kt_flush(s);
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// CG / PCG, host side.  spmv k decides the stop of update k-1, so every batch
// of launches ends with an spmv whose state slot the host reads; launches
// after the stop leave before their first write (after at most one group of
// loads) and are left out of the kernel timer (tags >= the stopping iteration).
// ---------------------------------------------------------------------------
struct CGSetup {
    Geom g;
    SolverShape sh;
    int nparts = 0;
    double *PA = nullptr, *PB = nullptr, *PZ = nullptr;
    float ih2[3] = {0, 0, 0};
    int precond = 0;
    bool mg = false, fused = false;
    float* mgx = nullptr;
    const float* zr = nullptr;
    const double* pz = nullptr;
    int nz = 0;
};

// The per-solve constants (after mg_setup: the fused V-cycle's partials are
// allocated with its levels).
static CGSetup cg_setup(fe_scene* s, int precond) {
    CGSetup c;
    c.g = s->g;
    c.sh = solver_shape(s->g);
    c.nparts = (int)c.sh.blocks;
    c.PA = s->partials;                  // 2 per block
    c.PB = s->partials + 2 * c.sh.blocks;  // 2 per block
    c.PZ = s->partials + 4 * c.sh.blocks;  // r.z of the V-cycle, 1 per block
    // BuildSystem's invH * invH (single_phase_pressure_solver.cpp:102-103); exact
    // in FP32 for power-of-two spacings
    for (int a = 0; a < 3; ++a) c.ih2[a] = (float)(s->g.ihd[a] * s->g.ihd[a]);
    c.precond = precond;
    c.mg = precond == 2;
    c.fused = c.mg && s->tune.mg == FE_MG_FUSED;
    if (c.fused) c.PZ = s->mg_pz;  // one r.z partial per workgroup of the level-0 ascent
    c.mgx = c.mg && !c.fused ? mg_first_sweep_target(s) : nullptr;  // the fused descent sweeps from b itself
    c.zr = precond ? s->z : s->r;
    c.pz = c.mg ? c.PZ : nullptr;
    c.nz = c.fused ? (int)s->mg_npz : c.nparts;
    return c;
}

static void cg_update(fe_scene* s, const CGSetup& c, CGRun& R, int kk) {
    R.dcur = 1 - R.dcur;
    launch_timed(s, FE_KERNEL_CG_UPDATE, kk, k_cg_update, dim3((unsigned)c.sh.blocks), dim3(kBlock), c.g,
                 (const uint16_t*)s->code, (const float*)s->b, (const double*)R.xa, R.xb, (const float*)s->d[R.dcur],
                 (const float*)s->q, s->r, s->z, c.mgx, s->sc + ((kk + 1) & 1), (const double*)c.PA, c.PB, c.nparts,
                 c.ih2[0], c.ih2[1], c.ih2[2], (const uint8_t*)s->chunk_fluid, c.sh.chunk, c.sh.blocks);
    std::swap(R.xa, R.xb);
    if (c.mg) (void)mg_apply(s, s->r, s->z, s->sc + ((kk + 1) & 1), c.PZ, kk, c.PB, c.nparts);
}

// nb spmv launches with an update (and V-cycle) between each two; R.k is the
// last spmv's iteration afterwards.
static void cg_batch(fe_scene* s, const CGSetup& c, CGRun& R, int nb) {
    for (int t = 0; t < nb; ++t) {
        launch_timed(s, FE_KERNEL_CG_SPMV, R.k, k_cg_spmv, dim3((unsigned)c.sh.blocks), dim3(kBlock), c.g,
                     (const uint16_t*)s->code, c.zr, (const float*)s->d[R.dcur], s->d[1 - R.dcur], s->q,
                     (const SolverScalars*)(s->sc + (R.k & 1)), s->sc + ((R.k + 1) & 1), (const double*)c.PB, c.pz,
                     c.nz, c.PA, c.nparts, c.ih2[0], c.ih2[1], c.ih2[2], (const uint8_t*)s->chunk_fluid, c.sh.chunk,
                     c.sh.blocks);
        if (t == nb - 1) break;
        cg_update(s, c, R, R.k);
        ++R.k;
    }
}

// The iterations' start: the solver vectors' zeros, the initial residual and
// the first V-cycle.  The CG and level-0 multigrid kernels skip the chunks
// without FLUID cells (s->chunk_fluid): every vector they would write there
// starts at 0 and, FLUID rows coupling only to FLUID cells, stays 0
// (k_cg_init writes r, d, the first sweep and the solution's zeros outside
// FLUID itself).
static hipError_t cg_start(fe_scene* s, const CGSetup& c, CGRun& R) {
    const int64_t nc = cell_count(c.g);
    hipError_t e;
    ZeroList L;
    zero_add(L, s->z, 4 * nc);
    zero_add(L, s->q, 4 * nc);
    zero_add(L, s->d[1], 4 * nc);
    if (c.mg) zero_add(L, s->mg[0].x2, 4 * nc);
    // the fused V-cycle skips level-0 tiles without FLUID cells: the coarse
    // right-hand side below them stays 0
    if (c.fused) zero_add(L, s->mg[1].b, 4 * s->mg[1].cells);
    if (c.mg) zero_add(L, s->mg[0].r, 4 * nc);
    zero_add(L, s->p2, 8 * nc);
    hipLaunchKernelGGL(k_zero_many, dim3((unsigned)L.first[L.n]), dim3(kBlock), 0, s->stream, L);
    hipLaunchKernelGGL(k_cg_init, dim3((unsigned)c.sh.blocks), dim3(kBlock), 0, s->stream, c.g, s->code, s->b, s->p,
                       s->r, s->z, s->d[0], c.mgx, s->sc, c.PA, c.PB, c.nparts, c.ih2[0], c.ih2[1], c.ih2[2],
                       c.sh.chunk, c.sh.blocks);
    if (c.mg && (e = mg_apply(s, s->r, s->z, s->sc, c.PZ, -1, nullptr, 0)) != hipSuccess) return e;
    R = CGRun{};
    R.xa = s->p;
    R.xb = s->p2;
    return hipGetLastError();
}

// Host-driven iterations: batches of spmv launches, the host reading the last
// one's state slot.  resume: the state slot read last said "not done" after a
// batch launched elsewhere (a captured substep), so the next update comes
// first.  first: spmv launches in the first batch.
static hipError_t cg_host_loop(fe_scene* s, const fe_solver_desc* d, const CGSetup& c, CGRun& R, int first,
                               bool resume) {
    hipError_t e;
    const int batch = c.mg ? 4 : 16;
    if (resume) {
        cg_update(s, c, R, R.k);
        ++R.k;
    }
    for (;;) {
        const int nb = std::min(first, d->max_iterations + 1 - R.k);
        first = c.mg ? 2 : batch;
        cg_batch(s, c, R, nb);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = read_scalars(s, (R.k + 1) & 1)) != hipSuccess) return e;
        kt_flush(s, s->sc_host->done ? s->sc_host->it : (1 << 30));
        if (s->sc_host->done || nb <= 0) break;
        cg_update(s, c, R, R.k);
        ++R.k;
    }
    return hipSuccess;
}

hipError_t launch_pressure_solve(fe_scene* s, const fe_solver_desc* d) {
    SolverScalars init;
    hipError_t e = solve_begin(s, d, init);
    if (e != hipSuccess) return e;
    if (d->kind == FE_SOLVER_JACOBI) return jacobi_solve(s, d);
    if (init.precond == 2 && (e = mg_setup(s)) != hipSuccess) return e;
    const CGSetup c = cg_setup(s, init.precond);
    CGRun R;
    if ((e = cg_start(s, c, R)) != hipSuccess) return e;
    // iterations per host check: fewer when each carries a V-cycle; the first
    // check comes where the previous solve of this kind stopped (warm-started
    // solves of consecutive substeps take similar counts), so a typical solve
    // waits for the device once
    const int prev = s->last_cg_iters[init.precond];
    const int first = prev > 0 ? std::min(prev + 1, 64) : (c.mg ? 4 : 16);
    if ((e = cg_host_loop(s, d, c, R, first, false)) != hipSuccess) return e;
    if (R.xa != s->p) {
        s->p2 = s->p;
        s->p = R.xa;
    }
    s->last_cg_iters[init.precond] = s->sc_host->done ? s->sc_host->it : d->max_iterations;
    return hipSuccess;
}

// A captured substep's part of the solve: the start and the first `iters`
// spmv launches, then the state slot of the last one copied into the pinned
// host mirror.  Returns the run's state after them.
hipError_t capture_pressure_head(fe_scene* s, const fe_solver_desc* d, int iters, CGRun* out) {
    SolverScalars init;
    hipError_t e = solve_begin(s, d, init);
    if (e != hipSuccess) return e;
    if (init.precond == 2 && (e = mg_setup(s)) != hipSuccess) return e;
    const CGSetup c = cg_setup(s, init.precond);
    CGRun R;
    if ((e = cg_start(s, c, R)) != hipSuccess) return e;
    cg_batch(s, c, R, std::max(1, std::min(iters, d->max_iterations + 1)));
    if ((e = hipMemcpyAsync(s->sc_host, s->sc + ((R.k + 1) & 1), sizeof(SolverScalars), hipMemcpyDeviceToHost,
                            s->stream)) != hipSuccess)
        return e;
    *out = R;
    return hipGetLastError();
}

// After a captured head has run (s->sc_host holds its last state slot): the
// host-driven iterations it left, if any, then the solution copied into s->p
// when the iteration count left it in the partner buffer (a captured tail
// reads s->p).  *resumed: whether iterations were added.
hipError_t finish_pressure_after_head(fe_scene* s, const fe_solver_desc* d, const CGRun& head, bool* resumed) {
    const int precond = d->kind == FE_SOLVER_PCG_JACOBI ? 1 : (d->kind == FE_SOLVER_MGPCG ? 2 : 0);
    const CGSetup c = cg_setup(s, precond);
    CGRun R = head;
    hipError_t e;
    *resumed = false;
    if (!s->sc_host->done && R.k + 1 <= d->max_iterations) {
        *resumed = true;
        if ((e = cg_host_loop(s, d, c, R, c.mg ? 2 : 16, true)) != hipSuccess) return e;
    }
    if (R.xa != s->p &&
        (e = hipMemcpyAsync(s->p, R.xa, sizeof(double) * cell_count(s->g), hipMemcpyDeviceToDevice, s->stream)) !=
            hipSuccess)
        return e;
    s->last_cg_iters[precond] = s->sc_host->done ? s->sc_host->it : d->max_iterations;
    return hipSuccess;
}

}  // namespace fe
