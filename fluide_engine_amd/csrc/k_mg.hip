// k_mg.hip -- geometric multigrid V-cycle used as the preconditioner of the
// pressure CG (no reference counterpart: the reference's CG has an identity
// preconditioner, src/linear_system/conjugate_gradient_solver.cpp:57-63).
//
// Levels: level 0 is the pressure grid; each coarser level halves every axis
// (rounding up).  A coarse cell is AIR if any child is AIR, else FLUID if any
// child is FLUID, else BOUNDRY (the free surface keeps its Dirichlet cells on
// every level, as in McAdams, Sifakis and Teran, "A parallel multigrid Poisson
// solver for fluids simulation on large grids", SCA 2010), and its operator is
// BuildSystem's 7-point rule on those markers with twice the spacing.  The V-cycle is symmetric, so the
// preconditioner is SPD and CG stays CG:
//   * damped Jacobi smoothing, the same number of sweeps before and after;
//   * prolongation: trilinear interpolation of the cell-centred coarse
//     correction (weights 3/4, 1/4 per axis; out-of-range coarse neighbours
//     clamp onto the boundary cell, which is the Neumann wall);
//   * restriction: exactly the transpose of that prolongation divided by 8;
//   * coarsest level (<= 1024 cells): damped Jacobi sweeps from zero in one
//     workgroup, in LDS.
// Non-FLUID cells carry no residual into the coarse levels and take no
// correction; on level 0 they are smoothed like the identity rows they are.
#include <algorithm>

#include "fe_stencil.hpp"

namespace fe {

constexpr float kOmega = 2.0f / 3.0f;       // smoother damping
constexpr float kOmegaCoarse = 0.9f;        // coarsest-level Jacobi damping
constexpr int kPreSweeps = 2;               // = post sweeps (symmetry)
constexpr int kCoarseSweeps = 64;
constexpr int kCoarseMax = 1024;            // cells of the coarsest level (one workgroup)

// ---------------------------------------------------------------------------
// Level construction
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_mg_coarsen(const uint8_t* __restrict__ ff, int fx, int fy, int fz,
                                                       uint8_t* __restrict__ cf, int cx, int cy, int cz) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t nc = (int64_t)cx * cy * cz;
    if (c >= nc) return;
    const uint32_t cu = (uint32_t)c;
    const int I = (int)(cu % (uint32_t)cx), J = (int)((cu / (uint32_t)cx) % (uint32_t)cy),
              K = (int)(cu / ((uint32_t)cx * (uint32_t)cy));
    bool fluid = false, air = false;
    for (int dk = 0; dk < 2; ++dk) {
        const int k = 2 * K + dk;
        if (k >= fz) continue;
        for (int dj = 0; dj < 2; ++dj) {
            const int j = 2 * J + dj;
            if (j >= fy) continue;
            for (int di = 0; di < 2; ++di) {
                const int i = 2 * I + di;
                if (i >= fx) continue;
                const uint8_t f = ff[(int64_t)i + (int64_t)fx * ((int64_t)j + (int64_t)fy * k)];
                fluid |= f == FE_FLUID;
                air |= f == FE_AIR;
            }
        }
    }
    cf[c] = air ? (uint8_t)FE_AIR : (fluid ? (uint8_t)FE_FLUID : (uint8_t)FE_BOUNDRY);
}

__global__ __launch_bounds__(kBlock) void k_mg_codes(const uint8_t* __restrict__ flags, int nx, int ny, int nz,
                                                     uint16_t* __restrict__ code) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t nc = (int64_t)nx * ny * nz;
    if (c >= nc) return;
    const uint32_t cu = (uint32_t)c;
    const uint32_t i = cu % (uint32_t)nx, j = (cu / (uint32_t)nx) % (uint32_t)ny,
                   k = cu / ((uint32_t)nx * (uint32_t)ny);
    code[c] = row_code(flags, c, i, j, k, (uint32_t)nx, (uint32_t)ny, (uint32_t)nz);
}

// ---------------------------------------------------------------------------
// Smoothing and residual
// ---------------------------------------------------------------------------
struct LevelArgs {
    int64_t nc;
    int64_t sy, sz;
    float ih2[3];
    int64_t chunk, padded;
};

// x_out = x_in + omega (b - A x_in) / D; x_in == nullptr means x_in = 0.
// With prz != nullptr (last level-0 sweep) also leaves partials of b . x_out,
// which is r . z of the CG.
__global__ __launch_bounds__(kBlock) void k_mg_smooth(LevelArgs L, const uint16_t* __restrict__ code,
                                                      const float* __restrict__ b, const float* __restrict__ xin,
                                                      float* __restrict__ xout, float omega,
                                                      const SolverScalars* __restrict__ st,
                                                      double* __restrict__ prz) {
    if (st && st->done) return;  // uniform: the solve has stopped
    const Span sp = block_span(L.nc, L.chunk, L.padded);
    const Stride stv{L.sy, L.sz};
    double v[1] = {0.0};
    auto X = [&](int64_t n) { return xin[n]; };
    for (int64_t c = sp.c0 + threadIdx.x; c < sp.c1; c += kBlock) {
        const uint16_t cd = code[c];
        const float dgn = code_center(cd, L.ih2);
        const float w = dgn > 0.0f ? omega / dgn : 0.0f;  // a FLUID cell walled in on all sides has no diagonal
        float xo;
        if (!xin) {
            xo = w * b[c];
        } else {
            const float xc = xin[c];
            const float ax = code_apply(cd, dgn, c, stv, L.ih2, xc, X);
            xo = xc + w * (b[c] - ax);
        }
        xout[c] = xo;
        if (prz) v[0] += (double)b[c] * (double)xo;
    }
    if (prz) write_partials<1>(v, prz);
}

__global__ __launch_bounds__(kBlock) void k_mg_residual(LevelArgs L, const uint16_t* __restrict__ code,
                                                        const float* __restrict__ b, const float* __restrict__ x,
                                                        float* __restrict__ r, const SolverScalars* __restrict__ st) {
    if (st && st->done) return;
    const Span sp = block_span(L.nc, L.chunk, L.padded);
    const Stride stv{L.sy, L.sz};
    auto X = [&](int64_t n) { return x[n]; };
    for (int64_t c = sp.c0 + threadIdx.x; c < sp.c1; c += kBlock) {
        const uint16_t cd = code[c];
        float rc = 0.0f;
        if (cd & kFluidBit) rc = b[c] - code_apply(cd, code_center(cd, L.ih2), c, stv, L.ih2, x[c], X);
        r[c] = rc;
    }
}

// ---------------------------------------------------------------------------
// Transfer operators
// ---------------------------------------------------------------------------
// Prolongation weights of fine index fi on one axis: two coarse neighbours.
__device__ inline void prolong_axis(int fi, int nc, int& a, float& wa, int& b, float& wb) {
    const int m = fi >> 1;
    if (fi & 1) {
        a = m;
        wa = 0.75f;
        b = min(m + 1, nc - 1);
        wb = 0.25f;
    } else {
        a = max(m - 1, 0);
        wa = 0.25f;
        b = m;
        wb = 0.75f;
    }
}

// Weight with which fine index fi feeds coarse index I (the transpose).
__device__ inline float restrict_w(int fi, int I, int nc) {
    int a, b;
    float wa, wb;
    prolong_axis(fi, nc, a, wa, b, wb);
    return (a == I ? wa : 0.0f) + (b == I ? wb : 0.0f);
}

// b_c = (1/8) sum_f w(f, c) r_f over the fine cells whose prolongation stencil
// reaches c: fine indices 2I-1 .. 2I+2 per axis.
__global__ __launch_bounds__(kBlock) void k_mg_restrict(const float* __restrict__ rf, int fx, int fy, int fz,
                                                        const uint16_t* __restrict__ ccode, float* __restrict__ bc,
                                                        int cx, int cy, int cz, const SolverScalars* __restrict__ st) {
    if (st && st->done) return;
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t nc = (int64_t)cx * cy * cz;
    if (c >= nc) return;
    if (!(ccode[c] & kFluidBit)) {
        bc[c] = 0.0f;
        return;
    }
    const uint32_t cu = (uint32_t)c;
    const int I = (int)(cu % (uint32_t)cx), J = (int)((cu / (uint32_t)cx) % (uint32_t)cy),
              K = (int)(cu / ((uint32_t)cx * (uint32_t)cy));
    float wx[4], wy[4], wz[4];
    for (int t = 0; t < 4; ++t) {
        const int fi = 2 * I - 1 + t, fj = 2 * J - 1 + t, fk = 2 * K - 1 + t;
        wx[t] = (fi >= 0 && fi < fx) ? restrict_w(fi, I, cx) : 0.0f;
        wy[t] = (fj >= 0 && fj < fy) ? restrict_w(fj, J, cy) : 0.0f;
        wz[t] = (fk >= 0 && fk < fz) ? restrict_w(fk, K, cz) : 0.0f;
    }
    float acc = 0.0f;
    for (int tk = 0; tk < 4; ++tk) {
        if (wz[tk] == 0.0f) continue;
        const int64_t kk = 2 * K - 1 + tk;
        for (int tj = 0; tj < 4; ++tj) {
            const float wyz = wy[tj] * wz[tk];
            if (wyz == 0.0f) continue;
            const int64_t row = (int64_t)fx * ((int64_t)(2 * J - 1 + tj) + (int64_t)fy * kk);
            for (int ti = 0; ti < 4; ++ti) {
                if (wx[ti] == 0.0f) continue;
                acc += (wx[ti] * wyz) * rf[row + (2 * I - 1 + ti)];
            }
        }
    }
    bc[c] = acc * 0.125f;
}

// x_f += sum_c w(f, c) x_c on FLUID fine cells.
__global__ __launch_bounds__(kBlock) void k_mg_prolong(const float* __restrict__ xc, int cx, int cy, int cz,
                                                       const uint16_t* __restrict__ fcode, float* __restrict__ xf,
                                                       int fx, int fy, int fz, const SolverScalars* __restrict__ st) {
    if (st && st->done) return;
    const int64_t f = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t nf = (int64_t)fx * fy * fz;
    if (f >= nf) return;
    if (!(fcode[f] & kFluidBit)) return;
    const uint32_t fu = (uint32_t)f;
    const int i = (int)(fu % (uint32_t)fx), j = (int)((fu / (uint32_t)fx) % (uint32_t)fy),
              k = (int)(fu / ((uint32_t)fx * (uint32_t)fy));
    int ia, ib, ja, jb, ka, kb;
    float wia, wib, wja, wjb, wka, wkb;
    prolong_axis(i, cx, ia, wia, ib, wib);
    prolong_axis(j, cy, ja, wja, jb, wjb);
    prolong_axis(k, cz, ka, wka, kb, wkb);
    auto X = [&](int a, int b, int c) { return xc[(int64_t)a + (int64_t)cx * ((int64_t)b + (int64_t)cy * c)]; };
    const float v = wka * (wja * (wia * X(ia, ja, ka) + wib * X(ib, ja, ka)) +
                           wjb * (wia * X(ia, jb, ka) + wib * X(ib, jb, ka))) +
                    wkb * (wja * (wia * X(ia, ja, kb) + wib * X(ib, ja, kb)) +
                           wjb * (wia * X(ia, jb, kb) + wib * X(ib, jb, kb)));
    xf[f] += v;
}

// Coarsest level: damped Jacobi from zero, entirely in LDS (one workgroup).
__global__ __launch_bounds__(kCoarseMax) void k_mg_coarse(int nx, int ny, int nz, const uint16_t* __restrict__ code,
                                                           const float* __restrict__ b, float* __restrict__ x,
                                                           float ihx2, float ihy2, float ihz2, float omega,
                                                           int sweeps, const SolverScalars* __restrict__ st) {
    if (st && st->done) return;
    __shared__ float xs[2][kCoarseMax];
    __shared__ float bs[kCoarseMax];
    __shared__ uint16_t cs[kCoarseMax];
    const int nc = nx * ny * nz;
    const int c = threadIdx.x;
    const float ih2[3] = {ihx2, ihy2, ihz2};
    if (c < nc) {
        cs[c] = code[c];
        bs[c] = b[c];
        xs[0][c] = 0.0f;
    }
    __syncthreads();
    const Stride stv{nx, (int64_t)nx * ny};
    int cur = 0;
    for (int it = 0; it < sweeps; ++it) {
        if (c < nc) {
            const uint16_t cd = cs[c];
            const float dgn = code_center(cd, ih2);
            auto X = [&](int64_t n) { return xs[cur][n]; };
            const float xc = xs[cur][c];
            const float ax = code_apply(cd, dgn, c, stv, ih2, xc, X);
            xs[1 - cur][c] = dgn > 0.0f ? xc + omega * (bs[c] - ax) / dgn : 0.0f;
        }
        __syncthreads();
        cur = 1 - cur;
    }
    if (c < nc) x[c] = xs[cur][c];
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static inline dim3 grid1(int64_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

static LevelArgs level_args(const MGLevel& l) {
    const SolverShape sh = solver_shape_n(l.cells);
    LevelArgs a;
    a.nc = l.cells;
    a.sy = l.n[0];
    a.sz = (int64_t)l.n[0] * l.n[1];
    for (int q = 0; q < 3; ++q) a.ih2[q] = l.ih2[q];
    a.chunk = sh.chunk;
    a.padded = sh.blocks;
    return a;
}

void mg_free(fe_scene* s);

static hipError_t mg_alloc(fe_scene* s) {
    const Geom& g = s->g;
    std::vector<MGLevel> lv;
    MGLevel l0;
    for (int a = 0; a < 3; ++a) {
        l0.n[a] = g.n[a];
        l0.ih2[a] = (float)(g.ihd[a] * g.ihd[a]);
    }
    l0.cells = cell_count(g);
    lv.push_back(l0);
    while (lv.back().cells > kCoarseMax) {
        const MGLevel& f = lv.back();
        MGLevel c;
        for (int a = 0; a < 3; ++a) {
            c.n[a] = (f.n[a] + 1) / 2;
            c.ih2[a] = f.ih2[a] * 0.25f;  // twice the spacing
        }
        c.cells = (int64_t)c.n[0] * c.n[1] * c.n[2];
        if (c.cells == f.cells) break;  // cannot coarsen further
        c.owns = true;
        lv.push_back(c);
    }
    if (lv.back().cells > kCoarseMax) return hipErrorInvalidValue;
    hipError_t e = hipSuccess;
    for (size_t q = 0; q < lv.size(); ++q) {
        MGLevel& l = lv[q];
        const size_t n = (size_t)l.cells;
        if (l.owns) {
            if ((e = hipMalloc((void**)&l.flags, n)) != hipSuccess) break;
            if ((e = hipMalloc((void**)&l.code, n * sizeof(uint16_t))) != hipSuccess) break;
            if ((e = hipMalloc((void**)&l.b, n * sizeof(float))) != hipSuccess) break;
        }
        if ((e = hipMalloc((void**)&l.x, n * sizeof(float))) != hipSuccess) break;
        if ((e = hipMalloc((void**)&l.x2, n * sizeof(float))) != hipSuccess) break;
        if ((e = hipMalloc((void**)&l.r, n * sizeof(float))) != hipSuccess) break;
    }
    s->mg = lv;
    if (e != hipSuccess) mg_free(s);  // leaves the hierarchy empty: the next solve retries
    return e;
}

void mg_free(fe_scene* s) {
    for (MGLevel& l : s->mg) {
        if (l.owns) {
            (void)hipFree(l.flags);
            (void)hipFree(l.code);
            (void)hipFree(l.b);
        }
        (void)hipFree(l.x);
        (void)hipFree(l.x2);
        (void)hipFree(l.r);
    }
    s->mg.clear();
}

// Per solve: coarse markers and row codes from the current level-0 markers
// (the level-0 codes come from k_build_system).
hipError_t mg_setup(fe_scene* s) {
    hipError_t e;
    if (s->mg.empty() && (e = mg_alloc(s)) != hipSuccess) return e;
    s->mg[0].flags = s->flags;
    s->mg[0].code = s->code;
    for (size_t q = 1; q < s->mg.size(); ++q) {
        const MGLevel& f = s->mg[q - 1];
        MGLevel& c = s->mg[q];
        hipLaunchKernelGGL(k_mg_coarsen, grid1(c.cells), dim3(kBlock), 0, s->stream, (const uint8_t*)f.flags, f.n[0],
                           f.n[1], f.n[2], c.flags, c.n[0], c.n[1], c.n[2]);
        hipLaunchKernelGGL(k_mg_codes, grid1(c.cells), dim3(kBlock), 0, s->stream, (const uint8_t*)c.flags, c.n[0],
                           c.n[1], c.n[2], c.code);
    }
    return hipGetLastError();
}

// z = M^-1 r with one V-cycle; leaves partials of r.z in prz.  Every kernel
// returns at once when `st` says the solve has stopped.
hipError_t mg_apply(fe_scene* s, const float* r, float* z, const SolverScalars* st, double* prz) {
    std::vector<MGLevel>& lv = s->mg;
    const int L = (int)lv.size();
    lv[0].b = const_cast<float*>(r);
    std::vector<float*> cur(L);
    for (int q = 0; q < L - 1; ++q) {
        MGLevel& l = lv[q];
        const LevelArgs a = level_args(l);
        // pre-smoothing from zero
        hipLaunchKernelGGL(k_mg_smooth, dim3((unsigned)a.padded), dim3(kBlock), 0, s->stream, a,
                           (const uint16_t*)l.code, (const float*)l.b, (const float*)nullptr, l.x, kOmega, st,
                           (double*)nullptr);
        float* xa = l.x;
        float* xb = l.x2;
        for (int t = 1; t < kPreSweeps; ++t) {
            hipLaunchKernelGGL(k_mg_smooth, dim3((unsigned)a.padded), dim3(kBlock), 0, s->stream, a,
                               (const uint16_t*)l.code, (const float*)l.b, (const float*)xa, xb, kOmega, st,
                               (double*)nullptr);
            std::swap(xa, xb);
        }
        cur[q] = xa;
        hipLaunchKernelGGL(k_mg_residual, dim3((unsigned)a.padded), dim3(kBlock), 0, s->stream, a,
                           (const uint16_t*)l.code, (const float*)l.b, (const float*)xa, l.r, st);
        MGLevel& c = lv[q + 1];
        hipLaunchKernelGGL(k_mg_restrict, grid1(c.cells), dim3(kBlock), 0, s->stream, (const float*)l.r, l.n[0],
                           l.n[1], l.n[2], (const uint16_t*)c.code, c.b, c.n[0], c.n[1], c.n[2], st);
    }
    {
        MGLevel& c = lv[L - 1];
        hipLaunchKernelGGL(k_mg_coarse, dim3(1), dim3(kCoarseMax), 0, s->stream, c.n[0], c.n[1], c.n[2],
                           (const uint16_t*)c.code, (const float*)c.b, c.x, c.ih2[0], c.ih2[1], c.ih2[2],
                           kOmegaCoarse, kCoarseSweeps, st);
        cur[L - 1] = c.x;
    }
    for (int q = L - 2; q >= 0; --q) {
        MGLevel& l = lv[q];
        const MGLevel& c = lv[q + 1];
        const LevelArgs a = level_args(l);
        float* xa = cur[q];
        float* xb = xa == l.x ? l.x2 : l.x;
        hipLaunchKernelGGL(k_mg_prolong, grid1(l.cells), dim3(kBlock), 0, s->stream, (const float*)cur[q + 1], c.n[0],
                           c.n[1], c.n[2], (const uint16_t*)l.code, xa, l.n[0], l.n[1], l.n[2], st);
        for (int t = 0; t < kPreSweeps; ++t) {
            const bool last0 = (q == 0 && t == kPreSweeps - 1);
            float* out = last0 ? z : xb;
            hipLaunchKernelGGL(k_mg_smooth, dim3((unsigned)a.padded), dim3(kBlock), 0, s->stream, a,
                               (const uint16_t*)l.code, (const float*)l.b, (const float*)xa, out, kOmega, st,
                               last0 ? prz : (double*)nullptr);
            if (!last0) std::swap(xa, xb);
        }
        cur[q] = xa;
    }
    return hipGetLastError();
}

}  // namespace fe
