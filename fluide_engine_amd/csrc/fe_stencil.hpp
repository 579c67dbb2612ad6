// fe_stencil.hpp -- matrix-free 7-point rows of the pressure system and the
// deterministic block reductions shared by the CG and multigrid kernels.
#pragma once

#include <algorithm>

#include "fe_kernels.hpp"

namespace fe {

// Grid-wide kernels with reductions run at most this many blocks over
// contiguous cell chunks (one FP64 partial per block and value).
constexpr int kMaxSolverBlocks = 1024;

struct SolverShape {
    int64_t blocks;  // multiple of 8 (XCD remap)
    int64_t chunk;   // cells per block, multiple of kBlock
};

inline SolverShape solver_shape_n(int64_t nc) {
    int64_t blocks = std::min<int64_t>(pad_blocks((nc + kBlock - 1) / kBlock), kMaxSolverBlocks);
    int64_t chunk = (nc + blocks - 1) / blocks;
    chunk = (chunk + kBlock - 1) / kBlock * kBlock;
    blocks = pad_blocks((nc + chunk - 1) / chunk);
    return {blocks, chunk};
}

// ---------------------------------------------------------------------------
// The 16-bit row code.
//   bits 0-5  : neighbour (i-1, i+1, j-1, j+1, k-1, k+1) is coupled (both FLUID)
//   bits 6-11 : per axis x, y, z, 2 bits: in-range non-BOUNDRY neighbour count
//   bit 12    : the cell is FLUID
// ---------------------------------------------------------------------------
constexpr uint16_t kFluidBit = 1u << 12;

template <typename I>
struct StrideT {
    I sy, sz;  // +-1 in j, k
    I last;    // cells - 1
};
using Stride = StrideT<int64_t>;
using Stride32 = StrideT<int>;  // grids held in LDS

__host__ __device__ inline Stride stride_of(int64_t nx, int64_t ny, int64_t nz) {
    return Stride{nx, nx * ny, nx * ny * nz - 1};
}

__device__ inline float code_center(uint16_t code, const float ih2[3]) {
    if (!(code & kFluidBit)) return 1.0f;
    // BuildSystem's accumulation order: +x, -x, +y, -y, +z, -z
    float c = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int n = (code >> (6 + 2 * a)) & 3;
        if (n >= 1) c += ih2[a];
        if (n >= 2) c += ih2[a];
    }
    return c;
}

__device__ inline double code_center64(uint16_t code, const double ih2[3]) {
    if (!(code & kFluidBit)) return 1.0;
    double c = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int n = (code >> (6 + 2 * a)) & 3;
        if (n >= 1) c += ih2[a];
        if (n >= 2) c += ih2[a];
    }
    return c;
}

// (A v)(c) in the reference SpMV order (src/linear_system/blas.cpp:72-83):
// center*v(c) + the coupled neighbours i-1, i+1, j-1, j+1, k-1, k+1.  All six
// neighbour values are fetched (at indices clamped into the grid) before the
// row code is consulted, so the loads never wait on it; an uncoupled
// neighbour contributes fma(0, v, acc) = acc (every vector holds finite values).
template <typename I, typename F, typename T>
__device__ inline void code_neighbours(I c, const StrideT<I>& st, F v, T (&nb)[6]) {
    nb[0] = v(c > 0 ? c - 1 : 0);
    nb[1] = v(c < st.last ? c + 1 : st.last);
    nb[2] = v(c >= st.sy ? c - st.sy : 0);
    nb[3] = v(c + st.sy <= st.last ? c + st.sy : st.last);
    nb[4] = v(c >= st.sz ? c - st.sz : 0);
    nb[5] = v(c + st.sz <= st.last ? c + st.sz : st.last);
}

__device__ inline float code_apply_nb(uint16_t code, float center, const float ih2[3], float vc, const float (&nb)[6]) {
    float acc = center * vc;
    acc = fmaf((code & 1) ? -ih2[0] : 0.0f, nb[0], acc);
    acc = fmaf((code & 2) ? -ih2[0] : 0.0f, nb[1], acc);
    acc = fmaf((code & 4) ? -ih2[1] : 0.0f, nb[2], acc);
    acc = fmaf((code & 8) ? -ih2[1] : 0.0f, nb[3], acc);
    acc = fmaf((code & 16) ? -ih2[2] : 0.0f, nb[4], acc);
    acc = fmaf((code & 32) ? -ih2[2] : 0.0f, nb[5], acc);
    return acc;
}

template <typename I, typename F>
__device__ inline float code_apply(uint16_t code, float center, I c, const StrideT<I>& st, const float ih2[3],
                                   float vc, F v) {
    float nb[6];
    code_neighbours(c, st, v, nb);
    return code_apply_nb(code, center, ih2, vc, nb);
}

// A thread's cells of a block span (c0 + threadIdx.x, then every kBlock-th)
// in groups of G: load(c) fetches what cell c needs into a T and use(c, t)
// consumes it.  A full group issues all G loads before the first use, so the
// group costs one memory latency instead of G; a partial group at the end of
// the span runs cell by cell.  Cells are used in ascending order either way.
// span_groups from a thread's first cell `first` (threadIdx.x included) on.
template <int G, typename T, typename Load, typename Use>
__device__ inline void span_groups_from(int64_t first, int64_t c1, Load load, Use use) {
    for (int64_t g0 = first; g0 < c1; g0 += (int64_t)G * kBlock) {
        if (g0 + (int64_t)(G - 1) * kBlock < c1) {
            T t[G];
#pragma unroll
            for (int u = 0; u < G; ++u) t[u] = load(g0 + (int64_t)u * kBlock);
#pragma unroll
            for (int u = 0; u < G; ++u) use(g0 + (int64_t)u * kBlock, t[u]);
        } else {
            for (int64_t c = g0; c < c1; c += kBlock) use(c, load(c));
        }
    }
}

template <int G, typename T, typename Load, typename Use>
__device__ inline void span_groups(int64_t c0, int64_t c1, Load load, Use use) {
    for (int64_t g0 = c0 + threadIdx.x; g0 < c1; g0 += (int64_t)G * kBlock) {
        if (g0 + (int64_t)(G - 1) * kBlock < c1) {
            T t[G];
#pragma unroll
            for (int u = 0; u < G; ++u) t[u] = load(g0 + (int64_t)u * kBlock);
#pragma unroll
            for (int u = 0; u < G; ++u) use(g0 + (int64_t)u * kBlock, t[u]);
        } else {
            for (int64_t c = g0; c < c1; c += kBlock) use(c, load(c));
        }
    }
}

// code_apply64 on neighbour values fetched beforehand (clamped indices, as
// code_neighbours), same arithmetic in the same order.
__device__ inline double code_apply64_nb(uint16_t code, const double ih2[3], double vc, const double (&nb)[6]) {
    if (!(code & kFluidBit)) return vc;
    double off = 0.0;
    if (code & 1) off -= ih2[0] * nb[0];
    if (code & 2) off -= ih2[0] * nb[1];
    if (code & 4) off -= ih2[1] * nb[2];
    if (code & 8) off -= ih2[1] * nb[3];
    if (code & 16) off -= ih2[2] * nb[4];
    if (code & 32) off -= ih2[2] * nb[5];
    return code_center64(code, ih2) * vc + off;
}

template <typename F>
__device__ inline double code_apply64(uint16_t code, int64_t c, const Stride& st, const double ih2[3], double vc,
                                      F v) {
    if (!(code & kFluidBit)) return vc;
    double off = 0.0;
    if (code & 1) off -= ih2[0] * v(c - 1);
    if (code & 2) off -= ih2[0] * v(c + 1);
    if (code & 4) off -= ih2[1] * v(c - st.sy);
    if (code & 8) off -= ih2[1] * v(c + st.sy);
    if (code & 16) off -= ih2[2] * v(c - st.sz);
    if (code & 32) off -= ih2[2] * v(c + st.sz);
    return code_center64(code, ih2) * vc + off;
}

// Damped Jacobi weight omega / A(c,c) of the multigrid smoother (0 for a FLUID
// cell with no diagonal, i.e. walled in by BOUNDRY cells on every side).
// omega = 6/7 minimises the smoothing factor of damped Jacobi on the 3-D
// 7-point Laplacian (the high-frequency modes' largest amplification); below
// 1, so the smoother stays convergent and the V-cycle SPD.  Measured on the
// 128^3 dam break: 11 multigrid-PCG iterations against 12 with omega = 2/3.
// Multigrid smoothing: two steps of Chebyshev iteration on D^-1 A over
// [lo, 2] per smoothing (2 bounds every eigenvalue of D^-1 A, each row's
// diagonal being at least the sum of its off-diagonal magnitudes):
//   x1 = x0 + omega D^-1 (b - A x0),
//   x2 = x1 + beta (x1 - x0) + w2 D^-1 (b - A x1),  w2 = w21 * (omega / diag)
// (MgSmooth; the finest level and the levels below it may use different
// intervals).
// The same polynomial before and after the coarse correction keeps the
// V-cycle symmetric, and its residual polynomial lies in (0, 1) on (0, 2], so
// the preconditioner is SPD.
constexpr double kMgSmHi = 2.0;
constexpr double kMgSmLo = 0.15;        // the finest level's interval [kMgSmLo, 2]
constexpr double kMgSmLoCoarse = 0.07;  // the levels below it
struct MgSmooth {
    float omega;  // the first step's weight (times 1/diag): 1/theta
    float beta;   // the second step's (x1 - x0) coefficient
    float w21;    // the second step's weight over the first's
};
constexpr MgSmooth mg_smooth_of(double lo) {
    const double theta = 0.5 * (lo + kMgSmHi), delta = 0.5 * (kMgSmHi - lo);
    const double rho0 = delta / theta, rho1 = 1.0 / (2.0 * theta / delta - rho0);
    return MgSmooth{(float)(1.0 / theta), (float)(rho1 * rho0), (float)(2.0 * rho1 / delta * theta)};
}
constexpr MgSmooth kMgFine = mg_smooth_of(kMgSmLo), kMgCoarse = mg_smooth_of(kMgSmLoCoarse);
constexpr float kMgOmega = kMgFine.omega;  // the finest level's first step (level 0's first sweep from zero)

// The second smoothing step from x1 (x0 the value before smoothing, r the
// residual b - A x1, w2 = w21 * the first step's weight); fixed rounding
// (no contraction), so every kernel that smooths gives the same bits.
__device__ inline float mg_cheb2(float x1, float x0, float w2, float r, float beta) {
#pragma clang fp contract(off)
    return fmaf(w2, r, fmaf(beta, x1 - x0, x1));
}

__device__ inline float mg_weight(uint16_t code, const float ih2[3], float omega = kMgOmega) {
    const float dg = code_center(code, ih2);
    return dg > 0.0f ? omega * __builtin_amdgcn_rcpf(dg) : 0.0f;  // hardware reciprocal: a smoother weight
}

// CG stopping rule: relative ||r|| <= tol ||b||, else the reference's sigma = r.z <= tol^2.
__device__ inline bool stop_rule(const SolverScalars* sc, double bnorm2, double sigma, double rr) {
    if (sc->relative) return sqrt(rr) <= sc->tol * sqrt(bnorm2);
    return !(sigma > sc->tol * sc->tol);
}

// ---------------------------------------------------------------------------
// Launch shape and block-level FP64 reductions.
// ---------------------------------------------------------------------------
struct Span {
    int64_t c0, c1;
    int64_t lb;  // the block's chunk index
};

__device__ inline Span block_span(int64_t nc, int64_t chunk, int64_t padded) {
    const int64_t lb = xcd_remap(blockIdx.x, padded);
    Span s;
    s.c0 = lb * chunk;
    s.c1 = min(s.c0 + chunk, nc);
    s.lb = lb;
    return s;
}

template <int NV>
__device__ inline void block_reduce(double (&v)[NV], double (*sm)[kBlock / 64]) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double x = v[q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if (lane == 0) sm[q][wid] = x;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) t += sm[q][w];
        v[q] = t;  // every thread holds the block total
    }
    __syncthreads();
}

// One FP64 partial per block and value.
template <int NV>
__device__ inline void write_partials(double (&v)[NV], double* __restrict__ part) {
    __shared__ double sm[NV][kBlock / 64];
    block_reduce<NV>(v, sm);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) part[(int64_t)blockIdx.x * NV + q] = v[q];
    }
}

// A thread's share of a partials array: entries b = first, first + step, ...
// below n, added in that order.  The loads of U entries are issued before the
// first add, so a share costs one memory latency per group
// instead of one per entry; the order of the adds is unchanged.
constexpr int kPartialGroup = 8;

template <int NV, int U = kPartialGroup>
__device__ inline void add_partials(const double* __restrict__ part, int stride, int off, int first, int step, int n,
                                    double (&tot)[NV]) {
    for (int b0 = first; b0 < n; b0 += U * step) {
        double v[U][NV];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int b = b0 + u * step;
#pragma unroll
            for (int q = 0; q < NV; ++q) v[u][q] = b < n ? part[(int64_t)b * stride + off + q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (b0 + u * step < n) {
#pragma unroll
                for (int q = 0; q < NV; ++q) tot[q] += v[u][q];
            }
        }
    }
}

// Sum of the previous kernel's partials, identical in every block.
template <int NV>
__device__ inline void sum_partials(const double* __restrict__ part, int nparts, double (&tot)[NV]) {
    __shared__ double sm[NV][kBlock / 64];
#pragma unroll
    for (int q = 0; q < NV; ++q) tot[q] = 0.0;
    add_partials<NV>(part, NV, 0, (int)threadIdx.x, kBlock, nparts, tot);
    block_reduce<NV>(tot, sm);
}

// Two sums in one pass: tot[0] over a[sa * b + oa] (b < na), tot[1] over
// c[sc * b + oc] (b < nc), each in sum_partials' order, so the values are the same.
__device__ inline void sum_partials_two(const double* __restrict__ a, int sa, int oa, int na,
                                        const double* __restrict__ c, int sc, int oc, int nc, double (&tot)[2]) {
    __shared__ double sm[2][kBlock / 64];
    double ta[1] = {0.0}, tc[1] = {0.0};
    add_partials<1>(a, sa, oa, (int)threadIdx.x, kBlock, na, ta);
    add_partials<1>(c, sc, oc, (int)threadIdx.x, kBlock, nc, tc);
    tot[0] = ta[0];
    tot[1] = tc[0];
    block_reduce<2>(tot, sm);
}

// Row code of cell (i,j,k) from the markers with BuildSystem's rules
// (src/fluid_solvers/single_phase_pressure_solver.cpp:116-161).
__device__ inline uint16_t row_code(const uint8_t* __restrict__ flags, int64_t c, uint32_t i, uint32_t j, uint32_t k,
                                    uint32_t nx, uint32_t ny, uint32_t nz) {
    if (flags[c] != FE_FLUID) return 0;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    const bool in[6] = {i > 0, i + 1 < nx, j > 0, j + 1 < ny, k > 0, k + 1 < nz};
    const int64_t nb[6] = {c - 1, c + 1, c - sy, c + sy, c - sz, c + sz};
    uint16_t cd = kFluidBit;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        if (!in[q]) continue;
        const uint8_t f = flags[nb[q]];
        if (f == FE_BOUNDRY) continue;
        cd = (uint16_t)(cd + (1u << (6 + 2 * (q >> 1))));
        if (f == FE_FLUID) cd |= (uint16_t)(1u << q);
    }
    return cd;
}

}  // namespace fe
