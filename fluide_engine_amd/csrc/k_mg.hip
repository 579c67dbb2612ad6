// k_mg.hip -- geometric multigrid V-cycle used as the preconditioner of the
// pressure CG (no reference counterpart: the reference's CG has an identity
// preconditioner, src/linear_system/conjugate_gradient_solver.cpp:57-63).
//
// Levels: level 0 is the pressure grid; each coarser level halves every axis
// (rounding up).  A coarse cell is AIR if any child is AIR, else FLUID if any
// child is FLUID, else BOUNDRY (the free surface keeps its Dirichlet cells on
// every level, as in McAdams, Sifakis and Teran, "A parallel multigrid Poisson
// solver for fluids simulation on large grids", SCA 2010), and its operator is
// BuildSystem's 7-point rule on those markers with twice the spacing.
//
// The V-cycle is symmetric, so the preconditioner is SPD and CG stays CG:
//   * damped Jacobi smoothing (omega = 2/3), two sweeps before and two after;
//   * prolongation: trilinear interpolation of the cell-centred coarse
//     correction (weights 3/4, 1/4 per axis; out-of-range coarse neighbours
//     clamp onto the boundary cell, which is the Neumann wall);
//   * restriction: the transpose of that prolongation divided by 8;
//   * coarsest level (<= 512 cells): 16 damped Jacobi sweeps (omega = 0.9)
//     from zero, a fixed polynomial in the operator.
// Non-FLUID cells carry no residual into the coarse levels and take no
// correction; on level 0 they are smoothed like the identity rows they are.
//
// Kernels per V-cycle on the large levels: the first sweep from zero is fused
// into the kernel that produces the level's right-hand side (the CG update on
// level 0, the restriction below it), then one sweep, the residual, the
// restriction; on the way up the prolongation and two sweeps.  All levels of
// at most 4096 cells run in one workgroup in LDS (k_mg_tail), which replaces
// about ten latency-bound launches.
#include <algorithm>

#include "fe_stencil.hpp"

namespace fe {

constexpr float kOmegaCoarse = 0.9f;  // coarsest-level Jacobi damping
constexpr int kCoarseSweeps = 16;
constexpr int kCoarseMax = 512;       // cells of the coarsest level
constexpr int kTailMax = 4096;        // levels of at most this many cells run in k_mg_tail
constexpr int kTailLds = 4608;        // cells of all tail levels together (3 floats + a row code each in LDS)
constexpr int kTailThreads = 1024;
constexpr int kMaxTail = 8;

// ---------------------------------------------------------------------------
// Level construction
// ---------------------------------------------------------------------------
__device__ inline void cell_ijk(int64_t c, int nx, int ny, int& i, int& j, int& k) {
    const uint32_t cu = (uint32_t)c;
    i = (int)(cu % (uint32_t)nx);
    j = (int)((cu / (uint32_t)nx) % (uint32_t)ny);
    k = (int)(cu / ((uint32_t)nx * (uint32_t)ny));
}

__global__ __launch_bounds__(kBlock) void k_mg_coarsen(const uint8_t* __restrict__ ff, int fx, int fy, int fz,
                                                       uint8_t* __restrict__ cf, int cx, int cy, int cz) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t nc = (int64_t)cx * cy * cz;
    if (c >= nc) return;
    int I, J, K;
    cell_ijk(c, cx, cy, I, J, K);
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
    int i, j, k;
    cell_ijk(c, nx, ny, i, j, k);
    code[c] = row_code(flags, c, (uint32_t)i, (uint32_t)j, (uint32_t)k, (uint32_t)nx, (uint32_t)ny, (uint32_t)nz);
}

// ---------------------------------------------------------------------------
// Transfer weights
// ---------------------------------------------------------------------------
// Prolongation of fine index fi on one axis: two coarse neighbours a, b.
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

// The fine indices 2I-1 .. 2I+2 whose prolongation reaches coarse index I,
// clamped into range, with their weights (0 outside the grid).
__device__ inline void restrict_taps(int I, int nf, int nc, int (&idx)[4], float (&w)[4]) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int fi = 2 * I - 1 + t;
        const bool in = fi >= 0 && fi < nf;
        idx[t] = min(max(fi, 0), nf - 1);
        w[t] = in ? restrict_w(fi, I, nc) : 0.0f;
    }
}

// (1/8) sum of the 4x4x4 fine window, 64 independent loads.
template <typename F>
__device__ inline float restrict_sum(int I, int J, int K, int fx, int fy, int fz, int cx, int cy, int cz, F R) {
    int ix[4], iy[4], iz[4];
    float wx[4], wy[4], wz[4];
    restrict_taps(I, fx, cx, ix, wx);
    restrict_taps(J, fy, cy, iy, wy);
    restrict_taps(K, fz, cz, iz, wz);
    float acc = 0.0f;
#pragma unroll
    for (int tk = 0; tk < 4; ++tk) {
#pragma unroll
        for (int tj = 0; tj < 4; ++tj) {
            const float wjk = wy[tj] * wz[tk];
            const int64_t row = (int64_t)fx * ((int64_t)iy[tj] + (int64_t)fy * iz[tk]);
#pragma unroll
            for (int ti = 0; ti < 4; ++ti) acc += (wx[ti] * wjk) * R(row + ix[ti]);
        }
    }
    return acc * 0.125f;
}

// Trilinear prolongation of the coarse correction at fine cell (i, j, k).
template <typename F>
__device__ inline float prolong_at(int i, int j, int k, int cx, int cy, int cz, F X) {
    int ia, ib, ja, jb, ka, kb;
    float wia, wib, wja, wjb, wka, wkb;
    prolong_axis(i, cx, ia, wia, ib, wib);
    prolong_axis(j, cy, ja, wja, jb, wjb);
    prolong_axis(k, cz, ka, wka, kb, wkb);
    auto V = [&](int a, int b, int c) { return X((int64_t)a + (int64_t)cx * ((int64_t)b + (int64_t)cy * c)); };
    return wka * (wja * (wia * V(ia, ja, ka) + wib * V(ib, ja, ka)) + wjb * (wia * V(ia, jb, ka) + wib * V(ib, jb, ka))) +
           wkb * (wja * (wia * V(ia, ja, kb) + wib * V(ib, ja, kb)) + wjb * (wia * V(ia, jb, kb) + wib * V(ib, jb, kb)));
}

// ---------------------------------------------------------------------------
// Large levels: two kernels per level and V-cycle.
// A workgroup owns a 16x16x8 tile of fine cells (an 8x8x4 tile of the coarse
// level) and recomputes the halo it needs from LDS instead of making one pass
// over the level per step:
//   down: x1 = w b on the tile + 3 cells, x2 = x1 + w (b - A x1) on the tile + 2
//         (written out for the tile), r = b - A x2 on the tile + 1, and the
//         restriction of r onto the coarse tile (the coarse right-hand side);
//   up:   xp = x2 + P xc on the tile + 2, one sweep on the tile + 1 and one on
//         the tile, written out (on level 0 as z, with partials of r . z).
// Grid cells outside the domain hold 0 in the windows; the row codes never
// couple to them.
// ---------------------------------------------------------------------------
constexpr int kRtx = 8, kRty = 8, kRtz = 4;  // coarse tile, one cell per thread
static_assert(kRtx * kRty * kRtz == kBlock, "one coarse cell per thread");
constexpr int kFtx = 2 * kRtx, kFty = 2 * kRty, kFtz = 2 * kRtz;  // fine tile

// Fine window of the tile widened by H cells on every side.
template <int H>
struct Win {
    static constexpr int x = kFtx + 2 * H, y = kFty + 2 * H, z = kFtz + 2 * H, n = x * y * z;
    static constexpr Stride32 st{x, x * y, n - 1};
    __device__ static int at(int wx, int wy, int wz) { return wx + x * (wy + y * wz); }
};

struct FineArgs {
    int n[3];         // fine level
    float ih2[3];
    int c[3];         // coarse level
    int tiles[3];     // tile grid
    int64_t padded;   // tile count padded for the XCD remap
};

// Tile coordinates of this workgroup (consecutive tiles share an XCD) and
// whether it maps past the last tile.
__device__ inline bool tile_of(const FineArgs& a, int& tx, int& ty, int& tz) {
    const int64_t t = xcd_remap(blockIdx.x, a.padded);
    if (t >= (int64_t)a.tiles[0] * a.tiles[1] * a.tiles[2]) return false;
    tx = (int)(t % a.tiles[0]);
    ty = (int)((t / a.tiles[0]) % a.tiles[1]);
    tz = (int)(t / ((int64_t)a.tiles[0] * a.tiles[1]));
    return true;
}

__device__ inline bool fine_inside(const FineArgs& a, int i, int j, int k) {
    return i >= 0 && j >= 0 && k >= 0 && i < a.n[0] && j < a.n[1] && k < a.n[2];
}

__device__ inline int64_t fine_index(const FineArgs& a, int i, int j, int k) {
    return (int64_t)i + (int64_t)a.n[0] * ((int64_t)j + (int64_t)a.n[1] * k);
}

__global__ __launch_bounds__(kBlock) void k_mg_down(FineArgs a, const float* __restrict__ b,
                                                    const uint16_t* __restrict__ code, float* __restrict__ x2out,
                                                    const uint16_t* __restrict__ ccode, float* __restrict__ bc,
                                                    const SolverScalars* __restrict__ st) {
    if (st->done) return;  // uniform: the solve has stopped
    int tx, ty, tz;
    if (!tile_of(a, tx, ty, tz)) return;
    __shared__ float s1[Win<3>::n];  // x1
    __shared__ float s2[Win<2>::n];  // x2
    __shared__ float s3[Win<1>::n];  // r
    const int o[3] = {tx * kFtx, ty * kFty, tz * kFtz};
    const int tid = threadIdx.x;
    // x1 = w b on the tile + 3
    for (int e = tid; e < Win<3>::n; e += kBlock) {
        const int wx = e % Win<3>::x, wy = (e / Win<3>::x) % Win<3>::y, wz = e / (Win<3>::x * Win<3>::y);
        const int i = o[0] - 3 + wx, j = o[1] - 3 + wy, k = o[2] - 3 + wz;
        float v = 0.0f;
        if (fine_inside(a, i, j, k)) {
            const int64_t f = fine_index(a, i, j, k);
            v = mg_weight(code[f], a.ih2) * b[f];
        }
        s1[e] = v;
    }
    __syncthreads();
    // x2 on the tile + 2
    for (int e = tid; e < Win<2>::n; e += kBlock) {
        const int wx = e % Win<2>::x, wy = (e / Win<2>::x) % Win<2>::y, wz = e / (Win<2>::x * Win<2>::y);
        const int i = o[0] - 2 + wx, j = o[1] - 2 + wy, k = o[2] - 2 + wz;
        float v = 0.0f;
        if (fine_inside(a, i, j, k)) {
            const int64_t f = fine_index(a, i, j, k);
            const uint16_t cd = code[f];
            const float dgn = code_center(cd, a.ih2);
            const int c3 = Win<3>::at(wx + 1, wy + 1, wz + 1);
            const float xc = s1[c3];
            v = xc + mg_weight(cd, a.ih2) *
                         (b[f] - code_apply(cd, dgn, c3, Win<3>::st, a.ih2, xc, [&](int n) { return s1[n]; }));
            if (wx >= 2 && wx < 2 + kFtx && wy >= 2 && wy < 2 + kFty && wz >= 2 && wz < 2 + kFtz) x2out[f] = v;
        }
        s2[e] = v;
    }
    __syncthreads();
    // r = b - A x2 on FLUID cells of the tile + 1
    for (int e = tid; e < Win<1>::n; e += kBlock) {
        const int wx = e % Win<1>::x, wy = (e / Win<1>::x) % Win<1>::y, wz = e / (Win<1>::x * Win<1>::y);
        const int i = o[0] - 1 + wx, j = o[1] - 1 + wy, k = o[2] - 1 + wz;
        float v = 0.0f;
        if (fine_inside(a, i, j, k)) {
            const int64_t f = fine_index(a, i, j, k);
            const uint16_t cd = code[f];
            if (cd & kFluidBit) {
                const int c2 = Win<2>::at(wx + 1, wy + 1, wz + 1);
                v = b[f] - code_apply(cd, code_center(cd, a.ih2), c2, Win<2>::st, a.ih2, s2[c2],
                                      [&](int n) { return s2[n]; });
            }
        }
        s3[e] = v;
    }
    __syncthreads();
    // restriction onto the coarse tile: fine index 2I-1+t sits at window
    // coordinate 2*(I - I0) + t (clamped taps carry weight 0)
    const int lx = tid % kRtx, ly = (tid / kRtx) % kRty, lz = tid / (kRtx * kRty);
    const int I = tx * kRtx + lx, J = ty * kRty + ly, K = tz * kRtz + lz;
    if (I >= a.c[0] || J >= a.c[1] || K >= a.c[2]) return;
    const int64_t cidx = (int64_t)I + (int64_t)a.c[0] * ((int64_t)J + (int64_t)a.c[1] * K);
    float v = 0.0f;
    if (ccode[cidx] & kFluidBit) {
        int ix[4], iy[4], iz[4];
        float wx[4], wy[4], wz[4];
        restrict_taps(I, a.n[0], a.c[0], ix, wx);
        restrict_taps(J, a.n[1], a.c[1], iy, wy);
        restrict_taps(K, a.n[2], a.c[2], iz, wz);
        float acc = 0.0f;
#pragma unroll
        for (int tk = 0; tk < 4; ++tk) {
#pragma unroll
            for (int tj = 0; tj < 4; ++tj) {
                const float wjk = wy[tj] * wz[tk];
#pragma unroll
                for (int ti = 0; ti < 4; ++ti)
                    acc += (wx[ti] * wjk) * s3[Win<1>::at(2 * lx + ti, 2 * ly + tj, 2 * lz + tk)];
            }
        }
        v = acc * 0.125f;
    }
    bc[cidx] = v;
}

__global__ __launch_bounds__(kBlock) void k_mg_up(FineArgs a, const float* __restrict__ x2, const float* __restrict__ xc,
                                                  const float* __restrict__ b, const uint16_t* __restrict__ code,
                                                  float* __restrict__ out, double* __restrict__ prz,
                                                  const SolverScalars* __restrict__ st) {
    if (st->done) return;
    int tx, ty, tz;
    const bool real = tile_of(a, tx, ty, tz);
    __shared__ float sp[Win<2>::n];  // x2 + P xc
    __shared__ float s1[Win<1>::n];  // first sweep
    double v[1] = {0.0};
    if (real) {
        const int o[3] = {tx * kFtx, ty * kFty, tz * kFtz};
        const int tid = threadIdx.x;
        for (int e = tid; e < Win<2>::n; e += kBlock) {
            const int wx = e % Win<2>::x, wy = (e / Win<2>::x) % Win<2>::y, wz = e / (Win<2>::x * Win<2>::y);
            const int i = o[0] - 2 + wx, j = o[1] - 2 + wy, k = o[2] - 2 + wz;
            float val = 0.0f;
            if (fine_inside(a, i, j, k)) {
                const int64_t f = fine_index(a, i, j, k);
                val = x2[f];
                if (code[f] & kFluidBit)
                    val += prolong_at(i, j, k, a.c[0], a.c[1], a.c[2], [&](int64_t n) { return xc[n]; });
            }
            sp[e] = val;
        }
        __syncthreads();
        for (int e = tid; e < Win<1>::n; e += kBlock) {
            const int wx = e % Win<1>::x, wy = (e / Win<1>::x) % Win<1>::y, wz = e / (Win<1>::x * Win<1>::y);
            const int i = o[0] - 1 + wx, j = o[1] - 1 + wy, k = o[2] - 1 + wz;
            float val = 0.0f;
            if (fine_inside(a, i, j, k)) {
                const int64_t f = fine_index(a, i, j, k);
                const uint16_t cd = code[f];
                const float dgn = code_center(cd, a.ih2);
                const int c2 = Win<2>::at(wx + 1, wy + 1, wz + 1);
                const float xv = sp[c2];
                val = xv + mg_weight(cd, a.ih2) *
                               (b[f] - code_apply(cd, dgn, c2, Win<2>::st, a.ih2, xv, [&](int n) { return sp[n]; }));
            }
            s1[e] = val;
        }
        __syncthreads();
        for (int e = tid; e < Win<0>::n; e += kBlock) {
            const int wx = e % Win<0>::x, wy = (e / Win<0>::x) % Win<0>::y, wz = e / (Win<0>::x * Win<0>::y);
            const int i = o[0] + wx, j = o[1] + wy, k = o[2] + wz;
            if (!fine_inside(a, i, j, k)) continue;
            const int64_t f = fine_index(a, i, j, k);
            const uint16_t cd = code[f];
            const float dgn = code_center(cd, a.ih2);
            const int c1 = Win<1>::at(wx + 1, wy + 1, wz + 1);
            const float xv = s1[c1];
            const float bf = b[f];
            const float val =
                xv + mg_weight(cd, a.ih2) *
                         (bf - code_apply(cd, dgn, c1, Win<1>::st, a.ih2, xv, [&](int n) { return s1[n]; }));
            out[f] = val;
            v[0] += (double)bf * (double)val;
        }
    }
    if (prz) write_partials<1>(v, prz);  // every workgroup, real or padding, leaves its partial
}

// ---------------------------------------------------------------------------
// Small levels: the rest of the V-cycle in one workgroup
// ---------------------------------------------------------------------------
struct TailArgs {
    int levels;  // tail levels; the last is the coarsest
    int n[kMaxTail][3];
    float ih2[kMaxTail][3];
    const uint16_t* code[kMaxTail];
    int off[kMaxTail];  // offset of each level in the LDS arrays
};

// Tail level t in LDS: values at V + o, row codes at Cd + o.
struct TailLevel {
    int nx, ny, nz, nc, o;
    Stride32 st;
    float ih2[3];
};

__device__ inline TailLevel tail_level(const TailArgs& a, int t) {
    TailLevel l;
    l.nx = a.n[t][0];
    l.ny = a.n[t][1];
    l.nz = a.n[t][2];
    l.nc = l.nx * l.ny * l.nz;
    l.o = a.off[t];
    l.st = Stride32{l.nx, l.nx * l.ny, l.nc - 1};
    for (int q = 0; q < 3; ++q) l.ih2[q] = a.ih2[t][q];
    return l;
}

// One Jacobi sweep: out = in + w (b - A in), or out = w b when `in` is null.
__device__ inline void tail_sweep(const TailLevel& l, const uint16_t* code, const float* in, float* out,
                                  const float* b, float omega) {
    for (int c = threadIdx.x; c < l.nc; c += kTailThreads) {
        const uint16_t cd = code[c];
        const float dgn = code_center(cd, l.ih2);
        const float w = mg_weight(cd, l.ih2, omega);
        float xo;
        if (in) {
            const float xc = in[c];
            xo = xc + w * (b[c] - code_apply(cd, dgn, c, l.st, l.ih2, xc, [&](int n) { return in[n]; }));
        } else {
            xo = w * b[c];
        }
        out[c] = xo;
    }
    __syncthreads();
}

// Solves level T's right-hand side bT approximately with the V-cycle of the
// tail levels; the result goes to xT.  Everything lives in LDS: three values
// and the row code per cell of every tail level (at most 4608 cells, 63 KiB).
__global__ __launch_bounds__(kTailThreads) void k_mg_tail(TailArgs a, const float* __restrict__ bT,
                                                          float* __restrict__ xT,
                                                          const SolverScalars* __restrict__ st) {
    if (st->done) return;
    __shared__ float X[kTailLds], Tm[kTailLds], B[kTailLds];
    __shared__ uint16_t Cd[kTailLds];
    for (int t = 0; t < a.levels; ++t) {
        const int nc = a.n[t][0] * a.n[t][1] * a.n[t][2];
        for (int c = threadIdx.x; c < nc; c += kTailThreads) Cd[a.off[t] + c] = a.code[t][c];
    }
    const int n0 = a.n[0][0] * a.n[0][1] * a.n[0][2];
    for (int c = threadIdx.x; c < n0; c += kTailThreads) B[c] = bT[c];
    __syncthreads();
    // down: the solution of every level ends in Tm after its pre-smoothing
    for (int t = 0; t < a.levels - 1; ++t) {
        const TailLevel l = tail_level(a, t), lc = tail_level(a, t + 1);
        const int o = l.o;
        const uint16_t* code = Cd + o;
        tail_sweep(l, code, nullptr, X + o, B + o, kMgOmega);
        tail_sweep(l, code, X + o, Tm + o, B + o, kMgOmega);
        // residual into X (free now)
        {
            const float* xs = Tm + o;
            for (int c = threadIdx.x; c < l.nc; c += kTailThreads) {
                const uint16_t cd = code[c];
                float rc = 0.0f;
                if (cd & kFluidBit)
                    rc = B[o + c] - code_apply(cd, code_center(cd, l.ih2), c, l.st, l.ih2, xs[c],
                                               [&](int n) { return xs[n]; });
                X[o + c] = rc;
            }
            __syncthreads();
        }
        // restriction into the next level's B
        {
            const float* rs = X + o;
            for (int c = threadIdx.x; c < lc.nc; c += kTailThreads) {
                float v = 0.0f;
                if (Cd[lc.o + c] & kFluidBit) {
                    int I, J, K;
                    cell_ijk(c, lc.nx, lc.ny, I, J, K);
                    v = restrict_sum(I, J, K, l.nx, l.ny, l.nz, lc.nx, lc.ny, lc.nz,
                                     [&](int64_t n) { return rs[n]; });
                }
                B[lc.o + c] = v;
            }
            __syncthreads();
        }
    }
    // coarsest level: Jacobi sweeps from zero, ending in Tm
    {
        const TailLevel l = tail_level(a, a.levels - 1);
        const int o = l.o;
        const uint16_t* code = Cd + o;
        static_assert(kCoarseSweeps % 2 == 0, "the sweeps ping-pong X -> Tm");
        tail_sweep(l, code, nullptr, X + o, B + o, kOmegaCoarse);
        tail_sweep(l, code, X + o, Tm + o, B + o, kOmegaCoarse);
        for (int s = 2; s < kCoarseSweeps; s += 2) {
            tail_sweep(l, code, Tm + o, X + o, B + o, kOmegaCoarse);
            tail_sweep(l, code, X + o, Tm + o, B + o, kOmegaCoarse);
        }
    }
    // up: prolongation, two sweeps Tm -> X -> Tm
    for (int t = a.levels - 2; t >= 0; --t) {
        const TailLevel l = tail_level(a, t), lc = tail_level(a, t + 1);
        const int o = l.o;
        const uint16_t* code = Cd + o;
        const float* xc = Tm + lc.o;
        for (int f = threadIdx.x; f < l.nc; f += kTailThreads) {
            if (!(code[f] & kFluidBit)) continue;
            int i, j, k;
            cell_ijk(f, l.nx, l.ny, i, j, k);
            Tm[o + f] += prolong_at(i, j, k, lc.nx, lc.ny, lc.nz, [&](int64_t n) { return xc[n]; });
        }
        __syncthreads();
        tail_sweep(l, code, Tm + o, X + o, B + o, kMgOmega);
        tail_sweep(l, code, X + o, Tm + o, B + o, kMgOmega);
    }
    for (int c = threadIdx.x; c < n0; c += kTailThreads) xT[c] = Tm[c];
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static inline dim3 grid1(int64_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

static FineArgs fine_args(const MGLevel& l, const MGLevel& c) {
    FineArgs a;
    for (int q = 0; q < 3; ++q) {
        a.n[q] = l.n[q];
        a.ih2[q] = l.ih2[q];
        a.c[q] = c.n[q];
    }
    a.tiles[0] = (c.n[0] + kRtx - 1) / kRtx;
    a.tiles[1] = (c.n[1] + kRty - 1) / kRty;
    a.tiles[2] = (c.n[2] + kRtz - 1) / kRtz;
    a.padded = pad_blocks((int64_t)a.tiles[0] * a.tiles[1] * a.tiles[2]);
    return a;
}

void mg_free(fe_scene* s);

// First level of the tail: the first level below the finest with at most
// kTailMax cells whose tail fits the LDS arrays and the level table.
static int tail_start(const std::vector<MGLevel>& lv) {
    const int L = (int)lv.size();
    for (int q = 1; q < L; ++q) {
        int64_t sum = 0;
        for (int t = q; t < L; ++t) sum += lv[t].cells;
        if (lv[q].cells <= kTailMax && sum <= kTailLds && L - q <= kMaxTail) return q;
    }
    return L - 1;
}

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
    // every axis has at least 2 cells, so the first coarsening always shrinks
    while (lv.back().cells > kCoarseMax || lv.size() < 2) {
        const MGLevel& f = lv.back();
        MGLevel c;
        for (int a = 0; a < 3; ++a) {
            c.n[a] = (f.n[a] + 1) / 2;
            c.ih2[a] = f.ih2[a] * 0.25f;  // twice the spacing
        }
        c.cells = (int64_t)c.n[0] * c.n[1] * c.n[2];
        if (c.cells == f.cells) break;
        c.owns = true;
        lv.push_back(c);
    }
    if (lv.size() < 2 || lv.back().cells > kCoarseMax) return hipErrorInvalidValue;
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
    }
    s->mg_npz = fine_args(lv[0], lv[1]).padded;
    if (e == hipSuccess) e = hipMalloc((void**)&s->mg_pz, (size_t)s->mg_npz * sizeof(double));
    s->mg = lv;
    s->mg_tail = tail_start(lv);
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
    }
    s->mg.clear();
    (void)hipFree(s->mg_pz);
    s->mg_pz = nullptr;
    s->mg_npz = 0;
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

// z = M^-1 r with one V-cycle; leaves s->mg_npz partials of r.z in prz.
// Every kernel returns at once when `st` says the solve has stopped.
hipError_t mg_apply(fe_scene* s, const float* r, float* z, const SolverScalars* st, double* prz, int tag) {
    std::vector<MGLevel>& lv = s->mg;
    const int L = (int)lv.size();
    const int T = s->mg_tail;
    lv[0].b = const_cast<float*>(r);
    for (int q = 0; q < T; ++q) {
        const MGLevel& l = lv[q];
        MGLevel& c = lv[q + 1];
        const FineArgs a = fine_args(l, c);
        const dim3 grid((unsigned)a.padded);
        if (q == 0)
            launch_timed(s, FE_KERNEL_MG_DOWN, tag, k_mg_down, grid, dim3(kBlock), a, (const float*)l.b,
                         (const uint16_t*)l.code, l.x2, (const uint16_t*)c.code, c.b, st);
        else
            hipLaunchKernelGGL(k_mg_down, grid, dim3(kBlock), 0, s->stream, a, (const float*)l.b,
                               (const uint16_t*)l.code, l.x2, (const uint16_t*)c.code, c.b, st);
    }
    {
        TailArgs ta{};
        ta.levels = L - T;
        int off = 0;
        for (int t = 0; t < ta.levels; ++t) {
            const MGLevel& l = lv[T + t];
            for (int q = 0; q < 3; ++q) {
                ta.n[t][q] = l.n[q];
                ta.ih2[t][q] = l.ih2[q];
            }
            ta.code[t] = l.code;
            ta.off[t] = off;
            off += (int)l.cells;
        }
        hipLaunchKernelGGL(k_mg_tail, dim3(1), dim3(kTailThreads), 0, s->stream, ta, (const float*)lv[T].b, lv[T].x,
                           st);
    }
    // the solution of every level ends in its x
    for (int q = T - 1; q >= 0; --q) {
        MGLevel& l = lv[q];
        const MGLevel& c = lv[q + 1];
        const FineArgs a = fine_args(l, c);
        const dim3 grid((unsigned)a.padded);
        if (q == 0)
            launch_timed(s, FE_KERNEL_MG_UP, tag, k_mg_up, grid, dim3(kBlock), a, (const float*)l.x2,
                         (const float*)c.x, (const float*)l.b, (const uint16_t*)l.code, z, prz, st);
        else
            hipLaunchKernelGGL(k_mg_up, grid, dim3(kBlock), 0, s->stream, a, (const float*)l.x2, (const float*)c.x,
                               (const float*)l.b, (const uint16_t*)l.code, l.x, (double*)nullptr, st);
    }
    return hipGetLastError();
}

}  // namespace fe
