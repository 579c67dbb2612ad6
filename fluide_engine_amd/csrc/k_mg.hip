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
//   * smoothing: two steps of Chebyshev iteration on D^-1 A over [0.15, 2]
//     on the finest level and [0.07, 2] below it, before and the same two after (mg_cheb2; levels inside the
//     one-workgroup tail on small grids: two damped Jacobi sweeps with the
//     first step's weight before and after);
//   * prolongation: trilinear interpolation of the cell-centred coarse
//     correction (weights 3/4, 1/4 per axis; out-of-range coarse neighbours
//     clamp onto the boundary cell, which is the Neumann wall);
//   * restriction: the transpose of that prolongation divided by 8;
//   * coarsest level (<= 512 cells): 8 steps of Chebyshev iteration on D^-1 A over [0.05, 2]
//     from zero, a fixed polynomial in the operator.
// Non-FLUID cells carry no residual into the coarse levels and take no
// correction; on level 0 they are smoothed like the identity rows they are.
//
// Kernels per V-cycle on the large levels: the first sweep from zero is fused
// into the kernel that produces the level's right-hand side (the CG update on
// level 0, the restriction below it), then one sweep, the residual, the
// restriction; on the way up the prolongation and two sweeps.  The coarsest
// level (at most 512 cells) runs in one workgroup in LDS (k_mg_tail).  The
// tail once also took the next level up (16^3 at 128^3): one workgroup on one
// CU then did 4096 cells' sweeps, about 15 us per V-cycle, where the fused
// kernels' 8 workgroups take 11 us for that level's descent and ascent.
#include <algorithm>

#include "fe_stencil.hpp"

namespace fe {

// Coarsest level: kCoarseSweeps steps of Chebyshev iteration from zero on
// D^-1 A over [kChebLo, kChebHi].  Gershgorin bounds every eigenvalue of D^-1 A
// (diagonal >= the sum of the off-diagonal magnitudes on every row) by 2, so
// with kChebHi = 2 the residual polynomial's Chebyshev form lies in (0, 1] at
// every eigenvalue, below kChebLo included: the step is a fixed symmetric
// positive-definite polynomial of D^-1 A, as PCG needs.
constexpr int kCoarseSweeps = 8;
constexpr float kChebLo = 0.05f, kChebHi = 2.0f;
constexpr int kCoarseMax = 512;       // cells of the coarsest level
constexpr int kTailMax = kCoarseMax;  // levels of at most this many cells run in k_mg_tail
constexpr int kTailLds = kCoarseMax;  // cells of all tail levels together (5 floats + a row code each in LDS)
constexpr int kTailThreads = kTailMax;  // one cell per thread
constexpr int kMaxTail = 8;
constexpr int kTailPer = (kTailMax + kTailThreads - 1) / kTailThreads;  // cells per thread on a tail level

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

// The coarse markers of levels 1 .. m (m <= kPyr) in one launch: a
// workgroup takes a block of 8 x 8 x 8 level-1 cells and the pyramid above
// it (4^3, 2^3 and 1 cells of levels 2, 3, 4; a level-q cell's children are
// level-(q-1) cells of the same block), each level from the one below in LDS
// by k_mg_coarsen's rule -- AIR if a child is AIR, else FLUID if a child is
// FLUID, else BOUNDRY, over the children that exist.  It also leaves the
// fine-tile flags of the fused V-cycle's finest level (a block's level-0
// region is two 16 x 16 x 8 tiles): whether the tile holds a FLUID cell.
constexpr int kPyr = 4;
constexpr int kPyrThreads = 512;
constexpr uint8_t kNoCell = 255;  // a cell beyond its level's dims

struct PyrArgs {
    int levels;           // m, levels 1 .. m computed here
    int n[kPyr + 1][3];   // dims of levels 0 .. m
    uint8_t* flags[kPyr + 1];
    uint8_t* tfl;         // tile flags of level 0 (null: not wanted)
    int tiles[3];         // tile counts of level 0's fused tiles
};

__global__ __launch_bounds__(kPyrThreads) void k_mg_pyramid(PyrArgs a) {
    __shared__ uint8_t L[2][512];
    const int t = (int)threadIdx.x;
    const int bx = (a.n[1][0] + 7) / 8, by = (a.n[1][1] + 7) / 8;
    const int b = (int)blockIdx.x;
    const int B[3] = {b % bx, (b / bx) % by, b / (bx * by)};
    // level 1 from level 0, plus the tile flags
    int fluid0 = 0;
    {
        const int I = 8 * B[0] + t % 8, J = 8 * B[1] + (t / 8) % 8, K = 8 * B[2] + t / 64;
        uint8_t out = kNoCell;
        if (I < a.n[1][0] && J < a.n[1][1] && K < a.n[1][2]) {
            const int fx = a.n[0][0], fy = a.n[0][1], fz = a.n[0][2];
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
                        const uint8_t f = a.flags[0][(int64_t)i + (int64_t)fx * ((int64_t)j + (int64_t)fy * k)];
                        fluid |= f == FE_FLUID;
                        air |= f == FE_AIR;
                    }
                }
            }
            out = air ? (uint8_t)FE_AIR : (fluid ? (uint8_t)FE_FLUID : (uint8_t)FE_BOUNDRY);
            a.flags[1][(int64_t)I + (int64_t)a.n[1][0] * ((int64_t)J + (int64_t)a.n[1][1] * K)] = out;
            fluid0 = fluid ? 1 : 0;
        }
        L[0][t] = out;
    }
    if (a.tfl) {
        // tiles of 8 x 8 x 4 level-1 cells: this block's lower and upper halves in z
        const int lower = __syncthreads_or(t < 256 ? fluid0 : 0);
        const int upper = __syncthreads_or(t >= 256 ? fluid0 : 0);
        if (t < 2) {
            const int tz = 2 * B[2] + t;
            if (B[0] < a.tiles[0] && B[1] < a.tiles[1] && tz < a.tiles[2])
                a.tfl[(int64_t)B[0] + (int64_t)a.tiles[0] * ((int64_t)B[1] + (int64_t)a.tiles[1] * tz)] =
                    (uint8_t)(t == 0 ? lower : upper);
        }
    } else {
        __syncthreads();
    }
    // levels 2 .. m from the level below, in LDS
    int side = 8;
    for (int q = 2; q <= a.levels; ++q) {
        const uint8_t* below = L[q & 1];
        uint8_t* here = L[(q + 1) & 1];
        const int cs = side / 2;  // this level's block side
        if (t < cs * cs * cs) {
            const int tx = t % cs, ty = (t / cs) % cs, tz = t / (cs * cs);
            const int I = cs * B[0] + tx, J = cs * B[1] + ty, K = cs * B[2] + tz;
            uint8_t out = kNoCell;
            if (I < a.n[q][0] && J < a.n[q][1] && K < a.n[q][2]) {
                bool fluid = false, air = false;
                for (int d = 0; d < 8; ++d) {
                    const uint8_t f = below[(2 * tx + (d & 1)) + side * ((2 * ty + ((d >> 1) & 1)) + side * (2 * tz + (d >> 2)))];
                    fluid |= f == FE_FLUID;
                    air |= f == FE_AIR;
                }
                out = air ? (uint8_t)FE_AIR : (fluid ? (uint8_t)FE_FLUID : (uint8_t)FE_BOUNDRY);
                a.flags[q][(int64_t)I + (int64_t)a.n[q][0] * ((int64_t)J + (int64_t)a.n[q][1] * K)] = out;
            }
            here[t] = out;
        }
        __syncthreads();
        side = cs;
    }
}

// The row codes of levels 1 .. L-1 in one launch: thread c of the flattened
// cell range, level by level (offsets in off[]).
constexpr int kMaxLevels = 16;
struct CodesArgs {
    int levels;                  // coarse levels (1 .. levels)
    int64_t off[kMaxLevels + 1]; // first flattened cell of each coarse level
    int n[kMaxLevels][3];
    const uint8_t* flags[kMaxLevels];
    uint16_t* code[kMaxLevels];
};

__global__ __launch_bounds__(kBlock) void k_mg_codes_all(CodesArgs a) {
    const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (g >= a.off[a.levels]) return;
    int q = 0;
    while (q + 1 < a.levels && g >= a.off[q + 1]) ++q;
    const int64_t c = g - a.off[q];
    const int nx = a.n[q][0], ny = a.n[q][1], nz = a.n[q][2];
    int i, j, k;
    cell_ijk(c, nx, ny, i, j, k);
    a.code[q][c] = row_code(a.flags[q], c, (uint32_t)i, (uint32_t)j, (uint32_t)k, (uint32_t)nx, (uint32_t)ny,
                            (uint32_t)nz);
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

// Trilinear prolongation of the coarse correction at fine cell (i, j, k);
// V(a, b, c) is the coarse value at coarse index (a, b, c).
template <typename F>
__device__ inline float prolong_at3(int i, int j, int k, int cx, int cy, int cz, F V) {
    int ia, ib, ja, jb, ka, kb;
    float wia, wib, wja, wjb, wka, wkb;
    prolong_axis(i, cx, ia, wia, ib, wib);
    prolong_axis(j, cy, ja, wja, jb, wjb);
    prolong_axis(k, cz, ka, wka, kb, wkb);
    return wka * (wja * (wia * V(ia, ja, ka) + wib * V(ib, ja, ka)) + wjb * (wia * V(ia, jb, ka) + wib * V(ib, jb, ka))) +
           wkb * (wja * (wia * V(ia, ja, kb) + wib * V(ib, ja, kb)) + wjb * (wia * V(ia, jb, kb) + wib * V(ib, jb, kb)));
}

// The same with X(n) the coarse value at linear index n.
template <typename F>
__device__ inline float prolong_at(int i, int j, int k, int cx, int cy, int cz, F X) {
    return prolong_at3(i, j, k, cx, cy, cz,
                       [&](int a, int b, int c) { return X((int64_t)a + (int64_t)cx * ((int64_t)b + (int64_t)cy * c)); });
}

// One damped-Jacobi update of a cell, x + w (b - A x): the smoother of every
// level and kernel variant (the same expression, so the same rounding).
__device__ inline float mg_sweep_cell(uint16_t cd, float xc, float bc, const float (&nb)[6], const float ih2[3],
                                     float omega) {
    const float dgn = code_center(cd, ih2);
    const float w = mg_weight(cd, ih2, omega);
    return xc + w * (bc - code_apply_nb(cd, dgn, ih2, xc, nb));
}
// The second smoothing step at cell c (mg_cheb2 with the cell's weights).
__device__ inline float mg_cheb2_cell(uint16_t cd, float xc, float x0, float bc, const float (&nb)[6],
                                      const float ih2[3], const MgSmooth& sm) {
    const float dgn = code_center(cd, ih2);
    const float w = mg_weight(cd, ih2, sm.omega);
    return mg_cheb2(xc, x0, sm.w21 * w, bc - code_apply_nb(cd, dgn, ih2, xc, nb), sm.beta);
}

// The residual b - A x of a FLUID cell, 0 elsewhere.
__device__ inline float mg_residual_cell(uint16_t cd, float xc, float bc, const float (&nb)[6], const float ih2[3]) {
    return (cd & kFluidBit) ? bc - code_apply_nb(cd, code_center(cd, ih2), ih2, xc, nb) : 0.0f;
}

// A level's diagonal and smoother weight for FLUID rows, one entry per lane:
// lane n + 3 m + 9 p of every wavefront holds code_center and mg_weight of a
// FLUID row with n, m and p non-BOUNDRY neighbours along x, y and z (the only
// bits of a FLUID row's code that the two functions read), so the values are
// theirs.  Each wavefront builds its own copy (no LDS, no barrier) and reads it
// with a cross-lane permute.
struct CodeLanes {
    float cen, wgt;
};

__device__ inline CodeLanes code_lanes(const float ih2[3], float omega) {
    const int L = (int)(threadIdx.x & 63) % 27;  // lanes 27..63 repeat entries, never read
    const uint16_t cd = (uint16_t)(kFluidBit | ((L % 3) << 6) | (((L / 3) % 3) << 8) | ((L / 9) << 10));
    return CodeLanes{code_center(cd, ih2), mg_weight(cd, ih2, omega)};
}

// Lane of a FLUID row's entry.  Every lane of the wavefront must execute the
// permutes that read it (call outside divergent code).
__device__ inline int code_lane(uint32_t cd) {
    return (int)(((cd >> 6) & 3) + 3 * ((cd >> 8) & 3) + 9 * ((cd >> 10) & 3));
}

__device__ inline float lane_read(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(lane << 2, __float_as_int(v)));
}

// mg_weight of any row: a non-FLUID row's weight is omega (diagonal 1).
__device__ inline float lanes_weight(uint32_t cd, const CodeLanes& t, float omega) {
    const float w = lane_read(t.wgt, code_lane(cd));
    return (cd & kFluidBit) ? w : omega;
}

// A x at a FLUID row whose neighbours outside FLUID hold 0: a FLUID row
// couples to exactly its FLUID neighbours, and every vector of the V-cycle is
// 0 outside FLUID cells (the CG start puts the solution's non-FLUID entries at
// 0, the residual is 0 there from then on, restriction and prolongation skip
// those cells), so each coupling term is taken without consulting the code:
// an uncoupled neighbour adds -ih2 * 0.  The terms and their order are
// code_apply_nb's, so the value is too.  cen = code_center of the row; nih2 =
// -1/h^2 per axis.
__device__ inline float mg_apply_fluid_row(float cen, float xc, const float (&nb)[6], const float nih2[3]) {
    float acc = cen * xc;
    acc = fmaf(nih2[0], nb[0], acc);
    acc = fmaf(nih2[0], nb[1], acc);
    acc = fmaf(nih2[1], nb[2], acc);
    acc = fmaf(nih2[1], nb[3], acc);
    acc = fmaf(nih2[2], nb[4], acc);
    acc = fmaf(nih2[2], nb[5], acc);
    return acc;
}

// mg_sweep_cell at a FLUID row under the same condition (w = mg_weight).  At a
// non-FLUID cell (x and b 0) mg_sweep_cell gives 0, which the callers write
// without reading the neighbours.
__device__ inline float mg_sweep_fluid_row(float cen, float w, float xc, float bc, const float (&nb)[6],
                                           const float nih2[3]) {
    return xc + w * (bc - mg_apply_fluid_row(cen, xc, nb, nih2));
}

// ---------------------------------------------------------------------------
// Large levels: one kernel per step
// ---------------------------------------------------------------------------
struct LevelArgs {
    int64_t nc;
    Stride st;
    float ih2[3];
    int64_t chunk, padded;
};

// What a smoothing sweep or a residual reads for cell c.
struct StencilIn {
    uint16_t cd;
    float xc, bc, nb[6];
};

template <typename F>
__device__ inline StencilIn stencil_in(int64_t c, const uint16_t* __restrict__ code, const float* __restrict__ x,
                                       const float* __restrict__ b, const Stride& st, F X) {
    StencilIn t;
    t.cd = code[c];
    t.xc = x[c];
    t.bc = b[c];
    code_neighbours(c, st, X, t.nb);
    return t;
}

// x_out = x_in + w (b - A x_in).  With prz != nullptr (the last level-0 sweep)
// also leaves partials of b . x_out, which is r . z of the CG.  With pconf !=
// nullptr (the V-cycle's first kernel) and a pending confirmation, the
// residual the CG update just recomputed as b - A x is tested first: when it
// meets the relative stop the solve ends here (iteration count of that
// update), without this V-cycle and the next spmv.  cfl (level 0 only): the
// solver chunks holding a FLUID cell; the others hold zeros and are skipped.
__global__ __launch_bounds__(kBlock) void k_mg_smooth(LevelArgs L, const uint16_t* __restrict__ code,
                                                      const float* __restrict__ b, const float* __restrict__ xin,
                                                      float* xout, const SolverScalars* st,
                                                      double* __restrict__ prz, const double* __restrict__ pconf,
                                                      int nparts, const uint8_t* __restrict__ cfl, int second,
                                                      const float* x0, int fine) {
    // The stop flag is read with the data and tested before the first write,
    // so its load adds no latency of its own (the loads before it are in bounds).
    const int done = st->done;
    if (pconf && done) return;
    if (pconf && st->confirm) {
        double t[2];
        sum_partials<2>(pconf, nparts, t);
        if (stop_rule(st, st->bnorm2, t[0], t[1])) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                SolverScalars* w = const_cast<SolverScalars*>(st);
                w->rr = t[1];
                w->it = st->it + 1;
                w->converged = 1;
                w->done = 1;
            }
            return;
        }
    }
    const Span sp = block_span(L.nc, L.chunk, L.padded);
    if (cfl && !cfl[sp.lb]) {  // level 0 chunk without FLUID cells: x stays 0
        if (prz && threadIdx.x == 0) prz[blockIdx.x] = 0.0;
        return;
    }
    const Stride stv = L.st;
    double v[1] = {0.0};
    auto X = [&](int64_t n) { return xin[n]; };
    span_groups<8, StencilIn>(
        sp.c0, sp.c1, [&](int64_t c) { return stencil_in(c, code, xin, b, stv, X); },
        [&](int64_t c, const StencilIn& t) {
            if (done) return;  // uniform: the solve has stopped
            // second: the second smoothing step, from x0 (null: 0); x0 may be
            // xout (each cell reads its own x0 before writing it)
            const MgSmooth& sm = fine ? kMgFine : kMgCoarse;
            const float xo = second ? mg_cheb2_cell(t.cd, t.xc, x0 ? x0[c] : 0.0f, t.bc, t.nb, L.ih2, sm)
                                    : mg_sweep_cell(t.cd, t.xc, t.bc, t.nb, L.ih2, sm.omega);
            xout[c] = xo;
            v[0] += (double)t.bc * (double)xo;
        });
    if (done) return;
    if (prz) write_partials<1>(v, prz);
}

__global__ __launch_bounds__(kBlock) void k_mg_residual(LevelArgs L, const uint16_t* __restrict__ code,
                                                        const float* __restrict__ b, const float* __restrict__ x,
                                                        float* __restrict__ r, const SolverScalars* __restrict__ st,
                                                        const uint8_t* __restrict__ cfl) {
    const int done = st->done;  // tested before the first write (see k_mg_smooth)
    const Span sp = block_span(L.nc, L.chunk, L.padded);
    if (cfl && !cfl[sp.lb]) return;  // level 0 chunk without FLUID cells: r stays 0
    const Stride stv = L.st;
    auto X = [&](int64_t n) { return x[n]; };
    span_groups<8, StencilIn>(
        sp.c0, sp.c1, [&](int64_t c) { return stencil_in(c, code, x, b, stv, X); },
        [&](int64_t c, const StencilIn& t) {
            if (done) return;
            r[c] = mg_residual_cell(t.cd, t.xc, t.bc, t.nb, L.ih2);
        });
}

// b_c = R r_f on coarse FLUID cells (0 elsewhere); with x1 != nullptr also the
// coarse level's first sweep from zero, x1 = w b_c.  A workgroup covers an
// 8x8x4 coarse tile and first stages the 18x18x10 fine window it reads in LDS.
constexpr int kRtx = 8, kRty = 8, kRtz = 4;
constexpr int kRwx = 2 * kRtx + 2, kRwy = 2 * kRty + 2, kRwz = 2 * kRtz + 2;
static_assert(kRtx * kRty * kRtz == kBlock, "one coarse cell per thread");

__global__ __launch_bounds__(kBlock) void k_mg_restrict(const float* __restrict__ rf, int fx, int fy, int fz,
                                                        const uint16_t* __restrict__ ccode, float* __restrict__ bc,
                                                        float* __restrict__ x1, int cx, int cy, int cz, float cihx2,
                                                        float cihy2, float cihz2,
                                                        const SolverScalars* __restrict__ st) {
    const int done = st->done;  // tested before the first write (see k_mg_smooth)
    __shared__ float win[kRwz][kRwy][kRwx];
    const int I0 = blockIdx.x * kRtx, J0 = blockIdx.y * kRty, K0 = blockIdx.z * kRtz;
    const int fi0 = 2 * I0 - 1, fj0 = 2 * J0 - 1, fk0 = 2 * K0 - 1;
    const int tx = threadIdx.x % kRtx, ty = (threadIdx.x / kRtx) % kRty, tz = threadIdx.x / (kRtx * kRty);
    const int I = I0 + tx, J = J0 + ty, K = K0 + tz;
    const bool inside = I < cx && J < cy && K < cz;
    const int64_t c = (int64_t)I + (int64_t)cx * ((int64_t)J + (int64_t)cy * K);
    const uint16_t cd = inside ? ccode[c] : (uint16_t)0;  // issued with the window loads
    // All of a thread's window loads are issued before the first LDS store, so
    // the window costs one memory latency rather than one per element.
    constexpr int kWin = kRwx * kRwy * kRwz;
    constexpr int kPer = (kWin + kBlock - 1) / kBlock;
    float wv[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int e = threadIdx.x + u * kBlock;
        wv[u] = 0.0f;
        if (e < kWin) {
            const int wi = e % kRwx, wj = (e / kRwx) % kRwy, wk = e / (kRwx * kRwy);
            const int fi = min(max(fi0 + wi, 0), fx - 1);
            const int fj = min(max(fj0 + wj, 0), fy - 1);
            const int fk = min(max(fk0 + wk, 0), fz - 1);
            wv[u] = rf[(int64_t)fi + (int64_t)fx * ((int64_t)fj + (int64_t)fy * fk)];
        }
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int e = threadIdx.x + u * kBlock;
        if (e < kWin) (&win[0][0][0])[e] = wv[u];
    }
    __syncthreads();
    if (done || !inside) return;
    float v = 0.0f;
    if (cd & kFluidBit) {
        // window coordinates of fine index 2I-1+t are 2*tx+t (clamped taps carry weight 0)
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
#pragma unroll
                for (int ti = 0; ti < 4; ++ti) acc += (wx[ti] * wjk) * win[2 * tz + tk][2 * ty + tj][2 * tx + ti];
            }
        }
        v = acc * 0.125f;
    }
    bc[c] = v;
    if (x1) {
        const float ih2[3] = {cihx2, cihy2, cihz2};
        x1[c] = mg_weight(cd, ih2, kMgCoarse.omega) * v;
    }
}

// x_f += P x_c on FLUID fine cells.
__global__ __launch_bounds__(kBlock) void k_mg_prolong(const float* __restrict__ xc, int cx, int cy, int cz,
                                                       const uint16_t* __restrict__ fcode, float* __restrict__ xf,
                                                       int fx, int fy, int fz, const SolverScalars* __restrict__ st) {
    const int done = st->done;  // tested before the write (see k_mg_smooth)
    const int64_t f = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t nf = (int64_t)fx * fy * fz;
    if (f >= nf) return;
    // the marker, the fine value and the coarse values are loaded together
    const uint16_t cd = fcode[f];
    const float x0 = xf[f];
    int i, j, k;
    cell_ijk(f, fx, fy, i, j, k);
    const float add = prolong_at(i, j, k, cx, cy, cz, [&](int64_t n) { return xc[n]; });
    if (done || !(cd & kFluidBit)) return;
    xf[f] = x0 + add;
}

// ---------------------------------------------------------------------------
// Large levels, fused (FE_MG_FUSED): one descent and one ascent kernel per
// level.  A workgroup owns a coarse tile of T::rx x T::ry x T::rz cells and the
// fine tile of twice that size below it.  It stages what the tile's sweeps
// read, with halos, in LDS and recomputes the halo cells of each sweep itself
// instead of waiting for a neighbour tile, so a level takes two launches
// instead of six.  Every value is the split kernels' (the same terms in the
// same order; only the source of a neighbour value differs, and a neighbour
// outside the grid, which enters with weight 0, reads 0 here).  The sweeps and
// residuals take the FLUID-row forms (mg_sweep_fluid_row: diagonal and weight
// from a per-wavefront lane table, coupling terms without the code's selects,
// non-FLUID cells written as the 0 they are without reading their
// neighbours): the finest level's phases are VALU-bound (16 waves of a tile
// share one CU's four SIMDs), and these forms take about half the
// instructions of the general ones.
//   descent:  x1 = w b (first sweep from zero), x2 = x1 + w (b - A x1) (kept
//             for the ascent), r = b - A x2, coarse b = R r;
//   ascent:   y = x2 + P xc, x' = y + w (b - A y), out = x' + w (b - A x'),
//             with level 0's out = z and the partials of r.z.
// Window of halo H around the fine tile: coordinates 0 .. T::fx + 2H - 1
// stand for fine indices f0 - H ...
// Two tile shapes: the finest level's work is throughput-bound and takes large
// tiles (little halo per cell); the levels below are latency-bound (a 32^3
// level has 16 large tiles for 256 CUs, and each workgroup's chain of LDS
// phases sets the kernel's time), so they take small tiles: 2.5x fewer cells
// per thread in every phase for more halo work in all.  A cell's value does
// not depend on the tile it is computed in.
// ---------------------------------------------------------------------------
template <int RX, int RY, int RZ>
struct FusedTile {
    static constexpr int rx = RX, ry = RY, rz = RZ;              // coarse tile
    static constexpr int fx = 2 * RX, fy = 2 * RY, fz = 2 * RZ;  // fine tile
    static constexpr int coarse = RX * RY * RZ;
};
using TileFinest = FusedTile<8, 8, 4>;
using TileCoarse = FusedTile<4, 4, 4>;
template <class T, int H>
struct FWin {
    static constexpr int x = T::fx + 2 * H, y = T::fy + 2 * H, z = T::fz + 2 * H;
    static constexpr int sy = x, sz = x * y, n = x * y * z;
};
// coarse window of the ascent: coarse indices I0 - 2 .. I0 + T::rx + 1 per axis
template <class T>
struct CWin {
    static constexpr int x = T::rx + 4, y = T::ry + 4, z = T::rz + 4;
    static constexpr int sy = x, sz = x * y, n = x * y * z;
};

struct FusedArgs {
    int n[3];        // the fine level
    int c[3];        // the coarse level
    float ih2[3];    // the fine level's 1/h^2
    int tiles[3];    // coarse tiles per axis
    int64_t padded;  // launched workgroups (XCD remap), >= the tile count
};

__device__ inline int64_t fused_tile(const FusedArgs& a) { return xcd_remap(blockIdx.x, a.padded); }

__device__ inline bool fine_in(const FusedArgs& a, int i, int j, int k) {
    return i >= 0 && j >= 0 && k >= 0 && i < a.n[0] && j < a.n[1] && k < a.n[2];
}

// Linear index of an in-grid cell of a level with dims n: 32-bit arithmetic
// (a scene holds at most 2^30 cells).
__device__ inline uint32_t fused_index(int i, int j, int k, const int n[3]) {
    return (uint32_t)i + (uint32_t)n[0] * ((uint32_t)j + (uint32_t)n[1] * (uint32_t)k);
}

// tfl (level 0 only): per tile, whether its fine tile holds a FLUID cell.  A
// tile without one has every output 0 (its right-hand side and its coarse
// cells' are 0), and those stay as the solve's start left them.
constexpr int kFusedThreads = 1024;  // 16 waves per tile

// sum_partials for a kFusedThreads workgroup: every workgroup adds the NV-wide
// partials in the same fixed order (threads strided over the partials, lanes
// by the shuffle tree, waves in order), so all derive identical totals.
template <int NV>
__device__ inline void sum_partials_fused(const double* __restrict__ part, int nparts, double (&tot)[NV]) {
    __shared__ double sm[NV][kFusedThreads / 64];
#pragma unroll
    for (int q = 0; q < NV; ++q) tot[q] = 0.0;
    add_partials<NV, 1>(part, NV, 0, (int)threadIdx.x, kFusedThreads, nparts, tot);
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double x = tot[q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if ((threadIdx.x & 63) == 0) sm[q][threadIdx.x >> 6] = x;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double t = 0.0;
        for (int w = 0; w < kFusedThreads / 64; ++w) t += sm[q][w];
        tot[q] = t;
    }
    __syncthreads();
}

template <bool L0, class T>
__global__ __launch_bounds__(kFusedThreads) void k_mg_down(FusedArgs a, const float* __restrict__ b,
                                                    const uint16_t* __restrict__ code, float* __restrict__ x2out,
                                                    const uint16_t* __restrict__ ccode, float* __restrict__ cb,
                                                    const SolverScalars* st, const double* __restrict__ pconf,
                                                    int nconf, const uint8_t* __restrict__ tfl) {
    using W3 = FWin<T, 3>;
    using W2 = FWin<T, 2>;
    using W1 = FWin<T, 1>;
    __shared__ float sX1[W3::n];  // x1; later, over W1, the residual
    __shared__ float sB[W2::n];
    __shared__ uint16_t sC[W2::n];
    __shared__ float sX2[W2::n];
    const int done = st->done;  // read with the data, tested before the first write
    if (L0 && pconf) {
        // the V-cycle's first kernel: after a stop it returns; a relative stop
        // pending confirmation is tested on the residual the CG update just
        // recomputed (as k_mg_smooth), by every workgroup (workgroup 0 records
        // a stop).  Tested before the tile's loads: with the loads' registers
        // live across the test the kernel needs more than 64 VGPRs, and two
        // tiles no longer share a CU.
        if (done) return;
        if (st->confirm) {
            double tt[2];
            sum_partials_fused<2>(pconf, nconf, tt);
            if (stop_rule(st, st->bnorm2, tt[0], tt[1])) {
                if (blockIdx.x == 0 && threadIdx.x == 0) {
                    SolverScalars* w = const_cast<SolverScalars*>(st);
                    w->rr = tt[1];
                    w->it = st->it + 1;
                    w->converged = 1;
                    w->done = 1;
                }
                return;
            }
        }
    }
    const uint32_t t = (uint32_t)fused_tile(a);
    const bool real = t < (uint32_t)(a.tiles[0] * a.tiles[1] * a.tiles[2]) && !(L0 && tfl && !tfl[t]);
    if (!real) return;
    const uint32_t tx0 = (uint32_t)a.tiles[0], txy = tx0 * (uint32_t)a.tiles[1];
    const int TX = (int)(t % tx0), TY = (int)((t % txy) / tx0), TZ = (int)(t / txy);
    const int I0 = TX * T::rx, J0 = TY * T::ry, K0 = TZ * T::rz;
    const int fx0 = 2 * I0, fy0 = 2 * J0, fz0 = 2 * K0;
    // this thread's coarse cell (the restriction; the first T::coarse threads)
    const int tx = threadIdx.x % T::rx, ty = (threadIdx.x / T::rx) % T::ry, tz = threadIdx.x / (T::rx * T::ry);
    const int I = I0 + tx, J = J0 + ty, K = K0 + tz;
    const bool cin = threadIdx.x < T::coarse && I < a.c[0] && J < a.c[1] && K < a.c[2];
    const uint32_t ci = fused_index(I, J, K, a.c);
    // b and the row codes on the tile + 3: every load issued before any LDS
    // store (the codes in whole registers, so no load waits for another)
    constexpr int kPer = (W3::n + kFusedThreads - 1) / kFusedThreads;
    float vb[kPer];
    uint32_t vc[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int e = threadIdx.x + u * kFusedThreads;
        vb[u] = 0.0f;
        vc[u] = 0;
        if (e < W3::n) {
            const int wx = e % W3::x, wy = (e / W3::x) % W3::y, wz = e / W3::sz;
            const int i = fx0 - 3 + wx, j = fy0 - 3 + wy, k = fz0 - 3 + wz;
            if (fine_in(a, i, j, k)) {
                const uint32_t f = fused_index(i, j, k, a.n);
                vb[u] = b[f];
                vc[u] = code[f];
            }
        }
    }
    const uint32_t ccd = cin ? (uint32_t)ccode[ci] : 0u;
    constexpr MgSmooth SM = L0 ? kMgFine : kMgCoarse;
    const CodeLanes lanes = code_lanes(a.ih2, SM.omega);
    const float nih2[3] = {-a.ih2[0], -a.ih2[1], -a.ih2[2]};
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int e = threadIdx.x + u * kFusedThreads;
        const float w = lanes_weight(vc[u], lanes, SM.omega);  // every lane: a cross-lane read
        if (e < W3::n) {
            sX1[e] = w * vb[u];  // the first sweep from zero
            const int wx = e % W3::x, wy = (e / W3::x) % W3::y, wz = e / W3::sz;
            if (wx >= 1 && wx < W3::x - 1 && wy >= 1 && wy < W3::y - 1 && wz >= 1 && wz < W3::z - 1) {
                const int e2 = (wx - 1) + W2::sy * (wy - 1) + W2::sz * (wz - 1);
                sB[e2] = vb[u];
                sC[e2] = (uint16_t)vc[u];
            }
        }
    }
    __syncthreads();
    if (done) return;  // uniform: the solve has stopped
    // x2 on the tile + 2; the tile's own cells also go to global memory
    for (int e0 = 0; e0 < W2::n; e0 += kFusedThreads) {
        const int e = e0 + (int)threadIdx.x;
        const bool on = e < W2::n;
        const uint32_t cd = on ? sC[e] : 0u;
        const int li = code_lane(cd);
        const float cen = lane_read(lanes.cen, li), w = lane_read(lanes.wgt, li);
        if (!on) continue;
        const int wx = e % W2::x, wy = (e / W2::x) % W2::y, wz = e / W2::sz;
        const int e3 = (wx + 1) + W3::sy * (wy + 1) + W3::sz * (wz + 1);
        float x2 = 0.0f;
        if (cd & kFluidBit) {
            const float nb[6] = {sX1[e3 - 1],       sX1[e3 + 1],       sX1[e3 - W3::sy],
                                 sX1[e3 + W3::sy], sX1[e3 - W3::sz], sX1[e3 + W3::sz]};
            const float x1 = sX1[e3];
            x2 = mg_cheb2(x1, 0.0f, SM.w21 * w, sB[e] - mg_apply_fluid_row(cen, x1, nb, nih2), SM.beta);
        }
        sX2[e] = x2;
        if (wx >= 2 && wx < 2 + T::fx && wy >= 2 && wy < 2 + T::fy && wz >= 2 && wz < 2 + T::fz) {
            const int i = fx0 + wx - 2, j = fy0 + wy - 2, k = fz0 + wz - 2;
            if (fine_in(a, i, j, k)) x2out[fused_index(i, j, k, a.n)] = x2;
        }
    }
    __syncthreads();
    // r on the tile + 1 (into sX1, whose x1 is no longer read)
    float* sR = sX1;
    for (int e0 = 0; e0 < W1::n; e0 += kFusedThreads) {
        const int e = e0 + (int)threadIdx.x;
        const bool on = e < W1::n;
        const int wx = e % W1::x, wy = (e / W1::x) % W1::y, wz = e / W1::sz;
        const int e2 = (wx + 1) + W2::sy * (wy + 1) + W2::sz * (wz + 1);
        const uint32_t cd = on ? sC[e2] : 0u;
        const float cen = lane_read(lanes.cen, code_lane(cd));
        if (!on) continue;
        float rv = 0.0f;
        if (cd & kFluidBit) {
            const float nb[6] = {sX2[e2 - 1],       sX2[e2 + 1],       sX2[e2 - W2::sy],
                                 sX2[e2 + W2::sy], sX2[e2 - W2::sz], sX2[e2 + W2::sz]};
            rv = sB[e2] - mg_apply_fluid_row(cen, sX2[e2], nb, nih2);
        }
        sR[e] = rv;
    }
    __syncthreads();
    // coarse b = R r: fine index 2I-1+t sits at window coordinate 2*(I - I0) + t
    if (!cin) return;
    float v = 0.0f;
    if (ccd & kFluidBit) {
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
                    acc += (wx[ti] * wjk) * sR[(2 * tx + ti) + W1::sy * (2 * ty + tj) + W1::sz * (2 * tz + tk)];
            }
        }
        v = acc * 0.125f;
    }
    cb[ci] = v;
}

// out = the level's corrected and post-smoothed solution.  With prz (level 0)
// every launched workgroup leaves one partial of b . out (r.z).
template <bool L0, class T>
__global__ __launch_bounds__(kFusedThreads) void k_mg_up(FusedArgs a, const float* __restrict__ x2,
                                                  const float* __restrict__ xc, const float* __restrict__ b,
                                                  const uint16_t* __restrict__ code, float* __restrict__ out,
                                                  double* __restrict__ prz, const SolverScalars* __restrict__ st,
                                                  const uint8_t* __restrict__ tfl) {
    using W2 = FWin<T, 2>;
    using W1 = FWin<T, 1>;
    __shared__ float sY[W2::n];
    __shared__ uint16_t sC[W2::n];
    __shared__ float sB[W1::n];
    __shared__ float sXp[W1::n];
    __shared__ float sXc[CWin<T>::n];
    const int done = st->done;  // read with the data, tested before the first write
    const uint32_t t = (uint32_t)fused_tile(a);
    const bool real = t < (uint32_t)(a.tiles[0] * a.tiles[1] * a.tiles[2]) && !(L0 && tfl && !tfl[t]);
    int fx0 = 0, fy0 = 0, fz0 = 0, cx0 = 0, cy0 = 0, cz0 = 0;
    if (real) {
        const uint32_t tx0 = (uint32_t)a.tiles[0], txy = tx0 * (uint32_t)a.tiles[1];
        const int TX = (int)(t % tx0), TY = (int)((t % txy) / tx0), TZ = (int)(t / txy);
        fx0 = 2 * TX * T::rx;
        fy0 = 2 * TY * T::ry;
        fz0 = 2 * TZ * T::rz;
        cx0 = TX * T::rx - 2;
        cy0 = TY * T::ry - 2;
        cz0 = TZ * T::rz - 2;
        // x2 and the row codes on the tile + 2, b on the tile + 1, the coarse
        // correction on its window: every load issued before any LDS store
        // (the codes in whole registers, so no load waits for another)
        constexpr int kP2 = (W2::n + kFusedThreads - 1) / kFusedThreads;
        constexpr int kPc = (CWin<T>::n + kFusedThreads - 1) / kFusedThreads;
        float vx[kP2], vb[kP2];
        uint32_t vc[kP2];
        float vxc[kPc];
#pragma unroll
        for (int u = 0; u < kP2; ++u) {
            const int e = threadIdx.x + u * kFusedThreads;
            vx[u] = vb[u] = 0.0f;
            vc[u] = 0;
            if (e < W2::n) {
                const int wx = e % W2::x, wy = (e / W2::x) % W2::y, wz = e / W2::sz;
                const int i = fx0 - 2 + wx, j = fy0 - 2 + wy, k = fz0 - 2 + wz;
                if (fine_in(a, i, j, k)) {
                    const uint32_t f = fused_index(i, j, k, a.n);
                    vx[u] = x2[f];
                    vc[u] = code[f];
                    const bool w1 = wx >= 1 && wx < W2::x - 1 && wy >= 1 && wy < W2::y - 1 && wz >= 1 && wz < W2::z - 1;
                    if (w1) vb[u] = b[f];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kPc; ++u) {
            const int e = threadIdx.x + u * kFusedThreads;
            vxc[u] = 0.0f;
            if (e < CWin<T>::n) {
                const int wx = e % CWin<T>::x, wy = (e / CWin<T>::x) % CWin<T>::y, wz = e / CWin<T>::sz;
                const int I = min(max(cx0 + wx, 0), a.c[0] - 1), J = min(max(cy0 + wy, 0), a.c[1] - 1),
                          K = min(max(cz0 + wz, 0), a.c[2] - 1);
                vxc[u] = xc[fused_index(I, J, K, a.c)];
            }
        }
#pragma unroll
        for (int u = 0; u < kP2; ++u) {
            const int e = threadIdx.x + u * kFusedThreads;
            if (e < W2::n) {
                sY[e] = vx[u];
                sC[e] = (uint16_t)vc[u];
                const int wx = e % W2::x, wy = (e / W2::x) % W2::y, wz = e / W2::sz;
                if (wx >= 1 && wx < W2::x - 1 && wy >= 1 && wy < W2::y - 1 && wz >= 1 && wz < W2::z - 1)
                    sB[(wx - 1) + W1::sy * (wy - 1) + W1::sz * (wz - 1)] = vb[u];
            }
        }
#pragma unroll
        for (int u = 0; u < kPc; ++u) {
            const int e = threadIdx.x + u * kFusedThreads;
            if (e < CWin<T>::n) sXc[e] = vxc[u];
        }
    }
    constexpr MgSmooth SM = L0 ? kMgFine : kMgCoarse;
    const CodeLanes lanes = code_lanes(a.ih2, SM.omega);
    const float nih2[3] = {-a.ih2[0], -a.ih2[1], -a.ih2[2]};
    __syncthreads();
    if (done) return;  // uniform: the solve has stopped
    double v[1] = {0.0};
    if (real) {
        // y = x2 + P xc on FLUID cells (each from its own x2: in place)
        for (int e = threadIdx.x; e < W2::n; e += kFusedThreads) {
            if (!(sC[e] & kFluidBit)) continue;
            const int wx = e % W2::x, wy = (e / W2::x) % W2::y, wz = e / W2::sz;
            const float add = prolong_at3(fx0 - 2 + wx, fy0 - 2 + wy, fz0 - 2 + wz, a.c[0], a.c[1], a.c[2],
                                          [&](int I, int J, int K) {
                                              return sXc[(I - cx0) + CWin<T>::sy * (J - cy0) + CWin<T>::sz * (K - cz0)];
                                          });
            sY[e] = sY[e] + add;
        }
    }
    __syncthreads();
    if (real) {
        // x' on the tile + 1
        for (int e0 = 0; e0 < W1::n; e0 += kFusedThreads) {
            const int e = e0 + (int)threadIdx.x;
            const bool on = e < W1::n;
            const int wx = e % W1::x, wy = (e / W1::x) % W1::y, wz = e / W1::sz;
            const int e2 = (wx + 1) + W2::sy * (wy + 1) + W2::sz * (wz + 1);
            const uint32_t cd = on ? sC[e2] : 0u;
            const int li = code_lane(cd);
            const float cen = lane_read(lanes.cen, li), w = lane_read(lanes.wgt, li);
            if (!on) continue;
            float xp = 0.0f;
            if (cd & kFluidBit) {
                const float nb[6] = {sY[e2 - 1],       sY[e2 + 1],       sY[e2 - W2::sy],
                                     sY[e2 + W2::sy], sY[e2 - W2::sz], sY[e2 + W2::sz]};
                xp = mg_sweep_fluid_row(cen, w, sY[e2], sB[e], nb, nih2);
            }
            sXp[e] = xp;
        }
    }
    __syncthreads();
    if (real) {
        // out on the tile
        constexpr int kTile = T::fx * T::fy * T::fz;
        for (int e0 = 0; e0 < kTile; e0 += kFusedThreads) {
            const int e = e0 + (int)threadIdx.x;
            const bool on = e < kTile;
            const int wx = e % T::fx, wy = (e / T::fx) % T::fy, wz = e / (T::fx * T::fy);
            const int i = fx0 + wx, j = fy0 + wy, k = fz0 + wz;
            const int e1 = (wx + 1) + W1::sy * (wy + 1) + W1::sz * (wz + 1);
            const int e2 = (wx + 2) + W2::sy * (wy + 2) + W2::sz * (wz + 2);
            const uint32_t cd = on ? sC[e2] : 0u;
            const int li = code_lane(cd);
            const float cen = lane_read(lanes.cen, li), w = lane_read(lanes.wgt, li);
            if (!on || !fine_in(a, i, j, k)) continue;
            float o = 0.0f;
            if (cd & kFluidBit) {
                const float nb[6] = {sXp[e1 - 1],       sXp[e1 + 1],       sXp[e1 - W1::sy],
                                     sXp[e1 + W1::sy], sXp[e1 - W1::sz], sXp[e1 + W1::sz]};
                const float bc = sB[e1];
                const float x1 = sXp[e1];
                o = mg_cheb2(x1, sY[e2], SM.w21 * w, bc - mg_apply_fluid_row(cen, x1, nb, nih2), SM.beta);
                if (L0) v[0] += (double)bc * (double)o;
            }
            out[fused_index(i, j, k, a.n)] = o;
        }
    }
    if (L0 && prz) {  // one partial per workgroup: waves in order, lanes by the shuffle tree
        __shared__ double sm[kFusedThreads / 64];
        double x = v[0];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = x;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int w = 0; w < kFusedThreads / 64; ++w) t += sm[w];
            prz[blockIdx.x] = t;
        }
    }
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
// dg, wt: the level's diagonal and smoother weight per cell (precomputed).
// One Chebyshev step on the coarsest level (wt = 1/diagonal): d = c1 d +
// c2 D^-1 (b - A in), out = in + d; from zero (in == nullptr) d = c2 D^-1 b.
// d is per cell, so it is updated in place.
__device__ inline void tail_cheb(const TailLevel& l, const uint16_t* code, const float* dg, const float* wt,
                                 const float* in, float* out, const float* b, float* d, float c1, float c2) {
#pragma unroll
    for (int u = 0; u < kTailPer; ++u) {
        const int c = threadIdx.x + u * kTailThreads;
        if (c >= l.nc) break;
        const uint16_t cd = code[c];
        const float w = wt[c];
        float xo, dn;
        if (in) {
            const float xc = in[c];
            const float r = w * (b[c] - code_apply(cd, dg[c], c, l.st, l.ih2, xc, [&](int n) { return in[n]; }));
            dn = c1 * d[c] + c2 * r;
            xo = xc + dn;
        } else {
            dn = c2 * (w * b[c]);
            xo = dn;
        }
        d[c] = dn;
        out[c] = xo;
    }
    __syncthreads();
}

__device__ inline void tail_sweep(const TailLevel& l, const uint16_t* code, const float* dg, const float* wt,
                                  const float* in, float* out, const float* b) {
    // a tail level holds at most kTailMax cells: a fixed trip count lets the
    // LDS reads of a thread's cells overlap
#pragma unroll
    for (int u = 0; u < kTailPer; ++u) {
        const int c = threadIdx.x + u * kTailThreads;
        if (c >= l.nc) break;
        const uint16_t cd = code[c];
        const float w = wt[c];
        float xo;
        if (in) {
            const float xc = in[c];
            xo = xc + w * (b[c] - code_apply(cd, dg[c], c, l.st, l.ih2, xc, [&](int n) { return in[n]; }));
        } else {
            xo = w * b[c];
        }
        out[c] = xo;
    }
    __syncthreads();
}

// Solves level T's right-hand side bT approximately with the V-cycle of the
// tail levels; the result goes to xT.  Everything lives in LDS: five values
// (two iterates, right-hand side, diagonal, smoother weight) and the row code
// per cell of every tail level (at most kTailLds cells).
__global__ __launch_bounds__(kTailThreads) void k_mg_tail(TailArgs a, const float* __restrict__ bT,
                                                          float* __restrict__ xT,
                                                          const SolverScalars* __restrict__ st) {
    const int done = st->done;  // tested after the loads below (see k_mg_smooth)
    __shared__ float X[kTailLds], Tm[kTailLds], B[kTailLds], Dg[kTailLds], Wt[kTailLds];
    __shared__ float Dd[kCoarseMax];  // the coarsest level's Chebyshev direction
    __shared__ uint16_t Cd[kTailLds];
    // Row codes and the right-hand side: every load is issued before the
    // first LDS store (one memory latency in all).
    constexpr int kPer = kTailPer;
    const int n0 = a.n[0][0] * a.n[0][1] * a.n[0][2];
    {
        float bv[kPer];
        uint16_t cv[kMaxTail][kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int c = threadIdx.x + u * kTailThreads;
            bv[u] = c < n0 ? bT[c] : 0.0f;
        }
#pragma unroll
        for (int t = 0; t < kMaxTail; ++t) {
            const int nc = t < a.levels ? a.n[t][0] * a.n[t][1] * a.n[t][2] : 0;
#pragma unroll
            for (int u = 0; u < kPer; ++u) {
                const int c = threadIdx.x + u * kTailThreads;
                cv[t][u] = c < nc ? a.code[t][c] : (uint16_t)0;
            }
        }
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int c = threadIdx.x + u * kTailThreads;
            if (c < n0) B[c] = bv[u];
        }
#pragma unroll
        for (int t = 0; t < kMaxTail; ++t) {
            const int nc = t < a.levels ? a.n[t][0] * a.n[t][1] * a.n[t][2] : 0;
#pragma unroll
            for (int u = 0; u < kPer; ++u) {
                const int c = threadIdx.x + u * kTailThreads;
                if (c < nc) Cd[a.off[t] + c] = cv[t][u];
            }
        }
    }
    __syncthreads();
    if (done) return;
    // per cell: the diagonal and the damped-Jacobi weight (the coarsest level
    // with its own damping)
    for (int t = 0; t < a.levels; ++t) {
        const TailLevel l = tail_level(a, t);
        const float omega = t == a.levels - 1 ? 1.0f : kMgCoarse.omega;  // the coarsest: 1/diagonal for Chebyshev
#pragma unroll
        for (int u = 0; u < kTailPer; ++u) {
            const int c = threadIdx.x + u * kTailThreads;
            if (c >= l.nc) break;
            const uint16_t cd = Cd[l.o + c];
            Dg[l.o + c] = code_center(cd, l.ih2);
            Wt[l.o + c] = mg_weight(cd, l.ih2, omega);
        }
    }
    __syncthreads();
    // down: the solution of every level ends in Tm after its pre-smoothing
    for (int t = 0; t < a.levels - 1; ++t) {
        const TailLevel l = tail_level(a, t), lc = tail_level(a, t + 1);
        const int o = l.o;
        const uint16_t* code = Cd + o;
        tail_sweep(l, code, Dg + o, Wt + o, nullptr, X + o, B + o);
        tail_sweep(l, code, Dg + o, Wt + o, X + o, Tm + o, B + o);
        // residual into X (free now)
        {
            const float* xs = Tm + o;
#pragma unroll
            for (int u = 0; u < kTailPer; ++u) {
                const int c = threadIdx.x + u * kTailThreads;
                if (c >= l.nc) break;
                const uint16_t cd = code[c];
                float rc = 0.0f;
                if (cd & kFluidBit)
                    rc = B[o + c] - code_apply(cd, Dg[o + c], c, l.st, l.ih2, xs[c],
                                               [&](int n) { return xs[n]; });
                X[o + c] = rc;
            }
            __syncthreads();
        }
        // restriction into the next level's B
        {
            const float* rs = X + o;
#pragma unroll
            for (int u = 0; u < kTailPer; ++u) {
                const int c = threadIdx.x + u * kTailThreads;
                if (c >= lc.nc) break;
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
    // coarsest level: Chebyshev iteration from zero, ending in Tm (the
    // three-term recurrence with rho_k = 1 / (2 sigma - rho_(k-1)), rho_0 = 1 / sigma)
    {
        const TailLevel l = tail_level(a, a.levels - 1);
        const int o = l.o;
        const uint16_t* code = Cd + o;
        static_assert(kCoarseSweeps % 2 == 0, "the steps ping-pong X -> Tm");
        const float theta = 0.5f * (kChebHi + kChebLo), delta = 0.5f * (kChebHi - kChebLo);
        const float sigma = theta / delta;
        float rho = 1.0f / sigma;
        tail_cheb(l, code, Dg + o, Wt + o, nullptr, X + o, B + o, Dd, 0.0f, 1.0f / theta);
        for (int k = 1; k < kCoarseSweeps; ++k) {
            const float rn = 1.0f / (2.0f * sigma - rho);
            const float c1 = rn * rho, c2 = 2.0f * rn / delta;
            if (k & 1)
                tail_cheb(l, code, Dg + o, Wt + o, X + o, Tm + o, B + o, Dd, c1, c2);
            else
                tail_cheb(l, code, Dg + o, Wt + o, Tm + o, X + o, B + o, Dd, c1, c2);
            rho = rn;
        }
    }
    // up: prolongation, two sweeps Tm -> X -> Tm
    for (int t = a.levels - 2; t >= 0; --t) {
        const TailLevel l = tail_level(a, t), lc = tail_level(a, t + 1);
        const int o = l.o;
        const uint16_t* code = Cd + o;
        const float* xc = Tm + lc.o;
#pragma unroll
        for (int u = 0; u < kTailPer; ++u) {
            const int f = threadIdx.x + u * kTailThreads;
            if (f >= l.nc) break;
            if (!(code[f] & kFluidBit)) continue;
            int i, j, k;
            cell_ijk(f, l.nx, l.ny, i, j, k);
            Tm[o + f] += prolong_at(i, j, k, lc.nx, lc.ny, lc.nz, [&](int64_t n) { return xc[n]; });
        }
        __syncthreads();
        tail_sweep(l, code, Dg + o, Wt + o, Tm + o, X + o, B + o);
        tail_sweep(l, code, Dg + o, Wt + o, X + o, Tm + o, B + o);
    }
#pragma unroll
    for (int u = 0; u < kTailPer; ++u) {
        const int c = threadIdx.x + u * kTailThreads;
        if (c < n0) xT[c] = Tm[c];
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static inline dim3 grid1(int64_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

static LevelArgs level_args(const MGLevel& l) {
    const SolverShape sh = solver_shape_n(l.cells);
    LevelArgs a;
    a.nc = l.cells;
    a.st = stride_of(l.n[0], l.n[1], l.n[2]);
    for (int q = 0; q < 3; ++q) a.ih2[q] = l.ih2[q];
    a.chunk = sh.chunk;
    a.padded = sh.blocks;
    return a;
}

template <class T>
static FusedArgs fused_args(const MGLevel& l, const MGLevel& c) {
    FusedArgs a;
    for (int q = 0; q < 3; ++q) {
        a.n[q] = l.n[q];
        a.c[q] = c.n[q];
        a.ih2[q] = l.ih2[q];
    }
    a.tiles[0] = (c.n[0] + T::rx - 1) / T::rx;
    a.tiles[1] = (c.n[1] + T::ry - 1) / T::ry;
    a.tiles[2] = (c.n[2] + T::rz - 1) / T::rz;
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
        if ((e = hipMalloc((void**)&l.r, n * sizeof(float))) != hipSuccess) break;
    }
    // the fused V-cycle's level-0 tile flags and r.z partials
    const FusedArgs fa = fused_args<TileFinest>(lv[0], lv[1]);
    s->mg_npz = fa.padded;
    if (e == hipSuccess) e = hipMalloc((void**)&s->mg_tfl, (size_t)fa.padded);
    if (e == hipSuccess) e = hipMalloc((void**)&s->mg_pz, sizeof(double) * (size_t)fa.padded);
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
        (void)hipFree(l.r);
    }
    s->mg.clear();
    (void)hipFree(s->mg_tfl);
    (void)hipFree(s->mg_pz);
    s->mg_tfl = nullptr;
    s->mg_pz = nullptr;
    s->mg_npz = 0;
}

// Per solve: coarse markers and row codes from the current level-0 markers
// (the level-0 codes come from k_build_system).
hipError_t mg_setup(fe_scene* s) {
    hipError_t e;
    if (s->mg.empty() && (e = mg_alloc(s)) != hipSuccess) return e;
    std::vector<MGLevel>& lv = s->mg;
    lv[0].flags = s->flags;
    lv[0].code = s->code;
    const int L = (int)lv.size();
    // markers of levels 1 .. m and the finest tiles' flags in one launch,
    // deeper levels (grids above 128^3) one launch each
    const int m = std::min(L - 1, kPyr);
    PyrArgs pa{};
    pa.levels = m;
    for (int q = 0; q <= m; ++q) {
        for (int a = 0; a < 3; ++a) pa.n[q][a] = lv[q].n[a];
        pa.flags[q] = lv[q].flags;
    }
    if (s->tune.mg == FE_MG_FUSED) {
        const FusedArgs fa = fused_args<TileFinest>(lv[0], lv[1]);
        pa.tfl = s->mg_tfl;
        for (int a = 0; a < 3; ++a) pa.tiles[a] = fa.tiles[a];
    }
    const unsigned nb = (unsigned)(((lv[1].n[0] + 7) / 8) * ((lv[1].n[1] + 7) / 8) * ((lv[1].n[2] + 7) / 8));
    hipLaunchKernelGGL(k_mg_pyramid, dim3(nb), dim3(kPyrThreads), 0, s->stream, pa);
    for (int q = m + 1; q < L; ++q) {
        const MGLevel& f = lv[q - 1];
        MGLevel& c = lv[q];
        hipLaunchKernelGGL(k_mg_coarsen, grid1(c.cells), dim3(kBlock), 0, s->stream, (const uint8_t*)f.flags, f.n[0],
                           f.n[1], f.n[2], c.flags, c.n[0], c.n[1], c.n[2]);
    }
    // row codes of every coarse level in one launch
    if (L - 1 <= kMaxLevels) {
        CodesArgs ca{};
        ca.levels = L - 1;
        ca.off[0] = 0;
        for (int q = 1; q < L; ++q) {
            for (int a = 0; a < 3; ++a) ca.n[q - 1][a] = lv[q].n[a];
            ca.flags[q - 1] = lv[q].flags;
            ca.code[q - 1] = lv[q].code;
            ca.off[q] = ca.off[q - 1] + lv[q].cells;
        }
        hipLaunchKernelGGL(k_mg_codes_all, grid1(ca.off[L - 1]), dim3(kBlock), 0, s->stream, ca);
    } else {
        for (int q = 1; q < L; ++q)
            hipLaunchKernelGGL(k_mg_codes, grid1(lv[q].cells), dim3(kBlock), 0, s->stream, (const uint8_t*)lv[q].flags,
                               lv[q].n[0], lv[q].n[1], lv[q].n[2], lv[q].code);
    }
    return hipGetLastError();
}

float* mg_first_sweep_target(fe_scene* s) { return s->mg.empty() ? nullptr : s->mg[0].x; }

// The one-workgroup tail of a V-cycle (z = M^-1 r, mg_apply below).  Once `st`
// says the solve has stopped, every multigrid kernel still issues its first
// loads (the stop flag is read together with them) and then returns before its
// first write, except the finest-level descent, which tests the flag before
// its loads; the CG spmv and update issue their first group of cells' loads
// and the partial-sum loads before testing it.
static void launch_tail(fe_scene* s, SolverScalars* st) {
    std::vector<MGLevel>& lv = s->mg;
    const int L = (int)lv.size();
    const int T = s->mg_tail;
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
    hipLaunchKernelGGL(k_mg_tail, dim3(1), dim3(kTailThreads), 0, s->stream, ta, (const float*)lv[T].b, lv[T].x2, st);
}

// The fused V-cycle: level q's descent writes its pre-smoothed x2 and level
// q+1's right-hand side; the tail leaves level T's result in lv[T].x2; level
// q's ascent reads x2 and the coarse result and writes its own result to x
// (level 0: z).  x is never read before the ascent writes it, so no level
// holds a first sweep in memory.
static hipError_t mg_apply_fused(fe_scene* s, const float* r, float* z, SolverScalars* st, double* prz, int tag,
                                 const double* pconf, int nparts) {
    std::vector<MGLevel>& lv = s->mg;
    const int T = s->mg_tail;
    lv[0].b = const_cast<float*>(r);
    for (int q = 0; q < T; ++q) {
        const MGLevel& l = lv[q];
        const MGLevel& c = lv[q + 1];
        const FusedArgs a = q == 0 ? fused_args<TileFinest>(l, c) : fused_args<TileCoarse>(l, c);
        const dim3 grid((unsigned)a.padded);
        if (q == 0)
            launch_timed(s, FE_KERNEL_MG_DOWN, tag, k_mg_down<true, TileFinest>, grid, dim3(kFusedThreads), a,
                         (const float*)l.b, (const uint16_t*)l.code, l.x2, (const uint16_t*)c.code, c.b,
                         (const SolverScalars*)st, pconf, nparts, (const uint8_t*)s->mg_tfl);
        else
            hipLaunchKernelGGL((k_mg_down<false, TileCoarse>), grid, dim3(kFusedThreads), 0, s->stream, a, (const float*)l.b,
                               (const uint16_t*)l.code, l.x2, (const uint16_t*)c.code, c.b, (const SolverScalars*)st,
                               (const double*)nullptr, 0, (const uint8_t*)nullptr);
    }
    launch_tail(s, st);
    for (int q = T - 1; q >= 0; --q) {
        const MGLevel& l = lv[q];
        const MGLevel& c = lv[q + 1];
        const FusedArgs a = q == 0 ? fused_args<TileFinest>(l, c) : fused_args<TileCoarse>(l, c);
        const dim3 grid((unsigned)a.padded);
        const float* xc = q + 1 == T ? c.x2 : c.x;
        if (q == 0)
            launch_timed(s, FE_KERNEL_MG_UP, tag, k_mg_up<true, TileFinest>, grid, dim3(kFusedThreads), a,
                         (const float*)l.x2, xc, (const float*)l.b, (const uint16_t*)l.code, z, prz,
                         (const SolverScalars*)st, (const uint8_t*)s->mg_tfl);
        else
            hipLaunchKernelGGL((k_mg_up<false, TileCoarse>), grid, dim3(kFusedThreads), 0, s->stream, a, (const float*)l.x2, xc,
                               (const float*)l.b, (const uint16_t*)l.code, l.x, (double*)nullptr,
                               (const SolverScalars*)st, (const uint8_t*)nullptr);
    }
    return hipGetLastError();
}

// z = M^-1 r with one V-cycle, leaving partials of r.z in prz.  With the split
// kernels the caller has already written level 0's first sweep (x1 = w r)
// into mg[0].x; the fused descent sweeps from r itself.
hipError_t mg_apply(fe_scene* s, const float* r, float* z, SolverScalars* st, double* prz, int tag,
                    const double* pconf, int nparts) {
    if (s->tune.mg == FE_MG_FUSED) return mg_apply_fused(s, r, z, st, prz, tag, pconf, nparts);
    std::vector<MGLevel>& lv = s->mg;
    const int T = s->mg_tail;
    lv[0].b = const_cast<float*>(r);
    const uint8_t* cfl0 = s->chunk_fluid;
    for (int q = 0; q < T; ++q) {
        MGLevel& l = lv[q];
        MGLevel& c = lv[q + 1];
        const LevelArgs a = level_args(l);
        const dim3 grid((unsigned)a.padded);
        if (q == 0)
            launch_timed(s, FE_KERNEL_MG_SMOOTH, tag, k_mg_smooth, grid, dim3(kBlock), a, (const uint16_t*)l.code,
                         (const float*)l.b, (const float*)l.x, l.x2, st, (double*)nullptr, pconf, nparts, cfl0, 1,
                         (const float*)nullptr, 1);
        else
            hipLaunchKernelGGL(k_mg_smooth, grid, dim3(kBlock), 0, s->stream, a, (const uint16_t*)l.code,
                               (const float*)l.b, (const float*)l.x, l.x2, (const SolverScalars*)st, (double*)nullptr,
                               (const double*)nullptr, 0, (const uint8_t*)nullptr, 1, (const float*)nullptr, 0);
        hipLaunchKernelGGL(k_mg_residual, grid, dim3(kBlock), 0, s->stream, a, (const uint16_t*)l.code,
                           (const float*)l.b, (const float*)l.x2, l.r, st, q == 0 ? cfl0 : (const uint8_t*)nullptr);
        const dim3 rgrid((unsigned)((c.n[0] + kRtx - 1) / kRtx), (unsigned)((c.n[1] + kRty - 1) / kRty),
                         (unsigned)((c.n[2] + kRtz - 1) / kRtz));
        hipLaunchKernelGGL(k_mg_restrict, rgrid, dim3(kBlock), 0, s->stream, (const float*)l.r, l.n[0],
                           l.n[1], l.n[2], (const uint16_t*)c.code, c.b, q + 1 < T ? c.x : (float*)nullptr, c.n[0],
                           c.n[1], c.n[2], c.ih2[0], c.ih2[1], c.ih2[2], st);
    }
    launch_tail(s, st);
    for (int q = T - 1; q >= 0; --q) {
        MGLevel& l = lv[q];
        const MGLevel& c = lv[q + 1];
        const LevelArgs a = level_args(l);
        const dim3 grid((unsigned)a.padded);
        hipLaunchKernelGGL(k_mg_prolong, grid1(l.cells), dim3(kBlock), 0, s->stream, (const float*)c.x2, c.n[0],
                           c.n[1], c.n[2], (const uint16_t*)l.code, l.x2, l.n[0], l.n[1], l.n[2], st);
        if (q == 0) {
            launch_timed(s, FE_KERNEL_MG_SMOOTH, tag, k_mg_smooth, grid, dim3(kBlock), a, (const uint16_t*)l.code,
                         (const float*)l.b, (const float*)l.x2, l.x, st, (double*)nullptr, (const double*)nullptr, 0,
                         cfl0, 0, (const float*)nullptr, 1);
            launch_timed(s, FE_KERNEL_MG_SMOOTH, tag, k_mg_smooth, grid, dim3(kBlock), a, (const uint16_t*)l.code,
                         (const float*)l.b, (const float*)l.x, z, st, prz, (const double*)nullptr, 0, cfl0, 1,
                         (const float*)l.x2, 1);
        } else {
            hipLaunchKernelGGL(k_mg_smooth, grid, dim3(kBlock), 0, s->stream, a, (const uint16_t*)l.code,
                               (const float*)l.b, (const float*)l.x2, l.x, st, (double*)nullptr, (const double*)nullptr, 0,
                               (const uint8_t*)nullptr, 0, (const float*)nullptr, 0);
            hipLaunchKernelGGL(k_mg_smooth, grid, dim3(kBlock), 0, s->stream, a, (const uint16_t*)l.code,
                               (const float*)l.b, (const float*)l.x, l.x2, st, (double*)nullptr, (const double*)nullptr, 0,
                               (const uint8_t*)nullptr, 1, (const float*)l.x2, 0);
        }
    }
    return hipGetLastError();
}

}  // namespace fe
