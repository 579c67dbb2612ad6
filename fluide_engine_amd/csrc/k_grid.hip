// k_grid.hip -- grid-side kernels of the gfx950 substep: fluid SDF and cell
// markers (bit-exact FP64 decisions), velocity extrapolation, wall boundary
// conditions, gravity, pressure-gradient application and the CFL reduction.
// All are HBM-bound element or 7-point stencil sweeps; launches remap blocks
// so neighbouring tiles share an XCD's L2.
#include <algorithm>

#include "fe_kernels.hpp"

namespace fe {

static inline dim3 grid1(int64_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

__global__ __launch_bounds__(kBlock) void k_fill_u8(uint8_t* __restrict__ a, uint8_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_fill_f32(float* __restrict__ a, float v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) a[i] = v;
}

hipError_t launch_fill_u8(hipStream_t st, uint8_t* a, uint8_t v, int64_t n) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u8, grid1(n), dim3(kBlock), 0, st, a, v, n);
    return hipGetLastError();
}

hipError_t launch_fill_f32(hipStream_t st, float* a, float v, int64_t n) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_f32, grid1(n), dim3(kBlock), 0, st, a, v, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// BuildSignedDistanceField (src/pic_simulator.cpp:452-475).
// sdf(c) = min(2r, min |p - c|) - r with r = 1.2*max(h)/sqrt(2), over the
// particles with |p-c|^2 <= (2r)^2.  The reference searches 8 hash buckets of
// width 2r (hash_grid_particle_searcher.cpp:154-204), which always contain
// every particle closer than r; this kernel scans the cells within r of the
// centre (same guarantee), so sdf < 0 -- the only thing the substep uses -- and
// its value there are bit-exact: FP64 differences, no FMA contraction and one
// correctly rounded sqrt of the minimum squared distance (sqrt is monotone).
// Positive values may differ from the reference's where its buckets reach
// particles farther than r that this scan does not (they only feed the
// surface mesher).
// Every candidate's squared distance is first formed in FP32 (one 16-byte
// load of its position); the FP64 distance is computed only when that value
// is within `slack` of the running FP32 minimum.  slack is twice a bound E on
// |d2_fp32 - d2_fp64| for candidates within 2*band, so the candidate with the
// smallest FP64 distance is never skipped (its FP32 value is at most the FP32
// minimum + 2E) and the result is the same FP64 minimum.
//
// Buckets = true restates the reference's candidate set exactly, so that every
// value (positive ones included) is the reference's: a particle is a candidate
// iff its bucket floor(p / 2r) lies, per axis, in {o, o + d} with o the bucket
// of the centre and d = +1 when (o + 0.5f) * 2r <= centre else -1
// (hash_grid_particle_searcher.cpp:154-198; the 64-bucket wrap aliases only
// buckets farther than 2r, which the distance test drops).  The scan covers the
// cells overlapping that bucket block within 2r of the centre.
// ---------------------------------------------------------------------------
// Without buckets a workgroup takes a block of kSdfBX x kSdfBY x kSdfBZ cells
// and first checks whether any particle lies in the cells its scans could
// reach (the block grown by the scan extent): if none does, every cell of the
// block gets the value of an empty scan, band - radius, without scanning.
constexpr int kSdfBX = 8, kSdfBY = 8, kSdfBZ = 4;
static_assert(kSdfBX * kSdfBY * kSdfBZ == kBlock, "one cell per thread");

template <bool Buckets>
__global__ __launch_bounds__(kBlock) void k_sdf(Geom g, const float4* __restrict__ pos4,
                                                const int32_t* __restrict__ start, double* __restrict__ sdf,
                                                int ex, int ey, int ez, double radius, double band, double qr2,
                                                float slack, int64_t padded) {
#pragma clang fp contract(off)
    const int64_t blk = xcd_remap(blockIdx.x, padded);
    const int64_t nc = cell_count(g);
    int i, j, k;
    int64_t c;
    if (Buckets) {
        c = blk * kBlock + threadIdx.x;
        if (c >= nc) return;
        split_index(c, g.n[0], g.n[1], i, j, k);
    } else {
        const int tbx = (g.n[0] + kSdfBX - 1) / kSdfBX, tby = (g.n[1] + kSdfBY - 1) / kSdfBY,
                  tbz = (g.n[2] + kSdfBZ - 1) / kSdfBZ;
        if (blk >= (int64_t)tbx * tby * tbz) return;  // the whole workgroup
        int bi, bj, bk;
        split_index(blk, tbx, tby, bi, bj, bk);
        const int t = (int)threadIdx.x;
        i = bi * kSdfBX + t % kSdfBX;
        j = bj * kSdfBY + (t / kSdfBX) % kSdfBY;
        k = bk * kSdfBZ + t / (kSdfBX * kSdfBY);
        const int xa = max(bi * kSdfBX - ex, 0), xb = min(bi * kSdfBX + kSdfBX - 1 + ex, g.n[0] - 1);
        const int ya = max(bj * kSdfBY - ey, 0), yb = min(bj * kSdfBY + kSdfBY - 1 + ey, g.n[1] - 1);
        const int za = max(bk * kSdfBZ - ez, 0), zb = min(bk * kSdfBZ + kSdfBZ - 1 + ez, g.n[2] - 1);
        const int ylen = yb - ya + 1, rows = ylen * (zb - za + 1);
        int any = 0;
        for (int r = t; r < rows; r += kBlock) {
            const int64_t row = (int64_t)g.n[0] * ((int64_t)(ya + r % ylen) + (int64_t)g.n[1] * (za + r / ylen));
            any |= start[row + xb + 1] > start[row + xa] ? 1 : 0;
        }
        any = __syncthreads_or(any);
        if (i >= g.n[0] || j >= g.n[1] || k >= g.n[2]) return;
        c = (int64_t)i + (int64_t)g.n[0] * ((int64_t)j + (int64_t)g.n[1] * k);
        if (!any) {
            sdf[c] = band - radius;  // minD = band, nothing found
            return;
        }
    }
    // Grid3D::GridIndexToPosition on the SDF grid of origin o + h/2 (src/hybrid_simulator.cpp:13-15)
    const double gx = (g.od[0] + g.hd[0] / 2.0) + (double)i * g.hd[0];
    const double gy = (g.od[1] + g.hd[1] / 2.0) + (double)j * g.hd[1];
    const double gz = (g.od[2] + g.hd[2] / 2.0) + (double)k * g.hd[2];
    int i0 = max(i - ex, 0), i1 = min(i + ex, g.n[0] - 1);
    int j0 = max(j - ey, 0), j1 = min(j + ey, g.n[1] - 1);
    int k0 = max(k - ez, 0), k1 = min(k + ez, g.n[2] - 1);
    const double sp = 2.0 * radius;
    int blo[3] = {0, 0, 0}, bhi[3] = {0, 0, 0};
    if (Buckets) {
        const double gp[3] = {gx, gy, gz};
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            const int o = (int)floor(gp[a] / sp);
            const int d = ((double)((float)o + 0.5f) * sp <= gp[a]) ? 1 : -1;
            blo[a] = min(o, o + d);
            bhi[a] = max(o, o + d);
            // cells overlapping [blo*sp, (bhi+1)*sp) and [gp - band, gp + band], one cell of margin
            const double x0 = fmax((double)blo[a] * sp, gp[a] - band);
            const double x1 = fmin((double)(bhi[a] + 1) * sp, gp[a] + band);
            lo[a] = max((int)floor((x0 - g.od[a]) / g.hd[a]) - 1, 0);
            hi[a] = min((int)floor((x1 - g.od[a]) / g.hd[a]) + 1, g.n[a] - 1);
        }
        i0 = lo[0], i1 = hi[0], j0 = lo[1], j1 = hi[1], k0 = lo[2], k1 = hi[2];
    }
    const float gx32 = (float)gx, gy32 = (float)gy, gz32 = (float)gz;
    float best32 = __builtin_inff();
    double best = qr2;
    bool found = false;
    auto consider = [&](const float4 p) {
        const float fx = p.x - gx32, fy = p.y - gy32, fz = p.z - gz32;
        const float d2f = fx * fx + fy * fy + fz * fz;
        if (d2f > best32 + slack) return;
        if (Buckets) {
            const int bx = (int)floor((double)p.x / sp), by = (int)floor((double)p.y / sp),
                      bz = (int)floor((double)p.z / sp);
            if (bx < blo[0] || bx > bhi[0] || by < blo[1] || by > bhi[1] || bz < blo[2] || bz > bhi[2]) return;
        }
        best32 = fminf(best32, d2f);
        const double dx = (double)p.x - gx;
        const double dy = (double)p.y - gy;
        const double dz = (double)p.z - gz;
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 <= qr2 && (!found || d2 < best)) {
            best = d2;
            found = true;
        }
    };
    auto scan = [&](int s0, int s1) {
        constexpr int kB = 8;  // candidate loads issued together
        int s = s0;
        for (; s + kB <= s1; s += kB) {
            float4 p[kB];
#pragma unroll
            for (int q = 0; q < kB; ++q) p[q] = pos4[s + q];
#pragma unroll
            for (int q = 0; q < kB; ++q) consider(p[q]);
        }
        for (; s < s1; ++s) consider(pos4[s]);
    };
    if (Buckets) {
        for (int kk = k0; kk <= k1; ++kk) {
            for (int jj = j0; jj <= j1; ++jj) {
                const int64_t row = (int64_t)g.n[0] * ((int64_t)jj + (int64_t)g.n[1] * kk);
                scan(start[row + i0], start[row + i1 + 1]);
            }
        }
    } else {
        // The centre's own cell first; a neighbour cell is then skipped when
        // no point of it is nearer than the best distance so far (or than the
        // band while none is found): a particle binned in cell ii lies in
        // [o + ii h, o + (ii + 1) h) on each axis, so its distance to the
        // centre of cell i is at least (|ii - i| - 1/2) h there.  The bound is
        // shrunk by 1e-9 (far beyond FP64 rounding), so a skipped cell holds no
        // particle whose computed squared distance is below the best, and the
        // minimum -- the result -- is the full scan's.  Inside the fluid the
        // own cell's nearest particle (about 0.43 h away for 2 x 2 x 2 seeding)
        // usually rules out all 26 neighbours (they are at least h/2 away).
        // Per row of cells (fixed jj, kk) the bound grows with |ii - i|, so
        // the cells left after skipping form one interval around i, scanned as
        // one particle range (the centre's row: the two sides of the centre).
        scan(start[c], start[c + 1]);
        auto bound = [&](int d, double h) { return d == 0 ? 0.0 : ((double)d - 0.5) * h; };
        for (int kk = k0; kk <= k1; ++kk) {
            const double lz = bound(abs(kk - k), g.hd[2]);
            for (int jj = j0; jj <= j1; ++jj) {
                const double ly = bound(abs(jj - j), g.hd[1]);
                const double lyz2 = ly * ly + lz * lz;
                // best is qr2 until a particle is found
                auto skip = [&](int ii) {
                    const double lx = bound(abs(ii - i), g.hd[0]);
                    return (lx * lx + lyz2) * (1.0 - 1e-9) > best;
                };
                if (skip(i)) continue;
                int a = i0, b = i1;
                while (a < i && skip(a)) ++a;
                while (b > i && skip(b)) --b;
                const int64_t row = (int64_t)g.n[0] * ((int64_t)jj + (int64_t)g.n[1] * kk);
                if (jj == j && kk == k) {
                    if (a < i) scan(start[row + a], start[row + i]);
                    if (b > i) scan(start[row + i + 1], start[row + b + 1]);
                } else {
                    scan(start[row + a], start[row + b + 1]);
                }
            }
        }
    }
    double minD = band;
    if (found) minD = fmin(minD, sqrt(best));
    sdf[c] = minD - radius;
}

hipError_t launch_sdf(fe_scene* s, bool buckets) {
    const Geom& g = s->g;
    const double maxH = fmax(g.hd[0], fmax(g.hd[1], g.hd[2]));
    const double radius = 1.2 * maxH / std::sqrt(2.0);  // src/pic_simulator.cpp:459
    const double band = 2.0 * radius;
    const double qr2 = band * band;
    int ext[3];
    for (int a = 0; a < 3; ++a) ext[a] = (int)std::floor(radius / g.hd[a] + 0.5);
    // Bound E on |d2_fp32 - d2_fp64| for candidates within D = 2*band of the
    // centre: each FP32 difference is within delta = eps*(G + D) + eps^2*G of
    // the exact one (G bounds every coordinate), and the squares and sums add
    // at most 3*eps of 3*D^2 more.  slack = 2E, doubled again for margin.
    const double eps = std::ldexp(1.0, -24), D = 2.0 * band;
    double G = 0.0;
    for (int a = 0; a < 3; ++a) G = fmax(G, fabs(g.od[a]) + g.Ld[a] + g.hd[a]);
    const double delta = eps * (G + D) + eps * eps * G;
    const double E = 3.0 * (2.0 * D * delta + delta * delta) + 3.0 * eps * 3.0 * D * D;
    const float slack = (float)(4.0 * E);
    const int64_t nc = cell_count(g);
    buckets = buckets || s->has_collider || s->sdf_buckets;
    const int64_t nblk = buckets ? (nc + kBlock - 1) / kBlock
                                 : (int64_t)((g.n[0] + kSdfBX - 1) / kSdfBX) * ((g.n[1] + kSdfBY - 1) / kSdfBY) *
                                       ((g.n[2] + kSdfBZ - 1) / kSdfBZ);
    const int64_t padded = pad_blocks(nblk);
    if (buckets)
        launch_timed(s, FE_KERNEL_SDF, 0, k_sdf<true>, dim3((unsigned)padded), dim3(kBlock), g,
                     (const float4*)s->pos4, (const int32_t*)s->cell_start, s->sdf, ext[0], ext[1], ext[2], radius,
                     band, qr2, slack, padded);
    else
        launch_timed(s, FE_KERNEL_SDF, 0, k_sdf<false>, dim3((unsigned)padded), dim3(kBlock), g,
                     (const float4*)s->pos4, (const int32_t*)s->cell_start, s->sdf, ext[0], ext[1], ext[2], radius,
                     band, qr2, slack, padded);
    s->sdf_from_cells = buckets ? 0 : 1;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_sdf_into_collider(s);
}

// ---------------------------------------------------------------------------
// SinglePhasePressureSolver::BuildMarkers (single_phase_pressure_solver.cpp:30-48):
// FLUID iff fluidSdf.Sample(GetCellCenterPos(i,j,k)) < 0, evaluated with the
// reference's FP64 arithmetic (ScalarGrid3D::Sample, src/3d/scalar_grid3d.cpp:33-58);
// else BOUNDRY iff the collider SDF samples negative there, else AIR.  Without
// a collider the collider SDF is positive everywhere and the other cells are AIR.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_markers(Geom g, const double* __restrict__ sdf,
                                                    const double* __restrict__ csdf, uint8_t* __restrict__ flags,
                                                    int64_t padded) {
#pragma clang fp contract(off)
    const int64_t blk = xcd_remap(blockIdx.x, padded);
    const int64_t c = blk * kBlock + threadIdx.x;
    const int64_t nc = cell_count(g);
    if (c >= nc) return;
    int idx[3];
    split_index(c, g.n[0], g.n[1], idx[0], idx[1], idx[2]);
    double cc[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) cc[a] = (g.od[a] + 0.5 * g.hd[a]) + g.hd[a] * (double)idx[a];  // GetCellCenterPos
    uint8_t f = FE_AIR;
    if (sample_sdf64(g, sdf, cc) < 0)
        f = FE_FLUID;
    else if (csdf && sample_collider64(g, csdf, cc) < 0)
        f = FE_BOUNDRY;
    flags[c] = f;
}

hipError_t launch_markers(fe_scene* s) {
    const int64_t nc = cell_count(s->g);
    const int64_t padded = pad_blocks((nc + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_markers, dim3((unsigned)padded), dim3(kBlock), 0, s->stream, s->g, (const double*)s->sdf,
                       (const double*)(s->has_collider ? s->csdf : nullptr), s->flags, padded);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// ExtrapolateToRegion (src/common/math_utils.hpp:186-253), one iteration over
// all three face components.  In place: a thread writes its face only when the
// face is invalid and reads neighbours only when they are valid, and valid
// faces are never written in the same sweep, so there is no race.
// ---------------------------------------------------------------------------
struct ValidSet {
    const uint8_t* in[3];
    uint8_t* out[3];
};

// Four consecutive bytes a[off .. off+3] of the byte array a[0, n), off
// possibly outside [0, n): two aligned 32-bit loads, each clamped to the words
// that hold bytes of the array (device allocations carry slack past their
// end), so bytes outside [0, n) come back arbitrary and the loads stay in
// bounds.
__device__ inline uint32_t bytes4(const uint8_t* __restrict__ a, int64_t n, int64_t off) {
    const int64_t base = (int64_t)(uintptr_t)a;
    const int64_t A = base + off;
    const int64_t wlo = base >> 2, whi = (base + n - 1) >> 2;
    const int64_t w0 = min(max(A >> 2, wlo), whi), w1 = min(max((A >> 2) + 1, wlo), whi);
    const uint32_t lo = *(const uint32_t*)(uintptr_t)(w0 << 2), hi = *(const uint32_t*)(uintptr_t)(w1 << 2);
    return __builtin_amdgcn_alignbyte(hi, lo, (uint32_t)(A & 3));
}

__device__ inline uint32_t byte_of(uint32_t w, int q) { return (w >> (8 * q)) & 255u; }

// Each block covers kBlock quads (4 consecutive faces) of one component
// (blocks [0, b0) u, [b0, b1) v, [b1, ...) w), so the component and its arrays
// are uniform per block.  A thread reads the validity bytes of its quad and of
// the quad's six neighbour quads as 32-bit words (12 loads for 4 faces instead
// of 28 byte loads) before it uses any of them.
__global__ __launch_bounds__(kBlock) void k_extrapolate(Geom g, FaceSet a, ValidSet vs, int64_t b0, int64_t b1) {
    const int64_t blk = blockIdx.x;
    const int c = blk < b0 ? 0 : (blk < b1 ? 1 : 2);
    const int64_t lb = blk - (c == 0 ? 0 : (c == 1 ? b0 : b1));
    const uint8_t* __restrict__ vin = c == 0 ? vs.in[0] : (c == 1 ? vs.in[1] : vs.in[2]);
    uint8_t* __restrict__ vout = c == 0 ? vs.out[0] : (c == 1 ? vs.out[1] : vs.out[2]);
    float* __restrict__ A = c == 0 ? a.a[0] : (c == 1 ? a.a[1] : a.a[2]);
    const int sx = face_dim(g, c, 0), sy = face_dim(g, c, 1), sz = face_dim(g, c, 2);
    const int64_t nf = (int64_t)sx * sy * sz;
    const int64_t st[3] = {1, sx, (int64_t)sx * sy};
    const int64_t f0 = 4 * (lb * kBlock + threadIdx.x);
    if (f0 >= nf) return;
    // neighbour order of ExtrapolateToRegion: i+1, i-1, j+1, j-1, k+1, k-1
    const uint32_t self = bytes4(vin, nf, f0);
    const uint32_t xm = bytes4(vin, nf, f0 - 1), xp = bytes4(vin, nf, f0 + 1);
    const uint32_t ym = bytes4(vin, nf, f0 - st[1]), yp = bytes4(vin, nf, f0 + st[1]);
    const uint32_t zm = bytes4(vin, nf, f0 - st[2]), zp = bytes4(vin, nf, f0 + st[2]);
    int i, j, k;
    split_index(f0, sx, sy, i, j, k);
    uint32_t out = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int64_t f = f0 + q;
        if (f < nf) {
            if (byte_of(self, q)) {
                out |= 1u << (8 * q);
            } else {
                const bool in[6] = {i + 1 < sx, i > 0, j + 1 < sy, j > 0, k + 1 < sz, k > 0};
                const uint32_t nb[6] = {byte_of(xp, q), byte_of(xm, q), byte_of(yp, q),
                                        byte_of(ym, q), byte_of(zp, q), byte_of(zm, q)};
                float sum = 0.0f;
                unsigned count = 0;
#pragma unroll
                for (int d = 0; d < 6; ++d) {
                    if (!in[d] || !nb[d]) continue;
                    sum += A[(d & 1) ? f - st[d >> 1] : f + st[d >> 1]];
                    ++count;
                }
                if (count > 0) {
                    A[f] = sum / (float)count;
                    out |= 1u << (8 * q);
                }
            }
        }
        if (++i == sx) {
            i = 0;
            if (++j == sy) {
                j = 0;
                ++k;
            }
        }
    }
    uint8_t* o = vout + f0;
    if (f0 + 4 <= nf && ((uintptr_t)o & 3) == 0) {
        *(uint32_t*)o = out;
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (f0 + q < nf) o[q] = (uint8_t)byte_of(out, q);
    }
}

hipError_t launch_extrapolate(fe_scene* s) {
    const uint8_t* init[3] = {s->mark[0], s->mark[1], s->mark[2]};
    return launch_extrapolate_from(s, init);
}

// depth sweeps from the validity init[c]; the sweeps ping-pong through
// s->valid[0] and s->valid[1] (init may be s->valid[1])
hipError_t launch_extrapolate_from(fe_scene* s, const uint8_t* const init[3]) {
    const Geom& g = s->g;
    const int64_t per = kBlock * 4;  // faces per block
    const int64_t nb0 = (face_count(g, 0) + per - 1) / per, nb1 = (face_count(g, 1) + per - 1) / per,
                  nb2 = (face_count(g, 2) + per - 1) / per;
    FaceSet a{{s->u[0], s->u[1], s->u[2]}};
    const int64_t off[3] = {0, face_count(g, 0), face_count(g, 0) + face_count(g, 1)};
    for (int it = 0; it < s->depth; ++it) {
        ValidSet vs;
        for (int c = 0; c < 3; ++c) {
            vs.in[c] = it == 0 ? init[c] : s->valid[(it - 1) & 1] + off[c];
            vs.out[c] = s->valid[it & 1] + off[c];
        }
        hipLaunchKernelGGL(k_extrapolate, dim3((unsigned)(nb0 + nb1 + nb2)), dim3(kBlock), 0, s->stream, g, a, vs, nb0,
                           nb0 + nb1);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Wall layers of BlockedBoundryConditionSolver::ConstrainVelocity
// (src/fluid_solvers/blocked_boundry_condition_solver.cpp:198-244): u at i=0
// and i=nx, v at j=0 and j=ny, w at k=0 and k=nz.  The collider passes of the
// same function are identities when no collider is present.
// ---------------------------------------------------------------------------
__device__ inline bool is_wall(const Geom& g, int c, int i, int j, int k) {
    const int q = c == 0 ? i : (c == 1 ? j : k);
    return q == 0 || q == g.n[c];
}

__global__ __launch_bounds__(kBlock) void k_walls(Geom g, FaceSet a) {
    // flat index over the 6 wall layers
    const int64_t nyz = (int64_t)g.n[1] * g.n[2], nxz = (int64_t)g.n[0] * g.n[2], nxy = (int64_t)g.n[0] * g.n[1];
    int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t < 2 * nyz) {
        const int side = (int)(t / nyz);
        const int64_t r = t % nyz;
        const int j = (int)(r % g.n[1]), k = (int)(r / g.n[1]);
        const int i = side ? g.n[0] : 0;
        a.a[0][(int64_t)i + (int64_t)(g.n[0] + 1) * ((int64_t)j + (int64_t)g.n[1] * k)] = 0.0f;
        return;
    }
    t -= 2 * nyz;
    if (t < 2 * nxz) {
        const int side = (int)(t / nxz);
        const int64_t r = t % nxz;
        const int i = (int)(r % g.n[0]), k = (int)(r / g.n[0]);
        const int j = side ? g.n[1] : 0;
        a.a[1][(int64_t)i + (int64_t)g.n[0] * ((int64_t)j + (int64_t)(g.n[1] + 1) * k)] = 0.0f;
        return;
    }
    t -= 2 * nxz;
    if (t < 2 * nxy) {
        const int side = (int)(t / nxy);
        const int64_t r = t % nxy;
        const int i = (int)(r % g.n[0]), j = (int)(r / g.n[0]);
        const int k = side ? g.n[2] : 0;
        a.a[2][(int64_t)i + (int64_t)g.n[0] * ((int64_t)j + (int64_t)g.n[1] * k)] = 0.0f;
    }
}

hipError_t launch_walls(fe_scene* s) {
    const Geom& g = s->g;
    const int64_t tot = 2 * ((int64_t)g.n[1] * g.n[2] + (int64_t)g.n[0] * g.n[2] + (int64_t)g.n[0] * g.n[1]);
    FaceSet a{{s->u[0], s->u[1], s->u[2]}};
    hipLaunchKernelGGL(k_walls, grid1(tot), dim3(kBlock), 0, s->stream, g, a);
    return hipGetLastError();
}

// ExternalForce::ApplyExternalForce with a DirectionalField (src/forces/external_force.cpp:4-30):
// u += dt*g on every face, followed (optionally) by the wall layers of the
// ApplyBoundryCondition that ComputeExternalForces runs next (src/pic_simulator.cpp:203-212).
// Four consecutive faces of one component per thread, as one 16-byte vector
// (face arrays are separate allocations, so f0 = 4t is 16-byte aligned).
// Blocks [0, b0) take u, [b0, b1) v, [b1, ...) w.
__global__ __launch_bounds__(kBlock) void k_gravity(Geom g, FaceSet a, float gx, float gy, float gz,
                                                    const double* __restrict__ dtp, double g0, double g1, double g2,
                                                    int walls, int64_t b0, int64_t b1) {
    const int64_t blk = blockIdx.x;
    const int c = blk < b0 ? 0 : (blk < b1 ? 1 : 2);
    const int64_t f0 = 4 * ((blk - (c == 0 ? 0 : (c == 1 ? b0 : b1))) * kBlock + threadIdx.x);
    const int64_t nf = face_count(g, c);
    if (f0 >= nf) return;
    float* __restrict__ A = c == 0 ? a.a[0] : (c == 1 ? a.a[1] : a.a[2]);
    // dt * g by value, or formed from dt in device memory in a captured substep
    // (the same FP64 product rounded to FP32)
    const float add = dtp ? (float)(*dtp * (c == 0 ? g0 : (c == 1 ? g1 : g2))) : (c == 0 ? gx : (c == 1 ? gy : gz));
    const int sx = face_dim(g, c, 0), sy = face_dim(g, c, 1);
    // index of face f on the component's normal axis (the wall layers are 0 and n)
    auto normal_index = [&](int64_t f) -> int {
        const uint32_t u = (uint32_t)f;
        return c == 0 ? (int)(u % (uint32_t)sx)
                      : (c == 1 ? (int)((u / (uint32_t)sx) % (uint32_t)sy) : (int)(u / ((uint32_t)sx * sy)));
    };
    if (f0 + 4 <= nf) {
        float4 v = *(const float4*)(A + f0);
        float r[4] = {v.x + add, v.y + add, v.z + add, v.w + add};
        if (walls) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int nq = normal_index(f0 + q);
                if (nq == 0 || nq == g.n[c]) r[q] = 0.0f;
            }
        }
        *(float4*)(A + f0) = make_float4(r[0], r[1], r[2], r[3]);
    } else {
        for (int64_t f = f0; f < nf; ++f) {
            float v = A[f] + add;
            if (walls) {
                const int nq = normal_index(f);
                if (nq == 0 || nq == g.n[c]) v = 0.0f;
            }
            A[f] = v;
        }
    }
}

hipError_t launch_gravity(fe_scene* s, double dt, bool walls_after) {
    const Geom& g = s->g;
    const int64_t per = 4 * (int64_t)kBlock;  // faces per block
    const int64_t nb0 = (face_count(g, 0) + per - 1) / per, nb1 = (face_count(g, 1) + per - 1) / per,
                  nb2 = (face_count(g, 2) + per - 1) / per;
    FaceSet a{{s->u[0], s->u[1], s->u[2]}};
    hipLaunchKernelGGL(k_gravity, dim3((unsigned)(nb0 + nb1 + nb2)), dim3(kBlock), 0, s->stream, g, a,
                       (float)(dt * s->desc.gravity[0]), (float)(dt * s->desc.gravity[1]),
                       (float)(dt * s->desc.gravity[2]), (const double*)(s->capturing ? s->dt_dev : nullptr),
                       s->desc.gravity[0], s->desc.gravity[1], s->desc.gravity[2], walls_after ? 1 : 0, nb0, nb0 + nb1);
    return hipGetLastError();
}

// SinglePhasePressureSolver::ApplyPressure (single_phase_pressure_solver.cpp:56-92):
// only the +side face of each FLUID cell is corrected, when the +side cell is
// in range and not BOUNDRY: u(i+1) += (p(i+1) - p(i)) / h.  Optionally fused
// with the wall layers of the ApplyBoundryCondition that follows
// (src/pic_simulator.cpp:224-229).
// Four consecutive faces of one component per thread (u read and written as
// one 16-byte vector; face arrays are separate allocations).  Blocks [0, b0)
// take u, [b0, b1) v, [b1, ...) w.
__global__ __launch_bounds__(kBlock) void k_apply_pressure(Geom g, FaceSet a, const double* __restrict__ p,
                                                           const uint8_t* __restrict__ flags, int walls, int64_t b0,
                                                           int64_t b1) {
    const int64_t blk = blockIdx.x;
    const int c = blk < b0 ? 0 : (blk < b1 ? 1 : 2);
    const int64_t f0 = 4 * ((blk - (c == 0 ? 0 : (c == 1 ? b0 : b1))) * kBlock + threadIdx.x);
    const int64_t nf = face_count(g, c);
    if (f0 >= nf) return;
    float* __restrict__ A = c == 0 ? a.a[0] : (c == 1 ? a.a[1] : a.a[2]);
    const int sx = face_dim(g, c, 0), sy = face_dim(g, c, 1);
    const int64_t step = c == 0 ? 1 : (c == 1 ? (int64_t)g.n[0] : (int64_t)g.n[0] * g.n[1]);
    const double ih = g.ihd[c];
    auto face = [&](int64_t f, float v) {
        int i, j, k;
        split_index(f, sx, sy, i, j, k);
        const int q = c == 0 ? i : (c == 1 ? j : k);
        if (q >= 1 && q <= g.n[c] - 1) {
            const int64_t cr = (int64_t)i + (int64_t)g.n[0] * ((int64_t)j + (int64_t)g.n[1] * k);  // + side cell
            const int64_t cl = cr - step;
            // the pressures are read with the markers (both cells exist), not after them
            const uint8_t fl = flags[cl], fr = flags[cr];
            const double pr = p[cr], pl = p[cl];
            if (fl == FE_FLUID && fr != FE_BOUNDRY) v = (float)((double)v + ih * (pr - pl));
        }
        if (walls && (q == 0 || q == g.n[c])) v = 0.0f;
        return v;
    };
    if (f0 + 4 <= nf) {
        const float4 v = *(const float4*)(A + f0);
        *(float4*)(A + f0) = make_float4(face(f0, v.x), face(f0 + 1, v.y), face(f0 + 2, v.z), face(f0 + 3, v.w));
    } else {
        for (int64_t f = f0; f < nf; ++f) A[f] = face(f, A[f]);
    }
}

hipError_t launch_apply_pressure(fe_scene* s, bool walls_after) {
    const Geom& g = s->g;
    const int64_t per = 4 * (int64_t)kBlock;  // faces per block
    const int64_t nb0 = (face_count(g, 0) + per - 1) / per, nb1 = (face_count(g, 1) + per - 1) / per,
                  nb2 = (face_count(g, 2) + per - 1) / per;
    FaceSet a{{s->u[0], s->u[1], s->u[2]}};
    hipLaunchKernelGGL(k_apply_pressure, dim3((unsigned)(nb0 + nb1 + nb2)), dim3(kBlock), 0, s->stream, g, a, s->p,
                       s->flags, walls_after ? 1 : 0, nb0, nb0 + nb1);
    return hipGetLastError();
}

// PICSimulator::Cfl (src/pic_simulator.cpp:80-97): max over cells of
// |ValueAtCellCenter| + dt*(0,-9.8,0), in FP64; max is order independent, so
// an atomic max on the (non-negative) FP64 bit pattern is deterministic.
// out[0]: the maximum as FP64 bits (atomicMax on non-negative doubles), out[1]:
// set when a face value is not finite (NaN is otherwise ignored by fmax, as by
// the reference's std::max).  The launch also clears the other slot, which the
// next launch uses, so no memset is needed between launches.
__global__ __launch_bounds__(kBlock) void k_cfl(Geom g, FaceSet a, double dt, unsigned long long* __restrict__ out,
                                                unsigned long long* __restrict__ next) {
    __shared__ double smem[kBlock / 64];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        next[0] = 0ull;
        next[1] = 0ull;
    }
    double m = 0.0;
    int bad = 0;
    const int64_t nc = cell_count(g);
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    // eight cells per step, all their face loads issued before any is used
    constexpr int kG = 8;
    for (int64_t c0 = (int64_t)blockIdx.x * kBlock + threadIdx.x; c0 < nc; c0 += kG * stride) {
        float f[kG][6];
#pragma unroll
        for (int u = 0; u < kG; ++u) {
            const int64_t c = c0 + u * stride;
            if (c >= nc) continue;
            int i, j, k;
            split_index(c, g.n[0], g.n[1], i, j, k);
            const int64_t iu = (int64_t)i + (int64_t)(g.n[0] + 1) * ((int64_t)j + (int64_t)g.n[1] * k);
            const int64_t iv = (int64_t)i + (int64_t)g.n[0] * ((int64_t)j + (int64_t)(g.n[1] + 1) * k);
            const int64_t iw = (int64_t)i + (int64_t)g.n[0] * ((int64_t)j + (int64_t)g.n[1] * k);
            f[u][0] = a.a[0][iu];
            f[u][1] = a.a[0][iu + 1];
            f[u][2] = a.a[1][iv];
            f[u][3] = a.a[1][iv + g.n[0]];
            f[u][4] = a.a[2][iw];
            f[u][5] = a.a[2][iw + (int64_t)g.n[0] * g.n[1]];
        }
#pragma unroll
        for (int u = 0; u < kG; ++u) {
            if (c0 + u * stride >= nc) continue;
#pragma unroll
            for (int q = 0; q < 6; ++q) bad |= !isfinite(f[u][q]);
            const double cx = 0.5 * ((double)f[u][0] + (double)f[u][1]);
            const double cy = 0.5 * ((double)f[u][2] + (double)f[u][3]);
            const double cz = 0.5 * ((double)f[u][4] + (double)f[u][5]);
            m = fmax(m, fabs(cx) + dt * 0.0);
            m = fmax(m, fabs(cy) + dt * -9.8);
            m = fmax(m, fabs(cz) + dt * 0.0);
        }
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
    if ((threadIdx.x & 63) == 0) smem[threadIdx.x >> 6] = m;
    bad = __syncthreads_or(bad);
    if (threadIdx.x == 0) {
        double b = smem[0];
        for (int w = 1; w < kBlock / 64; ++w) b = fmax(b, smem[w]);
        atomicMax(out, (unsigned long long)__double_as_longlong(b));
        if (bad) atomicOr(out + 1, 1ull);
    }
}

hipError_t launch_cfl(fe_scene* s, double dt) {
    const int64_t nc = cell_count(s->g);
    FaceSet a{{s->u[0], s->u[1], s->u[2]}};
    // slots of two words alternate between launches (the scene starts with both clear)
    unsigned long long* out = s->cflmax + 2 * (s->cfl_slot & 1);
    unsigned long long* next = s->cflmax + 2 * ((s->cfl_slot + 1) & 1);
    // few blocks with a grid-stride loop: one atomic per block on a single address
    const unsigned blocks = (unsigned)std::min<int64_t>((nc + kBlock - 1) / kBlock, 512);
    hipLaunchKernelGGL(k_cfl, dim3(blocks), dim3(kBlock), 0, s->stream, s->g, a, dt, out, next);
    return hipGetLastError();
}

}  // namespace fe
