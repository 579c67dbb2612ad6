// k_particles.hip -- particle-side kernels of the gfx950 substep:
//   * counting sort of particles by cell (deterministic: bins hold their
//     particles in ascending previous order),
//   * P2G as a gather over 2x2x2 tiles of faces from the sorted cell bins (no float atomics),
//   * fused G2P + midpoint advection (+ the next substep's bin keys).
#include <algorithm>
#include <functional>
#include <vector>

#include "fe_kernels.hpp"

namespace fe {

// ---------------------------------------------------------------------------
// Binning
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_fill_i32(int32_t* __restrict__ a, int32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) a[i] = v;
}

// Runs of equal keys in consecutive lanes of a wavefront.  Particles arrive
// sorted by cell, so a wavefront holds few runs: one atomic per run instead of
// one per particle keeps same-address atomics from serialising.  The active
// lanes of a wavefront are a prefix (one particle per lane, in index order);
// every lane, active or not, must call these.
struct KeyRun {
    bool head;  // first lane of its run
    int first;  // lane of this lane's run head
    int len;    // run length (valid on the head lane)
};

__device__ inline KeyRun key_run(uint32_t k, bool active) {
    const int lane = threadIdx.x & 63;
    const uint32_t kp = __shfl_up(k, 1, 64);
    const bool head = active && (lane == 0 || kp != k);
    const uint64_t heads = __ballot(head);
    const uint64_t act = __ballot(active);
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);  // lanes 0..lane
    KeyRun r;
    r.head = head;
    r.first = 63 - __clzll(heads & upto);
    const uint64_t later = heads & ~upto;
    const int end = later ? __ffsll((long long)later) - 1 : 64 - __clzll(act);
    r.len = end - lane;
    return r;
}

// Adds 1 per active lane to count[k] with one atomic per run.
__device__ inline void count_runs(int32_t* __restrict__ count, uint32_t k, bool active) {
    const KeyRun r = key_run(k, active);
    if (r.head) atomicAdd(&count[k], r.len);
}

// Bin key of each particle + histogram.
__global__ __launch_bounds__(kBlock) void k_keys(Geom g, const float* __restrict__ X, int64_t cap, int64_t np,
                                                 uint32_t* __restrict__ key, int32_t* __restrict__ count) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool active = p < np;
    uint32_t k = 0;
    if (active) {
        const int cx = cell_axis(g, 0, X[p]);
        const int cy = cell_axis(g, 1, X[cap + p]);
        const int cz = cell_axis(g, 2, X[2 * cap + p]);
        k = (uint32_t)cx + (uint32_t)g.n[0] * ((uint32_t)cy + (uint32_t)g.n[1] * (uint32_t)cz);
        key[p] = k;
    }
    if (count) count_runs(count, k, active);
}

// Exclusive scan of int32 counts, three passes of 1024 items per block.
constexpr int kScanItems = 4;
constexpr int kScanTile = kBlock * kScanItems;

__device__ inline int block_exclusive_scan(int val, int* smem, int& total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = val;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) smem[wid] = x;
    __syncthreads();
    int pre = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        int s = smem[w];
        if (w < wid) pre += s;
        total += s;
    }
    __syncthreads();
    return x - val + pre;
}

__global__ __launch_bounds__(kBlock) void k_scan_tiles(const int32_t* __restrict__ in, int64_t n,
                                                       int32_t* __restrict__ tile_sums) {
    __shared__ int smem[kBlock / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int s = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q)
        if (base + q < n) s += in[base + q];
    int total;
    block_exclusive_scan(s, smem, total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kBlock) void k_scan_sums(int32_t* __restrict__ sums, int64_t ntiles) {
    __shared__ int smem[kBlock / 64];
    int carry = 0;
    for (int64_t off = 0; off < ntiles; off += kBlock) {
        const int64_t i = off + threadIdx.x;
        int v = i < ntiles ? sums[i] : 0;
        int total;
        int ex = block_exclusive_scan(v, smem, total);
        if (i < ntiles) sums[i] = ex + carry;
        carry += total;
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_final(const int32_t* __restrict__ in, int64_t n,
                                                       const int32_t* __restrict__ tile_off,
                                                       int32_t* __restrict__ out) {
    __shared__ int smem[kBlock / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        v[q] = base + q < n ? in[base + q] : 0;
        s += v[q];
    }
    int total;
    int ex = block_exclusive_scan(s, smem, total) + tile_off[blockIdx.x];
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        if (base + q < n) out[base + q] = ex;
        ex += v[q];
    }
    if (base + kScanItems >= n && base < n) out[n] = ex;  // total after the last item
}

// Slots in the bins: the head of each run reserves the run's slots with one
// atomic and the run's lanes take them in lane order.  The reservations count
// the histogram down (a bin's slots are taken from its end), so it is all
// zero afterwards, ready for the next histogram.  Workgroup 0 also clears
// P2G's value-bound slots, which k_gather raises, and the count of
// k_sort_bins' large-bin list.
__global__ __launch_bounds__(kBlock) void k_scatter(const uint32_t* __restrict__ key, int64_t np,
                                                    const int32_t* __restrict__ start, int32_t* __restrict__ count,
                                                    int32_t* __restrict__ order, uint32_t* __restrict__ bound,
                                                    int32_t* __restrict__ big) {
    static_assert(3 * kP2GBoundSlots <= kBlock, "one workgroup clears the bound slots");
    if (blockIdx.x == 0) {
        if (threadIdx.x < 3 * kP2GBoundSlots) bound[threadIdx.x] = 0u;
        if (threadIdx.x == 0) big[0] = 0;
    }
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool active = p < np;
    const uint32_t k = active ? key[p] : 0u;
    const KeyRun r = key_run(k, active);
    int base = 0;
    if (r.head) base = start[k] + atomicSub(&count[k], r.len) - r.len;
    base = __shfl(base, r.first, 64);
    if (active) order[base + ((int)(threadIdx.x & 63) - r.first)] = (int32_t)p;
}

// Bins of at most N entries sort in registers: all loads first, then an
// odd-even transposition network (N rounds of compare-exchanges at fixed
// positions, so no register array is indexed at run time); padding entries
// hold INT32_MAX and sort to the end.
template <int N>
__device__ inline void sort_bin_regs(int32_t* __restrict__ order, int s0, int n) {
    int32_t a[N];
#pragma unroll
    for (int u = 0; u < N; ++u) a[u] = u < n ? order[s0 + u] : INT32_MAX;
#pragma unroll
    for (int r = 0; r < N; ++r) {
#pragma unroll
        for (int j = r & 1; j + 1 < N; j += 2) {
            const int32_t lo = min(a[j], a[j + 1]), hi = max(a[j], a[j + 1]);
            a[j] = lo;
            a[j + 1] = hi;
        }
    }
#pragma unroll
    for (int u = 0; u < N; ++u)
        if (u < n) order[s0 + u] = a[u];
}

// Atomic slots give each bin its particles in arbitrary order; sorting every
// bin by previous index restores a deterministic, stable order.  Bins of up to
// 32 particles sort in registers here; larger ones (particles pile up in a
// few cells where the fluid is compressed: several hundred have been seen in a
// 128^3 dam break) go to a list for k_sort_big_bins, one workgroup per bin.
// The list lives in `cursor` (free once k_scatter has taken its slots):
// cursor[0] counts, cursor[1 ..] holds the cells.  A bin that finds the list
// full (more large bins than cells - 1, possible only on tiny grids) sorts by
// insertion here.
constexpr int kRegBin = 32;
__global__ __launch_bounds__(kBlock) void k_sort_bins(const int32_t* __restrict__ start, int64_t ncells,
                                                      int32_t* __restrict__ order, int32_t* __restrict__ big) {
    int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= ncells) return;
    const int s0 = start[c], s1 = start[c + 1];
    const int n = s1 - s0;
    if (n <= 1) return;
    if (n <= 8) {
        sort_bin_regs<8>(order, s0, n);
        return;
    }
    if (n <= 16) {
        sort_bin_regs<16>(order, s0, n);
        return;
    }
    if (n <= kRegBin) {
        sort_bin_regs<kRegBin>(order, s0, n);
        return;
    }
    const int slot = atomicAdd(&big[0], 1);
    if (slot < ncells - 1) {
        big[1 + slot] = (int32_t)c;
        return;
    }
    for (int a = s0 + 1; a < s1; ++a) {
        const int32_t v = order[a];
        int b = a - 1;
        while (b >= s0 && order[b] > v) {
            order[b + 1] = order[b];
            --b;
        }
        order[b + 1] = v;
    }
}

// The listed bins, one workgroup at a time each (the grid loops over the
// list; every workgroup reads the same count, so all leave together).  The
// entries of a bin are distinct previous indices, so each one's place is the
// number of entries below it: ranks are counted against chunks of the bin
// staged in LDS, the entries are written by rank into `scratch` and copied
// back.
constexpr int kRankChunk = 2048;
__global__ __launch_bounds__(kBlock) void k_sort_big_bins(const int32_t* __restrict__ start,
                                                          int32_t* __restrict__ order,
                                                          const int32_t* __restrict__ big, int64_t ncells,
                                                          int32_t* __restrict__ scratch) {
    __shared__ int32_t chunk[kRankChunk];
    const int nbig = min((int64_t)big[0], ncells - 1);
    for (int i = blockIdx.x; i < nbig; i += gridDim.x) {
        const int c = big[1 + i];
        const int s0 = start[c], n = start[c + 1] - s0;
        for (int e0 = 0; e0 < n; e0 += kBlock) {
            const int e = e0 + threadIdx.x;
            const int32_t v = e < n ? order[s0 + e] : 0;
            int rank = 0;
            for (int c0 = 0; c0 < n; c0 += kRankChunk) {
                const int m = min(kRankChunk, n - c0);
                __syncthreads();  // the previous chunk has been read
                for (int u = threadIdx.x; u < m; u += kBlock) chunk[u] = order[s0 + c0 + u];
                __syncthreads();
                if (e < n)
                    for (int u = 0; u < m; ++u) rank += chunk[u] < v ? 1 : 0;
            }
            if (e < n) scratch[s0 + rank] = v;
        }
        __syncthreads();  // every rank written before the copy reads them
        for (int u = threadIdx.x; u < n; u += kBlock) order[s0 + u] = scratch[s0 + u];
        __syncthreads();  // the copy done before the next bin's first chunk is staged
    }
}

// The stencil of GetCooridnatesAndWeights for face component c at position
// pp, with the scheme's stencil size per axis: base node index and fraction
// per axis (the same stencil_axis values on every path).  P2G evaluates it per
// candidate from the binned position, so all P2G variants use the same values.
__device__ inline void p2g_stencil(const Geom& g, int c, float4 pp, int base[3], float fr[3]) {
    const float p[3] = {pp.x, pp.y, pp.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) stencil_axis(g, c, a, p[a], stencil_size(g, c, a), base[a], fr[a]);
}

// Besides the reorder, k_gather writes the binned position as one float4 (the
// one load per candidate of P2G and the SDF).
// It also raises P2G's bound on the values it transfers (the fixed-point
// scale below): per component c the largest |v_c| and, for APIC, |v_c| plus
// sum_a |C_ca| times the larger distance from the (clamped) particle to the
// two face planes of its stencil on axis a, which bounds every affine value
// v_c + C_c . (x_face - x_p) the particle hands to a face.  Workgroup maxima
// go to slot (block mod kP2GBoundSlots) with atomicMax on the FP32 bits
// (monotonic for non-negative floats; NaN and infinity sort above every
// finite bound, so a non-finite velocity shows as a non-finite bound).
__global__ __launch_bounds__(kBlock) void k_gather(Geom g, const int32_t* __restrict__ order, int64_t np,
                                                   int64_t cap, Particles src, Particles dst, int with_c,
                                                   float4* __restrict__ pos4, uint32_t* __restrict__ bound) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t bb[3] = {0u, 0u, 0u};
    if (s < np) {
        const int64_t o = order[s];
        float p[3], v[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            p[a] = src.x[a * cap + o];
            v[a] = src.v[a * cap + o];
            dst.x[a * cap + s] = p[a];
            dst.v[a * cap + s] = v[a];
        }
        const float4 pp = make_float4(p[0], p[1], p[2], 0.0f);
        pos4[s] = pp;
#pragma unroll
        for (int c = 0; c < 3; ++c) bb[c] = __float_as_uint(fabsf(v[c]));
        if (with_c) {
            float cm[9];
#pragma unroll
            for (int a = 0; a < 9; ++a) {
                cm[a] = src.C[a * cap + o];
                dst.C[a * cap + s] = cm[a];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                int base[3];
                float fr[3];
                p2g_stencil(g, c, pp, base, fr);
                float t = fabsf(v[c]);
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    float pc = p[a];
                    if (a != c) {
                        const float clo = g.o[a] + 0.5f * g.h[a], chi = g.o[a] + g.L[a] - 0.5f * g.h[a];
                        pc = pc < clo ? clo : (pc > chi ? chi : pc);
                    }
                    const float g0 = g.co[c][a] + g.h[a] * (float)base[a];
                    const float g1 = g.co[c][a] + g.h[a] * (float)(base[a] + 1);
                    t += fabsf(cm[3 * c + a]) * fmaxf(fabsf(g0 - pc), fabsf(g1 - pc));
                }
                bb[c] = __float_as_uint(t);
            }
        }
        dst.pid[s] = src.pid[o];
    }
    __shared__ uint32_t wmax[kBlock / 64][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) bb[c] = max(bb[c], (uint32_t)__shfl_xor((int)bb[c], off, 64));
        if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6][c] = bb[c];
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        uint32_t m = 0u;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) m = max(m, wmax[w][threadIdx.x]);
        if (m) atomicMax(&bound[threadIdx.x * kP2GBoundSlots + (blockIdx.x % kP2GBoundSlots)], m);
    }
}

// ---------------------------------------------------------------------------
// P2G: TransferParticles2Grid, u_F = sum_p ve_pF w_pF / sum_p w_pF over the
// particles p whose trilinear stencil has face F as a node (PIC:
// src/pic_simulator.cpp:356-437; FLIP: src/flip_simulator.cpp:25-119; APIC:
// src/apic_simulator.cpp:15-118; ve_pF is v_p, or v_p + C_p . (x_F - x_p)
// for APIC).
//
// The two sums are formed in 64-bit fixed point.  Every term is rounded once
// to an integer and the integers are added; integer addition is exact and
// associative, so a face's sums do not depend on the order its terms arrive
// in.  That is what lets the scatter variant add terms with LDS atomics and
// still give the bits of the gather variants, run after run.  Scales: a term
// of the weight sum is w 2^H and one of the value sum ve w 2^(H-E), where
// |ve| < 2^E over all particles (k_gather's bound; it is exact up to FP32
// rounding) and H = min(61 - ceil(log2(np + 1)), 49), so every term is below
// 2^(H+1) <= 2^50 in magnitude and a face's sum over at most np terms stays
// below 2^62.  A term is the exact FP64 product of the FP32 ve and w, scaled
// by a power of two (exact) and rounded to an integer by adding 1.5 * 2^52
// (2^52 for the non-negative weight terms) in one fma: the integer is the
// low mantissa bits.  The quotient is formed in FP64 and rounded to FP32.
// Each term's rounding is 2^-H of the bound (2^-39 at 4M particles), far
// below the FP32 rounding of ve and w themselves.  A non-finite velocity
// (bound not finite) makes every face that has weight NaN.
// ---------------------------------------------------------------------------
struct P2GFixed {
    double sv;    // 2^(H-E): value-term scale
    double sw;    // 2^H: weight-term scale
    double back;  // 2^E: (value sum / weight sum) * back is the face value
    bool bad;     // the bound is not finite
};

// The bound of component c: the largest of its kP2GBoundSlots slots, read by
// every lane of the wavefront (call before any lane leaves the kernel).
__device__ inline P2GFixed p2g_fixed(const uint32_t* __restrict__ bound, int c, int head) {
    static_assert(kP2GBoundSlots == 64, "one slot per lane");
    uint32_t m = bound[c * kP2GBoundSlots + (threadIdx.x & 63)];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, off, 64));
    P2GFixed fx;
    fx.bad = m >= 0x7f800000u;
    int e = 0;
    if (m && !fx.bad) (void)frexpf(__uint_as_float(m), &e);  // bound < 2^e
    fx.sv = ldexp(1.0, head - e);
    fx.sw = ldexp(1.0, head);
    fx.back = ldexp(1.0, e);
    return fx;
}

// H of the scales for np particles (host side).
inline int p2g_head(int64_t np) {
    int c = 0;
    while (c < 62 && (int64_t(1) << c) < np + 1) ++c;
    return std::min(61 - c, 49);
}

// Value term round(ve w 2^(H-E)) from vs = ve * 2^(H-E) (exact in FP64).
__device__ inline int64_t p2g_value_term(double vs, float w) {
    const double r = fma(vs, (double)w, 6755399441055744.0);  // 1.5 * 2^52
    return (int64_t)((uint64_t)__double_as_longlong(r) - 0x4338000000000000ull);
}

// Weight term round(w 2^H), w >= 0.
__device__ inline int64_t p2g_weight_term(float w, double sw) {
    const double r = fma((double)w, sw, 4503599627370496.0);  // 2^52
    return (int64_t)((uint64_t)__double_as_longlong(r) - 0x4330000000000000ull);
}

__device__ inline float p2g_quotient(int64_t num, int64_t den, const P2GFixed& fx) {
    if (den <= 0) return 0.0f;
    if (fx.bad) return __int_as_float(0x7fc00000);
    return (float)((double)num / (double)den * fx.back);
}

// APIC's value for the face at gp: val + C_c . (gp - pc), pc the particle
// position clamped on the transverse axes (apic_simulator.cpp:54-63).
__device__ inline float apic_value(float val, float c0, float c1, float c2, float gx, float gy, float gz,
                                   const float pc[3]) {
#pragma clang fp contract(off)
    float t = c0 * (gx - pc[0]);
    t = fmaf(c1, gy - pc[1], t);
    t = fmaf(c2, gz - pc[2], t);
    return val + t;
}

// APIC: the particle position clamped to [o + h/2, o + L - h/2] on the axes
// transverse to component C.
template <int C>
__device__ inline void apic_clamped(const Geom& g, float4 pp, float pc[3]) {
#pragma clang fp contract(off)
    pc[0] = pp.x;
    pc[1] = pp.y;
    pc[2] = pp.z;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (a == C) continue;
        const float lo = g.o[a] + 0.5f * g.h[a], hi = g.o[a] + g.L[a] - 0.5f * g.h[a];
        pc[a] = pc[a] < lo ? lo : (pc[a] > hi ? hi : pc[a]);
    }
}

// Position of face index F on axis a of component C (GetXPos/GetYPos/GetZPos,
// face_centered_grid3d.cpp:303-316); rounded as a product and a sum in every
// variant (no contraction), so the variants' APIC values agree bit for bit.
__device__ inline float face_pos(const Geom& g, int C, int a, int F) {
#pragma clang fp contract(off)
    return g.co[C][a] + g.h[a] * (float)F;
}

// Candidate cells of the faces [F0, F1] of component C on axis a: a face's
// contributors lie in cells [F-1, F+1] on the transverse axes and [F-1, F]
// on the normal axis, plus cell n-1 for face n-2 when PIC/APIC clamp the
// normal axis's stencil at n-2 (their stencil size is the cell count).
template <int METHOD>
__device__ inline void candidate_cells(const Geom& g, int C, int a, int F0, int F1, int& lo, int& hi) {
    lo = max(F0 - 1, 0);
    if (a == C) {
        hi = F1;
        if (METHOD != FE_FLIP && F1 >= g.n[a] - 2) hi = g.n[a] - 1;
    } else {
        hi = F1 + 1;
    }
    hi = min(hi, g.n[a] - 1);
}

// ---------------------------------------------------------------------------
// P2G gather over 2x2x2 tiles of faces from the sorted cell bins.  A tile's
// candidate cells are the union of its faces' (4 x 4 x 3 cells, against
// 3 x 3 x 2 per face), so each candidate is read once for 8 faces.  Four
// lanes share a tile, lane q scanning the candidate cells of one z plane (one
// y row when z is the normal axis); cells are x-fastest and particles sorted
// by cell, so each (j, k) row of candidate cells is one contiguous particle
// range.  For every candidate the reference's own stencil (GetCooridnatesAndWeights
// with the scheme's stencil size, precomputed per particle by k_gather)
// decides: the particle contributes to a face iff the face is one of its 8
// nodes; the face is marked even for a zero weight.
// ---------------------------------------------------------------------------
constexpr int kP2GLanes = 4;  // lanes per tile (one candidate plane each)

template <int C, int METHOD>
__global__ __launch_bounds__(kBlock) void k_p2g(Geom g, const float* __restrict__ V, const float* __restrict__ Cm,
                                                const float4* __restrict__ pos4, int64_t cap,
                                                const int32_t* __restrict__ start,
                                                const uint32_t* __restrict__ bound, int head, float* __restrict__ out,
                                                uint8_t* __restrict__ mark, float* __restrict__ snap,
                                                int64_t padded) {
    const P2GFixed fx = p2g_fixed(bound, C, head);
    const int64_t blk = xcd_remap(blockIdx.x, padded);
    const int sx = face_dim(g, C, 0), sy = face_dim(g, C, 1), sz = face_dim(g, C, 2);
    const int tx = (sx + 1) >> 1, ty = (sy + 1) >> 1, tz = (sz + 1) >> 1;
    const int64_t t = blk * (kBlock / kP2GLanes) + (threadIdx.x / kP2GLanes);
    const int q = threadIdx.x % kP2GLanes;
    // Whole tiles (four aligned lanes) leave together, so the lane exchanges
    // below only ever read lanes of the same tile.
    if (t >= (int64_t)tx * ty * tz) return;
    int ti, tj, tk;
    split_index(t, tx, ty, ti, tj, tk);
    const int F0[3] = {2 * ti, 2 * tj, 2 * tk};
    int lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) candidate_cells<METHOD>(g, C, a, F0[a], F0[a] + 1, lo[a], hi[a]);
    // Face e = ex + 2 ey + 4 ez of the tile is (F0x + ex, F0y + ey, F0z + ez).
    int64_t num[8], den[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) num[e] = den[e] = 0;
    uint32_t marked = 0;
    // One candidate, from its stencil (base node, fraction per axis): face F
    // takes it iff F - base is 0 or 1 on every axis, with weight factor 1 - f
    // at offset 0 and f at offset 1.
    auto consider = [&](int s, float4 pp, float val) {
        int bs[3];
        float fr[3];
        p2g_stencil(g, C, pp, bs, fr);
        float w0[3], w1[3];
        uint32_t ok = 0;  // bit 2a: face offset 0 on axis a accepts; bit 2a+1: offset 1
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int d0 = F0[a] - bs[a], d1 = d0 + 1;
            const bool a0 = (unsigned)d0 <= 1u, a1 = (unsigned)d1 <= 1u;
            w0[a] = d0 ? fr[a] : 1.0f - fr[a];
            w1[a] = d1 ? fr[a] : 1.0f - fr[a];
            ok |= (a0 ? 1u : 0u) << (2 * a);
            ok |= (a1 ? 1u : 0u) << (2 * a + 1);
        }
        // faces accepted on all three axes, as an 8-bit mask over e
        const uint32_t mx = (ok & 3u) * 0x55u;
        const uint32_t my = ((ok & 4u) ? 0x33u : 0u) | ((ok & 8u) ? 0xCCu : 0u);
        const uint32_t mz = ((ok & 16u) ? 0x0Fu : 0u) | ((ok & 32u) ? 0xF0u : 0u);
        const uint32_t m = mx & my & mz;
        if (!m) return;
        float pc[3];
        float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
        if (METHOD == FE_APIC) {
            apic_clamped<C>(g, pp, pc);
            const float* cr = Cm + (int64_t)(3 * C) * cap + s;
            c0 = cr[0];
            c1 = cr[cap];
            c2 = cr[2 * cap];
        }
        const double vs = (double)val * fx.sv;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            if (!((m >> e) & 1u)) continue;
            const float w = (((e & 1) ? w1[0] : w0[0]) * ((e & 2) ? w1[1] : w0[1])) * ((e & 4) ? w1[2] : w0[2]);
            if (METHOD == FE_APIC) {
                const float ve = apic_value(val, c0, c1, c2, face_pos(g, C, 0, F0[0] + (e & 1)),
                                            face_pos(g, C, 1, F0[1] + ((e >> 1) & 1)),
                                            face_pos(g, C, 2, F0[2] + (e >> 2)), pc);
                num[e] += p2g_value_term((double)ve * fx.sv, w);
            } else {
                num[e] += p2g_value_term(vs, w);
            }
            den[e] += p2g_weight_term(w, fx.sw);
        }
        marked |= m;
    };
    // Lane q takes the candidate cells with coordinate lo[LA] + q on the lane
    // axis LA: the z planes (4 of them) when z is transverse, the y rows when
    // z is the normal axis (C = 2: 3 planes but 4 rows), so all four lanes work.
    constexpr int LA = C == 2 ? 1 : 2, OA = 3 - LA;
    const int lc = lo[LA] + q;
    if (lc <= hi[LA]) {
        const float* __restrict__ vc = V + (int64_t)C * cap;
        constexpr int kBatch = 8;  // candidates whose loads are issued together
        for (int oc = lo[OA]; oc <= hi[OA]; ++oc) {
            const int jc = LA == 1 ? lc : oc, kc = LA == 2 ? lc : oc;
            const int64_t row = (int64_t)g.n[0] * ((int64_t)jc + (int64_t)g.n[1] * kc);
            const int s0 = start[row + lo[0]], s1 = start[row + hi[0] + 1];
            int s = s0;
            for (; s + kBatch <= s1; s += kBatch) {
                float4 rr[kBatch];
                float pv[kBatch];
#pragma unroll
                for (int u = 0; u < kBatch; ++u) {
                    rr[u] = pos4[s + u];
                    pv[u] = vc[s + u];
                }
#pragma unroll
                for (int u = 0; u < kBatch; ++u) consider(s + u, rr[u], pv[u]);
            }
            for (; s < s1; ++s) consider(s, pos4[s], vc[s]);
        }
    }
    // the tile's four lanes' sums (exact integers: any order gives the same)
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        num[e] += __shfl_xor(num[e], 1, kP2GLanes);
        den[e] += __shfl_xor(den[e], 1, kP2GLanes);
        num[e] += __shfl_xor(num[e], 2, kP2GLanes);
        den[e] += __shfl_xor(den[e], 2, kP2GLanes);
    }
    marked |= __shfl_xor(marked, 1, kP2GLanes);
    marked |= __shfl_xor(marked, 2, kP2GLanes);
    // Lane q writes faces e = 2q and 2q + 1 (ey = q & 1, ez = q >> 1).
    const int J = F0[1] + (q & 1), K = F0[2] + (q >> 1);
    if (J >= sy || K >= sz) return;
#pragma unroll
    for (int ex = 0; ex < 2; ++ex) {
        const int I = F0[0] + ex;
        if (I >= sx) continue;
        const int e = ex + 2 * q;
        const int64_t f = (int64_t)I + (int64_t)sx * ((int64_t)J + (int64_t)sy * K);
        int64_t ne = 0, de = 0;
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (u == e) {
                ne = num[u];
                de = den[u];
            }
        const float res = p2g_quotient(ne, de, fx);
        out[f] = res;
        mark[f] = (uint8_t)((marked >> e) & 1u);
        if (METHOD == FE_FLIP) snap[f] = res;  // FLIP snapshot right after P2G (flip_simulator.cpp:107-118)
    }
}

// The same transfer with one thread per face (more threads, each face's
// candidates read by the face itself): the variant that wins on small grids.
template <int C, int METHOD>
__global__ __launch_bounds__(kBlock) void k_p2g_face(Geom g, const float* __restrict__ V, const float* __restrict__ Cm,
                                                     const float4* __restrict__ pos4, int64_t cap,
                                                     const int32_t* __restrict__ start,
                                                     const uint32_t* __restrict__ bound, int head,
                                                     float* __restrict__ out, uint8_t* __restrict__ mark,
                                                     float* __restrict__ snap, int64_t padded) {
    const P2GFixed fx = p2g_fixed(bound, C, head);
    const int64_t blk = xcd_remap(blockIdx.x, padded);
    const int sx = face_dim(g, C, 0), sy = face_dim(g, C, 1), sz = face_dim(g, C, 2);
    const int64_t f = blk * kBlock + threadIdx.x;
    if (f >= (int64_t)sx * sy * sz) return;
    int I, J, K;
    split_index(f, sx, sy, I, J, K);
    const int F[3] = {I, J, K};
    int lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) candidate_cells<METHOD>(g, C, a, F[a], F[a], lo[a], hi[a]);
    float gp[3];
    if (METHOD == FE_APIC) {
#pragma unroll
        for (int a = 0; a < 3; ++a) gp[a] = face_pos(g, C, a, F[a]);
    }
    int64_t num = 0, den = 0;
    uint32_t marked = 0;
    auto consider = [&](int s, float4 pp, float val) {
        int bs[3];
        float fr[3];
        p2g_stencil(g, C, pp, bs, fr);
        const uint32_t dx = (uint32_t)(I - bs[0]), dy = (uint32_t)(J - bs[1]), dz = (uint32_t)(K - bs[2]);
        if ((dx | dy | dz) > 1u) return;
        const float wx = dx ? fr[0] : 1.0f - fr[0];
        const float wy = dy ? fr[1] : 1.0f - fr[1];
        const float wz = dz ? fr[2] : 1.0f - fr[2];
        const float w = (wx * wy) * wz;
        if (METHOD == FE_APIC) {
            float pc[3];
            apic_clamped<C>(g, pp, pc);
            const float* cr = Cm + (int64_t)(3 * C) * cap + s;
            val = apic_value(val, cr[0], cr[cap], cr[2 * cap], gp[0], gp[1], gp[2], pc);
        }
        marked = 1;
        num += p2g_value_term((double)val * fx.sv, w);
        den += p2g_weight_term(w, fx.sw);
    };
    const float* __restrict__ vc = V + (int64_t)C * cap;
    constexpr int kBatch = 8;  // candidates whose loads are issued together
    for (int kc = lo[2]; kc <= hi[2]; ++kc) {
        for (int jc = lo[1]; jc <= hi[1]; ++jc) {
            const int64_t row = (int64_t)g.n[0] * ((int64_t)jc + (int64_t)g.n[1] * kc);
            const int s0 = start[row + lo[0]], s1 = start[row + hi[0] + 1];
            int s = s0;
            for (; s + kBatch <= s1; s += kBatch) {
                float4 rr[kBatch];
                float pv[kBatch];
#pragma unroll
                for (int u = 0; u < kBatch; ++u) {
                    rr[u] = pos4[s + u];
                    pv[u] = vc[s + u];
                }
#pragma unroll
                for (int u = 0; u < kBatch; ++u) consider(s + u, rr[u], pv[u]);
            }
            for (; s < s1; ++s) consider(s, pos4[s], vc[s]);
        }
    }
    const float res = p2g_quotient(num, den, fx);
    out[f] = res;
    mark[f] = (uint8_t)marked;
    if (METHOD == FE_FLIP) snap[f] = res;  // FLIP snapshot right after P2G (flip_simulator.cpp:107-118)
}

// ---------------------------------------------------------------------------
// P2G as a scatter over blocks of faces: the variant whose loads coalesce.  A
// workgroup owns an 8 x 8 x 8 block of faces of component C and walks the
// particles of the block's candidate cells (candidate_cells over the block's
// faces): each x row of candidate cells is one contiguous range of sorted
// particles, and the workgroup takes the concatenated ranges 1024 candidates
// at a time, the lanes of a wavefront 16 candidates apart (a thread finds its
// row by a binary search of the ranges' prefix sums in LDS).  Every stencil node of a
// candidate that lies in the block adds the node's two fixed-point terms to
// the block's LDS sums with 64-bit integer atomics; nodes outside the block
// belong to the neighbouring block, which reads the same particle.  A face is
// marked iff one of the 8 stencil bases that have it as a node holds a
// particle (one bit per base of the block in LDS), the reference's rule
// (marked even for a zero weight).  A particle is read by about 1.8 blocks
// (10 x 10 x 9 candidate cells per 8 x 8 x 8 faces) against 6 tiles of the
// 2x2x2 gather.
// ---------------------------------------------------------------------------
constexpr int kScB = 8;                                        // faces per block edge
constexpr int kScFaces = kScB * kScB * kScB;                   // faces per block
// LDS slot of block face (x, y, z): x + 9 y + 76 z.  A 64-bit sum spans two
// of the 32 four-byte banks, so slot mod 16 picks its bank pair; the padded
// strides put the four (y, z) parities of a wavefront's nodes at bank pairs
// 0, 9, 12 and 5 instead of all at 0 and 8.
constexpr int kScSy = 9, kScSz = 76;
constexpr int kScRun = 1;    // consecutive candidates per lane group (see k_p2g_scatter)
constexpr int kScAhead = 1;  // chunks whose loads are in flight ahead of the atomics
constexpr int kScThreads = 1024;  // threads per workgroup, one chunk of candidates per iteration
constexpr int kScSlots = (kScB - 1) * (1 + kScSy + kScSz) + 1;
static_assert(kScSy >= kScB && kScSz >= (kScB - 1) * (1 + kScSy) + 1, "faces of a block never share a slot");
constexpr int kScRows = (kScB + 2) * (kScB + 2);               // x rows of candidate cells, at most
constexpr int kScBases = (kScB + 1) * (kScB + 1) * (kScB + 1); // stencil bases with a node in the block
static_assert(kScRows <= 128, "one wavefront scans the row lengths, two per lane");

template <int C, int METHOD>
__global__ __launch_bounds__(kScThreads) void k_p2g_scatter(Geom g, const float* __restrict__ V,
                                                        const float* __restrict__ Cm,
                                                        const float4* __restrict__ pos4, int64_t cap,
                                                        const int32_t* __restrict__ start,
                                                        const uint32_t* __restrict__ bound, int head,
                                                        float* __restrict__ out, uint8_t* __restrict__ mark,
                                                        float* __restrict__ snap, int nbx, int nby, int nbz,
                                                        int64_t padded) {
    __shared__ unsigned long long s_num[kScSlots], s_den[kScSlots];
    __shared__ uint32_t s_occ[(kScBases + 31) / 32];
    __shared__ int32_t s_first[kScRows], s_pre[kScRows + 1];
    const P2GFixed fx = p2g_fixed(bound, C, head);
    const int64_t blk = xcd_remap(blockIdx.x, padded);
    if (blk >= (int64_t)nbx * nby * nbz) return;  // the whole workgroup
    int bi, bj, bk;
    split_index(blk, nbx, nby, bi, bj, bk);
    const int F0[3] = {bi * kScB, bj * kScB, bk * kScB};
    const int sx = face_dim(g, C, 0), sy = face_dim(g, C, 1), sz = face_dim(g, C, 2);
    int lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) candidate_cells<METHOD>(g, C, a, F0[a], F0[a] + kScB - 1, lo[a], hi[a]);
    const int ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1, rows = ny * nz;
    for (int u = threadIdx.x; u < kScSlots; u += kScThreads) s_num[u] = s_den[u] = 0ull;
    for (int u = threadIdx.x; u < (kScBases + 31) / 32; u += kScThreads) s_occ[u] = 0u;
    if (threadIdx.x < 64) {
        // row lengths and their exclusive prefix sums (rows 2l and 2l + 1 on lane l)
        const int lane = threadIdx.x;
        int len[2] = {0, 0};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int r = 2 * lane + h;
            if (r < rows) {
                // rows y-fastest; z-fastest for the w faces, whose stencil z
                // base is the cell's own: consecutive rows then differ in that
                // base, which spreads a chunk's atomics over more slots (the
                // sums do not depend on the order)
                const int jc = C == 2 ? lo[1] + r / nz : lo[1] + r % ny;
                const int kc = C == 2 ? lo[2] + r % nz : lo[2] + r / ny;
                const int64_t row = (int64_t)g.n[0] * ((int64_t)jc + (int64_t)g.n[1] * kc);
                const int s0 = start[row + lo[0]];
                len[h] = start[row + hi[0] + 1] - s0;
                s_first[r] = s0;
            }
        }
        const int pair = len[0] + len[1];
        int inc = pair;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int o = __shfl_up(inc, off, 64);
            if (lane >= off) inc += o;
        }
        const int ex = inc - pair;
        if (2 * lane < rows) s_pre[2 * lane] = ex;
        if (2 * lane + 1 < rows) s_pre[2 * lane + 1] = ex + len[0];
        if (lane == 63) s_pre[rows] = inc;
    }
    __syncthreads();
    const int total = s_pre[rows];
    const float* __restrict__ vc = V + (int64_t)C * cap;
    // Thread (wavefront w of W = kScThreads / 64, lane l) takes candidate
    // kScRun (w + W (l / kScRun)) + l % kScRun of each chunk of kScThreads
    // (with kScRun = 1: w + W l), so a wavefront's atomics land on the nodes
    // of 64 / kScRun places spread over the chunk instead of the few cells
    // that 64 consecutive particles fill: same-address atomics in one
    // instruction serialise.  Spreading the lanes measured faster than runs of
    // 2, 4 or 8 consecutive particles (15, 23 and 60 % more time per launch
    // at 128^3 with 256 threads), and 1024 threads (lanes 16 apart) faster than
    // 512 and 256 (254, 268 and 339 us for the three launches).  The positions and values of the next
    // kScAhead chunks are loaded before this chunk's atomics (one chunk
    // ahead measured faster than two).
    const int lane = threadIdx.x & 63;
    const int slot = (lane % kScRun) + kScRun * ((threadIdx.x >> 6) + (kScThreads / 64) * (lane / kScRun));
    auto locate = [&](int q) {
        int r = 0;
#pragma unroll
        for (int step = 64; step > 0; step >>= 1)
            if (r + step < rows && s_pre[r + step] <= q) r += step;
        return s_first[r] + (q - s_pre[r]);
    };
    auto process = [&](int s, float4 pp, float val) {
        int d[3];  // stencil base - F0 per axis
        float fr[3];
        p2g_stencil(g, C, pp, d, fr);
#pragma unroll
        for (int a = 0; a < 3; ++a) d[a] -= F0[a];
        if ((unsigned)(d[0] + 1) > (unsigned)kScB || (unsigned)(d[1] + 1) > (unsigned)kScB ||
            (unsigned)(d[2] + 1) > (unsigned)kScB)
            return;  // no node in the block
        const int bit = (d[0] + 1) + (kScB + 1) * ((d[1] + 1) + (kScB + 1) * (d[2] + 1));
        atomicOr(&s_occ[bit >> 5], 1u << (bit & 31));
        float pc[3];
        float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
        if (METHOD == FE_APIC) {
            apic_clamped<C>(g, pp, pc);
            const float* cr = Cm + (int64_t)(3 * C) * cap + s;
            c0 = cr[0];
            c1 = cr[cap];
            c2 = cr[2 * cap];
        }
        const double vs = (double)val * fx.sv;
        const int base = d[0] + kScSy * d[1] + kScSz * d[2];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int ex = e & 1, ey = (e >> 1) & 1, ez = e >> 2;
            if ((unsigned)(d[0] + ex) >= (unsigned)kScB || (unsigned)(d[1] + ey) >= (unsigned)kScB ||
                (unsigned)(d[2] + ez) >= (unsigned)kScB)
                continue;
            const float w = ((ex ? fr[0] : 1.0f - fr[0]) * (ey ? fr[1] : 1.0f - fr[1])) * (ez ? fr[2] : 1.0f - fr[2]);
            int64_t tn;
            if (METHOD == FE_APIC) {
                const float ve = apic_value(val, c0, c1, c2, face_pos(g, C, 0, F0[0] + d[0] + ex),
                                            face_pos(g, C, 1, F0[1] + d[1] + ey), face_pos(g, C, 2, F0[2] + d[2] + ez),
                                            pc);
                tn = p2g_value_term((double)ve * fx.sv, w);
            } else {
                tn = p2g_value_term(vs, w);
            }
            const int li = base + ex + kScSy * ey + kScSz * ez;
            atomicAdd(&s_num[li], (unsigned long long)tn);
            atomicAdd(&s_den[li], (unsigned long long)p2g_weight_term(w, fx.sw));
        }
    };
    int sp[kScAhead + 1];
    float4 rec[kScAhead + 1];
    float val[kScAhead + 1];
#pragma unroll
    for (int u = 0; u <= kScAhead; ++u) {
        sp[u] = 0;
        rec[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        val[u] = 0.0f;
    }
#pragma unroll
    for (int u = 0; u < kScAhead; ++u) {
        const int qa = slot + u * kScThreads;
        if (qa < total) {
            sp[u] = locate(qa);
            rec[u] = pos4[sp[u]];
            val[u] = vc[sp[u]];
        }
    }
    for (int q = slot; q < total; q += kScThreads) {
        const int qa = q + kScAhead * kScThreads;
        if (qa < total) {
            sp[kScAhead] = locate(qa);
            rec[kScAhead] = pos4[sp[kScAhead]];
            val[kScAhead] = vc[sp[kScAhead]];
        }
        process(sp[0], rec[0], val[0]);
#pragma unroll
        for (int u = 0; u < kScAhead; ++u) {
            sp[u] = sp[u + 1];
            rec[u] = rec[u + 1];
            val[u] = val[u + 1];
        }
    }
    __syncthreads();
    for (int u = threadIdx.x; u < kScFaces; u += kScThreads) {
        const int l[3] = {u % kScB, (u / kScB) % kScB, u / (kScB * kScB)};
        const int I = F0[0] + l[0], J = F0[1] + l[1], K = F0[2] + l[2];
        if (I >= sx || J >= sy || K >= sz) continue;
        uint32_t m = 0;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            // base F - (ex, ey, ez), as a bit of the block's base range [F0 - 1, F0 + 7]
            const int bit = (l[0] - (e & 1) + 1) + (kScB + 1) * ((l[1] - ((e >> 1) & 1) + 1) +
                                                                 (kScB + 1) * (l[2] - (e >> 2) + 1));
            m |= (s_occ[bit >> 5] >> (bit & 31)) & 1u;
        }
        const int li = l[0] + kScSy * l[1] + kScSz * l[2];
        const float res = p2g_quotient((int64_t)s_num[li], (int64_t)s_den[li], fx);
        const int64_t f = (int64_t)I + (int64_t)sx * ((int64_t)J + (int64_t)sy * K);
        out[f] = res;
        mark[f] = (uint8_t)m;
        if (METHOD == FE_FLIP) snap[f] = res;  // FLIP snapshot right after P2G (flip_simulator.cpp:107-118)
    }
}

// ---------------------------------------------------------------------------
// G2P (PIC: src/pic_simulator.cpp:439-450; FLIP: src/flip_simulator.cpp:121-161;
// APIC: src/apic_simulator.cpp:120-176) fused with MoveParticles
// (src/pic_simulator.cpp:239-312) and the bin keys of the next substep.
// ---------------------------------------------------------------------------
// PIC sample of a and the FLIP delta sample of a - snapshot (node by node,
// flip_simulator.cpp:121-161) on one stencil: 8 pair loads instead of 24 loads.
// Face values for G2P: the plain arrays (four 8-byte pair loads per sample)
// or their quad layouts (two 16-byte loads); the same values either way.
template <bool Quads>
struct GridSrc {
    FaceSet a;   // plain arrays
    QuadSet q;   // quad layouts (Quads only)
    __device__ float sample(const Geom& g, int c, float px, float py, float pz) const {
        if (Quads) return sample_quad(q.a[c], g, c, px, py, pz);
        return sample_face(a.a[c], g, c, px, py, pz);
    }
    // the same sample from the point's shared axis coordinates
    __device__ float sample_at(const Geom& g, int c, const PointAxes& p) const {
        const FaceStencil32 t = stencil_at(g, c, p);
        if (Quads) return trilerp_quads(q.a[c][t.r00], q.a[c][t.r01], t);
        const float* b = a.a[c];
        return trilerp_pairs(load_pair(b + t.r00), load_pair(b + t.r10), load_pair(b + t.r01), load_pair(b + t.r11),
                             t);
    }
    template <typename S>
    __device__ void pairs(const S& t, int c, float2 (&v)[4]) const {
        if (Quads) {
            const float4 q0 = q.a[c][t.r00], q1 = q.a[c][t.r01];
            v[0] = make_float2(q0.x, q0.y);
            v[1] = make_float2(q0.z, q0.w);
            v[2] = make_float2(q1.x, q1.y);
            v[3] = make_float2(q1.z, q1.w);
        } else {
            const float* p = a.a[c];
            v[0] = load_pair(p + t.r00);
            v[1] = load_pair(p + t.r10);
            v[2] = load_pair(p + t.r01);
            v[3] = load_pair(p + t.r11);
        }
    }
};

template <bool Quads>
__device__ inline void sample_pic_delta(const GridSrc<Quads>& ga, const GridSrc<Quads>& gs, const Geom& g, int c,
                                        const PointAxes& p, float& pic, float& delta) {
    const FaceStencil32 t = stencil_at(g, c, p);
    float2 av[4], sv[4];
    ga.pairs(t, c, av);
    gs.pairs(t, c, sv);
    const float2 a00 = av[0], a10 = av[1], a01 = av[2], a11 = av[3];
    const float2 s00 = sv[0], s10 = sv[1], s01 = sv[2], s11 = sv[3];
    pic = trilerp_pairs(a00, a10, a01, a11, t);
    auto d = [](float2 x, float2 y) { return make_float2(x.x - y.x, x.y - y.y); };
    delta = trilerp_pairs(d(a00, s00), d(a10, s10), d(a01, s01), d(a11, s11), t);
}

// GetCooridnatesAndGradientWeights (src/common/math_utils.hpp:114-176) contracted
// with the face values: returns sum_q gradw_q * a(node_q).
__device__ inline void grad_contract(const float* __restrict__ a, const Geom& g, int c, const float p[3],
                                     float out[3]) {
    const int sx = face_dim(g, c, 0), sy = face_dim(g, c, 1);
    int i, j, k;
    float fx, fy, fz;
    fast_axis(p[0], g.co[c][0], g.ih[0], g.n[0], i, fx);  // APIC passes the cell dims (Q1)
    fast_axis(p[1], g.co[c][1], g.ih[1], g.n[1], j, fy);
    fast_axis(p[2], g.co[c][2], g.ih[2], g.n[2], k, fz);
    const int ip1 = min(i + 1, g.n[0] - 1), jp1 = min(j + 1, g.n[1] - 1), kp1 = min(k + 1, g.n[2] - 1);
    const float ix = g.ih[0], iy = g.ih[1], iz = g.ih[2];
    auto A = [&](int ii, int jj, int kk) { return a[(int64_t)ii + (int64_t)sx * ((int64_t)jj + (int64_t)sy * kk)]; };
    const float v0 = A(i, j, k), v1 = A(ip1, j, k), v2 = A(i, jp1, k), v3 = A(ip1, jp1, k);
    const float v4 = A(i, j, kp1), v5 = A(ip1, j, kp1), v6 = A(i, jp1, kp1), v7 = A(ip1, jp1, kp1);
    out[0] = (-ix * (1 - fy) * (1 - fz)) * v0 + (ix * (1 - fy) * (1 - fz)) * v1 + ((-ix) * fy * (1 - fz)) * v2 +
             (ix * fy * (1 - fz)) * v3 + ((-ix) * (1 - fy) * fz) * v4 + (ix * (1 - fy) * fz) * v5 +
             ((-ix) * fy * fz) * v6 + (ix * fy * fz) * v7;
    out[1] = (-iy * (1 - fx) * (1 - fz)) * v0 + (fx * (-iy) * (1 - fz)) * v1 + ((1 - fx) * iy * (1 - fz)) * v2 +
             (fx * iy * (1 - fz)) * v3 + ((1 - fx) * (-iy) * fz) * v4 + (fx * (-iy) * fz) * v5 +
             ((1 - fx) * iy * fz) * v6 + (fx * iy * fz) * v7;
    out[2] = (-iz * (1 - fx) * (1 - fy)) * v0 + (fx * (1 - fy) * (-iz)) * v1 + ((1 - fx) * fy * (-iz)) * v2 +
             (fx * fy * (-iz)) * v3 + ((1 - fx) * (1 - fy) * iz) * v4 + (fx * (1 - fy) * iz) * v5 +
             ((1 - fx) * fy * iz) * v6 + (fx * fy * iz) * v7;
}

template <int METHOD, bool Quads>
__global__ __launch_bounds__(kBlock) void k_g2p_move(Geom g, Particles P, int64_t cap, int64_t np, FaceSet grid,
                                                     GridSrc<Quads> gsrc, GridSrc<Quads> ssrc, float blend, float dt_arg,
                                                     const double* __restrict__ dtp, int nsub, int do_g2p, int do_move,
                                                     uint32_t* __restrict__ key, int32_t* __restrict__ count) {
    // dt comes by value, or from device memory in a captured substep (the same FP32 rounding)
    const float dt = dtp ? (float)*dtp : dt_arg;
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool active = s < np;
    uint32_t k = 0;
    if (active) {  // inactive lanes stay for the wavefront-wide key runs at the end
        float x = P.x[s], y = P.x[cap + s], z = P.x[2 * cap + s];
        float v[3] = {P.v[s], P.v[cap + s], P.v[2 * cap + s]};
        // the grid velocity at the particle: G2P's PIC sample, and the first
        // sample of the first midpoint step (the same point, the same grid)
        float pic[3];
        const PointAxes here = point_axes(g, x, y, z);
        if (do_g2p) {
            if (METHOD == FE_PIC) {
#pragma unroll
                for (int c = 0; c < 3; ++c) v[c] = pic[c] = gsrc.sample_at(g, c, here);
            } else if (METHOD == FE_FLIP) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    float delta;
                    sample_pic_delta(gsrc, ssrc, g, c, here, pic[c], delta);
                    float fl = v[c] + delta;
                    if (blend > 0.0f) fl = (1.0f - blend) * fl + blend * pic[c];
                    v[c] = fl;
                }
            } else {
#pragma unroll
                for (int c = 0; c < 3; ++c) pic[c] = gsrc.sample_at(g, c, here);
#pragma unroll
                for (int c = 0; c < 3; ++c) v[c] = pic[c];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    float pc[3] = {x, y, z};
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        if (a == c) continue;
                        const float lo = g.o[a] + 0.5f * g.h[a], hi = g.o[a] + g.L[a] - 0.5f * g.h[a];
                        pc[a] = pc[a] < lo ? lo : (pc[a] > hi ? hi : pc[a]);
                    }
                    float cr[3];
                    grad_contract(grid.a[c], g, c, pc, cr);
                    P.C[(int64_t)(3 * c) * cap + s] = cr[0];
                    P.C[(int64_t)(3 * c + 1) * cap + s] = cr[1];
                    P.C[(int64_t)(3 * c + 2) * cap + s] = cr[2];
                }
            }
        }
        if (do_move) {
            const float sdt = dt / (float)nsub;
            float p0[3] = {x, y, z}, p1[3] = {x, y, z};
            for (int t = 0; t < nsub; ++t) {
                float v0[3], mid[3], mv[3];
                if (t == 0 && do_g2p) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) v0[c] = pic[c];
                } else {
                    const PointAxes pa = t == 0 ? here : point_axes(g, p0[0], p0[1], p0[2]);
#pragma unroll
                    for (int c = 0; c < 3; ++c) v0[c] = gsrc.sample_at(g, c, pa);
                }
#pragma unroll
                for (int a = 0; a < 3; ++a) mid[a] = p0[a] + 0.5f * sdt * v0[a];
                const PointAxes pm = point_axes(g, mid[0], mid[1], mid[2]);
#pragma unroll
                for (int c = 0; c < 3; ++c) mv[c] = gsrc.sample_at(g, c, pm);
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    p1[a] = p0[a] + sdt * mv[a];
                    p0[a] = p1[a];
                }
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float lo = g.o[a], hi = g.o[a] + g.L[a];
                if (p1[a] <= lo) {
                    p1[a] = lo;
                    v[a] = 0.0f;
                }
                if (p1[a] >= hi) {
                    p1[a] = hi;
                    v[a] = 0.0f;
                }
            }
            x = p1[0];
            y = p1[1];
            z = p1[2];
            P.x[s] = x;
            P.x[cap + s] = y;
            P.x[2 * cap + s] = z;
        }
        P.v[s] = v[0];
        P.v[cap + s] = v[1];
        P.v[2 * cap + s] = v[2];
        if (key) {
            const int cx = cell_axis(g, 0, x), cy = cell_axis(g, 1, y), cz = cell_axis(g, 2, z);
            k = (uint32_t)cx + (uint32_t)g.n[0] * ((uint32_t)cy + (uint32_t)g.n[1] * (uint32_t)cz);
            key[s] = k;
        }
    }
    if (key) count_runs(count, k, active);
}

// Scatter sorted particles back to upload order (download path).
__global__ __launch_bounds__(kBlock) void k_unpermute(const float* __restrict__ src, int ncomp, int64_t cap,
                                                      int64_t np, const int32_t* __restrict__ pid,
                                                      float* __restrict__ dst) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= np) return;
    const int64_t o = pid[s];
    for (int c = 0; c < ncomp; ++c) dst[o * ncomp + c] = src[c * cap + s];
}

// Interleaved upload (xyz xyz ...) -> SoA planes.
__global__ __launch_bounds__(kBlock) void k_deinterleave(const float* __restrict__ src, int ncomp, int64_t cap,
                                                         int64_t np, float* __restrict__ dst) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= np) return;
    for (int c = 0; c < ncomp; ++c) dst[c * cap + s] = src[s * ncomp + c];
}

__global__ __launch_bounds__(kBlock) void k_iota(int32_t* __restrict__ a, int64_t n) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s < n) a[s] = (int32_t)s;
}

__global__ __launch_bounds__(kBlock) void k_cells_by_pid(const uint32_t* __restrict__ key,
                                                         const int32_t* __restrict__ pid, int64_t np,
                                                         int32_t* __restrict__ out) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s < np) out[pid[s]] = (int32_t)key[s];
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static inline dim3 grid1(int64_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

hipError_t launch_fill_i32(hipStream_t st, int32_t* a, int32_t v, int64_t n) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_i32, grid1(n), dim3(kBlock), 0, st, a, v, n);
    return hipGetLastError();
}

hipError_t launch_keys(fe_scene* s, bool histogram) {
    const Particles& P = s->P[s->cur];
    if (s->np == 0) return hipSuccess;
    hipLaunchKernelGGL(k_keys, grid1(s->np), dim3(kBlock), 0, s->stream, s->g, P.x, s->cap, s->np, s->key,
                       histogram ? s->cell_count : nullptr);
    return hipGetLastError();
}

int64_t scan_tiles(int64_t n) { return (n + kScanTile - 1) / kScanTile; }

hipError_t launch_scan_with(hipStream_t st, const int32_t* in, int32_t* out, int64_t n, int32_t* sums) {
    const int64_t tiles = scan_tiles(n);
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(kBlock), 0, st, in, n, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, st, sums, tiles);
    hipLaunchKernelGGL(k_scan_final, dim3((unsigned)tiles), dim3(kBlock), 0, st, in, n, sums, out);
    return hipGetLastError();
}

hipError_t launch_scan(fe_scene* s, const int32_t* in, int32_t* out, int64_t n) {
    return launch_scan_with(s->stream, in, out, n, s->scan_sums);
}

constexpr int kBigBinGroups = 256;  // workgroups looping over the large-bin list

// Bins the particles of P[cur] by cell into P[1-cur]; keys and counts must be
// ready when keys_ready (the fused advection kernel produced them).
hipError_t launch_sort(fe_scene* s, bool keys_ready) {
    const int64_t nc = cell_count(s->g);
    hipError_t e;
    if (!keys_ready) {
        if (!s->count_zero && (e = launch_fill_i32(s->stream, s->cell_count, 0, nc + 1)) != hipSuccess) return e;
        if ((e = launch_keys(s, true)) != hipSuccess) return e;
        s->count_zero = false;
    }
    if ((e = launch_scan(s, s->cell_count, s->cell_start, nc)) != hipSuccess) return e;
    if (s->np > 0) {
        // counts the histogram back to zero; clears the large-bin list's count (in cursor)
        hipLaunchKernelGGL(k_scatter, grid1(s->np), dim3(kBlock), 0, s->stream, s->key, s->np, s->cell_start,
                           s->cell_count, s->order, s->p2g_bound, s->cursor);
        s->count_zero = true;
        hipLaunchKernelGGL(k_sort_bins, grid1(nc), dim3(kBlock), 0, s->stream, s->cell_start, nc, s->order,
                           s->cursor);
        hipLaunchKernelGGL(k_sort_big_bins, dim3(kBigBinGroups), dim3(kBlock), 0, s->stream,
                           (const int32_t*)s->cell_start, s->order, (const int32_t*)s->cursor, nc,
                           reinterpret_cast<int32_t*>(s->key));
        hipLaunchKernelGGL(k_gather, grid1(s->np), dim3(kBlock), 0, s->stream, s->g, s->order, s->np, s->cap,
                           s->P[s->cur], s->P[1 - s->cur], s->g.method == FE_APIC ? 1 : 0, s->pos4, s->p2g_bound);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        s->cur = 1 - s->cur;
    }
    s->sorted = true;
    s->sdf_from_cells = 0;  // new bins: an SDF in memory no longer matches them
    return hipSuccess;
}

template <int METHOD>
static hipError_t p2g_method(fe_scene* s) {
    const Particles& P = s->P[s->cur];
    const Geom& g = s->g;
    const int head = p2g_head(s->np);
    for (int c = 0; c < 3; ++c) {
        if (s->tune.p2g == FE_P2G_SCATTER) {
            const int nb[3] = {(face_dim(g, c, 0) + kScB - 1) / kScB, (face_dim(g, c, 1) + kScB - 1) / kScB,
                               (face_dim(g, c, 2) + kScB - 1) / kScB};
            const int64_t padded = pad_blocks((int64_t)nb[0] * nb[1] * nb[2]);
            auto kern = c == 0 ? k_p2g_scatter<0, METHOD>
                               : (c == 1 ? k_p2g_scatter<1, METHOD> : k_p2g_scatter<2, METHOD>);
            launch_timed(s, FE_KERNEL_P2G, 0, kern, dim3((unsigned)padded), dim3(kScThreads), g, (const float*)P.v,
                         (const float*)P.C, (const float4*)s->pos4, s->cap,
                         (const int32_t*)s->cell_start, (const uint32_t*)s->p2g_bound, head, s->u[c], s->mark[c],
                         s->snap[c], nb[0], nb[1], nb[2], padded);
            continue;
        }
        if (s->tune.p2g == FE_P2G_FACES) {
            const int64_t padded = pad_blocks((face_count(g, c) + kBlock - 1) / kBlock);
            auto kern = c == 0 ? k_p2g_face<0, METHOD> : (c == 1 ? k_p2g_face<1, METHOD> : k_p2g_face<2, METHOD>);
            launch_timed(s, FE_KERNEL_P2G, 0, kern, dim3((unsigned)padded), dim3(kBlock), g, (const float*)P.v,
                         (const float*)P.C, (const float4*)s->pos4, s->cap,
                         (const int32_t*)s->cell_start, (const uint32_t*)s->p2g_bound, head, s->u[c], s->mark[c],
                         s->snap[c], padded);
            continue;
        }
        const int64_t tiles = (int64_t)((face_dim(g, c, 0) + 1) / 2) * ((face_dim(g, c, 1) + 1) / 2) *
                              ((face_dim(g, c, 2) + 1) / 2);
        const int64_t per_block = kBlock / kP2GLanes;
        const int64_t padded = pad_blocks((tiles + per_block - 1) / per_block);
        auto kern = c == 0 ? k_p2g<0, METHOD> : (c == 1 ? k_p2g<1, METHOD> : k_p2g<2, METHOD>);
        launch_timed(s, FE_KERNEL_P2G, 0, kern, dim3((unsigned)padded), dim3(kBlock), g, (const float*)P.v,
                     (const float*)P.C, (const float4*)s->pos4, s->cap,
                     (const int32_t*)s->cell_start, (const uint32_t*)s->p2g_bound, head, s->u[c], s->mark[c],
                     s->snap[c], padded);
    }
    return hipGetLastError();
}

hipError_t launch_p2g(fe_scene* s) {
    switch (s->g.method) {
        case FE_PIC: return p2g_method<FE_PIC>(s);
        case FE_FLIP: return p2g_method<FE_FLIP>(s);
        default: return p2g_method<FE_APIC>(s);
    }
}

// Quad layout of the three face arrays a[c] into q[c] (see sample_quad).
__global__ __launch_bounds__(kBlock) void k_quads(Geom g, const float* __restrict__ a0, const float* __restrict__ a1,
                                                  const float* __restrict__ a2, float4* __restrict__ q0,
                                                  float4* __restrict__ q1, float4* __restrict__ q2, int64_t b0,
                                                  int64_t b1) {
    const int64_t blk = blockIdx.x;
    const int c = blk < b0 ? 0 : (blk < b1 ? 1 : 2);
    const int64_t f = (blk - (c == 0 ? 0 : (c == 1 ? b0 : b1))) * kBlock + threadIdx.x;
    const float* __restrict__ a = c == 0 ? a0 : (c == 1 ? a1 : a2);
    float4* __restrict__ q = c == 0 ? q0 : (c == 1 ? q1 : q2);
    const int sx = face_dim(g, c, 0), sy = face_dim(g, c, 1), sz = face_dim(g, c, 2);
    if (f >= (int64_t)sx * sy * sz) return;
    int i, j, k;
    split_index(f, sx, sy, i, j, k);
    const int64_t di = i + 1 < sx ? 1 : 0, dj = j + 1 < sy ? sx : 0;
    q[f] = make_float4(a[f], a[f + di], a[f + dj], a[f + dj + di]);
}

hipError_t launch_quads(fe_scene* s, float* const a[3], float4* const q[3]) {
    const Geom& g = s->g;
    const int64_t nb0 = (face_count(g, 0) + kBlock - 1) / kBlock, nb1 = (face_count(g, 1) + kBlock - 1) / kBlock,
                  nb2 = (face_count(g, 2) + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_quads, dim3((unsigned)(nb0 + nb1 + nb2)), dim3(kBlock), 0, s->stream, g, (const float*)a[0],
                       (const float*)a[1], (const float*)a[2], q[0], q[1], q[2], nb0, nb0 + nb1);
    return hipGetLastError();
}

hipError_t launch_g2p_move(fe_scene* s, double dt, bool do_g2p, bool do_move, bool emit_keys) {
    if (s->np == 0) return hipSuccess;
    const Particles& P = s->P[s->cur];
    FaceSet grid{{s->u[0], s->u[1], s->u[2]}};
    FaceSet snaps{{s->snap[0], s->snap[1], s->snap[2]}};
    QuadSet gq{{s->quad[0], s->quad[1], s->quad[2]}};
    QuadSet sq{{s->squad[0], s->squad[1], s->squad[2]}};
    const bool quads = s->tune.g2p == FE_G2P_QUADS;
    const bool flip_g2p = do_g2p && s->g.method == FE_FLIP;
    hipError_t e;
    if (quads && (e = launch_quads(s, s->u, s->quad)) != hipSuccess) return e;
    if (quads && flip_g2p && (e = launch_quads(s, s->snap, s->squad)) != hipSuccess) return e;
    if (emit_keys && !s->count_zero) {
        if ((e = launch_fill_i32(s->stream, s->cell_count, 0, cell_count(s->g) + 1)) != hipSuccess) return e;
        s->count_zero = true;
    }
    uint32_t* key = emit_keys ? s->key : nullptr;
    int32_t* cnt = emit_keys ? s->cell_count : nullptr;
    const float blend = (float)s->desc.blend;
    const double* dtp = s->capturing ? s->dt_dev : nullptr;
    if (quads) {
        auto kern = s->g.method == FE_PIC ? k_g2p_move<FE_PIC, true>
                                          : (s->g.method == FE_FLIP ? k_g2p_move<FE_FLIP, true> : k_g2p_move<FE_APIC, true>);
        launch_timed(s, FE_KERNEL_G2P_MOVE, 0, kern, grid1(s->np), dim3(kBlock), s->g, P, s->cap, s->np, grid,
                     GridSrc<true>{grid, gq}, GridSrc<true>{snaps, sq}, blend, (float)dt, dtp, s->move_substeps,
                     (int)do_g2p, (int)do_move, key, cnt);
    } else {
        auto kern = s->g.method == FE_PIC
                        ? k_g2p_move<FE_PIC, false>
                        : (s->g.method == FE_FLIP ? k_g2p_move<FE_FLIP, false> : k_g2p_move<FE_APIC, false>);
        launch_timed(s, FE_KERNEL_G2P_MOVE, 0, kern, grid1(s->np), dim3(kBlock), s->g, P, s->cap, s->np, grid,
                     GridSrc<false>{grid, gq}, GridSrc<false>{snaps, sq}, blend, (float)dt, dtp, s->move_substeps,
                     (int)do_g2p, (int)do_move, key, cnt);
    }
    if (emit_keys) s->count_zero = false;  // the histogram of the next bins
    if (do_move) s->sorted = false;
    return hipGetLastError();
}

hipError_t launch_unpermute(fe_scene* s, const float* src, int ncomp, float* dst) {
    if (s->np == 0) return hipSuccess;
    hipLaunchKernelGGL(k_unpermute, grid1(s->np), dim3(kBlock), 0, s->stream, src, ncomp, s->cap, s->np,
                       s->P[s->cur].pid, dst);
    return hipGetLastError();
}

hipError_t launch_deinterleave(fe_scene* s, const float* src, int ncomp, float* dst) {
    if (s->np == 0) return hipSuccess;
    hipLaunchKernelGGL(k_deinterleave, grid1(s->np), dim3(kBlock), 0, s->stream, src, ncomp, s->cap, s->np, dst);
    return hipGetLastError();
}

hipError_t launch_iota(fe_scene* s, int32_t* a, int64_t n) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_iota, grid1(n), dim3(kBlock), 0, s->stream, a, n);
    return hipGetLastError();
}

hipError_t launch_cells_by_pid(fe_scene* s, int32_t* out) {
    if (s->np == 0) return hipSuccess;
    hipError_t e = launch_keys(s, false);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_cells_by_pid, grid1(s->np), dim3(kBlock), 0, s->stream, s->key, s->P[s->cur].pid, s->np,
                       out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Autotuner of the kernel variants.  Each variant runs `reps` times between
// HIP events on the scene's stream and the median decides.  P2G runs on the
// sorted particles as the substep would; G2P and advection run on a copy of
// the particles in the second particle buffer (free between sorts), so the
// scene's particles are not changed.  The variants of a stage are
// bit-identical, so the choice never changes results.
// ---------------------------------------------------------------------------
static hipError_t time_median(fe_scene* s, int reps, double* out, const std::function<hipError_t()>& prep,
                              const std::function<hipError_t()>& run) {
    std::vector<double> ms;
    hipEvent_t a, b;
    hipError_t e;
    if ((e = hipEventCreate(&a)) != hipSuccess) return e;
    if ((e = hipEventCreate(&b)) != hipSuccess) {
        (void)hipEventDestroy(a);
        return e;
    }
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        if (prep && (e = prep()) != hipSuccess) break;
        if ((e = hipEventRecord(a, s->stream)) != hipSuccess) break;
        if ((e = run()) != hipSuccess) break;
        if ((e = hipEventRecord(b, s->stream)) != hipSuccess) break;
        if ((e = hipEventSynchronize(b)) != hipSuccess) break;
        float t = 0.0f;
        if ((e = hipEventElapsedTime(&t, a, b)) != hipSuccess) break;
        ms.push_back(t);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (e != hipSuccess) return e;
    std::sort(ms.begin(), ms.end());
    *out = ms.empty() ? 0.0 : ms[ms.size() / 2];
    return hipSuccess;
}

hipError_t autotune(fe_scene* s, int reps, fe_tuning_times* times) {
    fe_tuning_times t{};
    hipError_t e;
    if (s->np == 0) {
        if (times) *times = t;
        return hipSuccess;
    }
    if (!s->sorted && (e = launch_sort(s, false)) != hipSuccess) return e;
    const fe_tuning keep = s->tune;
    for (int v = 0; v < kP2GVariants; ++v) {
        s->tune.p2g = v;
        if ((e = time_median(s, reps, &t.p2g_ms[v], nullptr, [&] { return launch_p2g(s); })) != hipSuccess) break;
    }
    // G2P + advection on a copy of the particles in the other buffer
    const int cur = s->cur;
    const Particles& src = s->P[cur];
    const Particles& dst = s->P[1 - cur];
    auto copy = [&]() -> hipError_t {
        hipError_t r = hipMemcpyAsync(dst.x, src.x, sizeof(float) * 3 * s->cap, hipMemcpyDeviceToDevice, s->stream);
        if (r == hipSuccess)
            r = hipMemcpyAsync(dst.v, src.v, sizeof(float) * 3 * s->cap, hipMemcpyDeviceToDevice, s->stream);
        if (r == hipSuccess && src.C)
            r = hipMemcpyAsync(dst.C, src.C, sizeof(float) * 9 * s->cap, hipMemcpyDeviceToDevice, s->stream);
        return r;
    };
    const bool sorted = s->sorted;
    s->cur = 1 - cur;
    for (int v = 0; v < 2 && e == hipSuccess; ++v) {
        s->tune.g2p = v;
        e = time_median(s, reps, &t.g2p_ms[v], copy, [&] { return launch_g2p_move(s, 1.0 / 480.0, true, true, false); });
    }
    s->cur = cur;
    s->sorted = sorted;
    if (e != hipSuccess) {
        s->tune = keep;
        return e;
    }
    s->tune.p2g = FE_P2G_TILES;
    for (int v = 1; v < kP2GVariants; ++v)
        if (t.p2g_ms[v] < t.p2g_ms[s->tune.p2g]) s->tune.p2g = v;
    s->tune.g2p = t.g2p_ms[FE_G2P_PAIRS] < t.g2p_ms[FE_G2P_QUADS] ? FE_G2P_PAIRS : FE_G2P_QUADS;
    // leave the grid holding the chosen P2G of the current particles
    if ((e = launch_p2g(s)) != hipSuccess) return e;
    if (times) *times = t;
    return hipStreamSynchronize(s->stream);
}

}  // namespace fe
