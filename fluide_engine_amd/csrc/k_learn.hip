// k_learn.hip -- the pressure warm-start network on the device: forward,
// backward and Adam, BF16 operands on MFMA (v_mfma_f32_16x16x16_bf16) with
// FP32 accumulation.  No counterpart in the reference.
//
// Convolutions are implicit GEMMs over channel-last BF16 activations.  One
// wavefront computes 16 consecutive voxels x all output channels: for each of
// the 27 taps and each 16-channel chunk of the input it loads its A fragment
// (4 BF16 of one voxel, 8 bytes per lane) and issues one MFMA per 16 output
// channels against the packed weights (8 bytes per lane, contiguous).  The
// MFMA's lane layout (checked on the device): lane l holds A[l % 16][4 (l / 16)
// + j], B[4 (l / 16) + j][l % 16] and returns C[4 (l / 16) + i][l % 16].
// The backward data pass is the same kernel on the flipped, transposed weights;
// the weight gradient is a GEMM over voxels (M = input channel, N = output
// channel, K = voxel), one partial per wavefront summed in a fixed order, so
// every gradient is reproducible run to run.
#include <algorithm>
#include <cmath>
#include <vector>

#include "fe_learn.hpp"
#include "fe_stencil.hpp"

namespace fe {

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

struct Lv {
    int nx, ny, nz;
    int64_t V;
};

inline Lv lv_of(const NetLevel& l) { return Lv{l.n[0], l.n[1], l.n[2], l.V}; }

__device__ inline uint16_t to_bf16(float f) {  // round to nearest, ties to even
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline void decode(const Lv& d, int64_t v, int& x, int& y, int& z) {
    const int64_t nxy = (int64_t)d.nx * d.ny;
    z = (int)(v / nxy);
    const int64_t r = v - (int64_t)z * nxy;
    y = (int)(r / d.nx);
    x = (int)(r - (int64_t)y * d.nx);
}

inline dim3 g1(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

// ---------------------------------------------------------------------------
// Convolution (forward, and backward data on the flipped weights)
// ---------------------------------------------------------------------------
template <int NT>
__global__ __launch_bounds__(256) void k_conv(Lv d, const uint16_t* __restrict__ X, int cin,
                                              const s4* __restrict__ Wp, const float* __restrict__ bias, int relu,
                                              uint16_t* __restrict__ Ybf, float* __restrict__ Yf) {
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int64_t tile = (int64_t)blockIdx.x * 4 + wave;
    if (tile * 16 >= d.V) return;  // uniform per wavefront
    const int row = l & 15, kq = l >> 4;
    const int64_t va = tile * 16 + row;
    const bool rowok = va < d.V;
    int x = 0, y = 0, z = 0;
    if (rowok) decode(d, va, x, y, z);
    const int KC = cin >> 4;
    f4 acc[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[nt] = f4{0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < 27; ++t) {
        const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
        const bool ok = rowok && (unsigned)(x + dx) < (unsigned)d.nx && (unsigned)(y + dy) < (unsigned)d.ny &&
                        (unsigned)(z + dz) < (unsigned)d.nz;
        const int64_t u = ok ? va + dx + (int64_t)dy * d.nx + (int64_t)dz * d.nx * d.ny : 0;
        const s4* xa = reinterpret_cast<const s4*>(X + u * cin) + kq;
        const s4* wb = Wp + (int64_t)t * KC * NT * 64 + l;
        for (int kc = 0; kc < KC; ++kc) {
            s4 a = s4{0, 0, 0, 0};
            if (ok) a = xa[kc * 4];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, wb[(kc * NT + nt) * 64], acc[nt], 0, 0, 0);
        }
    }
    const int cout = NT * 16;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t v = tile * 16 + 4 * kq + i;
        if (v >= d.V) continue;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int co = nt * 16 + row;
            float val = acc[nt][i] + (bias ? bias[co] : 0.0f);
            if (relu) val = fmaxf(val, 0.0f);
            if (Ybf) Ybf[v * cout + co] = to_bf16(val);
            if (Yf) Yf[v * cout + co] = val;
        }
    }
}

// Packed BF16 weights: forward [t][kc][nt][lane][j] = W[t][16 kc + 4 (lane / 16) + j][16 nt + lane % 16];
// backward data (cin' = cout, cout' = cin): W'[t][a][b] = W[26 - t][b][a], packed the same way.
__global__ __launch_bounds__(256) void k_pack(int cin, int cout, const float* __restrict__ w,
                                              uint16_t* __restrict__ wp, uint16_t* __restrict__ wpt) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= 27LL * cin * cout) return;
    {
        const int KC = cin / 16, NT = cout / 16;
        int64_t r = i;
        const int j = (int)(r % 4);
        r /= 4;
        const int l = (int)(r % 64);
        r /= 64;
        const int nt = (int)(r % NT);
        r /= NT;
        const int kc = (int)(r % KC);
        const int t = (int)(r / KC);
        const int ci = kc * 16 + 4 * (l / 16) + j, co = nt * 16 + (l % 16);
        wp[i] = to_bf16(w[((int64_t)t * cin + ci) * cout + co]);
    }
    {
        const int KC = cout / 16, NT = cin / 16;
        int64_t r = i;
        const int j = (int)(r % 4);
        r /= 4;
        const int l = (int)(r % 64);
        r /= 64;
        const int nt = (int)(r % NT);
        r /= NT;
        const int kc = (int)(r % KC);
        const int t = (int)(r / KC);
        const int a = kc * 16 + 4 * (l / 16) + j;  // an output channel of W
        const int b = nt * 16 + (l % 16);          // an input channel of W
        wpt[i] = to_bf16(w[((int64_t)(26 - t) * cin + b) * cout + a]);
    }
}

// ---------------------------------------------------------------------------
// Weight gradient: dW[t][ci][co] = sum_v X[v + off(t)][ci] * G[v][co]
// ---------------------------------------------------------------------------
template <int KC, int NT>
__global__ __launch_bounds__(256) void k_wgrad(Lv d, const uint16_t* __restrict__ X, const uint16_t* __restrict__ G,
                                               float* __restrict__ part, int64_t vpw) {
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int t = blockIdx.y;
    const int64_t pw = (int64_t)blockIdx.x * 4 + wave;
    const int64_t v0 = pw * vpw, v1 = min(v0 + vpw, d.V);
    const int row = l & 15, kq = l >> 4;
    constexpr int cin = KC * 16, cout = NT * 16;
    const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
    const int64_t off = dx + (int64_t)dy * d.nx + (int64_t)dz * d.nx * d.ny;
    f4 acc[KC][NT];
#pragma unroll
    for (int kc = 0; kc < KC; ++kc)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[kc][nt] = f4{0.f, 0.f, 0.f, 0.f};
    for (int64_t vs = v0; vs < v1; vs += 16) {
        s4 a[KC], b[NT];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t v = vs + 4 * kq + j;
            const bool vin = v < v1;
            int x = 0, y = 0, z = 0;
            if (vin) decode(d, v, x, y, z);
            const bool ok = vin && (unsigned)(x + dx) < (unsigned)d.nx && (unsigned)(y + dy) < (unsigned)d.ny &&
                            (unsigned)(z + dz) < (unsigned)d.nz;
            const int64_t u = ok ? v + off : 0;
#pragma unroll
            for (int kc = 0; kc < KC; ++kc) a[kc][j] = ok ? (short)X[u * cin + kc * 16 + row] : (short)0;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) b[nt][j] = vin ? (short)G[v * cout + nt * 16 + row] : (short)0;
        }
#pragma unroll
        for (int kc = 0; kc < KC; ++kc)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                acc[kc][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[kc], b[nt], acc[kc][nt], 0, 0, 0);
    }
    float* P = part + (pw * 27 + t) * (int64_t)cin * cout;
#pragma unroll
    for (int kc = 0; kc < KC; ++kc)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int i = 0; i < 4; ++i) P[(kc * 16 + 4 * kq + i) * cout + nt * 16 + row] = acc[kc][nt][i];
}

// out[i] = sum_p part[p * size + i], p ascending
__global__ __launch_bounds__(256) void k_reduce_parts(int64_t nparts, int64_t size, const float* __restrict__ part,
                                                      float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= size) return;
    float s = 0.0f;
    for (int64_t p = 0; p < nparts; ++p) s += part[p * size + i];
    out[i] = s;
}

// per-block column sums of a BF16 [V][C] array (bias gradients)
__global__ __launch_bounds__(256) void k_colsum(int64_t V, int C, const uint16_t* __restrict__ G,
                                                float* __restrict__ part, int64_t chunk) {
    __shared__ float sm[256];
    const int R = 256 / C;
    const int tid = threadIdx.x;
    const int64_t b0 = (int64_t)blockIdx.x * chunk, b1 = min(b0 + chunk, V);
    float s = 0.0f;
    if (tid < R * C) {
        const int c = tid % C, r = tid / C;
        for (int64_t v = b0 + r; v < b1; v += R) s += from_bf16(G[v * C + c]);
    }
    sm[tid] = s;
    __syncthreads();
    if (tid < C) {
        float t = 0.0f;
        for (int r = 0; r < R; ++r) t += sm[r * C + tid];
        part[(int64_t)blockIdx.x * C + tid] = t;
    }
}

// ---------------------------------------------------------------------------
// Elementwise and resampling passes
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_features(int64_t V, const float* __restrict__ b,
                                                  const uint8_t* __restrict__ flags, double rms,
                                                  uint16_t* __restrict__ X0) {
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= V) return;
    uint16_t* o = X0 + v * kNetIn;
    o[0] = to_bf16((float)((double)b[v] / rms));
    o[1] = flags[v] == FE_FLUID ? (uint16_t)0x3F80 : (uint16_t)0;  // 1.0 in BF16
    o[2] = flags[v] == FE_AIR ? (uint16_t)0x3F80 : (uint16_t)0;
    for (int c = 3; c < kNetIn; ++c) o[c] = 0;
}

__global__ __launch_bounds__(256) void k_pool(Lv f, Lv c, int C, const uint16_t* __restrict__ in,
                                              uint16_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= c.V * C) return;
    const int64_t vc = i / C;
    const int ch = (int)(i - vc * C);
    int X, Y, Z;
    decode(c, vc, X, Y, Z);
    float s = 0.0f;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const int64_t v = (2 * X + dx) + (int64_t)f.nx * ((2 * Y + dy) + (int64_t)f.ny * (2 * Z + dz));
                s += from_bf16(in[v * C + ch]);
            }
    out[i] = to_bf16(s / 8.0f);
}

// out[v] = [coarse[parent(v)] (Cc channels), skip[v] (Cs channels)]
__global__ __launch_bounds__(256) void k_upcat(Lv f, Lv c, int Cc, const uint16_t* __restrict__ coarse, int Cs,
                                               const uint16_t* __restrict__ skip, uint16_t* __restrict__ out) {
    const int Ct = Cc + Cs;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= f.V * Ct) return;
    const int64_t v = i / Ct;
    const int ch = (int)(i - v * Ct);
    if (ch < Cc) {
        int x, y, z;
        decode(f, v, x, y, z);
        const int64_t p = (x / 2) + (int64_t)c.nx * ((y / 2) + (int64_t)c.ny * (z / 2));
        out[i] = coarse[p * Cc + ch];
    } else {
        out[i] = skip[v * Cs + ch - Cc];
    }
}

__global__ __launch_bounds__(256) void k_head(int64_t V, const float* __restrict__ D0, const float* __restrict__ hw,
                                              const float* __restrict__ hb, float* __restrict__ out) {
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= V) return;
    float s = 0.0f;
    for (int c = 0; c < 16; ++c) s += D0[v * 16 + c] * hw[c];
    out[v] = s + hb[0];
}

// dout = 2 (out - target) / nf on FLUID cells; per-block FP64 sums of the squares
__global__ __launch_bounds__(256) void k_loss(int64_t V, const float* __restrict__ out,
                                              const float* __restrict__ target, const uint8_t* __restrict__ flags,
                                              float nf, float* __restrict__ dout, double* __restrict__ part) {
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double s = 0.0;
    if (v < V) {
        const float diff = flags[v] == FE_FLUID ? out[v] - target[v] : 0.0f;
        dout[v] = 2.0f * diff / nf;
        s = (double)diff * diff;
    }
    double sv[1] = {s};
    write_partials<1>(sv, part);
}

// gD0 = dout * hw; per-block FP64 sums of D0 * dout (16) and dout (1)
__global__ __launch_bounds__(256) void k_head_bwd(int64_t V, const float* __restrict__ D0,
                                                  const float* __restrict__ dout, const float* __restrict__ hw,
                                                  float* __restrict__ gD0, double* __restrict__ part) {
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double s[17];
    for (int q = 0; q < 17; ++q) s[q] = 0.0;
    if (v < V) {
        const float g = dout[v];
        for (int c = 0; c < 16; ++c) {
            gD0[v * 16 + c] = g * hw[c];
            s[c] = (double)D0[v * 16 + c] * g;
        }
        s[16] = g;
    }
    write_partials<17>(s, part);
}

template <int NV>
__global__ __launch_bounds__(256) void k_sum_to(const double* __restrict__ part, int nparts, float* __restrict__ out,
                                                double* __restrict__ out64, double scale) {
    double tot[NV];
    sum_partials<NV>(part, nparts, tot);
    if (threadIdx.x == 0)
        for (int q = 0; q < NV; ++q) {
            if (out) out[q] = (float)(tot[q] * scale);
            if (out64) out64[q] = tot[q] * scale;
        }
}

// Energy loss: with q = A p (p zero outside FLUID), per-block FP64 sums of
// b.p and p.q over FLUID cells; q kept for the gradient.
__global__ __launch_bounds__(256) void k_energy_stats(Lv d, const uint16_t* __restrict__ code, double ihx, double ihy,
                                                      double ihz, const float* __restrict__ p,
                                                      const float* __restrict__ b, float* __restrict__ aq,
                                                      double* __restrict__ part) {
#pragma clang fp contract(off)
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double s[2] = {0.0, 0.0};
    if (c < d.V) {
        const uint16_t cd = code[c];
        double q = 0.0;
        if (cd & kFluidBit) {
            const double ih2[3] = {ihx, ihy, ihz};
            const Stride st = stride_of(d.nx, d.ny, d.nz);
            auto val = [&](int64_t k) { return (code[k] & kFluidBit) ? (double)p[k] : 0.0; };
            const double pc = (double)p[c];
            q = code_apply64(cd, c, st, ih2, pc, val);
            s[0] = (double)b[c] * pc;
            s[1] = pc * q;
        }
        aq[c] = (float)q;
    }
    write_partials<2>(s, part);
}

// loss = 1 - s^2 / (qq E); scal[0] = loss, scal[1] = 2 s / (qq E), scal[2] = s / qq
__global__ __launch_bounds__(256) void k_energy_finish(const double* __restrict__ part, int nparts, double energy,
                                                       float* __restrict__ loss_slot, double* __restrict__ scal) {
    double tot[2];
    sum_partials<2>(part, nparts, tot);
    if (threadIdx.x == 0) {
        const double sv = tot[0], qq = tot[1];
        const bool ok = qq > 0 && energy > 0;
        const double loss = ok ? 1.0 - sv * sv / (qq * energy) : 1.0;
        scal[0] = loss;
        scal[1] = ok ? 2.0 * sv / (qq * energy) : 0.0;
        scal[2] = ok ? sv / qq : 0.0;
        if (loss_slot) *loss_slot = (float)loss;
    }
}

// dout = 2 s / (qq E) (s / qq A p - b) on FLUID cells (FP64, stored FP32)
__global__ __launch_bounds__(256) void k_energy_grad(int64_t V, const uint16_t* __restrict__ code,
                                                     const float* __restrict__ aq, const float* __restrict__ b,
                                                     const double* __restrict__ scal, float* __restrict__ dout) {
#pragma clang fp contract(off)
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= V) return;
    dout[c] = (code[c] & kFluidBit) ? (float)(scal[1] * (scal[2] * (double)aq[c] - (double)b[c])) : 0.0f;
}

__global__ __launch_bounds__(256) void k_codes(Lv d, const uint8_t* __restrict__ flags, uint16_t* __restrict__ code) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= d.V) return;
    int x, y, z;
    decode(d, c, x, y, z);
    code[c] = row_code(flags, c, (uint32_t)x, (uint32_t)y, (uint32_t)z, (uint32_t)d.nx, (uint32_t)d.ny,
                       (uint32_t)d.nz);
}

// dpre = bf16(y > 0 ? g : 0)
__global__ __launch_bounds__(256) void k_mask(int64_t n, const float* __restrict__ g, const uint16_t* __restrict__ y,
                                              uint16_t* __restrict__ dpre) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint16_t h = y[i];
    dpre[i] = (h != 0 && !(h & 0x8000)) ? to_bf16(g[i]) : (uint16_t)0;
}

// gCoarse[p][c] = sum over the 8 children of gC[child][c], c < Cc
__global__ __launch_bounds__(256) void k_up_bwd(Lv f, Lv c, int Cc, int Ct, const float* __restrict__ gC,
                                                float* __restrict__ gCoarse) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= c.V * Cc) return;
    const int64_t p = i / Cc;
    const int ch = (int)(i - p * Cc);
    int X, Y, Z;
    decode(c, p, X, Y, Z);
    float s = 0.0f;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const int64_t v = (2 * X + dx) + (int64_t)f.nx * ((2 * Y + dy) + (int64_t)f.ny * (2 * Z + dz));
                s += gC[v * Ct + ch];
            }
    gCoarse[i] = s;
}

// gSkip[v][c] = gC[v][Cc + c]
__global__ __launch_bounds__(256) void k_skip_bwd(int64_t V, int Cc, int Cs, const float* __restrict__ gC,
                                                  float* __restrict__ gSkip) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= V * Cs) return;
    const int64_t v = i / Cs;
    gSkip[i] = gC[v * (Cc + Cs) + Cc + (i - v * Cs)];
}

// gFine[v][c] += gCoarse[parent(v)][c] / 8
__global__ __launch_bounds__(256) void k_pool_bwd(Lv f, Lv c, int C, const float* __restrict__ gCoarse,
                                                  float* __restrict__ gFine) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= f.V * C) return;
    const int64_t v = i / C;
    const int ch = (int)(i - v * C);
    int x, y, z;
    decode(f, v, x, y, z);
    const int64_t p = (x / 2) + (int64_t)c.nx * ((y / 2) + (int64_t)c.ny * (z / 2));
    gFine[i] += gCoarse[p * C + ch] / 8.0f;
}

__global__ __launch_bounds__(256) void k_adam(int64_t n, float* __restrict__ w, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, float lr, float bc1,
                                              float bc2) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// FP64 sums of b^2 over all cells, and of p^2, the count and b.p over FLUID cells
__global__ __launch_bounds__(256) void k_sample_stats(int64_t V, const float* __restrict__ b,
                                                      const double* __restrict__ p, const uint8_t* __restrict__ flags,
                                                      double* __restrict__ part) {
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    if (v < V) {
        s[0] = (double)b[v] * b[v];
        if (flags[v] == FE_FLUID) {
            s[1] = p ? p[v] * p[v] : 0.0;
            s[2] = 1.0;
            s[3] = p ? (double)b[v] * p[v] : 0.0;
        }
    }
    write_partials<4>(s, part);
}

__global__ __launch_bounds__(256) void k_target(int64_t V, const double* __restrict__ p,
                                                const uint8_t* __restrict__ flags, double rms,
                                                float* __restrict__ t) {
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= V) return;
    t[v] = flags[v] == FE_FLUID ? (float)(p[v] / rms) : 0.0f;
}

// Scale of the warm start: with q = A phat (phat zero outside FLUID), per-block
// FP64 sums of b.q, q.q (residual-optimal alpha = b.q / q.q) and b.phat,
// phat.q (energy-optimal alpha = b.phat / phat.A.phat, the A-norm error's
// minimiser, which is what CG reduces).
__global__ __launch_bounds__(256) void k_ls(Lv d, const uint16_t* __restrict__ code, double ihx, double ihy,
                                            double ihz, const float* __restrict__ phat, const float* __restrict__ b,
                                            double* __restrict__ part) {
#pragma clang fp contract(off)
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    if (c < d.V) {
        const uint16_t cd = code[c];
        if (cd & kFluidBit) {
            const double ih2[3] = {ihx, ihy, ihz};
            const Stride st = stride_of(d.nx, d.ny, d.nz);
            auto val = [&](int64_t q) { return (code[q] & kFluidBit) ? (double)phat[q] : 0.0; };
            const double ph = (double)phat[c];
            const double q = code_apply64(cd, c, st, ih2, ph, val);
            s[0] = (double)b[c] * q;
            s[1] = q * q;
            s[2] = (double)b[c] * ph;
            s[3] = ph * q;
        }
    }
    write_partials<4>(s, part);
}

__global__ __launch_bounds__(256) void k_set_guess(int64_t V, const uint16_t* __restrict__ code,
                                                   const float* __restrict__ phat, double alpha,
                                                   double* __restrict__ p) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= V) return;
    p[c] = (code[c] & kFluidBit) ? alpha * (double)phat[c] : 0.0;
}

__global__ __launch_bounds__(256) void k_zero_f(int64_t n, float* __restrict__ a) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) a[i] = 0.0f;
}

// ---------------------------------------------------------------------------
// Host drivers
// ---------------------------------------------------------------------------
hipError_t launch_conv(hipStream_t st, const Lv& d, const uint16_t* X, int cin, int cout, const uint16_t* wp,
                       const float* bias, int relu, uint16_t* ybf, float* yf) {
    const int64_t tiles = (d.V + 15) / 16;
    const dim3 grid((unsigned)((tiles + 3) / 4));
    const s4* W = reinterpret_cast<const s4*>(wp);
    switch (cout / 16) {
        case 1: hipLaunchKernelGGL(k_conv<1>, grid, dim3(256), 0, st, d, X, cin, W, bias, relu, ybf, yf); break;
        case 2: hipLaunchKernelGGL(k_conv<2>, grid, dim3(256), 0, st, d, X, cin, W, bias, relu, ybf, yf); break;
        case 3: hipLaunchKernelGGL(k_conv<3>, grid, dim3(256), 0, st, d, X, cin, W, bias, relu, ybf, yf); break;
        case 4: hipLaunchKernelGGL(k_conv<4>, grid, dim3(256), 0, st, d, X, cin, W, bias, relu, ybf, yf); break;
        case 5: hipLaunchKernelGGL(k_conv<5>, grid, dim3(256), 0, st, d, X, cin, W, bias, relu, ybf, yf); break;
        case 7: hipLaunchKernelGGL(k_conv<7>, grid, dim3(256), 0, st, d, X, cin, W, bias, relu, ybf, yf); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int64_t wgrad_vpw(int64_t V) {
    int64_t vpw = (V + 255) / 256;
    vpw = std::max<int64_t>(16, (vpw + 15) / 16 * 16);
    return vpw;
}
int64_t wgrad_parts(int64_t V) {
    const int64_t waves = (V + wgrad_vpw(V) - 1) / wgrad_vpw(V);
    return (waves + 3) / 4 * 4;
}

hipError_t launch_wgrad(hipStream_t st, const Lv& d, const uint16_t* X, int cin, const uint16_t* G, int cout,
                        float* part) {
    const int64_t vpw = wgrad_vpw(d.V);
    const dim3 grid((unsigned)(wgrad_parts(d.V) / 4), 27);
    const int key = (cin / 16) * 10 + cout / 16;
    switch (key) {
        case 11: hipLaunchKernelGGL((k_wgrad<1, 1>), grid, dim3(256), 0, st, d, X, G, part, vpw); break;
        case 12: hipLaunchKernelGGL((k_wgrad<1, 2>), grid, dim3(256), 0, st, d, X, G, part, vpw); break;
        case 23: hipLaunchKernelGGL((k_wgrad<2, 3>), grid, dim3(256), 0, st, d, X, G, part, vpw); break;
        case 34: hipLaunchKernelGGL((k_wgrad<3, 4>), grid, dim3(256), 0, st, d, X, G, part, vpw); break;
        case 73: hipLaunchKernelGGL((k_wgrad<7, 3>), grid, dim3(256), 0, st, d, X, G, part, vpw); break;
        case 52: hipLaunchKernelGGL((k_wgrad<5, 2>), grid, dim3(256), 0, st, d, X, G, part, vpw); break;
        case 31: hipLaunchKernelGGL((k_wgrad<3, 1>), grid, dim3(256), 0, st, d, X, G, part, vpw); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int red_blocks(int64_t V) { return (int)std::min<int64_t>((V + 255) / 256, 1 << 30); }

template <typename T>
hipError_t alloc(T** p, int64_t count) {
    *p = nullptr;
    return hipMalloc((void**)p, sizeof(T) * (size_t)std::max<int64_t>(count, 1));
}

template <typename T>
void release(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

void free_buffers(fe_net* n) {
    release(n->X0); release(n->E0); release(n->P1); release(n->E1); release(n->P2); release(n->E2);
    release(n->P3); release(n->E3); release(n->C2); release(n->D2); release(n->C1); release(n->D1);
    release(n->C0); release(n->D0); release(n->dpre); release(n->out); release(n->dout); release(n->aq);
    release(n->D0f);
    release(n->gD0); release(n->gC0); release(n->gE0); release(n->gD1); release(n->gC1); release(n->gE1);
    release(n->gP1); release(n->gD2); release(n->gC2); release(n->gE2); release(n->gP2); release(n->gE3);
    release(n->gP3); release(n->wpart); release(n->red);
    n->wpart_cap = n->red_cap = 0;
    for (auto& l : n->lv) l = NetLevel{};
}

// layer_back: dpre = mask(gY, Y); dW, db; gX = conv_bwd_data(dpre) when asked
hipError_t layer_back(fe_net* n, int li, const Lv& d, const uint16_t* X, const uint16_t* Y, const float* gY,
                      float* gX) {
    NetLayer& L = n->L[li];
    hipStream_t st = n->stream;
    hipError_t e;
    const int64_t ny = d.V * L.cout;
    hipLaunchKernelGGL(k_mask, g1(ny), dim3(256), 0, st, ny, gY, Y, n->dpre);
    if ((e = launch_wgrad(st, d, X, L.cin, n->dpre, L.cout, n->wpart)) != hipSuccess) return e;
    const int64_t wsz = 27LL * L.cin * L.cout;
    hipLaunchKernelGGL(k_reduce_parts, g1(wsz), dim3(256), 0, st, wgrad_parts(d.V), wsz, (const float*)n->wpart, L.gw);
    const int nb = (int)std::min<int64_t>(256, (d.V + 255) / 256);
    const int64_t chunk = (d.V + nb - 1) / nb;
    hipLaunchKernelGGL(k_colsum, dim3(nb), dim3(256), 0, st, d.V, L.cout, (const uint16_t*)n->dpre, n->wpart, chunk);
    hipLaunchKernelGGL(k_reduce_parts, g1(L.cout), dim3(256), 0, st, (int64_t)nb, (int64_t)L.cout,
                       (const float*)n->wpart, L.gb);
    if (gX) return launch_conv(st, d, n->dpre, L.cout, L.cin, L.wpt, nullptr, 0, nullptr, gX);
    return hipGetLastError();
}

hipError_t pack_all(fe_net* n) {
    for (auto& L : n->L) {
        const int64_t sz = 27LL * L.cin * L.cout;
        hipLaunchKernelGGL(k_pack, g1(sz), dim3(256), 0, n->stream, L.cin, L.cout, (const float*)L.w, L.wp, L.wpt);
    }
    return hipGetLastError();
}

uint64_t splitmix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

int64_t net_param_count() {
    int64_t c = 0;
    for (int i = 0; i < kNetLayers; ++i) c += 27LL * kNetLayerIn[i] * kNetLayerOut[i] + kNetLayerOut[i];
    return c + 17;
}

hipError_t net_alloc(fe_net* n, uint64_t seed) {
    hipError_t e;
    if ((e = hipStreamCreate(&n->stream)) != hipSuccess) return e;
    uint64_t state = seed;
    auto normal = [&]() {  // Box-Muller on SplitMix64 uniforms
        const double u1 = ((splitmix(state++) >> 11) + 1.0) * 0x1.0p-53, u2 = (splitmix(state++) >> 11) * 0x1.0p-53;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    for (int i = 0; i < kNetLayers; ++i) {
        NetLayer& L = n->L[i];
        L.cin = kNetLayerIn[i];
        L.cout = kNetLayerOut[i];
        const int64_t wsz = 27LL * L.cin * L.cout;
        for (float** p : {&L.w, &L.gw, &L.mw, &L.vw})
            if ((e = alloc(p, wsz)) != hipSuccess) return e;
        for (float** p : {&L.b, &L.gb, &L.mb, &L.vb})
            if ((e = alloc(p, L.cout)) != hipSuccess) return e;
        if ((e = alloc(&L.wp, wsz)) != hipSuccess || (e = alloc(&L.wpt, wsz)) != hipSuccess) return e;
        std::vector<float> w((size_t)wsz), z((size_t)wsz, 0.0f), zb((size_t)L.cout, 0.0f);
        const double sd = std::sqrt(2.0 / (27.0 * L.cin));  // He initialisation
        for (auto& x : w) x = (float)(normal() * sd);
        if ((e = hipMemcpy(L.w, w.data(), sizeof(float) * wsz, hipMemcpyHostToDevice)) != hipSuccess) return e;
        for (float* p : {L.mw, L.vw})
            if ((e = hipMemcpy(p, z.data(), sizeof(float) * wsz, hipMemcpyHostToDevice)) != hipSuccess) return e;
        for (float* p : {L.b, L.mb, L.vb})
            if ((e = hipMemcpy(p, zb.data(), sizeof(float) * L.cout, hipMemcpyHostToDevice)) != hipSuccess) return e;
    }
    for (float** p : {&n->head, &n->ghead, &n->mhead, &n->vhead})
        if ((e = alloc(p, 17)) != hipSuccess) return e;
    if ((e = alloc(&n->scal, 8)) != hipSuccess) return e;
    std::vector<float> hd(17, 0.0f), z17(17, 0.0f);
    for (int c = 0; c < 16; ++c) hd[c] = (float)(normal() * 0.25);
    if ((e = hipMemcpy(n->head, hd.data(), sizeof(float) * 17, hipMemcpyHostToDevice)) != hipSuccess) return e;
    for (float* p : {n->mhead, n->vhead})
        if ((e = hipMemcpy(p, z17.data(), sizeof(float) * 17, hipMemcpyHostToDevice)) != hipSuccess) return e;
    if ((e = pack_all(n)) != hipSuccess) return e;
    return hipStreamSynchronize(n->stream);
}

void net_free(fe_net* n) {
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    free_buffers(n);
    for (auto& L : n->L) {
        release(L.w); release(L.b); release(L.gw); release(L.gb); release(L.mw); release(L.vw);
        release(L.mb); release(L.vb); release(L.wp); release(L.wpt);
    }
    release(n->head); release(n->ghead); release(n->mhead); release(n->vhead);
    release(n->scal); release(n->losses);
    for (auto& s : n->samples) {
        release(s.b); release(s.flags); release(s.target); release(s.code);
    }
    n->samples.clear();
    if (n->stream) (void)hipStreamDestroy(n->stream);
    n->stream = nullptr;
}

hipError_t net_get_params(fe_net* n, float* host) {
    hipError_t e;
    for (auto& L : n->L) {
        const int64_t wsz = 27LL * L.cin * L.cout;
        if ((e = hipMemcpy(host, L.w, sizeof(float) * wsz, hipMemcpyDeviceToHost)) != hipSuccess) return e;
        host += wsz;
        if ((e = hipMemcpy(host, L.b, sizeof(float) * L.cout, hipMemcpyDeviceToHost)) != hipSuccess) return e;
        host += L.cout;
    }
    return hipMemcpy(host, n->head, sizeof(float) * 17, hipMemcpyDeviceToHost);
}

hipError_t net_set_params(fe_net* n, const float* host) {
    hipError_t e;
    for (auto& L : n->L) {
        const int64_t wsz = 27LL * L.cin * L.cout;
        if ((e = hipMemcpy(L.w, host, sizeof(float) * wsz, hipMemcpyHostToDevice)) != hipSuccess) return e;
        host += wsz;
        if ((e = hipMemcpy(L.b, host, sizeof(float) * L.cout, hipMemcpyHostToDevice)) != hipSuccess) return e;
        host += L.cout;
    }
    if ((e = hipMemcpy(n->head, host, sizeof(float) * 17, hipMemcpyHostToDevice)) != hipSuccess) return e;
    if ((e = pack_all(n)) != hipSuccess) return e;
    return hipStreamSynchronize(n->stream);
}

hipError_t net_ensure(fe_net* n, const int nn[3]) {
    if (n->lv[0].n[0] == nn[0] && n->lv[0].n[1] == nn[1] && n->lv[0].n[2] == nn[2] && n->X0) return hipSuccess;
    (void)hipStreamSynchronize(n->stream);
    free_buffers(n);
    for (int l = 0; l < 4; ++l) {
        for (int a = 0; a < 3; ++a) n->lv[l].n[a] = nn[a] >> l;
        n->lv[l].V = (int64_t)n->lv[l].n[0] * n->lv[l].n[1] * n->lv[l].n[2];
    }
    const int64_t V0 = n->lv[0].V, V1 = n->lv[1].V, V2 = n->lv[2].V, V3 = n->lv[3].V;
    hipError_t e;
#define NA(p, cnt) \
    if ((e = alloc(&n->p, (cnt))) != hipSuccess) return e
    NA(X0, V0 * 16); NA(E0, V0 * 16); NA(P1, V1 * 16); NA(E1, V1 * 32); NA(P2, V2 * 32); NA(E2, V2 * 48);
    NA(P3, V3 * 48); NA(E3, V3 * 64); NA(C2, V2 * 112); NA(D2, V2 * 48); NA(C1, V1 * 80); NA(D1, V1 * 32);
    NA(C0, V0 * 48); NA(D0, V0 * 16);
    NA(dpre, std::max({V0 * 16, V1 * 32, V2 * 48, V3 * 64}));
    NA(out, V0); NA(dout, V0); NA(aq, V0); NA(D0f, V0 * 16);
    NA(gD0, V0 * 16); NA(gC0, V0 * 48); NA(gE0, V0 * 16); NA(gD1, V1 * 32); NA(gC1, V1 * 80); NA(gE1, V1 * 32);
    NA(gP1, V1 * 16); NA(gD2, V2 * 48); NA(gC2, V2 * 112); NA(gE2, V2 * 48); NA(gP2, V2 * 32); NA(gE3, V3 * 64);
    NA(gP3, V3 * 48);
    const int level_of[kNetLayers] = {0, 1, 2, 3, 2, 1, 0};
    int64_t wp = 256LL * 64;  // column-sum partials
    for (int i = 0; i < kNetLayers; ++i)
        wp = std::max<int64_t>(wp, wgrad_parts(n->lv[level_of[i]].V) * 27LL * kNetLayerIn[i] * kNetLayerOut[i]);
    NA(wpart, wp);
    n->wpart_cap = wp;
    NA(red, (int64_t)red_blocks(V0) * 17);  // at most 17 sums per block
    n->red_cap = (int64_t)red_blocks(V0) * 17;
#undef NA
    return hipSuccess;
}

hipError_t net_forward(fe_net* n, const float* b, const uint8_t* flags, double rms_b) {
    hipStream_t st = n->stream;
    const Lv l0 = lv_of(n->lv[0]), l1 = lv_of(n->lv[1]), l2 = lv_of(n->lv[2]), l3 = lv_of(n->lv[3]);
    hipError_t e;
    hipLaunchKernelGGL(k_features, g1(l0.V), dim3(256), 0, st, l0.V, b, flags, rms_b, n->X0);
    auto conv = [&](int li, const Lv& d, const uint16_t* X, uint16_t* Y) {
        const NetLayer& L = n->L[li];
        return launch_conv(st, d, X, L.cin, L.cout, L.wp, L.b, 1, Y, nullptr);
    };
    if ((e = conv(0, l0, n->X0, n->E0)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_pool, g1(l1.V * 16), dim3(256), 0, st, l0, l1, 16, (const uint16_t*)n->E0, n->P1);
    if ((e = conv(1, l1, n->P1, n->E1)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_pool, g1(l2.V * 32), dim3(256), 0, st, l1, l2, 32, (const uint16_t*)n->E1, n->P2);
    if ((e = conv(2, l2, n->P2, n->E2)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_pool, g1(l3.V * 48), dim3(256), 0, st, l2, l3, 48, (const uint16_t*)n->E2, n->P3);
    if ((e = conv(3, l3, n->P3, n->E3)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_upcat, g1(l2.V * 112), dim3(256), 0, st, l2, l3, 64, (const uint16_t*)n->E3, 48,
                       (const uint16_t*)n->E2, n->C2);
    if ((e = conv(4, l2, n->C2, n->D2)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_upcat, g1(l1.V * 80), dim3(256), 0, st, l1, l2, 48, (const uint16_t*)n->D2, 32,
                       (const uint16_t*)n->E1, n->C1);
    if ((e = conv(5, l1, n->C1, n->D1)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_upcat, g1(l0.V * 48), dim3(256), 0, st, l0, l1, 32, (const uint16_t*)n->D1, 16,
                       (const uint16_t*)n->E0, n->C0);
    {
        const NetLayer& L = n->L[6];
        if ((e = launch_conv(st, l0, n->C0, L.cin, L.cout, L.wp, L.b, 1, n->D0, n->D0f)) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_head, g1(l0.V), dim3(256), 0, st, l0.V, (const float*)n->D0f, (const float*)n->head,
                       (const float*)(n->head + 16), n->out);
    return hipGetLastError();
}

hipError_t net_backward(fe_net* n, const LossIn& in, float* loss_slot) {
    hipStream_t st = n->stream;
    const Lv l0 = lv_of(n->lv[0]), l1 = lv_of(n->lv[1]), l2 = lv_of(n->lv[2]), l3 = lv_of(n->lv[3]);
    hipError_t e;
    const int nb = red_blocks(l0.V);
    if (n->loss_mode == 0) {
        const double nf = (double)std::max<int64_t>(in.nfluid, 1);
        hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, st, l0.V, (const float*)n->out, in.target, in.flags,
                           (float)nf, n->dout, n->red);
        hipLaunchKernelGGL(k_sum_to<1>, dim3(1), dim3(256), 0, st, (const double*)n->red, nb, loss_slot, n->scal,
                           1.0 / nf);
    } else {
        hipLaunchKernelGGL(k_energy_stats, dim3(nb), dim3(256), 0, st, l0, in.code, in.ih2[0], in.ih2[1], in.ih2[2],
                           (const float*)n->out, in.b, n->aq, n->red);
        hipLaunchKernelGGL(k_energy_finish, dim3(1), dim3(256), 0, st, (const double*)n->red, nb, in.energy, loss_slot,
                           n->scal);
        hipLaunchKernelGGL(k_energy_grad, g1(l0.V), dim3(256), 0, st, l0.V, in.code, (const float*)n->aq, in.b,
                           (const double*)n->scal, n->dout);
    }
    hipLaunchKernelGGL(k_head_bwd, dim3(nb), dim3(256), 0, st, l0.V, (const float*)n->D0f, (const float*)n->dout,
                       (const float*)n->head, n->gD0, n->red);
    hipLaunchKernelGGL(k_sum_to<17>, dim3(1), dim3(256), 0, st, (const double*)n->red, nb, n->ghead, (double*)nullptr,
                       1.0);
    // decoder
    if ((e = layer_back(n, 6, l0, n->C0, n->D0, n->gD0, n->gC0)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_up_bwd, g1(l1.V * 32), dim3(256), 0, st, l0, l1, 32, 48, (const float*)n->gC0, n->gD1);
    hipLaunchKernelGGL(k_skip_bwd, g1(l0.V * 16), dim3(256), 0, st, l0.V, 32, 16, (const float*)n->gC0, n->gE0);
    if ((e = layer_back(n, 5, l1, n->C1, n->D1, n->gD1, n->gC1)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_up_bwd, g1(l2.V * 48), dim3(256), 0, st, l1, l2, 48, 80, (const float*)n->gC1, n->gD2);
    hipLaunchKernelGGL(k_skip_bwd, g1(l1.V * 32), dim3(256), 0, st, l1.V, 48, 32, (const float*)n->gC1, n->gE1);
    if ((e = layer_back(n, 4, l2, n->C2, n->D2, n->gD2, n->gC2)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_up_bwd, g1(l3.V * 64), dim3(256), 0, st, l2, l3, 64, 112, (const float*)n->gC2, n->gE3);
    hipLaunchKernelGGL(k_skip_bwd, g1(l2.V * 48), dim3(256), 0, st, l2.V, 64, 48, (const float*)n->gC2, n->gE2);
    // encoder
    if ((e = layer_back(n, 3, l3, n->P3, n->E3, n->gE3, n->gP3)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_pool_bwd, g1(l2.V * 48), dim3(256), 0, st, l2, l3, 48, (const float*)n->gP3, n->gE2);
    if ((e = layer_back(n, 2, l2, n->P2, n->E2, n->gE2, n->gP2)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_pool_bwd, g1(l1.V * 32), dim3(256), 0, st, l1, l2, 32, (const float*)n->gP2, n->gE1);
    if ((e = layer_back(n, 1, l1, n->P1, n->E1, n->gE1, n->gP1)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_pool_bwd, g1(l0.V * 16), dim3(256), 0, st, l0, l1, 16, (const float*)n->gP1, n->gE0);
    if ((e = layer_back(n, 0, l0, n->X0, n->E0, n->gE0, nullptr)) != hipSuccess) return e;
    return hipGetLastError();
}

hipError_t net_adam(fe_net* n, float lr) {
    ++n->adam_t;
    const float bc1 = (float)(1.0 - std::pow(0.9, (double)n->adam_t));
    const float bc2 = (float)(1.0 - std::pow(0.999, (double)n->adam_t));
    hipStream_t st = n->stream;
    for (auto& L : n->L) {
        const int64_t wsz = 27LL * L.cin * L.cout;
        hipLaunchKernelGGL(k_adam, g1(wsz), dim3(256), 0, st, wsz, L.w, (const float*)L.gw, L.mw, L.vw, lr, bc1, bc2);
        hipLaunchKernelGGL(k_adam, g1(L.cout), dim3(256), 0, st, (int64_t)L.cout, L.b, (const float*)L.gb, L.mb, L.vb,
                           lr, bc1, bc2);
    }
    hipLaunchKernelGGL(k_adam, dim3(1), dim3(256), 0, st, (int64_t)17, n->head, (const float*)n->ghead, n->mhead,
                       n->vhead, lr, bc1, bc2);
    return pack_all(n);
}

hipError_t net_add_sample(fe_net* n, fe_scene* s) {
    const int64_t V = cell_count(s->g);
    NetSample smp;
    for (int a = 0; a < 3; ++a) smp.n[a] = s->g.n[a];
    hipError_t e;
    if ((e = alloc(&smp.b, V)) != hipSuccess) return e;
    if ((e = alloc(&smp.flags, V)) != hipSuccess) {
        release(smp.b);
        return e;
    }
    if ((e = alloc(&smp.target, V)) != hipSuccess || (e = alloc(&smp.code, V)) != hipSuccess) {
        release(smp.b);
        release(smp.flags);
        release(smp.target);
        return e;
    }
    for (int a = 0; a < 3; ++a) smp.ih2[a] = s->g.ihd[a] * s->g.ihd[a];
    (void)hipStreamSynchronize(s->stream);
    const int nb = red_blocks(V);
    double* part = nullptr;
    if ((e = alloc(&part, (int64_t)nb * 4 + 4)) != hipSuccess) return e;
    double sums[4] = {0, 0, 0, 0};
    if ((e = hipMemcpyAsync(smp.b, s->b, sizeof(float) * V, hipMemcpyDeviceToDevice, n->stream)) == hipSuccess &&
        (e = hipMemcpyAsync(smp.flags, s->flags, V, hipMemcpyDeviceToDevice, n->stream)) == hipSuccess &&
        (e = hipMemcpyAsync(smp.code, s->code, sizeof(uint16_t) * V, hipMemcpyDeviceToDevice, n->stream)) ==
            hipSuccess) {
        hipLaunchKernelGGL(k_sample_stats, dim3(nb), dim3(256), 0, n->stream, V, (const float*)s->b,
                           (const double*)s->p, (const uint8_t*)s->flags, part);
        hipLaunchKernelGGL(k_sum_to<4>, dim3(1), dim3(256), 0, n->stream, (const double*)part, nb, (float*)nullptr,
                           part + (int64_t)nb * 4, 1.0);
        e = hipMemcpyAsync(sums, part + (int64_t)nb * 4, sizeof(sums), hipMemcpyDeviceToHost, n->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
    }
    if (e == hipSuccess) {
        smp.nfluid = (int64_t)sums[2];
        smp.energy = sums[3];
        smp.rms_b = std::sqrt(sums[0] / (double)V);
        const double rms_p = smp.nfluid > 0 ? std::sqrt(sums[1] / (double)smp.nfluid) : 0.0;
        hipLaunchKernelGGL(k_target, g1(V), dim3(256), 0, n->stream, V, (const double*)s->p,
                           (const uint8_t*)s->flags, rms_p > 0 ? rms_p : 1.0, smp.target);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
    }
    release(part);
    if (e != hipSuccess || !(smp.rms_b > 0) || smp.nfluid == 0 || !(smp.energy > 0)) {
        release(smp.b);
        release(smp.flags);
        release(smp.target);
        release(smp.code);
        return e != hipSuccess ? e : hipErrorInvalidValue;
    }
    n->samples.push_back(smp);
    return hipSuccess;
}

hipError_t net_train(fe_net* n, int steps, float lr, uint64_t seed) {
    if (n->samples.empty() || steps <= 0) return hipErrorInvalidValue;
    hipError_t e;
    if (n->losses_cap < steps) {
        release(n->losses);
        if ((e = alloc(&n->losses, steps)) != hipSuccess) return e;
        n->losses_cap = steps;
    }
    for (int it = 0; it < steps; ++it) {
        const NetSample& smp = n->samples[splitmix(seed + (uint64_t)it) % n->samples.size()];
        if ((e = net_ensure(n, smp.n)) != hipSuccess) return e;
        if ((e = net_forward(n, smp.b, smp.flags, smp.rms_b)) != hipSuccess) return e;
        LossIn in;
        in.target = smp.target;
        in.flags = smp.flags;
        in.nfluid = smp.nfluid;
        in.code = smp.code;
        for (int a = 0; a < 3; ++a) in.ih2[a] = smp.ih2[a];
        in.b = smp.b;
        in.energy = smp.energy;
        if ((e = net_backward(n, in, n->losses + it)) != hipSuccess) return e;
        if ((e = net_adam(n, lr)) != hipSuccess) return e;
    }
    return hipStreamSynchronize(n->stream);
}

hipError_t net_codes(fe_net* n, const int nn[3], const uint8_t* flags, uint16_t* code) {
    const Lv d{nn[0], nn[1], nn[2], (int64_t)nn[0] * nn[1] * nn[2]};
    hipLaunchKernelGGL(k_codes, g1(d.V), dim3(256), 0, n->stream, d, flags, code);
    return hipGetLastError();
}

hipError_t net_fluid_count(fe_net* n, const uint8_t* flags, int64_t V, int64_t* count) {
    const int nb = red_blocks(V);
    double* part = nullptr;
    hipError_t e;
    if ((e = alloc(&part, (int64_t)nb * 4 + 4)) != hipSuccess) return e;
    float* bz = nullptr;
    if ((e = alloc(&bz, V)) != hipSuccess) {
        release(part);
        return e;
    }
    hipLaunchKernelGGL(k_zero_f, g1(V), dim3(256), 0, n->stream, V, bz);
    hipLaunchKernelGGL(k_sample_stats, dim3(nb), dim3(256), 0, n->stream, V, (const float*)bz, (const double*)nullptr,
                       flags, part);
    hipLaunchKernelGGL(k_sum_to<4>, dim3(1), dim3(256), 0, n->stream, (const double*)part, nb, (float*)nullptr,
                       part + (int64_t)nb * 4, 1.0);
    double sums[4] = {0, 0, 0, 0};
    e = hipMemcpyAsync(sums, part + (int64_t)nb * 4, sizeof(sums), hipMemcpyDeviceToHost, n->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
    release(part);
    release(bz);
    *count = (int64_t)sums[2];
    return e;
}

hipError_t net_warm_start(fe_net* n, fe_scene* s) {
    std::lock_guard<std::mutex> lock(n->mu);
    const Geom& g = s->g;
    const int64_t V = cell_count(g);
    hipError_t e;
    if ((e = net_ensure(n, g.n)) != hipSuccess) return e;
    // the scene's stream has built b and the row codes; the network runs on its own stream after that
    if ((e = hipStreamSynchronize(s->stream)) != hipSuccess) return e;
    const int nb = red_blocks(V);
    hipLaunchKernelGGL(k_sample_stats, dim3(nb), dim3(256), 0, n->stream, V, (const float*)s->b,
                       (const double*)nullptr, (const uint8_t*)s->flags, n->red);
    hipLaunchKernelGGL(k_sum_to<4>, dim3(1), dim3(256), 0, n->stream, (const double*)n->red, nb, (float*)nullptr,
                       n->scal, 1.0);
    double sums[4] = {0, 0, 0, 0};
    if ((e = hipMemcpyAsync(sums, n->scal, sizeof(sums), hipMemcpyDeviceToHost, n->stream)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(n->stream)) != hipSuccess) return e;
    const double rms_b = std::sqrt(sums[0] / (double)V);
    double alpha = 0.0;
    if (rms_b > 0) {
        if ((e = net_forward(n, s->b, s->flags, rms_b)) != hipSuccess) return e;
        const Lv d{g.n[0], g.n[1], g.n[2], V};
        hipLaunchKernelGGL(k_ls, dim3(nb), dim3(256), 0, n->stream, d, (const uint16_t*)s->code,
                           g.ihd[0] * g.ihd[0], g.ihd[1] * g.ihd[1], g.ihd[2] * g.ihd[2], (const float*)n->out,
                           (const float*)s->b, n->red);
        hipLaunchKernelGGL(k_sum_to<4>, dim3(1), dim3(256), 0, n->stream, (const double*)n->red, nb, (float*)nullptr,
                           n->scal, 1.0);
        double sm[4] = {0, 0, 0, 0};
        if ((e = hipMemcpyAsync(sm, n->scal, sizeof(sm), hipMemcpyDeviceToHost, n->stream)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(n->stream)) != hipSuccess) return e;
        const double num = n->scale_mode == 0 ? sm[0] : sm[2], den = n->scale_mode == 0 ? sm[1] : sm[3];
        if (den > 0 && std::isfinite(num / den)) alpha = num / den;
    }
    s->net_alpha = alpha;
    hipLaunchKernelGGL(k_set_guess, g1(V), dim3(256), 0, n->stream, V, (const uint16_t*)s->code,
                       (const float*)n->out, alpha, s->p);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return hipStreamSynchronize(n->stream);
}

}  // namespace fe
