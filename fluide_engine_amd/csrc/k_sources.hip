// k_sources.hip -- scene sources on the device: PICSimulator::InitializeFromTriangleMesh
// + InitializeParticles (src/pic_simulator.cpp:33-39,521-548) and
// VolumeEmitter::Emitt (src/emmiter/volume_emmiter.cpp:58-85).
//
// Both append particlesPerBlock particles to each picked cell, at pos + h * u
// per axis (pos = the cell's GridIndexToPosition, u uniform), in ForEachIndex
// order (i outer, k inner).  The reference draws u from std::rand() / RAND_MAX
// on racing threads, which no run reproduces; u here is SplitMix64 (G. Steele,
// D. Lea, C. Flood, "Fast splittable pseudorandom number generators", OOPSLA
// 2014) keyed by (seed, call, particle, axis), (r >> 11) * 2^-53 in [0, 1).
// Positions are formed in FP64 and stored as FP32.
//
// One pass counts particlesPerBlock per picked cell in that order, an
// exclusive scan gives each cell its first particle, the host reads the total
// (and grows the particle storage when it would not fit: the reference's
// particle system grows without bound), and a second pass writes the
// particles behind the existing ones.
#include <algorithm>
#include <vector>

#include "fe_kernels.hpp"

namespace fe {

namespace {

__device__ inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ inline double uniform01(uint64_t seed, uint64_t call, uint64_t n, int axis) {
    const uint64_t r = splitmix64(splitmix64(splitmix64(seed) ^ call) ^ (3 * n + (uint64_t)axis));
    return (double)(r >> 11) * 0x1.0p-53;
}

// order index t (i outer, k inner) -> cell (i, j, k)
__device__ inline void order_cell(const Geom& g, int64_t t, int& i, int& j, int& k) {
    const int64_t jk = (int64_t)g.n[1] * g.n[2];
    i = (int)(t / jk);
    const int64_t r = t - (int64_t)i * jk;
    j = (int)(r / g.n[2]);
    k = (int)(r - (int64_t)j * g.n[2]);
}

__device__ inline int64_t cell_of(const Geom& g, int i, int j, int k) {
    return (int64_t)i + (int64_t)g.n[0] * ((int64_t)j + (int64_t)g.n[1] * k);
}

// Emitt's test at node (i, j, k): emitterSdf <= 0, fluidSdf >= 0 and
// colliderSdf > 0, all read at the same indices; cval is 0 before the first
// BuildCollider, else the collider SDF (MAX_DISTANCE without colliders).
__global__ __launch_bounds__(kBlock) void k_emit_count(Geom g, const double* __restrict__ esdf,
                                                       const double* __restrict__ fsdf,
                                                       const double* __restrict__ csdf, int built, int ppb,
                                                       int32_t* __restrict__ cnt) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= cell_count(g)) return;
    int i, j, k;
    order_cell(g, t, i, j, k);
    const int64_t c = cell_of(g, i, j, k);
    const double cval = !built ? 0.0 : (csdf ? csdf[c] : 1.7976931348623157E+30);
    cnt[t] = (esdf[c] <= 0 && fsdf[c] >= 0 && cval > 0) ? ppb : 0;
}

// InitializeParticles' test: the fluid SDF (cell centred) sampled at the
// cell's corner o + h*(i,j,k) is negative.
__global__ __launch_bounds__(kBlock) void k_init_count(Geom g, const double* __restrict__ fsdf, int ppb,
                                                       int32_t* __restrict__ cnt) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= cell_count(g)) return;
    int idx[3];
    order_cell(g, t, idx[0], idx[1], idx[2]);
    double p[3];
    for (int a = 0; a < 3; ++a) p[a] = g.od[a] + (double)idx[a] * g.hd[a];
    cnt[t] = sample_sdf64(g, fsdf, p) < 0 ? ppb : 0;
}

// Writes cell t's particles [off[t], off[t+1]) at np0 + n, pid np0 + n.
__global__ __launch_bounds__(kBlock) void k_seed_write(Geom g, NodeGrid3 E, const int32_t* __restrict__ off,
                                                       int64_t np0, int64_t cap, uint64_t seed, uint64_t call,
                                                       float vx, float vy, float vz, Particles P, int with_c) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= cell_count(g)) return;
    const int64_t n0 = off[t], n1 = off[t + 1];
    if (n1 == n0) return;
    int idx[3];
    order_cell(g, t, idx[0], idx[1], idx[2]);
    double pos[3];
    for (int a = 0; a < 3; ++a) pos[a] = E.o[a] + (double)idx[a] * E.h[a];  // GridIndexToPosition
    for (int64_t n = n0; n < n1; ++n) {
        const int64_t q = np0 + n;
        for (int a = 0; a < 3; ++a) P.x[a * cap + q] = (float)(E.h[a] * uniform01(seed, call, (uint64_t)n, a) + pos[a]);
        P.v[q] = vx;
        P.v[cap + q] = vy;
        P.v[2 * cap + q] = vz;
        if (with_c)
            for (int r = 0; r < 9; ++r) P.C[r * cap + q] = 0.0f;
        P.pid[q] = (int32_t)q;
    }
}

inline dim3 grid_n(int64_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

hipError_t ensure_scratch(fe_scene* s) {
    if (s->src_cnt) return hipSuccess;
    const int64_t nc = cell_count(s->g);
    hipError_t e;
    if ((e = hipMalloc((void**)&s->src_cnt, sizeof(int32_t) * (nc + 1))) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&s->src_off, sizeof(int32_t) * (nc + 1))) != hipSuccess) return e;
    return hipMalloc((void**)&s->src_sums, sizeof(int32_t) * (scan_tiles(nc) + 1));
}

// Scan of the counts in src_cnt, total to the host, storage grown to fit,
// write.  Returns hipErrorInvalidValue (and sets *over) past the int32 range
// of particle ids.
hipError_t seed_from_counts(fe_scene* s, const NodeGrid3& E, uint64_t seed, uint64_t call, const double vel[3],
                            bool* over) {
    const int64_t nc = cell_count(s->g);
    hipError_t e;
    if ((e = launch_scan_with(s->stream, s->src_cnt, s->src_off, nc, s->src_sums)) != hipSuccess) return e;
    int32_t total = 0;
    if ((e = hipMemcpyAsync(&total, s->src_off + nc, sizeof(int32_t), hipMemcpyDeviceToHost, s->stream)) !=
        hipSuccess)
        return e;
    if ((e = hipStreamSynchronize(s->stream)) != hipSuccess) return e;
    *over = false;
    if (total == 0) return hipSuccess;
    if (s->np + total > INT32_MAX) {
        *over = true;
        return hipErrorInvalidValue;
    }
    if (s->np + total > s->cap &&
        (e = reserve_particles(s, std::max<int64_t>(s->np + total, std::min<int64_t>(2 * s->cap, INT32_MAX)))) !=
            hipSuccess)
        return e;
    hipLaunchKernelGGL(k_seed_write, grid_n(nc), dim3(kBlock), 0, s->stream, s->g, E, (const int32_t*)s->src_off,
                       s->np, s->cap, seed, call, (float)vel[0], (float)vel[1], (float)vel[2], s->P[s->cur],
                       s->desc.method == FE_APIC ? 1 : 0);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    s->np += total;
    s->sorted = false;
    s->keys_ready = false;
    return hipSuccess;
}

}  // namespace

// Grows every per-particle array to `capacity`, keeping the current particles
// (positions, velocities, affine rows and ids of P[cur]); the derived arrays
// (binned positions, keys, order) are rebuilt by the next sort.
hipError_t reserve_particles(fe_scene* s, int64_t capacity) {
    if (capacity <= s->cap) return hipSuccess;
    const bool apic = s->desc.method == FE_APIC;
    const int64_t oc = s->cap, np = s->np;
    Particles NP[2];
    float4* pos4 = nullptr;
    uint32_t* key = nullptr;
    int32_t* order = nullptr;
    hipError_t e = hipSuccess;
    auto A = [&](auto** p, size_t count) {
        if (e == hipSuccess) e = hipMalloc((void**)p, sizeof(**p) * count);
    };
    for (int b = 0; b < 2; ++b) {
        A(&NP[b].x, 3 * capacity);
        A(&NP[b].v, 3 * capacity);
        A(&NP[b].pid, capacity);
        if (apic) A(&NP[b].C, 9 * capacity);
    }
    A(&pos4, capacity);
    A(&key, capacity);
    A(&order, capacity);
    auto release = [](Particles* P, float4* a, uint32_t* k, int32_t* o) {
        for (int q = 0; q < 2; ++q) {
            if (P[q].x) (void)hipFree(P[q].x);
            if (P[q].v) (void)hipFree(P[q].v);
            if (P[q].pid) (void)hipFree(P[q].pid);
            if (P[q].C) (void)hipFree(P[q].C);
        }
        if (a) (void)hipFree(a);
        if (k) (void)hipFree(k);
        if (o) (void)hipFree(o);
    };
    if (e != hipSuccess) {
        release(NP, pos4, key, order);
        return e;
    }
    const Particles& P = s->P[s->cur];
    Particles& Q = NP[s->cur];
    auto copy_soa = [&](float* dst, const float* src, int comps) {
        for (int c = 0; c < comps && e == hipSuccess && np > 0; ++c)
            e = hipMemcpyAsync(dst + c * capacity, src + c * oc, sizeof(float) * np, hipMemcpyDeviceToDevice,
                               s->stream);
    };
    copy_soa(Q.x, P.x, 3);
    copy_soa(Q.v, P.v, 3);
    if (apic) copy_soa(Q.C, P.C, 9);
    if (e == hipSuccess && np > 0)
        e = hipMemcpyAsync(Q.pid, P.pid, sizeof(int32_t) * np, hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        release(NP, pos4, key, order);
        return e;
    }
    release(s->P, s->pos4, s->key, s->order);
    s->P[0] = NP[0];
    s->P[1] = NP[1];
    s->pos4 = pos4;
    s->key = key;
    s->order = order;
    s->cap = capacity;
    s->sorted = false;
    s->keys_ready = false;
    return hipSuccess;
}

hipError_t launch_emit(fe_scene* s, bool* over) {
    *over = false;
    if (s->emitters.empty()) return hipSuccess;
    hipError_t e;
    if ((e = ensure_scratch(s)) != hipSuccess) return e;
    const int64_t nc = cell_count(s->g);
    for (Emitter& em : s->emitters) {
        hipLaunchKernelGGL(k_emit_count, grid_n(nc), dim3(kBlock), 0, s->stream, s->g, (const double*)em.sdf,
                           (const double*)s->sdf, (const double*)(s->has_collider ? s->csdf : nullptr),
                           s->substeps_run > 0 ? 1 : 0, em.ppb, s->src_cnt);
        if ((e = seed_from_counts(s, em.grid, em.seed, em.calls, em.vel, over)) != hipSuccess) return e;
        ++em.calls;
    }
    return hipSuccess;
}

hipError_t launch_initialize_from_sdf(fe_scene* s, int ppb, uint64_t seed, bool* over) {
    *over = false;
    hipError_t e;
    if ((e = ensure_scratch(s)) != hipSuccess) return e;
    const int64_t nc = cell_count(s->g);
    hipLaunchKernelGGL(k_init_count, grid_n(nc), dim3(kBlock), 0, s->stream, s->g, (const double*)s->sdf, ppb,
                       s->src_cnt);
    NodeGrid3 E{};
    for (int a = 0; a < 3; ++a) {
        E.n[a] = s->g.n[a];
        E.o[a] = s->g.od[a];
        E.h[a] = s->g.hd[a];
    }
    const double zero[3] = {0.0, 0.0, 0.0};
    return seed_from_counts(s, E, seed, 0, zero, over);
}

void sources_free(fe_scene* s) {
    for (Emitter& em : s->emitters)
        if (em.sdf) (void)hipFree(em.sdf);
    s->emitters.clear();
    auto F = [](int32_t*& p) {
        if (p) (void)hipFree(p);
        p = nullptr;
    };
    F(s->src_cnt);
    F(s->src_off);
    F(s->src_sums);
}

}  // namespace fe
