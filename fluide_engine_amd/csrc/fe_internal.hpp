// fe_internal.hpp -- shared definitions of the gfx950 backend: scene state in
// HBM, launch helpers and the device-side geometry routines every kernel uses.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fluide_hip.h"

namespace fe {

constexpr int kBlock = 256;        // threads per workgroup for element-wise kernels (4 waves of 64)
constexpr int kXcds = 8;           // MI355X: workgroups are dealt round-robin to 8 XCDs, one L2 each
constexpr int kMaxStages = 16;
constexpr int kP2GVariants = 3;       // FE_P2G_TILES, FE_P2G_FACES, FE_P2G_SCATTER

// ---------------------------------------------------------------------------
// Geometry of a MAC grid, in the form the kernels need.
// Face data origins follow FaceCenteredGrid3D::CalculateDataOrigins
// (src/grid_systems/face_centered_grid3d.cpp:324-332): u at origin + h*(0,.5,.5),
// v at origin + h*(.5,0,.5), w at origin + h*(.5,.5,0).
// ---------------------------------------------------------------------------
struct Geom {
    int n[3];
    int method;
    float o[3], h[3], ih[3], L[3];
    float co[3][3];        // co[c][a]: component c data origin on axis a (FP32)
    double od[3], hd[3], ihd[3], Ld[3];
    double cod[3][3];      // same in FP64 (the reference's values)
};

// Face-array dims of component c: the normal axis has n+1 entries.
__host__ __device__ inline int face_dim(const Geom& g, int c, int a) { return g.n[a] + (a == c ? 1 : 0); }
__host__ __device__ inline int64_t face_count(const Geom& g, int c) {
    return (int64_t)face_dim(g, c, 0) * face_dim(g, c, 1) * face_dim(g, c, 2);
}
__host__ __device__ inline int64_t cell_count(const Geom& g) { return (int64_t)g.n[0] * g.n[1] * g.n[2]; }

// Stencil size the reference passes to GetCooridnatesAndWeights for component c
// on axis a: FLIP passes the component's own dims (src/flip_simulator.cpp:58,66,74),
// PIC and APIC pass the cell dims (src/pic_simulator.cpp:389,397,405;
// src/apic_simulator.cpp:58,71,84).
__host__ __device__ inline int stencil_size(const Geom& g, int c, int a) {
    return g.method == FE_FLIP ? face_dim(g, c, a) : g.n[a];
}

// ---------------------------------------------------------------------------
// floor((p - o)/h) and its fraction, bit-exact with the reference's FP64
// expression.  The FP32 form is exact for power-of-two h with a dyadic origin;
// in general it is within 2^-22*((|p|+|o|)/h + |x|) of the FP64 value, so when
// the FP32 fraction lies inside a guard band 2^-19*(...) of an integer the
// index is recomputed in FP64 exactly as (x - origin) / gridSpacing
// (src/common/math_utils.hpp:85-89).
// ---------------------------------------------------------------------------
__device__ inline void exact_floor64(float p, double od, double hd, int& s, float& f) {
#pragma clang fp contract(off)
    double xd = ((double)p - od) / hd;
    xd = fmin(fmax(xd, -1.0e9), 1.0e9);
    const double sd = floor(xd);
    s = (int)sd;
    f = (float)(xd - sd);
}

__device__ inline void axis_floor(float p, float o, float ih, double od, double hd, int& s, float& f) {
    const float x = (p - o) * ih;
    const float fs = floorf(x);
    const float fr = x - fs;
    const float band = 1.9073486e-6f * ((fabsf(p) + fabsf(o)) * ih + fabsf(x) + 1.0f);  // 2^-19 * (...)
    if (fr >= band && fr <= 1.0f - band && fabsf(x) < 1.0e9f) {
        s = (int)fs;
        f = fr;
        return;
    }
    exact_floor64(p, od, hd, s, f);
}

// GetBarycentric(x, 0, size-1) clamp rules (src/common/math_utils.hpp:48-73).
__device__ inline void bary_clamp(int s, float fr, int size, int& i, float& f) {
    if (size <= 1) {
        i = 0;
        f = 0.0f;
    } else if (s < 0) {
        i = 0;
        f = 0.0f;
    } else if (s > size - 2) {
        i = size - 2;
        f = 1.0f;
    } else {
        i = s;
        f = fr;
    }
}

// Exact stencil base index + fraction of position p for component c, axis a.
__device__ inline void stencil_axis(const Geom& g, int c, int a, float p, int size, int& i, float& f) {
    int s;
    float fr;
    axis_floor(p, g.co[c][a], g.ih[a], g.cod[c][a], g.hd[a], s, fr);
    bary_clamp(s, fr, size, i, f);
}

// Cell containing p, clamped to the grid: the particle's sort bin.  Same
// arithmetic as the normal axis of the component stencils (data origin of the
// normal axis is the grid origin), so bin and stencil agree bit for bit.
__device__ inline int cell_axis(const Geom& g, int a, float p) {
    int s;
    float fr;
    axis_floor(p, g.o[a], g.ih[a], g.od[a], g.hd[a], s, fr);
    return s < 0 ? 0 : (s > g.n[a] - 1 ? g.n[a] - 1 : s);
}

// (i, j, k) of the linear index f of an x-fastest sx * sy * sz array.  Every
// array here holds fewer than 2^31 elements (a scene has at most 2^30 cells),
// so 32-bit division suffices (64-bit division is a long software sequence).
__device__ inline void split_index(int64_t f, int sx, int sy, int& i, int& j, int& k) {
    const uint32_t u = (uint32_t)f, ux = (uint32_t)sx, uy = (uint32_t)sy;
    const uint32_t t = u / ux;
    i = (int)(u - t * ux);
    j = (int)(t % uy);
    k = (int)(t / uy);
}

// GetBarycentric(x, 0, size-1) in FP64 (src/common/math_utils.hpp:48-73).
__device__ inline void bary64(double x, int size, int& i, double& f) {
#pragma clang fp contract(off)
    const double s = floor(x);
    i = (int)fmin(fmax(s, -1.0e9), 1.0e9);
    if (size <= 1) {
        i = 0;
        f = 0.0;
    } else if (i < 0) {
        i = 0;
        f = 0.0;
    } else if (i > size - 2) {
        i = size - 2;
        f = 1.0;
    } else {
        f = x - s;
    }
}

__device__ inline double lerp64(double a, double b, double f) {
#pragma clang fp contract(off)
    return (1 - f) * a + f * b;
}

// ScalarGrid3D::Sample (src/3d/scalar_grid3d.cpp:33-58) of the fluid SDF
// (cell centred, origin o + h/2) at p, in the reference's FP64 arithmetic.
__device__ inline double sample_sdf64(const Geom& g, const double* __restrict__ sdf, const double p[3]) {
#pragma clang fp contract(off)
    int ii[3];
    double ff[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double so = g.od[a] + g.hd[a] / 2.0;
        bary64((p[a] - so) / g.hd[a], g.n[a], ii[a], ff[a]);
    }
    const int ip1 = min(ii[0] + 1, g.n[0] - 1), jp1 = min(ii[1] + 1, g.n[1] - 1), kp1 = min(ii[2] + 1, g.n[2] - 1);
    auto S = [&](int a, int b, int cc) {
        return sdf[(int64_t)a + (int64_t)g.n[0] * ((int64_t)b + (int64_t)g.n[1] * cc)];
    };
    const int i = ii[0], j = ii[1], k = ii[2];
    return lerp64(lerp64(lerp64(S(i, j, k), S(ip1, j, k), ff[0]), lerp64(S(i, jp1, k), S(ip1, jp1, k), ff[0]), ff[1]),
                  lerp64(lerp64(S(i, j, kp1), S(ip1, j, kp1), ff[0]), lerp64(S(i, jp1, kp1), S(ip1, jp1, kp1), ff[0]), ff[1]),
                  ff[2]);
}

// ScalarGrid3D::Sample of the collider SDF (node centred at the domain origin,
// src/3d/collider_collection.cpp:5-8) at p, in FP64.
__device__ inline double sample_collider64(const Geom& g, const double* __restrict__ c, const double p[3]) {
#pragma clang fp contract(off)
    int ii[3];
    double ff[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) bary64((p[a] - g.od[a]) / g.hd[a], g.n[a], ii[a], ff[a]);
    const int ip1 = min(ii[0] + 1, g.n[0] - 1), jp1 = min(ii[1] + 1, g.n[1] - 1), kp1 = min(ii[2] + 1, g.n[2] - 1);
    auto S = [&](int a, int b, int cc) { return c[(int64_t)a + (int64_t)g.n[0] * ((int64_t)b + (int64_t)g.n[1] * cc)]; };
    const int i = ii[0], j = ii[1], k = ii[2];
    return lerp64(lerp64(lerp64(S(i, j, k), S(ip1, j, k), ff[0]), lerp64(S(i, jp1, k), S(ip1, jp1, k), ff[0]), ff[1]),
                  lerp64(lerp64(S(i, j, kp1), S(ip1, j, kp1), ff[0]), lerp64(S(i, jp1, kp1), S(ip1, jp1, kp1), ff[0]), ff[1]),
                  ff[2]);
}

// A node grid of its own (a TriangleMeshCollider's ScalarGrid3D): size,
// origin, spacing; values x fastest.
struct NodeGrid3 {
    int n[3];
    double o[3];
    double h[3];
};

// ScalarGrid3D::Sample on such a grid, in FP64.
__device__ inline double sample_grid64(const NodeGrid3& G, const double* __restrict__ a, const double p[3]) {
#pragma clang fp contract(off)
    int ii[3];
    double ff[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) bary64((p[q] - G.o[q]) / G.h[q], G.n[q], ii[q], ff[q]);
    const int ip1 = min(ii[0] + 1, G.n[0] - 1), jp1 = min(ii[1] + 1, G.n[1] - 1), kp1 = min(ii[2] + 1, G.n[2] - 1);
    auto S = [&](int x, int y, int z) { return a[(int64_t)x + (int64_t)G.n[0] * ((int64_t)y + (int64_t)G.n[1] * z)]; };
    const int i = ii[0], j = ii[1], k = ii[2];
    return lerp64(lerp64(lerp64(S(i, j, k), S(ip1, j, k), ff[0]), lerp64(S(i, jp1, k), S(ip1, jp1, k), ff[0]), ff[1]),
                  lerp64(lerp64(S(i, j, kp1), S(ip1, j, kp1), ff[0]), lerp64(S(i, jp1, kp1), S(ip1, jp1, kp1), ff[0]), ff[1]),
                  ff[2]);
}

// ScalarGrid3D::Gradient (src/3d/scalar_grid3d.cpp:67-90) of the collider SDF:
// the trilinear weights at p times clamped central differences at the nodes.
__device__ inline void collider_gradient64(const Geom& g, const double* __restrict__ c, const double p[3],
                                           double out[3]) {
#pragma clang fp contract(off)
    int ii[3];
    double ff[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) bary64((p[a] - g.od[a]) / g.hd[a], g.n[a], ii[a], ff[a]);
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    auto S = [&](int a, int b, int cc) { return c[(int64_t)a + (int64_t)n0 * ((int64_t)b + (int64_t)n1 * cc)]; };
    out[0] = out[1] = out[2] = 0.0;
    for (int q = 0; q < 8; ++q) {
        const int i = min(ii[0] + (q & 1), n0 - 1), j = min(ii[1] + ((q >> 1) & 1), n1 - 1),
                  k = min(ii[2] + ((q >> 2) & 1), n2 - 1);
        const double wx = (q & 1) ? ff[0] : 1.00 - ff[0], wy = ((q >> 1) & 1) ? ff[1] : 1.00 - ff[1],
                     wz = ((q >> 2) & 1) ? ff[2] : 1.00 - ff[2];
        const double w = wx * wy * wz;  // GetCooridnatesAndWeights' products
        const double d[3] = {S(i + 1 < n0 ? i + 1 : i, j, k) - S(i > 0 ? i - 1 : i, j, k),
                             S(i, j + 1 < n1 ? j + 1 : j, k) - S(i, j > 0 ? j - 1 : j, k),
                             S(i, j, k + 1 < n2 ? k + 1 : k) - S(i, j, k > 0 ? k - 1 : k)};
        for (int a = 0; a < 3; ++a) out[a] += (w * 0.5 * d[a]) / g.hd[a];
    }
}

// FP32 fast coordinates for interpolation (values, not parity-critical indices).
__device__ inline void fast_axis(float p, float o, float ih, int size, int& i, float& f) {
    const float x = (p - o) * ih;
    const float fs = floorf(x);
    int s = (int)fmaxf(fminf(fs, 1.0e9f), -1.0e9f);
    bary_clamp(s, x - fs, size, i, f);
}

__device__ inline float lerpf_ref(float a, float b, float f) { return (1.0f - f) * a + f * b; }

// Trilinear stencil of FaceCenteredGrid3D::SampleArray
// (src/grid_systems/face_centered_grid3d.cpp:339-371) on face array c, clamped
// with the array's own dims.  Every face dim is >= 2, so the clamp keeps
// i <= dim-2 and the x-neighbour i+1 always exists: each (i, i+1) pair is one
// 8-byte load, four loads per sample.
struct FaceStencil {
    int64_t r00, r10, r01, r11;  // row offsets (j,k), (j+1,k), (j,k+1), (j+1,k+1), plus i
    float fx, fy, fz;
};

__device__ inline FaceStencil face_stencil(const Geom& g, int c, float px, float py, float pz) {
    const int sx = face_dim(g, c, 0), sy = face_dim(g, c, 1), sz = face_dim(g, c, 2);
    int i, j, k;
    FaceStencil t;
    fast_axis(px, g.co[c][0], g.ih[0], sx, i, t.fx);
    fast_axis(py, g.co[c][1], g.ih[1], sy, j, t.fy);
    fast_axis(pz, g.co[c][2], g.ih[2], sz, k, t.fz);
    const int jp1 = min(j + 1, sy - 1), kp1 = min(k + 1, sz - 1);
    t.r00 = (int64_t)sx * ((int64_t)j + (int64_t)sy * k) + i;
    t.r10 = (int64_t)sx * ((int64_t)jp1 + (int64_t)sy * k) + i;
    t.r01 = (int64_t)sx * ((int64_t)j + (int64_t)sy * kp1) + i;
    t.r11 = (int64_t)sx * ((int64_t)jp1 + (int64_t)sy * kp1) + i;
    return t;
}

// The stencils of one point on the three face arrays share their axis
// coordinates: on axis a there are only the normal one (component a: origin
// o, size n + 1) and the transverse one (the other two components: origin
// o + h/2, size n), so a point takes six fast_axis evaluations instead of
// nine.  stencil_at(c) is face_stencil(c) of the same point, value for value,
// with 32-bit offsets (a face array holds fewer than 2^31 values).
struct PointAxes {
    int in[3], it[3];
    float fn[3], ft[3];
};

__device__ inline PointAxes point_axes(const Geom& g, float px, float py, float pz) {
    const float p[3] = {px, py, pz};
    PointAxes t;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        fast_axis(p[a], g.co[a][a], g.ih[a], g.n[a] + 1, t.in[a], t.fn[a]);
        fast_axis(p[a], g.co[(a + 1) % 3][a], g.ih[a], g.n[a], t.it[a], t.ft[a]);
    }
    return t;
}

struct FaceStencil32 {
    int r00, r10, r01, r11;  // as FaceStencil
    float fx, fy, fz;
};

__device__ inline FaceStencil32 stencil_at(const Geom& g, int c, const PointAxes& t) {
    const int sx = face_dim(g, c, 0), sy = face_dim(g, c, 1), sz = face_dim(g, c, 2);
    FaceStencil32 s;
    const int i = c == 0 ? t.in[0] : t.it[0], j = c == 1 ? t.in[1] : t.it[1], k = c == 2 ? t.in[2] : t.it[2];
    s.fx = c == 0 ? t.fn[0] : t.ft[0];
    s.fy = c == 1 ? t.fn[1] : t.ft[1];
    s.fz = c == 2 ? t.fn[2] : t.ft[2];
    const int jp1 = min(j + 1, sy - 1), kp1 = min(k + 1, sz - 1);
    s.r00 = sx * (j + sy * k) + i;
    s.r10 = sx * (jp1 + sy * k) + i;
    s.r01 = sx * (j + sy * kp1) + i;
    s.r11 = sx * (jp1 + sy * kp1) + i;
    return s;
}

__device__ inline float2 load_pair(const float* __restrict__ p) {
    float2 v;
    __builtin_memcpy(&v, p, sizeof(v));  // 4-byte aligned 8-byte load
    return v;
}

// The reference's lerp nesting: x first, then y, then z.
template <typename S>
__device__ inline float trilerp_pairs(float2 v00, float2 v10, float2 v01, float2 v11, const S& t) {
    return lerpf_ref(lerpf_ref(lerpf_ref(v00.x, v00.y, t.fx), lerpf_ref(v10.x, v10.y, t.fx), t.fy),
                     lerpf_ref(lerpf_ref(v01.x, v01.y, t.fx), lerpf_ref(v11.x, v11.y, t.fx), t.fy), t.fz);
}

__device__ inline float sample_face(const float* __restrict__ a, const Geom& g, int c, float px, float py, float pz) {
    const FaceStencil t = face_stencil(g, c, px, py, pz);
    return trilerp_pairs(load_pair(a + t.r00), load_pair(a + t.r10), load_pair(a + t.r01), load_pair(a + t.r11), t);
}

// The quad layout of a face array: Q[i + sx (j + sy k)] = (a(i, j, k),
// a(i1, j, k), a(i, j1, k), a(i1, j1, k)) with i1 = min(i+1, sx-1) and
// j1 = min(j+1, sy-1), the clamps of the stencil above.  A sample then takes
// two 16-byte loads (planes k and k+1) instead of four 8-byte ones, and
// interpolates the same values in the same order, so it is bit-identical.
template <typename S>
__device__ inline float trilerp_quads(float4 q0, float4 q1, const S& t) {
    return trilerp_pairs(make_float2(q0.x, q0.y), make_float2(q0.z, q0.w), make_float2(q1.x, q1.y),
                         make_float2(q1.z, q1.w), t);
}

__device__ inline float sample_quad(const float4* __restrict__ q, const Geom& g, int c, float px, float py,
                                    float pz) {
    const FaceStencil t = face_stencil(g, c, px, py, pz);
    return trilerp_quads(q[t.r00], q[t.r01], t);
}

// XCD-aware block remap: consecutive logical tiles land on the same XCD (and
// its L2) instead of being dealt round-robin.  nblocks is padded to a multiple
// of 8 by the launcher; blocks mapped past the real count exit.
__device__ inline int64_t xcd_remap(int64_t bid, int64_t padded) {
    const int64_t per = padded / kXcds;
    return (bid % kXcds) * per + bid / kXcds;
}

inline int64_t pad_blocks(int64_t nblocks) { return (nblocks + kXcds - 1) / kXcds * kXcds; }

// Slots per component of P2G's value bound: the sort's workgroups raise slot
// (block mod 64) with atomicMax and P2G takes the maximum over the slots, so
// no one address takes every workgroup's atomic.
constexpr int kP2GBoundSlots = 64;

// ---------------------------------------------------------------------------
// Device-resident solver scalars (FP64).  Reductions leave per-block partials;
// the next kernel's blocks each sum them in the same fixed order (so every
// block derives identical scalars) and block 0 records the state here, so the
// CG loop never waits on the host and uses no atomics.
// ---------------------------------------------------------------------------
struct SolverScalars {
    double sigma;      // r . z
    double sigma_old;
    double rr;         // r . r
    double dq;         // d . q
    double alpha;
    double beta;
    double bnorm2;     // b . b
    double tol;
    double rr_ref;     // r . r right after the last residual recomputation
    int32_t it;        // iterations done (-1 before the first spmv)
    int32_t done;      // stopping rule met or iteration cap reached
    int32_t trigger;   // sigma grew: recompute the residual next iteration
    int32_t relative;
    int32_t precond;
    int32_t max_it;
    int32_t replace;   // the next (or last) update recomputes r = b - A x
    int32_t converged; // the stopping rule (not the iteration cap) ended the solve
    int32_t confirm;   // a relative stop met on the recurrence awaits the recomputed residual
};

// ---------------------------------------------------------------------------
// Host-side scene.
// ---------------------------------------------------------------------------
// The state of a conjugate-gradient solve between its launches (the
// iterations' buffers rotate with the iteration count).
struct CGRun {
    int k = 0;        // spmv / update pairs launched
    int dcur = 0;     // d[dcur] holds the current search direction
    double* xa = nullptr;  // current solution
    double* xb = nullptr;  // its ping-pong partner
};

// One captured substep: graph `a` runs the stages up to the pressure solve's
// first `iters` spmv launches and copies the solver state to the host; graph
// `b` runs the stages from the pressure application on.  The key fields say
// which scene state the capture saw; the after-fields the host state that
// running it leaves.
struct SubstepGraph {
    bool valid = false;
    // key: what the capture baked in
    uint64_t gen = 0;
    fe_solver_desc desc{};
    double* p = nullptr;
    int64_t np = 0, cap = 0;
    float* px0 = nullptr;
    fe_tuning tune{};
    int iters = 0;  // spmv launches in graph a
    hipGraphExec_t a = nullptr, b = nullptr;
    // host state after graph a
    CGRun run_after_a{};
    int cur_after = 0;
    int sdf_from_cells_after = 0;
    // substeps in a row that needed host iterations / stopped > 4 iterations early
    int resume_streak = 0, idle_streak = 0;
};

struct Particles {
    float* x = nullptr;   // 3*cap: px | py | pz
    float* v = nullptr;   // 3*cap
    float* C = nullptr;   // 9*cap (APIC only)
    int32_t* pid = nullptr;
};

// One level of the multigrid preconditioner (level 0 is the pressure grid).
struct MGLevel {
    int n[3] = {0, 0, 0};
    int64_t cells = 0;
    float ih2[3] = {0, 0, 0};
    uint8_t* flags = nullptr;   // level 0: the scene's markers
    uint16_t* code = nullptr;   // level 0: the scene's row codes
    float* x = nullptr;         // correction and its ping-pong partner
    float* x2 = nullptr;
    float* b = nullptr;         // right-hand side (level 0: the CG residual, not owned)
    float* r = nullptr;         // residual scratch
    bool owns = false;          // flags/code/b allocated here (levels >= 1)
};

}  // namespace fe

namespace fe {
// Surface extraction buffers (allocated on first use): crossed-edge flags and
// vertex ids, per-cube triangle counts and offsets, scan tile sums, the case
// tables, the six boundary planes, the cube part of the mesh on the device
// and the caps on the host.
struct McState {
    int32_t *flag = nullptr, *vid = nullptr, *cnt = nullptr, *toff = nullptr, *sums = nullptr;
    void* tables = nullptr;      // FE_SURFACE_CLOSED cube tables
    void* tables_ref = nullptr;  // FE_SURFACE_REFERENCE cube tables
    double* planes = nullptr;
    int64_t plane_stride = 0;
    double *pos = nullptr, *nrm = nullptr;
    int32_t* tri = nullptr;
    double* comb = nullptr;  // fluid SDF combined with the collider's
    int64_t vcap = 0, ncap = 0, tcap = 0;
    int64_t nv3 = 0, nt3 = 0;  // cube-part vertices and triangles of the last build
    std::vector<double> cap_pos, cap_nrm;
    std::vector<int32_t> cap_tri;
    bool built = false;
};
}  // namespace fe

namespace fe {
// A VolumeEmitter: its Mesh2SDF on its node grid, particles per cell, start
// velocity, generator seed and Emitt call count.
struct Emitter {
    double* sdf = nullptr;
    NodeGrid3 grid{};
    int ppb = 0;
    double vel[3] = {0, 0, 0};
    uint64_t seed = 0, calls = 0;
};
}  // namespace fe

struct fe_net;  // the warm-start network (fe_learn.hpp)

struct fe_ctx {
    int device = 0;
    std::string err;
    std::mutex err_mu;  // scenes of one context may run on several host threads (fe_batch)
    char name[64] = {0};
};

struct fe_scene {
    fe_ctx* ctx = nullptr;
    fe_scene_desc desc{};
    fe::Geom g{};
    hipStream_t stream = nullptr;
    int depth = 3;            // ceil(maxCfl): extrapolation / BC depth
    int move_substeps = 3;    // (unsigned)max(maxCfl, 1)
    fe_tuning tune{FE_P2G_SCATTER, FE_G2P_QUADS, FE_MG_FUSED};  // kernel variants

    // grid (HBM)
    float* u[3] = {nullptr, nullptr, nullptr};       // face velocities
    float* snap[3] = {nullptr, nullptr, nullptr};    // FLIP snapshot
    float4* quad[3] = {nullptr, nullptr, nullptr};   // face values as 2x2 (x, y) quads for G2P sampling
    float4* squad[3] = {nullptr, nullptr, nullptr};  // the FLIP snapshot as quads
    uint8_t* mark[3] = {nullptr, nullptr, nullptr};  // P2G face markers
    uint8_t* valid[2] = {nullptr, nullptr};          // extrapolation validity ping-pong (all 3 comps)
    double* sdf = nullptr;
    uint8_t* flags = nullptr;
    uint16_t* code = nullptr;     // per-cell 7-point row code (BuildSystem)
    double* p = nullptr;          // pressure (LinearSystem::x), FP64: kept between solves (warm start)
    double* p2 = nullptr;         // pressure ping-pong partner
    float* b = nullptr;
    float* r = nullptr;
    float* z = nullptr;           // preconditioned residual
    float* d[2] = {nullptr, nullptr};
    float* q = nullptr;
    double* partials = nullptr;   // per-block FP64 partial sums (2 per block)
    uint8_t* chunk_fluid = nullptr;  // per solver block: its cells include a FLUID cell (from BuildSystem)
    int64_t npartials = 0;
    fe::SolverScalars* sc = nullptr;
    fe::SolverScalars* sc_host = nullptr;  // pinned mirror
    unsigned long long* cflmax = nullptr;  // two slots of (max |u| as FP64 bits, non-finite flag)
    int cfl_slot = 0;                      // the slot the next launch_cfl writes

    // particles (HBM, SoA, ping-pong for the reorder)
    int64_t np = 0, cap = 0;
    fe::Particles P[2];
    float4* pos4 = nullptr;  // positions of the sorted particles as (x, y, z, 0): one load per gathered candidate
    int cur = 0;
    uint32_t* key = nullptr;
    int32_t* cell_count = nullptr;   // n^3 + 1: bin histogram (all zero after the sort's k_scatter)
    bool count_zero = false;         // cell_count is known to be all zero (no fill needed before a histogram)
    int32_t* cell_start = nullptr;   // n^3 + 1
    int32_t* cursor = nullptr;       // n^3
    int32_t* order = nullptr;        // sorted slot -> old index
    int32_t* scan_sums = nullptr;
    uint32_t* p2g_bound = nullptr;   // 3 x kP2GBoundSlots: FP32 bits of P2G's |value| bound (set by the sort)
    bool sorted = false;
    bool keys_ready = false;  // bin keys + histogram already produced by the advection kernel

    // timing
    hipEvent_t ev[fe::kMaxStages + 1] = {};
    double stage_ms[fe::kMaxStages] = {};
    fe_solve_stats last_stats{};
    int last_cg_iters[3] = {0, 0, 0};  // iterations of the last CG solve per preconditioner (none, diagonal, multigrid)
    std::vector<fe::MGLevel> mg;  // multigrid hierarchy, built on first use
    int mg_tail = 0;              // first level handled by the one-workgroup tail kernel
    // fused V-cycle (FE_MG_FUSED): level 0's per-tile FLUID flags and the r.z
    // partials of its ascent kernel (one per launched workgroup)
    uint8_t* mg_tfl = nullptr;
    double* mg_pz = nullptr;
    int64_t mg_npz = 0;

    // BackwardEulerDiffusionSolver: its linear solver (default Jacobi(1000, 5, 1e-14)),
    // buffers allocated on first use (FP64 x ping-pong, row codes, face markers
    // per component) and the stats of the last solve per component
    fe_solver_desc diff_desc{FE_SOLVER_JACOBI, 1000, 5, 0, 0, 0, 1e-14};
    double* dx[2] = {nullptr, nullptr};
    uint16_t* dcode = nullptr;
    double* db = nullptr;
    uint8_t* dmark[3] = {nullptr, nullptr, nullptr};
    fe_solve_stats diff_stats[3] = {};

    fe::McState mc;  // surface extraction

    // colliders: the ColliderCollection SDF (node centred at the origin, n^3
    // FP64) and scratch for the tangential pass of the boundary condition
    double* csdf = nullptr;
    bool has_collider = false;
    bool sdf_buckets = false;  // FE_SDF_SEARCH_BUCKETS
    int sdf_from_cells = 0;    // the fluid SDF in memory came from the cell scan of the current bins
    int surface_mode = FE_SURFACE_REFERENCE;
    float* bc_tmp[3] = {nullptr, nullptr, nullptr};

    // scene sources: emitters, substeps run (the collider SDF is built from
    // the first one on), and count / offset scratch of the seeding passes
    std::vector<fe::Emitter> emitters;
    int64_t substeps_run = 0;

    // learned warm start: the attached network and the scale of its last guess
    fe_net* net = nullptr;
    double net_alpha = 0.0;
    int32_t *src_cnt = nullptr, *src_off = nullptr, *src_sums = nullptr;

    // per-kernel timer: HIP events around every launch of one selected kernel
    int32_t kt_id = 0;
    std::vector<hipEvent_t> kt_ev;  // pairs
    std::vector<int> kt_tag;        // one tag per pair
    int kt_used = 0;
    double kt_ms = 0.0;
    int64_t kt_launches = 0;

    // Substeps as HIP graphs (fe_scene_set_graph): one pair per parity of the
    // particle ping-pong at the substep's start, since the sort swaps the
    // buffers every substep.
    int graph_mode = 0;
    bool capturing = false;      // launches go into a graph: no timers, dt from dt_dev
    double* dt_dev = nullptr;    // the substep's dt on the device (graph substeps)
    double* res_host = nullptr;  // pinned: ||b - A p||^2 and ||b||^2 of the last graph substep
    uint64_t graph_gen = 1;      // bumped by every change a captured graph would not see
    fe::SubstepGraph graph[2];
    int64_t graph_substeps = 0, graph_captures = 0, graph_resumed = 0;
};

struct fe_batch {
    fe_ctx* ctx = nullptr;
    std::vector<fe_scene*> scenes;  // not owned
    std::vector<int32_t> order;     // scene indices, largest first
    int workers = 1;
};

namespace fe {

constexpr int kTimerPairs = 1024;

// Accumulate the finished pairs whose tag is below `limit` (call after the
// stream is synchronized); tags let a caller drop launches that did no work.
inline void kt_flush(fe_scene* s, int limit = 1 << 30) {
    for (int k = 0; k < s->kt_used; ++k) {
        if (s->kt_tag[k] >= limit) continue;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, s->kt_ev[2 * k], s->kt_ev[2 * k + 1]) == hipSuccess) {
            s->kt_ms += ms;
            s->kt_launches += 1;
        }
    }
    s->kt_used = 0;
}

// Launch `kernel` on the scene's stream; when the timer selects `id`, the
// start/stop events ride on the dispatch packet itself (hipExtLaunchKernelGGL),
// so they time the kernel's execution as rocprofv3's kernel trace does.
// Launches captured into a graph are never timed.
template <typename... Args, typename F>
inline void launch_timed(fe_scene* s, int32_t id, int tag, F kernel, dim3 grid, dim3 block, Args... args) {
    if (s->kt_id == id && !s->kt_ev.empty() && !s->capturing) {
        if (s->kt_used == kTimerPairs) {
            (void)hipStreamSynchronize(s->stream);
            kt_flush(s);
        }
        const int u = s->kt_used++;
        s->kt_tag[u] = tag;
        hipExtLaunchKernelGGL(kernel, grid, block, 0, s->stream, s->kt_ev[2 * u], s->kt_ev[2 * u + 1], 0, args...);
    } else {
        hipLaunchKernelGGL(kernel, grid, block, 0, s->stream, args...);
    }
}

}  // namespace fe
