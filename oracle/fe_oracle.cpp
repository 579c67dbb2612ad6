// fe_oracle.cpp -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// Clean-room FP64 CPU restatement of the Fluide_Engine hybrid particle-grid
// substep (PIC / FLIP / APIC transfers, fluid SDF + cell marking, velocity
// extrapolation, blocked boundary conditions with an SDF collider, gravity,
// backward-Euler viscosity, SinglePhasePressureSolver with Jacobi / CG, G2P
// and midpoint advection with collider collision, marching-cubes surface).
// It is the checker for the HIP library and the CPU baseline timed beside it.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
//
// Every function cites the reference file:line it restates (paths relative to
// the Fluide_Engine source tree).  Reference quirks are reproduced on purpose:
//   * PIC/APIC P2G pass the cell dims for every component (src/pic_simulator.cpp:389,397,405;
//     src/apic_simulator.cpp:58,71,84), FLIP passes each component's own dims
//     (src/flip_simulator.cpp:58,66,74).
//   * face markers are set for all 8 stencil nodes, weight or not (src/pic_simulator.cpp:394).
//   * APIC weights use the unclamped position, the affine term the clamped one
//     (src/apic_simulator.cpp:55-63,150-153).
//   * FLIP snapshot taken right after P2G normalisation (src/flip_simulator.cpp:107-118).
//   * ApplyPressure only corrects +side faces of FLUID cells
//     (src/fluid_solvers/single_phase_pressure_solver.cpp:74-91).
//   * RHS is the unscaled divergence, A = 1/h^2 stencil, gradient 1/h (same file :59-60,103-118).
//   * pressure warm start from the previous solve (src/common/array3.hpp:189-197).
//   * Cfl adds dt*(0,-9.8,0) to |u| (src/pic_simulator.cpp:88).
// P2G is serial in particle order (the reference's threaded P2G races, so its
// own output is not reproducible; the serial order is the deterministic one).
//
// Parity pinning (see tests/test_oracle_*.py): the reference's known-answer
// tests (test/blas_test.cpp, test/math_utils_test.cpp, test/face_centered_grid3d_test.cpp,
// test/scalar_grid3d_test.cpp, test/single_phase_pressuer_solver_test.cpp,
// test/hash_grid_particle_searcher_test.cpp), closed forms, dense direct
// solves and brute force on tiny inputs.
//
// Build: oracle/Makefile -> oracle/_build/libfe_oracle.so (C ABI, loaded by ctypes).

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <thread>
#include <vector>

#include "fe_oracle.h"
#include "../fluide_engine_amd/csrc/fe_mc_table.hpp"

namespace {

constexpr double kMaxDistance = 1.7976931348623157E+30;  // src/3d/collider_collection.cpp:3

// ---------------------------------------------------------------------------
// Thread model for the CPU baseline: the reference forks hardware_concurrency()
// std::threads per loop and chunks the slowest-varying index
// (src/common/parallel_utils.hpp:24-61,114-154).  The loops below are all
// race-free element-wise sweeps, so the result does not depend on threads.
// ---------------------------------------------------------------------------
void par_for(int nthreads, int64_t n, const std::function<void(int64_t, int64_t)>& body) {
    if (nthreads <= 1 || n < 4096) {
        body(0, n);
        return;
    }
    std::vector<std::thread> pool;
    int64_t chunk = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        int64_t a = t * chunk, b = std::min<int64_t>(n, a + chunk);
        if (a >= b) break;
        pool.emplace_back(body, a, b);
    }
    for (auto& th : pool) th.join();
}

struct Arr3 {
    int sx = 0, sy = 0, sz = 0;
    std::vector<double> d;
    void resize(int x, int y, int z, double v = 0.0) {
        sx = x; sy = y; sz = z;
        d.assign((size_t)x * y * z, v);
    }
    // src/common/array3.hpp:65-85 -- x-fastest layout i + sx*(j + sy*k)
    inline size_t at(int i, int j, int k) const { return (size_t)i + (size_t)sx * ((size_t)j + (size_t)sy * k); }
    inline double& operator()(int i, int j, int k) { return d[at(i, j, k)]; }
    inline double operator()(int i, int j, int k) const { return d[at(i, j, k)]; }
    size_t size() const { return d.size(); }
};

struct Arr3i {
    int sx = 0, sy = 0, sz = 0;
    std::vector<int> d;
    void resize(int x, int y, int z, int v = 0) {
        sx = x; sy = y; sz = z;
        d.assign((size_t)x * y * z, v);
    }
    inline size_t at(int i, int j, int k) const { return (size_t)i + (size_t)sx * ((size_t)j + (size_t)sy * k); }
    inline int& operator()(int i, int j, int k) { return d[at(i, j, k)]; }
    inline int operator()(int i, int j, int k) const { return d[at(i, j, k)]; }
};

// src/common/math_utils.hpp:13-28 -- Clamp(val, maxVal, minVal)
inline double clamp_ref(double val, double maxVal, double minVal) {
    if (val < minVal) return minVal;
    if (val > maxVal) return maxVal;
    return val;
}

// src/common/math_utils.hpp:30-34 -- Lerp(a, b, f) = (1-f)*a + f*b
inline double lerp(double a, double b, double f) { return (1 - f) * a + f * b; }

// src/common/math_utils.hpp:42-46
inline double trilerp(double x000, double x100, double x010, double x110, double x001, double x101, double x011,
                      double x111, double fx, double fy, double fz) {
    return lerp(lerp(lerp(x000, x100, fx), lerp(x010, x110, fx), fy),
                lerp(lerp(x001, x101, fx), lerp(x011, x111, fx), fy), fz);
}

// src/common/math_utils.hpp:48-73 -- GetBarycentric(x, iLow=0, iHigh, &i, &f)
inline void barycentric(double x, int iLow, int iHigh, int* i, double* f) {
    double s = std::floor(x);
    *i = static_cast<int>(s);
    int siLow = iLow, siHigh = iHigh;
    int offset = -siLow;
    siLow += offset;
    siHigh += offset;
    if (siLow == siHigh) {
        *i = siLow;
        *f = 0;
    } else if (*i < siLow) {
        *i = siLow;
        *f = 0;
    } else if (*i > siHigh - 1) {
        *i = siHigh - 1;
        *f = 1;
    } else {
        *f = x - s;
    }
    *i -= offset;
}

struct Stencil {
    int idx[8][3];
    double w[8];
};

// src/common/math_utils.hpp:75-112 -- GetCooridnatesAndWeights(size, origin, h, x, indexes, weights)
inline void coords_weights(const int size[3], const double origin[3], const double h[3], const double x[3],
                           Stencil* st) {
    int i = 0, j = 0, k = 0;
    double fx = 0, fy = 0, fz = 0;
    const double nx = (x[0] - origin[0]) / h[0];
    const double ny = (x[1] - origin[1]) / h[1];
    const double nz = (x[2] - origin[2]) / h[2];
    barycentric(nx, 0, size[0] - 1, &i, &fx);
    barycentric(ny, 0, size[1] - 1, &j, &fy);
    barycentric(nz, 0, size[2] - 1, &k, &fz);
    const int ip1 = std::min(i + 1, size[0] - 1);
    const int jp1 = std::min(j + 1, size[1] - 1);
    const int kp1 = std::min(k + 1, size[2] - 1);
    const int I[8] = {i, ip1, i, ip1, i, ip1, i, ip1};
    const int J[8] = {j, j, jp1, jp1, j, j, jp1, jp1};
    const int K[8] = {k, k, k, k, kp1, kp1, kp1, kp1};
    for (int c = 0; c < 8; ++c) {
        st->idx[c][0] = I[c];
        st->idx[c][1] = J[c];
        st->idx[c][2] = K[c];
    }
    st->w[0] = (1.00 - fx) * (1.00 - fy) * (1.00 - fz);
    st->w[1] = fx * (1.00 - fy) * (1.00 - fz);
    st->w[2] = (1.00 - fx) * fy * (1.00 - fz);
    st->w[3] = fx * fy * (1.00 - fz);
    st->w[4] = (1.00 - fx) * (1.00 - fy) * fz;
    st->w[5] = fx * (1.00 - fy) * fz;
    st->w[6] = (1.00 - fx) * fy * fz;
    st->w[7] = fx * fy * fz;
}

struct GradStencil {
    int idx[8][3];
    double g[8][3];
};

// src/common/math_utils.hpp:114-176 -- GetCooridnatesAndGradientWeights
inline void coords_grad_weights(const int size[3], const double origin[3], const double h[3], const double x[3],
                                GradStencil* st) {
    int i = 0, j = 0, k = 0;
    double fx = 0, fy = 0, fz = 0;
    const double nx = (x[0] - origin[0]) / h[0];
    const double ny = (x[1] - origin[1]) / h[1];
    const double nz = (x[2] - origin[2]) / h[2];
    const double ihx = 1.0 / h[0], ihy = 1.0 / h[1], ihz = 1.0 / h[2];
    barycentric(nx, 0, size[0] - 1, &i, &fx);
    barycentric(ny, 0, size[1] - 1, &j, &fy);
    barycentric(nz, 0, size[2] - 1, &k, &fz);
    const int ip1 = std::min(i + 1, size[0] - 1);
    const int jp1 = std::min(j + 1, size[1] - 1);
    const int kp1 = std::min(k + 1, size[2] - 1);
    const int I[8] = {i, ip1, i, ip1, i, ip1, i, ip1};
    const int J[8] = {j, j, jp1, jp1, j, j, jp1, jp1};
    const int K[8] = {k, k, k, k, kp1, kp1, kp1, kp1};
    for (int c = 0; c < 8; ++c) {
        st->idx[c][0] = I[c];
        st->idx[c][1] = J[c];
        st->idx[c][2] = K[c];
    }
    double(*g)[3] = st->g;
    g[0][0] = -ihx * (1 - fy) * (1 - fz); g[0][1] = -ihy * (1 - fx) * (1 - fz); g[0][2] = -ihz * (1 - fx) * (1 - fy);
    g[1][0] = ihx * (1 - fy) * (1 - fz);  g[1][1] = fx * (-ihy) * (1 - fz);     g[1][2] = fx * (1 - fy) * (-ihz);
    g[2][0] = (-ihx) * fy * (1 - fz);     g[2][1] = (1 - fx) * ihy * (1 - fz);  g[2][2] = (1 - fx) * fy * (-ihz);
    g[3][0] = ihx * fy * (1 - fz);        g[3][1] = fx * ihy * (1 - fz);        g[3][2] = fx * fy * (-ihz);
    g[4][0] = (-ihx) * (1 - fy) * fz;     g[4][1] = (1 - fx) * (-ihy) * fz;     g[4][2] = (1 - fx) * (1 - fy) * ihz;
    g[5][0] = ihx * (1 - fy) * fz;        g[5][1] = fx * (-ihy) * fz;           g[5][2] = fx * (1 - fy) * ihz;
    g[6][0] = (-ihx) * fy * fz;           g[6][1] = (1 - fx) * ihy * fz;        g[6][2] = (1 - fx) * fy * ihz;
    g[7][0] = ihx * fy * fz;              g[7][1] = fx * ihy * fz;              g[7][2] = fx * fy * ihz;
}

// src/grid_systems/face_centered_grid3d.cpp:339-371 -- SampleArray(position, origin, arr)
// (also src/3d/scalar_grid3d.cpp:33-58 for ScalarGrid3D::Sample with the grid origin)
inline double sample_array(const Arr3& a, const double origin[3], const double h[3], const double p[3]) {
    int i = 0, j = 0, k = 0;
    double fx = 0, fy = 0, fz = 0;
    const double nx = (p[0] - origin[0]) / h[0];
    const double ny = (p[1] - origin[1]) / h[1];
    const double nz = (p[2] - origin[2]) / h[2];
    barycentric(nx, 0, a.sx - 1, &i, &fx);
    barycentric(ny, 0, a.sy - 1, &j, &fy);
    barycentric(nz, 0, a.sz - 1, &k, &fz);
    const int ip1 = std::min(i + 1, a.sx - 1);
    const int jp1 = std::min(j + 1, a.sy - 1);
    const int kp1 = std::min(k + 1, a.sz - 1);
    return trilerp(a(i, j, k), a(ip1, j, k), a(i, jp1, k), a(ip1, jp1, k), a(i, j, kp1), a(ip1, j, kp1),
                   a(i, jp1, kp1), a(ip1, jp1, kp1), fx, fy, fz);
}

// src/common/math_utils.hpp:186-253 -- ExtrapolateToRegion(input, valid, iterations, output)
// Jacobi-style: an invalid node with valid 6-neighbours takes their mean and
// becomes valid for the next iteration.  Reads of `output` only touch nodes that
// are valid in valid0, which are never written in the same sweep, so the
// reference's in-place parallel sweep is deterministic.
void extrapolate_to_region(const Arr3& input, const Arr3i& valid, int iterations, Arr3& output, int nthreads) {
    const int sx = input.sx, sy = input.sy, sz = input.sz;
    Arr3i valid0 = valid, valid1 = valid;
    output = input;
    for (int it = 0; it < iterations; ++it) {
        par_for(nthreads, sz, [&](int64_t k0, int64_t k1) {
            for (int k = (int)k0; k < (int)k1; ++k)
                for (int j = 0; j < sy; ++j)
                    for (int i = 0; i < sx; ++i) {
                        if (!valid0(i, j, k)) {
                            double sum = 0;
                            unsigned count = 0;
                            if (i + 1 < sx && valid0(i + 1, j, k)) { sum += output(i + 1, j, k); ++count; }
                            if (i > 0 && valid0(i - 1, j, k)) { sum += output(i - 1, j, k); ++count; }
                            if (j + 1 < sy && valid0(i, j + 1, k)) { sum += output(i, j + 1, k); ++count; }
                            if (j > 0 && valid0(i, j - 1, k)) { sum += output(i, j - 1, k); ++count; }
                            if (k + 1 < sz && valid0(i, j, k + 1)) { sum += output(i, j, k + 1); ++count; }
                            if (k > 0 && valid0(i, j, k - 1)) { sum += output(i, j, k - 1); ++count; }
                            if (count > 0) {
                                output(i, j, k) = sum / count;
                                valid1(i, j, k) = 1;
                            }
                        } else {
                            valid1(i, j, k) = 1;
                        }
                    }
        });
        std::swap(valid0.d, valid1.d);  // src/common/math_utils.hpp:251 Swap
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Scene state: the Fluid3 struct (src/fluid3.hpp:10-26) plus the solver state
// the reference keeps in SinglePhasePressureSolver / FLIPSimulator / APICSimulator.
// ---------------------------------------------------------------------------
struct orc_scene {
    int n[3];
    double origin[3], size[3], h[3];
    int method;
    double blend;   // FLIP _blendingFactor = PIC weight (src/flip_simulator.cpp:5,155-159)
    double maxCfl;  // src/hybrid_simulator.cpp:21
    double gravity[3];
    int nthreads = 1;
    int has_collider = 0;
    int fast_dots = 0;  // orc_set_fast_dots: storage-order blocked dot products in the solvers
    int surface_mode = 0;  // ORC_SURFACE_*
    int planted = 0;  // ORC_PLANT_*: a deliberate restatement mistake, set only by the pinning tests
    double xo[3], yo[3], zo[3];  // face data origins, src/grid_systems/face_centered_grid3d.cpp:324-332

    Arr3 u, v, w;         // face velocities
    Arr3i mu, mv, mw;     // face markers from P2G (Fluid3::x/y/zMarkers)
    Arr3 ru, rv, rw;      // FLIP _velocityResiduals (snapshot, then delta)
    Arr3 sdf;             // fluid SDF, cell centred (origin + h/2), src/hybrid_simulator.cpp:13-15
    Arr3 csdf;            // collider SDF, node centred (origin), src/3d/collider_collection.cpp:8-11
    std::vector<uint8_t> flags;  // SinglePhasePressureSolver::_fluidMarkers
    Arr3 p, b;            // LinearSystem x (kept between solves: warm start) and b
    Arr3 Ac, Ar, Au, Af;  // LinearSystem A rows {center,right,up,front}
    std::vector<double> x, vel, C;  // particles (3N, 3N, 9N)
    int64_t np = 0;
    orc_solve_stats last{};
    // BackwardEulerDiffusionSolver state: viscosity (Fluid3::viscosity), its
    // linear solver (CudaJacobiIterationSolver(1000, 5, 1e-14),
    // src/fluid_solvers/backward_euler_diffusion_solver.cpp:10), the face
    // markers of the last solve per component (_fluidMarkers) and its stats.
    double viscosity = 0.0;
    orc_solver_desc diff{ORC_SOLVER_JACOBI, 1000, 5, 0, 0, 0, 1e-14};
    std::vector<uint8_t> dmark[3];
    orc_solve_stats diff_last[3]{};
    // VolumeEmitters (src/emmiter/volume_emmiter.cpp:12-85): the emitter's SDF
    // on its node grid, particles per cell, start velocity, and the seed and
    // call count of the generator that stands in for std::rand
    struct Emitter {
        Arr3 sdf;
        double origin[3], h[3];
        int ppb;
        double vel[3];
        uint64_t seed, calls;
    };
    std::vector<Emitter> emitters;
    int64_t substeps_run = 0;  // the collider SDF exists once a substep has run (Q19 below)
};

namespace {

void face_origins(orc_scene* s) {
    // src/grid_systems/face_centered_grid3d.cpp:324-332
    for (int a = 0; a < 3; ++a) {
        s->xo[a] = s->origin[a] + s->h[a] * (a == 0 ? 0.0 : 0.5);
        s->yo[a] = s->origin[a] + s->h[a] * (a == 1 ? 0.0 : 0.5);
        s->zo[a] = s->origin[a] + s->h[a] * (a == 2 ? 0.0 : 0.5);
    }
}

const double* comp_origin(const orc_scene* s, int c) { return c == 0 ? s->xo : (c == 1 ? s->yo : s->zo); }
Arr3& comp_arr(orc_scene* s, int c) { return c == 0 ? s->u : (c == 1 ? s->v : s->w); }
Arr3i& comp_mark(orc_scene* s, int c) { return c == 0 ? s->mu : (c == 1 ? s->mv : s->mw); }
Arr3& comp_res(orc_scene* s, int c) { return c == 0 ? s->ru : (c == 1 ? s->rv : s->rw); }

// FaceCenteredGrid3D::Sample (src/grid_systems/face_centered_grid3d.cpp:138-144)
inline void sample_vel(const orc_scene* s, const Arr3& u, const Arr3& v, const Arr3& w, const double p[3],
                       double out[3]) {
    out[0] = sample_array(u, s->xo, s->h, p);
    out[1] = sample_array(v, s->yo, s->h, p);
    out[2] = sample_array(w, s->zo, s->h, p);
}

// ScalarGrid3D::Sample on the collider SDF (node centred at the domain origin)
inline double sample_collider(const orc_scene* s, const double p[3]) { return sample_array(s->csdf, s->origin, s->h, p); }

// ScalarGrid3D::Gradient (src/3d/scalar_grid3d.cpp:67-90) on the collider SDF:
// the trilinear weights of the position times clamped central differences.
void collider_gradient(const orc_scene* s, const double p[3], double g[3]) {
    const int size[3] = {s->csdf.sx, s->csdf.sy, s->csdf.sz};
    Stencil st;
    coords_weights(size, s->origin, s->h, p, &st);
    g[0] = g[1] = g[2] = 0.0;
    for (int c = 0; c < 8; ++c) {
        const int i = st.idx[c][0], j = st.idx[c][1], k = st.idx[c][2];
        const double left = s->csdf(i > 0 ? i - 1 : i, j, k), right = s->csdf(i + 1 < size[0] ? i + 1 : i, j, k);
        const double down = s->csdf(i, j > 0 ? j - 1 : j, k), up = s->csdf(i, j + 1 < size[1] ? j + 1 : j, k);
        const double back = s->csdf(i, j, k > 0 ? k - 1 : k), front = s->csdf(i, j, k + 1 < size[2] ? k + 1 : k);
        const double d[3] = {right - left, up - down, front - back};
        for (int a = 0; a < 3; ++a) g[a] += (st.w[c] * 0.5 * d[a]) / s->h[a];
    }
}


// ---------------------------------------------------------------------------
// P2G.  PIC: src/pic_simulator.cpp:356-437.  FLIP: src/flip_simulator.cpp:25-119.
// APIC: src/apic_simulator.cpp:15-118.
// ---------------------------------------------------------------------------
void p2g(orc_scene* s) {
    const int64_t N = s->np;
    int cell[3] = {s->n[0], s->n[1], s->n[2]};
    Arr3 wsum[3];
    for (int c = 0; c < 3; ++c) {
        Arr3& a = comp_arr(s, c);
        std::fill(a.d.begin(), a.d.end(), 0.0);
        wsum[c].resize(a.sx, a.sy, a.sz, 0.0);
        Arr3i& m = comp_mark(s, c);
        m.resize(a.sx, a.sy, a.sz, 0);
    }
    if (s->method == ORC_APIC && (int64_t)s->C.size() != 9 * N) s->C.assign(9 * N, 0.0);  // _cX.resize (:22-24)
    const double hh[3] = {s->h[0] / 2.0, s->h[1] / 2.0, s->h[2] / 2.0};
    Stencil st;
    for (int64_t pi = 0; pi < N; ++pi) {  // serial: deterministic order
        const double* pos = &s->x[3 * pi];
        const double* vel = &s->vel[3 * pi];
        for (int c = 0; c < 3; ++c) {
            Arr3& a = comp_arr(s, c);
            Arr3i& m = comp_mark(s, c);
            // Q1: PIC/APIC pass the cell dims, FLIP each component's own dims
            int own[3] = {a.sx, a.sy, a.sz};
            const int* size = (s->method == ORC_FLIP) ? own : cell;
            coords_weights(size, comp_origin(s, c), s->h, pos, &st);
            if (s->method == ORC_APIC) {
                // src/apic_simulator.cpp:54-63 -- clamp the two transverse axes to [o+h/2, o+L-h/2]
                double pc[3] = {pos[0], pos[1], pos[2]};
                for (int ax = 0; ax < 3; ++ax)
                    if (ax != c)
                        pc[ax] = clamp_ref(pc[ax], s->origin[ax] + s->size[ax] - hh[ax], s->origin[ax] + hh[ax]);
                const double* cr = &s->C[9 * pi + 3 * c];
                const double* co = comp_origin(s, c);
                for (int q = 0; q < 8; ++q) {
                    // GetXPos: dataOrigin + h * (i,j,k)  (src/grid_systems/face_centered_grid3d.cpp:303-307)
                    double gp[3];
                    for (int ax = 0; ax < 3; ++ax) gp[ax] = co[ax] + s->h[ax] * (double)st.idx[q][ax];
                    // Vector3::Dot (src/common/vector3.hpp:136-139)
                    double apic = cr[0] * (gp[0] - pc[0]) + cr[1] * (gp[1] - pc[1]) + cr[2] * (gp[2] - pc[2]);
                    size_t id = a.at(st.idx[q][0], st.idx[q][1], st.idx[q][2]);
                    a.d[id] += (vel[c] + apic) * st.w[q];
                    wsum[c].d[id] += st.w[q];
                    m.d[id] = 1;
                }
            } else {
                for (int q = 0; q < 8; ++q) {
                    size_t id = a.at(st.idx[q][0], st.idx[q][1], st.idx[q][2]);
                    a.d[id] += vel[c] * st.w[q];
                    wsum[c].d[id] += st.w[q];
                    m.d[id] = 1;  // Q2
                }
            }
        }
    }
    for (int c = 0; c < 3; ++c) {
        Arr3& a = comp_arr(s, c);
        const Arr3& ws = wsum[c];
        par_for(s->nthreads, (int64_t)a.size(), [&](int64_t i0, int64_t i1) {
            for (int64_t i = i0; i < i1; ++i)
                if (ws.d[i] > 0.0) a.d[i] /= ws.d[i];
        });
    }
    if (s->method == ORC_FLIP) {  // src/flip_simulator.cpp:107-118
        s->ru = s->u;
        s->rv = s->v;
        s->rw = s->w;
    }
}

// ---------------------------------------------------------------------------
// G2P.  PIC: src/pic_simulator.cpp:439-450.  FLIP: src/flip_simulator.cpp:121-161.
// APIC: src/apic_simulator.cpp:120-176.
// ---------------------------------------------------------------------------
void g2p(orc_scene* s) {
    const int64_t N = s->np;
    if (s->method == ORC_FLIP) {
        for (int c = 0; c < 3; ++c) {  // delta = new - snapshot (:138-150)
            Arr3& a = comp_arr(s, c);
            Arr3& r = comp_res(s, c);
            if (s->planted == ORC_PLANT_FLIP_DELTA_SIGN)
                for (size_t i = 0; i < a.size(); ++i) r.d[i] = r.d[i] - a.d[i];
            else
                for (size_t i = 0; i < a.size(); ++i) r.d[i] = a.d[i] - r.d[i];
        }
    }
    if (s->method == ORC_APIC) s->C.assign(9 * N, 0.0);  // :129-134
    const double hh[3] = {s->h[0] / 2.0, s->h[1] / 2.0, s->h[2] / 2.0};
    int cell[3] = {s->n[0], s->n[1], s->n[2]};
    par_for(s->nthreads, N, [&](int64_t p0, int64_t p1) {
        for (int64_t pi = p0; pi < p1; ++pi) {
            const double* pos = &s->x[3 * pi];
            double* vel = &s->vel[3 * pi];
            double pic[3];
            sample_vel(s, s->u, s->v, s->w, pos, pic);
            if (s->method == ORC_PIC) {
                vel[0] = pic[0]; vel[1] = pic[1]; vel[2] = pic[2];
            } else if (s->method == ORC_FLIP) {
                double d[3];
                sample_vel(s, s->ru, s->rv, s->rw, pos, d);
                double fl[3] = {vel[0] + d[0], vel[1] + d[1], vel[2] + d[2]};
                if (s->blend > 0.0) {
                    // Lerp<Vector3>(flip, pic, blend): (1-b)*flip + b*pic (src/common/math_utils.hpp:30-34)
                    for (int a = 0; a < 3; ++a) fl[a] = (1 - s->blend) * fl[a] + s->blend * pic[a];
                }
                vel[0] = fl[0]; vel[1] = fl[1]; vel[2] = fl[2];
            } else {
                vel[0] = pic[0]; vel[1] = pic[1]; vel[2] = pic[2];
                GradStencil gs;
                for (int c = 0; c < 3; ++c) {
                    double pc[3] = {pos[0], pos[1], pos[2]};
                    for (int ax = 0; ax < 3; ++ax)
                        if (ax != c)
                            pc[ax] = clamp_ref(pc[ax], s->origin[ax] + s->size[ax] - hh[ax], s->origin[ax] + hh[ax]);
                    const Arr3& a = comp_arr(s, c);
                    int dims[3] = {cell[0], cell[1], cell[2]};
                    if (s->planted == ORC_PLANT_APIC_FACE_DIMS) dims[c] += 1;
                    coords_grad_weights(dims, comp_origin(s, c), s->h, pc, &gs);
                    // row c of C holds the gradient of component c (_cX, _cY, _cZ)
                    double* cr = &s->C[9 * pi + 3 * c];
                    const int64_t st = s->planted == ORC_PLANT_APIC_TRANSPOSE ? 3 : 1;
                    if (st == 3) cr = &s->C[9 * pi + c];
                    for (int q = 0; q < 8; ++q) {
                        double val = a(gs.idx[q][0], gs.idx[q][1], gs.idx[q][2]);
                        cr[0] += gs.g[q][0] * val;
                        cr[st] += gs.g[q][1] * val;
                        cr[2 * st] += gs.g[q][2] * val;
                    }
                }
            }
        }
    });
}

// ---------------------------------------------------------------------------
// MoveParticles, src/pic_simulator.cpp:239-312, with the collider resolution
// (:297-304) restated on the collider SDF (below).
// ---------------------------------------------------------------------------
void move_particles(orc_scene* s, double dt_total) {
    const int64_t N = s->np;
    unsigned numSub = static_cast<unsigned>(std::max(s->maxCfl, 1.0));
    double dt = dt_total / numSub;
    par_for(s->nthreads, N, [&](int64_t p0, int64_t p1) {
        for (int64_t pi = p0; pi < p1; ++pi) {
            double* pos = &s->x[3 * pi];
            double* vel = &s->vel[3 * pi];
            double pt0[3] = {pos[0], pos[1], pos[2]}, pt1[3] = {pos[0], pos[1], pos[2]};
            for (unsigned t = 0; t < numSub; ++t) {
                double v0[3], mid[3], mv[3];
                sample_vel(s, s->u, s->v, s->w, pt0, v0);
                for (int a = 0; a < 3; ++a) mid[a] = pt0[a] + 0.5 * dt * v0[a];
                sample_vel(s, s->u, s->v, s->w, mid, mv);
                if (s->planted == ORC_PLANT_EULER_ADVECTION)
                    for (int a = 0; a < 3; ++a) pt1[a] = pt0[a] + dt * v0[a];
                else
                    for (int a = 0; a < 3; ++a) pt1[a] = pt0[a] + dt * mv[a];
                for (int a = 0; a < 3; ++a) pt0[a] = pt1[a];
            }
            for (int a = 0; a < 3; ++a) {
                if (pt1[a] <= s->origin[a]) {
                    pt1[a] = s->origin[a];
                    vel[a] = 0.0;
                }
                if (pt1[a] >= s->origin[a] + s->size[a]) {
                    pt1[a] = s->origin[a] + s->size[a];
                    vel[a] = 0.0;
                }
            }
            // Particles that end inside a collider: the reference resolves them
            // against the collider mesh, indexing the particle by the
            // collider's loop index (src/pic_simulator.cpp:297-304, a shadowed
            // i).  Restated for an SDF collider on the particle itself: move it
            // to the zero level set along the normalised collider gradient,
            // x -= phi n, and drop an inward normal velocity, v -= min(v.n, 0) n
            // (restitution 0, friction 0, as the reference's ResolveCollision).
            if (s->has_collider && sample_collider(s, pt1) < 0) {
                double g[3];
                collider_gradient(s, pt1, g);
                const double len = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
                if (len * len > 0.0) {
                    const double n[3] = {g[0] / len, g[1] / len, g[2] / len};
                    const double phi = sample_collider(s, pt1);
                    for (int a = 0; a < 3; ++a) {
                        pt1[a] = pt1[a] - phi * n[a];
                        pt1[a] = std::min(std::max(pt1[a], s->origin[a]), s->origin[a] + s->size[a]);
                    }
                    const double vn = vel[0] * n[0] + vel[1] * n[1] + vel[2] * n[2];
                    if (vn < 0.0)
                        for (int a = 0; a < 3; ++a) vel[a] = vel[a] - vn * n[a];
                }
            }
            pos[0] = pt1[0]; pos[1] = pt1[1]; pos[2] = pt1[2];
        }
    });
}

// ---------------------------------------------------------------------------
// BuildSignedDistanceField, src/pic_simulator.cpp:452-475, with the searcher of
// src/particle_systems/hash_grid_particle_searcher.cpp:97-204 (64^3 wrapped
// buckets of width 2r, 8-bucket query).  Restated with an exact hash grid over
// the same bucket geometry so the candidate set equals the reference's.
// ---------------------------------------------------------------------------
struct HashGrid {
    double spacing;
    std::vector<std::vector<int64_t>> buckets;
    static constexpr int S = 64;  // src/particle_systems/particle_system.cpp:30-34
    static int wrap(int b) {
        // hash_grid_particle_searcher.cpp:142-152: `int % size_t` converts the int to
        // size_t first; for S = 64 (a power of two) that equals the mathematical mod.
        int r = (int)((size_t)(int64_t)b % (size_t)S);
        if (r < 0) r += S;
        return r;
    }
    size_t key(int bx, int by, int bz) const { return ((size_t)wrap(bz) * S + wrap(by)) * S + wrap(bx); }
    void bucket_index(const double p[3], int b[3]) const {
        for (int a = 0; a < 3; ++a) b[a] = static_cast<int>(std::floor(p[a] / spacing));
    }
    void build(const std::vector<double>& x, int64_t N, double sp) {
        spacing = sp;
        buckets.assign((size_t)S * S * S, {});
        for (int64_t i = 0; i < N; ++i) {
            int b[3];
            bucket_index(&x[3 * i], b);
            buckets[key(b[0], b[1], b[2])].push_back(i);
        }
    }
    void nearby_keys(const double p[3], size_t keys[8]) const {
        int o[3];
        bucket_index(p, o);
        int nb[8][3];
        for (int i = 0; i < 8; ++i) { nb[i][0] = o[0]; nb[i][1] = o[1]; nb[i][2] = o[2]; }
        // hash_grid_particle_searcher.cpp:164-198 (note the float 0.5f)
        int dx = ((o[0] + 0.5f) * spacing <= p[0]) ? 1 : -1;
        int dy = ((o[1] + 0.5f) * spacing <= p[1]) ? 1 : -1;
        int dz = ((o[2] + 0.5f) * spacing <= p[2]) ? 1 : -1;
        nb[4][0] += dx; nb[5][0] += dx; nb[6][0] += dx; nb[7][0] += dx;
        nb[2][1] += dy; nb[3][1] += dy; nb[6][1] += dy; nb[7][1] += dy;
        nb[1][2] += dz; nb[3][2] += dz; nb[5][2] += dz; nb[7][2] += dz;
        for (int i = 0; i < 8; ++i) keys[i] = key(nb[i][0], nb[i][1], nb[i][2]);
    }
};

void build_sdf(orc_scene* s) {
    const double maxH = std::max(s->h[0], std::max(s->h[1], s->h[2]));
    const double radius = 1.2 * maxH / std::sqrt(2.0);
    const double band = 2.0 * radius;
    HashGrid hg;
    hg.build(s->x, s->np, 2 * radius);
    const double so[3] = {s->origin[0] + s->h[0] / 2.0, s->origin[1] + s->h[1] / 2.0, s->origin[2] + s->h[2] / 2.0};
    const int nx = s->n[0], ny = s->n[1], nz = s->n[2];
    const double qr2 = band * band;
    par_for(s->nthreads, nz, [&](int64_t k0, int64_t k1) {
        for (int k = (int)k0; k < (int)k1; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    // Grid3D::GridIndexToPosition (src/common/grid3d.hpp:13-16)
                    double gp[3] = {so[0] + i * s->h[0], so[1] + j * s->h[1], so[2] + k * s->h[2]};
                    double minD = band;
                    size_t keys[8];
                    hg.nearby_keys(gp, keys);
                    for (int q = 0; q < 8; ++q) {
                        for (int64_t pi : hg.buckets[keys[q]]) {
                            const double* p = &s->x[3 * pi];
                            double dx = p[0] - gp[0], dy = p[1] - gp[1], dz = p[2] - gp[2];
                            double d2 = dx * dx + dy * dy + dz * dz;
                            if (d2 <= qr2) minD = std::min(minD, std::sqrt(dx * dx + dy * dy + dz * dz));
                        }
                    }
                    s->sdf(i, j, k) = minD - radius;
                }
    });
    // ExtrapolateIntoCollider (src/pic_simulator.cpp:497-519): valid = colliderSdf >= 0
    int depth = static_cast<int>(std::ceil(s->maxCfl));
    Arr3i valid;
    valid.resize(nx, ny, nz, 1);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double gp[3] = {so[0] + i * s->h[0], so[1] + j * s->h[1], so[2] + k * s->h[2]};
                valid(i, j, k) = sample_collider(s, gp) < 0 ? 0 : 1;
            }
    Arr3 prev = s->sdf;
    extrapolate_to_region(prev, valid, depth, s->sdf, s->nthreads);
}

// ExtrapolateVelocityToAir, src/pic_simulator.cpp:477-495 (validity = P2G face markers)
void extrapolate_to_air(orc_scene* s) {
    int iters = static_cast<int>(std::ceil(s->maxCfl));
    for (int c = 0; c < 3; ++c) {
        Arr3& a = comp_arr(s, c);
        Arr3 prev = a;
        extrapolate_to_region(prev, comp_mark(s, c), iters, a, s->nthreads);
    }
}

// FractionInsideSdf, src/fluid_solvers/blocked_boundry_condition_solver.cpp:304-322
double fraction_inside(double phi0, double phi1) {
    if (phi0 < 0.0 && phi1 < 0.0) return 1;
    if (phi0 < 0.0 && !(phi1 < 0.0)) return phi0 / (phi0 - phi1);
    if (!(phi0 < 0.0) && phi1 < 0.0) return phi1 / (phi1 - phi0);
    return 0;
}

// BlockedBoundryConditionSolver::ConstrainVelocity,
// src/fluid_solvers/blocked_boundry_condition_solver.cpp:18-291, collider
// velocity 0 and friction 0 as in the reference:
//  (1) faces whose collider fraction FractionInsideSdf(phi(x - h/2), phi(x + h/2))
//      is 1 are zeroed and marked invalid (:38-96);
//  (2) ExtrapolateToRegion from the valid faces, depth iterations (:98-103);
//  (3) faces inside the collider take the tangential part of the sampled
//      velocity, normal = normalised collider gradient, or 0 where the
//      gradient vanishes (:105-183; ApplyFriction with friction 0 is the
//      projection v - (v.n) n, src/common/vector3.hpp:160-163);
//  (4) the wall layers are zeroed (:198-244);
//  (5) faces between a collider node (colliderSdf(i,j,k) < 0, read at the
//      index) and a non-collider node are zeroed (:246-290).
// With no collider inside the domain (collider SDF >= 0 everywhere) steps 1,
// 2, 3 and 5 are identities and only the walls remain.
void constrain_velocity(orc_scene* s, int depth) {
    for (int c = 0; c < 3; ++c) {
        Arr3& a = comp_arr(s, c);
        Arr3i marker;
        marker.resize(a.sx, a.sy, a.sz, 1);
        const double* co = comp_origin(s, c);
        bool any_blocked = false;
        for (int k = 0; k < a.sz; ++k)
            for (int j = 0; j < a.sy; ++j)
                for (int i = 0; i < a.sx; ++i) {
                    double pos[3] = {co[0] + s->h[0] * i, co[1] + s->h[1] * j, co[2] + s->h[2] * k};
                    double p0[3] = {pos[0], pos[1], pos[2]}, p1[3] = {pos[0], pos[1], pos[2]};
                    p0[c] -= 0.5 * s->h[c];
                    p1[c] += 0.5 * s->h[c];
                    double frac = fraction_inside(sample_collider(s, p0), sample_collider(s, p1));
                    frac = 1 - clamp_ref(frac, 1.0, 0.0);
                    if (!(frac > 0.0)) {
                        a(i, j, k) = 0.0;
                        marker(i, j, k) = 0;
                        any_blocked = true;
                    }
                }
        if (any_blocked) {
            Arr3 prev = a;
            extrapolate_to_region(prev, marker, depth, a, s->nthreads);
        }
    }
    if (s->has_collider) {
        Arr3 tmp[3] = {s->u, s->v, s->w};
        for (int c = 0; c < 3; ++c) {
            const Arr3& a = comp_arr(s, c);
            const double* co = comp_origin(s, c);
            par_for(s->nthreads, a.sz, [&](int64_t k0, int64_t k1) {
                for (int k = (int)k0; k < (int)k1; ++k)
                    for (int j = 0; j < a.sy; ++j)
                        for (int i = 0; i < a.sx; ++i) {
                            const double pos[3] = {co[0] + s->h[0] * i, co[1] + s->h[1] * j, co[2] + s->h[2] * k};
                            if (!(sample_collider(s, pos) < 0)) continue;
                            double vel[3], g[3];
                            sample_vel(s, s->u, s->v, s->w, pos, vel);
                            collider_gradient(s, pos, g);
                            const double len = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
                            double out = 0.0;
                            if (len * len > 0.0) {
                                const double n[3] = {g[0] / len, g[1] / len, g[2] / len};
                                const double d = vel[0] * n[0] + vel[1] * n[1] + vel[2] * n[2];
                                out = vel[c] - n[c] * d;
                            }
                            tmp[c](i, j, k) = out;
                        }
            });
        }
        s->u = tmp[0];
        s->v = tmp[1];
        s->w = tmp[2];
    }
    // walls (:198-244)
    for (int k = 0; k < s->u.sz; ++k)
        for (int j = 0; j < s->u.sy; ++j) { s->u(0, j, k) = 0; s->u(s->u.sx - 1, j, k) = 0; }
    for (int k = 0; k < s->v.sz; ++k)
        for (int i = 0; i < s->v.sx; ++i) { s->v(i, 0, k) = 0; s->v(i, s->v.sy - 1, k) = 0; }
    for (int j = 0; j < s->w.sy; ++j)
        for (int i = 0; i < s->w.sx; ++i) { s->w(i, j, 0) = 0; s->w(i, j, s->w.sz - 1) = 0; }
    if (s->has_collider) {  // (:246-290)
        const int nx = s->n[0], ny = s->n[1], nz = s->n[2];
        auto solid = [&](int i, int j, int k) { return s->csdf(i, j, k) < 0; };
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    if (!solid(i, j, k)) continue;
                    if (i > 0 && !solid(i - 1, j, k)) s->u(i, j, k) = 0;
                    if (i < nx - 1 && !solid(i + 1, j, k)) s->u(i + 1, j, k) = 0;
                    if (j > 0 && !solid(i, j - 1, k)) s->v(i, j, k) = 0;
                    if (j < ny - 1 && !solid(i, j + 1, k)) s->v(i, j + 1, k) = 0;
                    if (k > 0 && !solid(i, j, k - 1)) s->w(i, j, k) = 0;
                    if (k < nz - 1 && !solid(i, j, k + 1)) s->w(i, j, k + 1) = 0;
                }
    }
}

// ExternalForce::ApplyExternalForce with a DirectionalField (src/forces/external_force.cpp:4-30,
// src/forces/directional_field.cpp:13-16)
void apply_gravity(orc_scene* s, double dt) {
    for (int c = 0; c < 3; ++c) {
        Arr3& a = comp_arr(s, c);
        const double g = s->gravity[c];
        par_for(s->nthreads, (int64_t)a.size(), [&](int64_t i0, int64_t i1) {
            for (int64_t i = i0; i < i1; ++i) a.d[i] += dt * g;
        });
    }
}

// ---------------------------------------------------------------------------
// SinglePhasePressureSolver, src/fluid_solvers/single_phase_pressure_solver.cpp
// ---------------------------------------------------------------------------
// BuildMarkers (:30-48): FLUID if fluidSdf(cell centre) < 0, else BOUNDRY if
// colliderSdf(cell centre) < 0, else AIR.  Both are trilinear samples
// (ScalarGrid3D::Sample) at GetCellCenterPos (face_centered_grid3d.cpp:318-321).
void build_markers(orc_scene* s) {
    const int nx = s->n[0], ny = s->n[1], nz = s->n[2];
    s->flags.assign((size_t)nx * ny * nz, ORC_AIR);
    const double so[3] = {s->origin[0] + s->h[0] / 2.0, s->origin[1] + s->h[1] / 2.0, s->origin[2] + s->h[2] / 2.0};
    par_for(s->nthreads, nz, [&](int64_t k0, int64_t k1) {
        for (int k = (int)k0; k < (int)k1; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    double cc[3];
                    for (int a = 0; a < 3; ++a) {
                        int id = a == 0 ? i : (a == 1 ? j : k);
                        cc[a] = s->origin[a] + 0.5 * s->h[a] + s->h[a] * id;
                    }
                    size_t id = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
                    if (sample_array(s->sdf, so, s->h, cc) < 0)
                        s->flags[id] = ORC_FLUID;
                    else if (sample_collider(s, cc) < 0)
                        s->flags[id] = ORC_BOUNDRY;
                    else
                        s->flags[id] = ORC_AIR;
                }
    });
}

inline uint8_t flag_at(const orc_scene* s, int i, int j, int k) {
    return s->flags[(size_t)i + (size_t)s->n[0] * ((size_t)j + (size_t)s->n[1] * k)];
}

// FaceCenteredGrid3D::DivergenceAtCallCenter (face_centered_grid3d.cpp:157-166)
inline double divergence(const orc_scene* s, int i, int j, int k) {
    double left = s->u(i, j, k), right = s->u(i + 1, j, k);
    double down = s->v(i, j, k), up = s->v(i, j + 1, k);
    double back = s->w(i, j, k), front = s->w(i, j, k + 1);
    return (right - left) / s->h[0] + (up - down) / s->h[1] + (front - back) / s->h[2];
}

// BuildSystem (:99-164)
void build_system(orc_scene* s) {
    const int nx = s->n[0], ny = s->n[1], nz = s->n[2];
    const double ihx2 = (1.0 / s->h[0]) * (1.0 / s->h[0]);
    const double ihy2 = (1.0 / s->h[1]) * (1.0 / s->h[1]);
    const double ihz2 = (1.0 / s->h[2]) * (1.0 / s->h[2]);
    // LinearSystem::Resize keeps x (warm start); A and b are rebuilt.
    if (s->p.sx != nx || s->p.sy != ny || s->p.sz != nz) s->p.resize(nx, ny, nz, 0.0);
    s->b.resize(nx, ny, nz, 0.0);
    s->Ac.resize(nx, ny, nz, 0.0); s->Ar.resize(nx, ny, nz, 0.0);
    s->Au.resize(nx, ny, nz, 0.0); s->Af.resize(nx, ny, nz, 0.0);
    par_for(s->nthreads, nz, [&](int64_t k0, int64_t k1) {
        for (int k = (int)k0; k < (int)k1; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    double c = 0, r = 0, u = 0, f = 0, bb = 0;
                    if (flag_at(s, i, j, k) == ORC_FLUID) {
                        bb = divergence(s, i, j, k);
                        if (i + 1 < nx && flag_at(s, i + 1, j, k) != ORC_BOUNDRY) {
                            c += ihx2;
                            if (flag_at(s, i + 1, j, k) == ORC_FLUID) r -= ihx2;
                        }
                        if (i > 0 && flag_at(s, i - 1, j, k) != ORC_BOUNDRY) c += ihx2;
                        if (j + 1 < ny && flag_at(s, i, j + 1, k) != ORC_BOUNDRY) {
                            c += ihy2;
                            if (flag_at(s, i, j + 1, k) == ORC_FLUID) u -= ihy2;
                        }
                        if (j > 0 && flag_at(s, i, j - 1, k) != ORC_BOUNDRY) c += ihy2;
                        if (k + 1 < nz && flag_at(s, i, j, k + 1) != ORC_BOUNDRY) {
                            c += ihz2;
                            if (flag_at(s, i, j, k + 1) == ORC_FLUID) f -= ihz2;
                        }
                        if (k > 0 && flag_at(s, i, j, k - 1) != ORC_BOUNDRY) c += ihz2;
                    } else {
                        c = 1.0;
                    }
                    s->Ac(i, j, k) = c; s->Ar(i, j, k) = r; s->Au(i, j, k) = u; s->Af(i, j, k) = f;
                    s->b(i, j, k) = bb;
                }
    });
}

// BLAS::MatrixVectorMultiplication (src/linear_system/blas.cpp:72-83)
void spmv(const orc_scene* s, const Arr3& v, Arr3& out, int nthreads) {
    const int nx = v.sx, ny = v.sy, nz = v.sz;
    par_for(nthreads, nz, [&](int64_t k0, int64_t k1) {
        for (int k = (int)k0; k < (int)k1; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    out(i, j, k) = s->Ac(i, j, k) * v(i, j, k) +
                                   ((i > 0) ? s->Ar(i - 1, j, k) * v(i - 1, j, k) : 0.0) +
                                   ((i + 1 < nx) ? s->Ar(i, j, k) * v(i + 1, j, k) : 0.0) +
                                   ((j > 0) ? s->Au(i, j - 1, k) * v(i, j - 1, k) : 0.0) +
                                   ((j + 1 < ny) ? s->Au(i, j, k) * v(i, j + 1, k) : 0.0) +
                                   ((k > 0) ? s->Af(i, j, k - 1) * v(i, j, k - 1) : 0.0) +
                                   ((k + 1 < nz) ? s->Af(i, j, k) * v(i, j, k + 1) : 0.0);
                }
    });
}

// BLAS::Residual (src/linear_system/blas.cpp:21-37)
void residual(const orc_scene* s, const Arr3& x, const Arr3& b, Arr3& out, int nthreads) {
    const int nx = x.sx, ny = x.sy, nz = x.sz;
    par_for(nthreads, nz, [&](int64_t k0, int64_t k1) {
        for (int k = (int)k0; k < (int)k1; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    out(i, j, k) = b(i, j, k) - s->Ac(i, j, k) * x(i, j, k) -
                                   ((i > 0) ? s->Ar(i - 1, j, k) * x(i - 1, j, k) : 0.0) -
                                   ((i + 1 < nx) ? s->Ar(i, j, k) * x(i + 1, j, k) : 0.0) -
                                   ((j > 0) ? s->Au(i, j - 1, k) * x(i, j - 1, k) : 0.0) -
                                   ((j + 1 < ny) ? s->Au(i, j, k) * x(i, j + 1, k) : 0.0) -
                                   ((k > 0) ? s->Af(i, j, k - 1) * x(i, j, k - 1) : 0.0) -
                                   ((k + 1 < nz) ? s->Af(i, j, k) * x(i, j, k + 1) : 0.0);
                }
    });
}

// BLAS::Dot (src/linear_system/blas.cpp:4-19): serial, i-outer / k-inner order.
double dot(const Arr3& a, const Arr3& b) {
    double r = 0;
    for (int i = 0; i < a.sx; ++i)
        for (int j = 0; j < a.sy; ++j)
            for (int k = 0; k < a.sz; ++k) r += a(i, j, k) * b(i, j, k);
    return r;
}

// The same sum in storage order over 64 fixed blocks added in block order:
// deterministic for any thread count, cache-friendly.  Only for tests that
// need a converged answer on large grids (orc_set_fast_dots); the default is
// the reference's order above.
double dot_blocked(const Arr3& a, const Arr3& b, int nthreads) {
    constexpr int kBlocks = 64;
    const int64_t n = (int64_t)a.size();
    double part[kBlocks] = {0};
    std::vector<std::thread> th;
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int q; (q = next.fetch_add(1)) < kBlocks;) {
            const int64_t i0 = n * q / kBlocks, i1 = n * (q + 1) / kBlocks;
            double r = 0;
            for (int64_t i = i0; i < i1; ++i) r += a.d[i] * b.d[i];
            part[q] = r;
        }
    };
    for (int t = 1; t < std::max(1, nthreads); ++t) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
    double r = 0;
    for (int q = 0; q < kBlocks; ++q) r += part[q];
    return r;
}

double sdot(const orc_scene* s, const Arr3& a, const Arr3& b) {
    return s->fast_dots ? dot_blocked(a, b, s->nthreads) : dot(a, b);
}

// JacobiIterationSolver (src/linear_system/jacobi_iteration_solver.cpp:14-76)
void jacobi(orc_scene* s, const orc_solver_desc* d, orc_solve_stats* st) {
    const int nx = s->n[0], ny = s->n[1], nz = s->n[2];
    Arr3 xt, res;
    xt.resize(nx, ny, nz, 0.0);
    res.resize(nx, ny, nz, 0.0);
    int it = 0;
    int executed = 0;
    for (it = 0; it < d->max_iterations; ++it) {
        Arr3& x = s->p;
        par_for(s->nthreads, nz, [&](int64_t k0, int64_t k1) {
            for (int k = (int)k0; k < (int)k1; ++k)
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i) {
                        double r = ((i > 0) ? s->Ar(i - 1, j, k) * x(i - 1, j, k) : 0.0) +
                                   ((i + 1 < nx) ? s->Ar(i, j, k) * x(i + 1, j, k) : 0.0) +
                                   ((j > 0) ? s->Au(i, j - 1, k) * x(i, j - 1, k) : 0.0) +
                                   ((j + 1 < ny) ? s->Au(i, j, k) * x(i, j + 1, k) : 0.0) +
                                   ((k > 0) ? s->Af(i, j, k - 1) * x(i, j, k - 1) : 0.0) +
                                   ((k + 1 < nz) ? s->Af(i, j, k) * x(i, j, k + 1) : 0.0);
                        xt(i, j, k) = (s->b(i, j, k) - r) / s->Ac(i, j, k);
                    }
        });
        std::swap(s->p.d, xt.d);
        executed = it + 1;
        if (it != 0 && d->check_interval > 0 && it % d->check_interval == 0) {
            residual(s, s->p, s->b, res, s->nthreads);
            double nrm = std::sqrt(sdot(s, res, res));
            if (nrm < d->tolerance) break;
        }
    }
    st->iterations = executed;
}

// ConjugateGradientSolver::Solve (src/linear_system/conjugate_gradient_solver.cpp:14-55),
// with the identity "preconditioner" (:57-63) or, as a build addition, the
// diagonal (Jacobi) preconditioner.  Stopping rule: the reference's absolute
// sigma <= tol^2, or (relative mode) ||r||_2 <= tol * ||b||_2.
void cg(orc_scene* s, const orc_solver_desc* d, orc_solve_stats* st) {
    const int nx = s->n[0], ny = s->n[1], nz = s->n[2];
    Arr3 r, dd, q, z;
    r.resize(nx, ny, nz); dd.resize(nx, ny, nz); q.resize(nx, ny, nz); z.resize(nx, ny, nz);
    const bool precond = d->kind == ORC_SOLVER_PCG_JACOBI;
    const int nt = s->nthreads;
    // BLAS::AXpY and Preconditioner's ParallelFill are parallel loops in the
    // reference (src/linear_system/blas.cpp:39-47, conjugate_gradient_solver.cpp:57-63)
    auto pfor = [&](const std::function<void(size_t)>& f) {
        par_for(nt, (int64_t)s->p.size(), [&](int64_t a, int64_t b) {
            for (int64_t i = a; i < b; ++i) f((size_t)i);
        });
    };
    auto apply_m = [&](const Arr3& in, Arr3& out) {
        if (!precond) {
            pfor([&](size_t i) { out.d[i] = in.d[i]; });
            return;
        }
        pfor([&](size_t i) { out.d[i] = in.d[i] / s->Ac.d[i]; });
    };
    if (!d->warm_start) std::fill(s->p.d.begin(), s->p.d.end(), 0.0);
    residual(s, s->p, s->b, r, s->nthreads);
    apply_m(r, dd);
    double sigmaNew = sdot(s, r, dd);
    const double bnorm = std::sqrt(sdot(s, s->b, s->b));
    auto done = [&](double sig, const Arr3& rr) {
        if (d->relative) {
            double rn = precond ? std::sqrt(sdot(s, rr, rr)) : std::sqrt(std::max(sig, 0.0));
            return rn <= d->tolerance * bnorm;
        }
        return !(sig > d->tolerance * d->tolerance);
    };
    int it = 0;
    bool trigger = false;
    while (!done(sigmaNew, r) && it < d->max_iterations) {
        spmv(s, dd, q, s->nthreads);
        double alpha = sigmaNew / sdot(s, dd, q);
        pfor([&](size_t i) { s->p.d[i] = alpha * dd.d[i] + s->p.d[i]; });
        if (trigger || (it % 50 == 0 && it > 0)) {
            residual(s, s->p, s->b, r, s->nthreads);
            trigger = false;
        } else {
            pfor([&](size_t i) { r.d[i] = -alpha * q.d[i] + r.d[i]; });
        }
        apply_m(r, z);
        double sigmaOld = sigmaNew;
        sigmaNew = sdot(s, r, z);
        if (sigmaNew > sigmaOld) trigger = true;
        double beta = sigmaNew / sigmaOld;
        pfor([&](size_t i) { dd.d[i] = beta * dd.d[i] + z.d[i]; });
        ++it;
    }
    st->iterations = it;
}

// ApplyPressure (:56-92)
void apply_pressure(orc_scene* s) {
    const int nx = s->n[0], ny = s->n[1], nz = s->n[2];
    const double scx = 1.0 / s->h[0], scy = 1.0 / s->h[1], scz = 1.0 / s->h[2];
    Arr3 iu = s->u, iv = s->v, iw = s->w;
    const Arr3& P = s->p;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                if (flag_at(s, i, j, k) != ORC_FLUID) continue;
                if (i + 1 < nx && flag_at(s, i + 1, j, k) != ORC_BOUNDRY)
                    s->u(i + 1, j, k) = iu(i + 1, j, k) + scx * (P(i + 1, j, k) - P(i, j, k));
                if (j + 1 < ny && flag_at(s, i, j + 1, k) != ORC_BOUNDRY)
                    s->v(i, j + 1, k) = iv(i, j + 1, k) + scy * (P(i, j + 1, k) - P(i, j, k));
                if (k + 1 < nz && flag_at(s, i, j, k + 1) != ORC_BOUNDRY)
                    s->w(i, j, k + 1) = iw(i, j, k + 1) + scz * (P(i, j, k + 1) - P(i, j, k));
            }
}

void pressure(orc_scene* s, const orc_solver_desc* d, orc_solve_stats* st) {
    auto t0 = std::chrono::steady_clock::now();
    build_markers(s);
    build_system(s);
    orc_solve_stats local{};
    if (d->kind == ORC_SOLVER_JACOBI)
        jacobi(s, d, &local);
    else
        cg(s, d, &local);
    // final residual in FP64
    Arr3 r;
    r.resize(s->n[0], s->n[1], s->n[2]);
    residual(s, s->p, s->b, r, s->nthreads);
    double rn = std::sqrt(sdot(s, r, r)), bn = std::sqrt(sdot(s, s->b, s->b));
    apply_pressure(s);
    auto t1 = std::chrono::steady_clock::now();
    local.abs_residual = rn;
    local.rel_residual = bn > 0 ? rn / bn : rn;
    local.ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    s->last = local;
    if (st) *st = local;
}

void apply_bc(orc_scene* s) { constrain_velocity(s, static_cast<int>(std::ceil(s->maxCfl))); }

// ---------------------------------------------------------------------------
// BackwardEulerDiffusionSolver::Solve (src/fluid_solvers/backward_euler_diffusion_solver.cpp:20-49),
// as PICSimulator::ComputeDiffusion calls it (src/pic_simulator.cpp:214-222).
// The call passes (dt, viscosity) into parameters named (viscosity, dt); the
// coefficient c = (dt * viscosity) / (h * h) is the same product either way.
// ---------------------------------------------------------------------------
// BuildX/Y/ZMarkers (:71-128): FLUID iff fluidSdf.Sample(face position) < 0,
// else BOUNDRY iff colliderSdf.Sample(face position) < 0, else AIR; face
// position GetX/Y/ZPos = data origin + h * (i, j, k) (face_centered_grid3d.cpp:303-316).
void diffusion_markers(orc_scene* s, int c) {
    const Arr3& a = comp_arr(s, c);
    const double* fo = comp_origin(s, c);
    const double so[3] = {s->origin[0] + s->h[0] / 2.0, s->origin[1] + s->h[1] / 2.0, s->origin[2] + s->h[2] / 2.0};
    std::vector<uint8_t>& m = s->dmark[c];
    m.assign(a.size(), ORC_AIR);
    par_for(s->nthreads, a.sz, [&](int64_t k0, int64_t k1) {
        for (int k = (int)k0; k < (int)k1; ++k)
            for (int j = 0; j < a.sy; ++j)
                for (int i = 0; i < a.sx; ++i) {
                    const double pos[3] = {fo[0] + s->h[0] * (double)i, fo[1] + s->h[1] * (double)j,
                                           fo[2] + s->h[2] * (double)k};
                    uint8_t f = ORC_AIR;
                    if (sample_array(s->sdf, so, s->h, pos) < 0)
                        f = ORC_FLUID;
                    else if (sample_collider(s, pos) < 0)
                        f = ORC_BOUNDRY;
                    m[a.at(i, j, k)] = f;
                }
    });
}

// BuildMatrix (:136-198) and BuildVectors (:200-242) into a row system on the
// component's face grid: a FLUID row has centre 1 + c_a per in-range non-AIR
// neighbour (+ neighbour first, then -, axes x, y, z), -c_a to each FLUID +
// neighbour (right/up/front), and b = u + c_a * u(neighbour) per BOUNDRY
// neighbour; every other row is the identity with b = u.  x starts at u.
void diffusion_system(orc_scene* s, int c, const double cc[3], orc_scene& sys) {
    const Arr3& a = comp_arr(s, c);
    const std::vector<uint8_t>& m = s->dmark[c];
    const int nx = a.sx, ny = a.sy, nz = a.sz;
    sys.n[0] = nx; sys.n[1] = ny; sys.n[2] = nz;
    sys.nthreads = s->nthreads;
    sys.Ac.resize(nx, ny, nz, 0.0); sys.Ar.resize(nx, ny, nz, 0.0);
    sys.Au.resize(nx, ny, nz, 0.0); sys.Af.resize(nx, ny, nz, 0.0);
    sys.b.resize(nx, ny, nz, 0.0); sys.p.resize(nx, ny, nz, 0.0);
    auto M = [&](int i, int j, int k) { return m[a.at(i, j, k)]; };
    par_for(s->nthreads, nz, [&](int64_t k0, int64_t k1) {
        for (int k = (int)k0; k < (int)k1; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    double center = 1.0, right = 0.0, up = 0.0, front = 0.0;
                    double b = a(i, j, k);
                    if (M(i, j, k) == ORC_FLUID) {
                        if (i + 1 < nx) {
                            if (M(i + 1, j, k) != ORC_AIR) center += cc[0];
                            if (M(i + 1, j, k) == ORC_FLUID) right -= cc[0];
                        }
                        if (i > 0 && M(i - 1, j, k) != ORC_AIR) center += cc[0];
                        if (j + 1 < ny) {
                            if (M(i, j + 1, k) != ORC_AIR) center += cc[1];
                            if (M(i, j + 1, k) == ORC_FLUID) up -= cc[1];
                        }
                        if (j > 0 && M(i, j - 1, k) != ORC_AIR) center += cc[1];
                        if (k + 1 < nz) {
                            if (M(i, j, k + 1) != ORC_AIR) center += cc[2];
                            if (M(i, j, k + 1) == ORC_FLUID) front -= cc[2];
                        }
                        if (k > 0 && M(i, j, k - 1) != ORC_AIR) center += cc[2];
                        if (i + 1 < nx && M(i + 1, j, k) == ORC_BOUNDRY) b += cc[0] * a(i + 1, j, k);
                        if (i > 0 && M(i - 1, j, k) == ORC_BOUNDRY) b += cc[0] * a(i - 1, j, k);
                        if (j + 1 < ny && M(i, j + 1, k) == ORC_BOUNDRY) b += cc[1] * a(i, j + 1, k);
                        if (j > 0 && M(i, j - 1, k) == ORC_BOUNDRY) b += cc[1] * a(i, j - 1, k);
                        if (k + 1 < nz && M(i, j, k + 1) == ORC_BOUNDRY) b += cc[2] * a(i, j, k + 1);
                        if (k > 0 && M(i, j, k - 1) == ORC_BOUNDRY) b += cc[2] * a(i, j, k - 1);
                    }
                    sys.Ac(i, j, k) = center; sys.Ar(i, j, k) = right;
                    sys.Au(i, j, k) = up; sys.Af(i, j, k) = front;
                    sys.b(i, j, k) = b;
                    sys.p(i, j, k) = a(i, j, k);
                }
    });
}

void diffusion(orc_scene* s, double dt) {
    const double cc[3] = {(dt * s->viscosity) / (s->h[0] * s->h[0]), (dt * s->viscosity) / (s->h[1] * s->h[1]),
                          (dt * s->viscosity) / (s->h[2] * s->h[2])};
    for (int c = 0; c < 3; ++c) {
        diffusion_markers(s, c);
        orc_scene sys{};
        diffusion_system(s, c, cc, sys);
        orc_solve_stats st{};
        if (s->diff.kind == ORC_SOLVER_JACOBI)
            jacobi(&sys, &s->diff, &st);
        else
            cg(&sys, &s->diff, &st);
        comp_arr(s, c).d = sys.p.d;  // output->x/y/z(i, j, k) = _system.x(i, j, k)
        s->diff_last[c] = st;
    }
}

// ---------------------------------------------------------------------------
// MarchingCubesSolver::BuildSurface (src/3d/marching_cubes_solver.cpp:25-51):
// cubes in the reference's i, j, k loop order with an edge-keyed vertex map
// (SolveSingleCube, :172-270), then caps on the boundary node planes, one
// vertex map per plane (SolveXY / SolveYZ / SolveXZ / SolveSingleSquare,
// :272-589).  The triangulation of a cube follows one face rule (restated
// here by tracing each cube's loops; the reference's 256-entry table is not
// used): on every cube face the crossed edges are joined in pairs, cutting
// diagonally opposite inside corners off on their own; the loops are wound
// counter-clockwise seen from outside and fanned from the first position none
// of whose triangles lies in one cube face.  The same rule caps the planes.
// ---------------------------------------------------------------------------
struct McMesh {
    std::vector<double> v, n;
    std::vector<int32_t> t;
};

// Cube corner c has offsets (c & 1, c >> 1 & 1, c >> 2 & 1); the corner
// pairs of the 12 edges, lower corner first, x edges then y then z.
void mc_edge_pair(int e, int& a, int& b) {
    static const int pairs[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3},
                                     {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
    a = pairs[e][0];
    b = pairs[e][1];
}

int mc_edge_between(int a, int b) {
    for (int e = 0; e < 12; ++e) {
        int p, q;
        mc_edge_pair(e, p, q);
        if ((p == a && q == b) || (p == b && q == a)) return e;
    }
    return -1;
}

// The four corners of cube face f (axis f/2, side f%2) in cyclic order in the
// plane of the two other axes, lower axis first.
void mc_face(int f, int q[4]) {
    const int axis = f / 2, side = f % 2;
    const int u = axis == 0 ? 1 : 0, w = axis == 2 ? 1 : 2;
    const int cyc[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
    for (int t = 0; t < 4; ++t) q[t] = (side << axis) | (cyc[t][0] << u) | (cyc[t][1] << w);
}

bool mc_in_face(int f, int e) {
    int q[4];
    mc_face(f, q);
    for (int t = 0; t < 4; ++t)
        if (mc_edge_between(q[t], q[(t + 1) % 4]) == e) return true;
    return false;
}

// Loops of the crossed edges of a cube with inside flags in[8], each as a
// triangle fan (edge indices, counter-clockwise from outside).
std::vector<int> mc_cube_triangles(const bool in[8]) {
    std::vector<int> link[12];
    for (int f = 0; f < 6; ++f) {
        int q[4];
        mc_face(f, q);
        int crossed[4], nc = 0;
        for (int t = 0; t < 4; ++t)
            if (in[q[t]] != in[q[(t + 1) % 4]]) crossed[nc++] = t;
        auto join = [&](int ta, int tb) {
            const int ea = mc_edge_between(q[ta], q[(ta + 1) % 4]), eb = mc_edge_between(q[tb], q[(tb + 1) % 4]);
            link[ea].push_back(eb);
            link[eb].push_back(ea);
        };
        if (nc == 2) join(crossed[0], crossed[1]);
        if (nc == 4)
            for (int m = 0; m < 4; ++m)
                if (in[q[m]]) join((m + 3) % 4, m);  // the inside corner's two edges
    }
    std::vector<int> tris;
    bool seen[12] = {false};
    for (int e0 = 0; e0 < 12; ++e0) {
        if (link[e0].size() != 2 || seen[e0]) continue;
        std::vector<int> loop;
        int prev = -1, cur = e0;
        do {
            loop.push_back(cur);
            seen[cur] = true;
            const int nx = link[cur][0] != prev ? link[cur][0] : link[cur][1];
            prev = cur;
            cur = nx;
        } while (cur != e0);
        const int L = (int)loop.size();
        double N[3] = {0, 0, 0}, D[3] = {0, 0, 0};
        std::vector<std::array<double, 3>> mid(L);
        for (int t = 0; t < L; ++t) {
            int a, b;
            mc_edge_pair(loop[t], a, b);
            const int ci = in[a] ? a : b, co = in[a] ? b : a;
            for (int x = 0; x < 3; ++x) {
                mid[t][x] = 0.5 * (((a >> x) & 1) + ((b >> x) & 1));
                D[x] += ((co >> x) & 1) - ((ci >> x) & 1);
            }
        }
        for (int t = 0; t < L; ++t) {  // Newell's normal
            const auto& p = mid[t];
            const auto& r = mid[(t + 1) % L];
            N[0] += (p[1] - r[1]) * (p[2] + r[2]);
            N[1] += (p[2] - r[2]) * (p[0] + r[0]);
            N[2] += (p[0] - r[0]) * (p[1] + r[1]);
        }
        if (N[0] * D[0] + N[1] * D[1] + N[2] * D[2] < 0) std::reverse(loop.begin() + 1, loop.end());
        int start = 0;
        for (int s0 = 0; s0 < L; ++s0) {
            bool ok = true;
            for (int t = 1; t + 1 < L && ok; ++t)
                for (int f = 0; f < 6 && ok; ++f)
                    ok = !(mc_in_face(f, loop[s0]) && mc_in_face(f, loop[(s0 + t) % L]) &&
                           mc_in_face(f, loop[(s0 + t + 1) % L]));
            if (ok) {
                start = s0;
                break;
            }
        }
        for (int t = 1; t + 1 < L; ++t) {
            tris.push_back(loop[start]);
            tris.push_back(loop[(start + t) % L]);
            tris.push_back(loop[(start + t + 1) % L]);
        }
    }
    return tris;
}

// MarchingCubesSolver::Gradient (:116-160)
void mc_gradient(const Arr3& sdf, const double ih[3], int i, int j, int k, double g[3]) {
    const int idx[3] = {i, j, k}, dim[3] = {sdf.sx, sdf.sy, sdf.sz};
    for (int a = 0; a < 3; ++a) {
        int p[3] = {i, j, k}, m[3] = {i, j, k};
        p[a] = idx[a] + 1;
        m[a] = idx[a] - 1;
        if (idx[a] > dim[a] - 2)
            p[a] = idx[a];
        else if (idx[a] == 0)
            m[a] = 0;
        g[a] = 0.5 * ih[a] * (sdf(p[0], p[1], p[2]) - sdf(m[0], m[1], m[2]));
    }
}

// The mesh in malloc'd arrays for the C ABI (freed by orc_free).
int mesh_out(const McMesh& M, double** vout, double** nout, int32_t** tout, int64_t* nv, int64_t* nt) {
    *nv = (int64_t)M.v.size() / 3;
    *nt = (int64_t)M.t.size() / 3;
    *vout = (double*)std::malloc(sizeof(double) * std::max<size_t>(M.v.size(), 1));
    *nout = (double*)std::malloc(sizeof(double) * std::max<size_t>(M.n.size(), 1));
    *tout = (int32_t*)std::malloc(sizeof(int32_t) * std::max<size_t>(M.t.size(), 1));
    std::copy(M.v.begin(), M.v.end(), *vout);
    std::copy(M.n.begin(), M.n.end(), *nout);
    std::copy(M.t.begin(), M.t.end(), *tout);
    return 0;
}

// ---------------------------------------------------------------------------
// MarchingCubesSolver::BuildSurface with the reference's own case tables
// (src/3d/marching_cubes_solver.cpp:25-609, tables :611-952 in
// fluide_engine_amd/csrc/fe_mc_table.hpp), restated step by step: cubes in
// i, j, k loop order (k innermost), the case from the corner order of
// SolveSingleCube, each crossed edge's vertex placed from the first cube that
// meets it (BoundingBox3D::Corner of that cube, the edge's table direction),
// triangles in table order; then the caps of SolveXY (back, front), SolveYZ
// (left, right) and SolveXZ (down, up), each plane with its own vertex map,
// corner order and loop order, triangulated by the square table.
// ---------------------------------------------------------------------------
int surface_reference(orc_scene* s, McMesh& M) {
    const int nx = s->n[0], ny = s->n[1], nz = s->n[2];
    const double so[3] = {s->origin[0] + s->h[0] / 2.0, s->origin[1] + s->h[1] / 2.0, s->origin[2] + s->h[2] / 2.0};
    const double ih[3] = {1.0 / s->h[0], 1.0 / s->h[1], 1.0 / s->h[2]};
    Arr3 sdf = s->sdf;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                if (s->sdf(i, j, k) < 0 && s->csdf(i, j, k) < 0) sdf(i, j, k) = s->csdf(i, j, k) * (-1);
    auto hexv = [](char c) { return c <= '9' ? c - '0' : c - 'a' + 10; };
    // the reference's cube corners as (x, y, z) offsets, and the bounding-box
    // corner each maps to (indexMap, :256)
    static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1},
                                  {0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}};
    static const int indexMap[8] = {0, 1, 5, 4, 2, 3, 7, 6};
    std::map<int64_t, int32_t> vmap;
    auto edge_key = [&](int axis, int i, int j, int k) {
        return (int64_t)axis + 3 * ((int64_t)i + (int64_t)(nx + 1) * ((int64_t)j + (int64_t)(ny + 1) * k));
    };
    for (int i = 0; i < nx - 1; ++i)
        for (int j = 0; j < ny - 1; ++j)
            for (int k = 0; k < nz - 1; ++k) {
                double data[8], nrm8[8][3];
                int cs = 0;
                for (int v = 0; v < 8; ++v) {
                    data[v] = sdf(i + off[v][0], j + off[v][1], k + off[v][2]);
                    mc_gradient(sdf, ih, i + off[v][0], j + off[v][1], k + off[v][2], nrm8[v]);
                    if (data[v] <= 0.0) cs |= 1 << v;
                }
                if (cs == 0 || cs == 255) continue;
                const double oc[3] = {so[0] + i * s->h[0], so[1] + j * s->h[1], so[2] + k * s->h[2]};
                const char* t = fe::kMcRefCube[cs];
                const int n = (int)std::strlen(t);
                for (int q = 0; q + 2 < n; q += 3) {
                    int32_t face[3];
                    for (int r = 0; r < 3; ++r) {
                        const int e = hexv(t[q + r]);
                        const int a = fe::kMcRefEdge[e][0], b = fe::kMcRefEdge[e][1];
                        int axis = 0;
                        for (int x = 0; x < 3; ++x)
                            if (off[a][x] != off[b][x]) axis = x;
                        const int64_t key = edge_key(axis, i + std::min(off[a][0], off[b][0]),
                                                     j + std::min(off[a][1], off[b][1]),
                                                     k + std::min(off[a][2], off[b][2]));
                        auto it = vmap.find(key);
                        if (it != vmap.end()) {
                            face[r] = it->second;
                            continue;
                        }
                        // DistanceToZeroLevelSet (:591-601) and the clamps (:263-268)
                        const double phi0 = data[a], phi1 = data[b];
                        double alpha = std::fabs(phi0) + std::fabs(phi1) > 0.000000001
                                           ? std::fabs(phi0) / (std::fabs(phi0) + std::fabs(phi1))
                                           : 0.5;
                        if (alpha < 0.000001) alpha = 0.000001;
                        if (alpha > 0.999999) alpha = 0.999999;
                        double nn[3];
                        for (int x = 0; x < 3; ++x) {
                            // BoundingBox3D::Corner (src/3d/bounding_box_3d.cpp:23-31)
                            const int ca = indexMap[a], cb = indexMap[b];
                            const double p0 = (s->h[x] * (((ca >> x) & 1) ? 0.5 : -0.5) + oc[x]) + 0.5 * s->h[x];
                            const double p1 = (s->h[x] * (((cb >> x) & 1) ? 0.5 : -0.5) + oc[x]) + 0.5 * s->h[x];
                            M.v.push_back((1.0 - alpha) * p0 + alpha * p1);
                            nn[x] = (1.0 - alpha) * nrm8[a][x] + alpha * nrm8[b][x];
                        }
                        const double len = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
                        for (int x = 0; x < 3; ++x) M.n.push_back(len * len > 0.0 ? nn[x] / len : nn[x]);
                        face[r] = (int32_t)(M.v.size() / 3 - 1);
                        vmap[key] = face[r];
                    }
                    M.t.insert(M.t.end(), face, face + 3);
                }
            }
    // caps: the reference's corner order per plane, as (i, j, k) offsets from
    // the square's low node, and its loops
    struct Plane {
        int axis, side;
        int corner[4][3];
    };
    static const Plane planes[6] = {
        {2, 0, {{1, 0, 0}, {0, 0, 0}, {0, 1, 0}, {1, 1, 0}}},  // back (SolveXY)
        {2, 1, {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}}},  // front
        {0, 0, {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0}}},  // left (SolveYZ)
        {0, 1, {{0, 0, 1}, {0, 0, 0}, {0, 1, 0}, {0, 1, 1}}},  // right
        {1, 0, {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}}},  // down (SolveXZ)
        {1, 1, {{1, 0, 0}, {0, 0, 0}, {0, 0, 1}, {1, 0, 1}}},  // up
    };
    const int dim[3] = {nx, ny, nz};
    for (const Plane& P : planes) {
        const int u = P.axis == 0 ? 1 : 0, w = P.axis == 2 ? 1 : 2;
        const double outward = P.side ? 1.0 : -1.0;
        std::map<int64_t, int32_t> pmap;
        // outer loop w (j for z planes, k otherwise), inner loop u
        for (int iw = 0; iw + 1 < dim[w]; ++iw)
            for (int iu = 0; iu + 1 < dim[u]; ++iu) {
                int node[4][3];
                double phi[4];
                int cs = 0;
                for (int t = 0; t < 4; ++t) {
                    int base[3];
                    base[P.axis] = P.side ? dim[P.axis] - 1 : 0;
                    base[u] = iu;
                    base[w] = iw;
                    for (int x = 0; x < 3; ++x) node[t][x] = base[x] + (x == P.axis ? 0 : P.corner[t][x]);
                    phi[t] = sdf(node[t][0], node[t][1], node[t][2]);
                    if (phi[t] <= 0.0) cs |= 1 << t;
                }
                if (cs == 0) continue;
                const char* t = fe::kMcRefSquare[cs];
                const int n = (int)std::strlen(t);
                for (int q = 0; q + 2 < n; q += 3) {
                    int32_t face[3];
                    for (int r = 0; r < 3; ++r) {
                        const int slot = hexv(t[q + r]);
                        int64_t key;
                        double p[3];
                        if (slot < 4) {
                            key = 4 * ((int64_t)node[slot][0] + (int64_t)(nx + 1) *
                                       ((int64_t)node[slot][1] + (int64_t)(ny + 1) * node[slot][2]));
                            for (int x = 0; x < 3; ++x) p[x] = so[x] + node[slot][x] * s->h[x];
                        } else {
                            const int e0 = slot - 4, e1 = (e0 + 1) % 4;  // _edgeConnection2D
                            int lo[3], axis = 0;
                            for (int x = 0; x < 3; ++x) {
                                lo[x] = std::min(node[e0][x], node[e1][x]);
                                if (node[e0][x] != node[e1][x]) axis = x;
                            }
                            key = 4 * ((int64_t)lo[0] + (int64_t)(nx + 1) * ((int64_t)lo[1] + (int64_t)(ny + 1) * lo[2])) +
                                  1 + axis;
                            // SolveSingleSquare (:524-546)
                            const double f0 = std::fabs(phi[e0]), f1 = std::fabs(phi[e1]);
                            double alpha = f0 + f1 > 1e-12 ? f0 / (f0 + f1) : 0.5f;
                            if (alpha < 0.000001f) alpha = 0.000001f;
                            if (alpha > 0.999999f) alpha = 0.999999f;
                            for (int x = 0; x < 3; ++x)
                                p[x] = (1.f - alpha) * (so[x] + node[e0][x] * s->h[x]) +
                                       alpha * (so[x] + node[e1][x] * s->h[x]);
                        }
                        auto it = pmap.find(key);
                        if (it == pmap.end()) {
                            for (int x = 0; x < 3; ++x) {
                                M.v.push_back(p[x]);
                                M.n.push_back(x == P.axis ? outward : 0.0);
                            }
                            it = pmap.emplace(key, (int32_t)(M.v.size() / 3 - 1)).first;
                        }
                        face[r] = it->second;
                    }
                    M.t.insert(M.t.end(), face, face + 3);
                }
            }
    }
    return 0;
}

int surface(orc_scene* s, double** vout, double** nout, int32_t** tout, int64_t* nv, int64_t* nt) {
    if (s->surface_mode == ORC_SURFACE_REFERENCE) {
        McMesh R;
        surface_reference(s, R);
        return mesh_out(R, vout, nout, tout, nv, nt);
    }
    const int nx = s->n[0], ny = s->n[1], nz = s->n[2];
    const double so[3] = {s->origin[0] + s->h[0] / 2.0, s->origin[1] + s->h[1] / 2.0, s->origin[2] + s->h[2] / 2.0};
    const double ih[3] = {1.0 / s->h[0], 1.0 / s->h[1], 1.0 / s->h[2]};
    // sdf = fluid SDF, replaced by -collider where both are negative (:29-37)
    Arr3 sdf = s->sdf;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                if (s->sdf(i, j, k) < 0 && s->csdf(i, j, k) < 0) sdf(i, j, k) = s->csdf(i, j, k) * (-1);
    auto node = [&](int i, int j, int k, int a) {  // GridIndexToPosition
        const int idx[3] = {i, j, k};
        return so[a] + idx[a] * s->h[a];
    };
    McMesh M;
    std::map<int64_t, int32_t> vmap;
    auto key3 = [&](int axis, int i, int j, int k) {
        return (int64_t)axis + 3 * ((int64_t)i + (int64_t)(nx + 1) * ((int64_t)j + (int64_t)(ny + 1) * k));
    };
    for (int i = 0; i < nx - 1; ++i)
        for (int j = 0; j < ny - 1; ++j)
            for (int k = 0; k < nz - 1; ++k) {
                bool in[8];
                double val[8];
                for (int c = 0; c < 8; ++c) {
                    val[c] = sdf(i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1));
                    in[c] = val[c] <= 0.0;
                }
                const std::vector<int> tris = mc_cube_triangles(in);
                if (tris.empty()) continue;
                const double oc[3] = {node(i, j, k, 0), node(i, j, k, 1), node(i, j, k, 2)};  // bbox origin
                for (size_t q = 0; q < tris.size(); q += 3) {
                    int32_t face[3];
                    for (int r = 0; r < 3; ++r) {
                        const int e = tris[q + r];
                        int a, b;
                        mc_edge_pair(e, a, b);
                        const int axis = e / 4;
                        const int64_t key = key3(axis, i + (a & 1), j + ((a >> 1) & 1), k + ((a >> 2) & 1));
                        auto it = vmap.find(key);
                        if (it != vmap.end()) {
                            face[r] = it->second;
                            continue;
                        }
                        // DistanceToZeroLevelSet (:591-601), clamps (:263-268)
                        const double phi0 = val[a], phi1 = val[b];
                        double alpha = std::fabs(phi0) + std::fabs(phi1) > 0.000000001
                                           ? std::fabs(phi0) / (std::fabs(phi0) + std::fabs(phi1))
                                           : 0.5;
                        if (alpha < 0.000001) alpha = 0.000001;
                        if (alpha > 0.999999) alpha = 0.999999;
                        double g0[3], g1[3], nrm[3];
                        mc_gradient(sdf, ih, i + (a & 1), j + ((a >> 1) & 1), k + ((a >> 2) & 1), g0);
                        mc_gradient(sdf, ih, i + (b & 1), j + ((b >> 1) & 1), k + ((b >> 2) & 1), g1);
                        for (int x = 0; x < 3; ++x) {
                            // BoundingBox3D::Corner (src/3d/bounding_box_3d.cpp:23-31)
                            const double p0 = (s->h[x] * (((a >> x) & 1) ? 0.5 : -0.5) + oc[x]) + 0.5 * s->h[x];
                            const double p1 = (s->h[x] * (((b >> x) & 1) ? 0.5 : -0.5) + oc[x]) + 0.5 * s->h[x];
                            M.v.push_back((1.0 - alpha) * p0 + alpha * p1);
                            nrm[x] = (1.0 - alpha) * g0[x] + alpha * g1[x];
                        }
                        const double len = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
                        for (int x = 0; x < 3; ++x) M.n.push_back(len * len > 0.0 ? nrm[x] / len : nrm[x]);
                        face[r] = (int32_t)(M.v.size() / 3 - 1);
                        vmap[key] = face[r];
                    }
                    M.t.insert(M.t.end(), face, face + 3);
                }
            }
    // caps: planes z (back, front), x (left, right), y (down, up)
    const int planes[6][2] = {{2, 0}, {2, 1}, {0, 0}, {0, 1}, {1, 0}, {1, 1}};
    for (const auto& pl : planes) {
        const int axis = pl[0], side = pl[1];
        const int u = axis == 0 ? 1 : 0, w = axis == 2 ? 1 : 2;
        const int dim[3] = {nx, ny, nz};
        const int fixed = side ? dim[axis] - 1 : 0;
        const double outward = side ? 1.0 : -1.0;
        std::map<int64_t, int32_t> pmap;
        auto nodeval = [&](int iu, int iw) {
            int idx[3];
            idx[axis] = fixed;
            idx[u] = iu;
            idx[w] = iw;
            return sdf(idx[0], idx[1], idx[2]);
        };
        auto nodepos = [&](int iu, int iw, double p[3]) {
            int idx[3];
            idx[axis] = fixed;
            idx[u] = iu;
            idx[w] = iw;
            for (int x = 0; x < 3; ++x) p[x] = node(idx[0], idx[1], idx[2], x);
        };
        auto add = [&](const double p[3]) {
            for (int x = 0; x < 3; ++x) {
                M.v.push_back(p[x]);
                M.n.push_back(x == axis ? outward : 0.0);
            }
            return (int32_t)(M.v.size() / 3 - 1);
        };
        for (int iw = 0; iw + 1 < dim[w]; ++iw)
            for (int iu = 0; iu + 1 < dim[u]; ++iu) {
                const int cu[4] = {iu, iu + 1, iu + 1, iu}, cw[4] = {iw, iw, iw + 1, iw + 1};
                double phi[4];
                bool in[4];
                for (int t = 0; t < 4; ++t) {
                    phi[t] = nodeval(cu[t], cw[t]);
                    in[t] = phi[t] <= 0.0;
                }
                // cap polygons: each inside corner alone when the four square
                // edges are crossed, else one run of inside corners and crossings
                std::vector<std::vector<int>> polys;  // slots: 0-3 corners, 4-7 crossings
                int ncross = 0;
                for (int t = 0; t < 4; ++t) ncross += in[t] != in[(t + 1) % 4];
                if (ncross == 4) {
                    for (int m = 0; m < 4; ++m)
                        if (in[m]) polys.push_back({4 + (m + 3) % 4, m, 4 + m});
                } else if (in[0] || in[1] || in[2] || in[3]) {
                    std::vector<int> p;
                    for (int t = 0; t < 4; ++t) {
                        if (in[t]) p.push_back(t);
                        if (in[t] != in[(t + 1) % 4]) p.push_back(4 + t);
                    }
                    polys.push_back(p);
                }
                for (const auto& poly : polys) {
                    std::vector<int32_t> ids;
                    for (int slot : poly) {
                        int64_t key;
                        double p[3];
                        if (slot < 4) {
                            key = 3 * ((int64_t)cu[slot] + (int64_t)(dim[u] + 1) * cw[slot]);
                            nodepos(cu[slot], cw[slot], p);
                        } else {
                            const int t = slot - 4, t1 = (t + 1) % 4;
                            const int lo = (cu[t] + cw[t] <= cu[t1] + cw[t1]) ? t : t1, hi = lo == t ? t1 : t;
                            key = 3 * ((int64_t)cu[lo] + (int64_t)(dim[u] + 1) * cw[lo]) + (cw[lo] == cw[hi] ? 1 : 2);
                            // SolveSingleSquare alpha (:536-549)
                            const double f0 = std::fabs(phi[lo]), f1 = std::fabs(phi[hi]);
                            double alpha = f0 + f1 > 1e-12 ? f0 / (f0 + f1) : 0.5f;
                            if (alpha < 0.000001f) alpha = 0.000001f;
                            if (alpha > 0.999999f) alpha = 0.999999f;
                            double p0[3], p1[3];
                            nodepos(cu[lo], cw[lo], p0);
                            nodepos(cu[hi], cw[hi], p1);
                            for (int x = 0; x < 3; ++x) p[x] = (1.0 - alpha) * p0[x] + alpha * p1[x];
                        }
                        auto it = pmap.find(key);
                        if (it == pmap.end()) it = pmap.emplace(key, add(p)).first;
                        ids.push_back(it->second);
                    }
                    // the cyclic (u, w) order turns about +x, -y, +z for x, y, z planes
                    const bool flip = (axis == 1 ? -1.0 : 1.0) != outward;
                    for (size_t t = 1; t + 1 < ids.size(); ++t) {
                        M.t.push_back(ids[0]);
                        M.t.push_back(flip ? ids[t + 1] : ids[t]);
                        M.t.push_back(flip ? ids[t] : ids[t + 1]);
                    }
                }
            }
    }
    return mesh_out(M, vout, nout, tout, nv, nt);
}

// ---------------------------------------------------------------------------
// Mesh colliders: TriangleMeshCollider (src/3d/triangle_mesh_collider.cpp:5-9)
// builds its SDF on its own grid with Mesh2SDF::Build (src/3d/mesh_2_sdf.cpp:14-95),
// and ColliderCollection::BuildSdf (src/3d/collider_collection.cpp:33-50) takes,
// at every node of the scene's collider grid, the minimum over colliders of
// that SDF sampled there.
// ---------------------------------------------------------------------------
struct V3 {
    double x, y, z;
};
inline V3 sub3(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot3(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }  // Vector3::Dot
inline double len3(const V3& a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }  // GetLength

// Collisions::ClossestPointOnTriangle (src/3d/collisions.cpp:21-213), which is
// D. Eberly's point-triangle distance ("Distance Between Point and Triangle in
// 3D", Geometric Tools): the closest point is p0 + s e0 + t e1 with (s, t)
// minimising the quadratic a00 s^2 + 2 a01 s t + a11 t^2 + 2 b0 s + 2 b1 t over
// the triangle; the unconstrained minimum's region of the (s, t) plane picks
// the edge or vertex it clamps to.  Same coefficients and branch tests as the
// reference, so the result is bit-identical.
V3 closest_on_triangle(const V3& x, const V3& p0, const V3& p1, const V3& p2) {
    const V3 d = sub3(p0, x), e0 = sub3(p1, p0), e1 = sub3(p2, p0);
    const double a00 = dot3(e0, e0), a01 = dot3(e0, e1), a11 = dot3(e1, e1);
    const double b0 = dot3(d, e0), b1 = dot3(d, e1);
    const double det = std::max(a00 * a11 - a01 * a01, 0.0);
    double s = a01 * b1 - a11 * b0;
    double t = a01 * b0 - a00 * b1;
    // minimiser along an edge that starts at the origin of (s, t): 0 when the
    // slope b is non-negative, 1 when -b reaches the curvature a, else -b / a
    auto along = [](double b, double a) { return b >= 0 ? 0.0 : (-b >= a ? 1.0 : -b / a); };
    // the edge s + t = 1 seen from the side where parameter u is 0:
    // tmp0 = a01 + bu, tmp1 = aw + bw with w the other parameter
    const double denom_h = a00 - 2 * a01 + a11;
    if (s + t <= det) {
        if (s < 0) {
            if (t < 0) {  // region 4
                if (b0 < 0) {
                    t = 0;
                    s = (-b0 >= a00) ? 1.0 : -b0 / a00;
                } else {
                    s = 0;
                    t = along(b1, a11);
                }
            } else {  // region 3
                s = 0;
                t = along(b1, a11);
            }
        } else if (t < 0) {  // region 5
            t = 0;
            s = along(b0, a00);
        } else {  // region 0: interior
            s /= det;
            t /= det;
        }
    } else if (s < 0) {  // region 2
        const double tmp0 = a01 + b0, tmp1 = a11 + b1;
        if (tmp1 > tmp0) {
            const double numer = tmp1 - tmp0;
            if (numer >= denom_h) {
                s = 1;
                t = 0;
            } else {
                s = numer / denom_h;
                t = 1 - s;
            }
        } else {
            s = 0;
            t = tmp1 <= 0 ? 1.0 : (b1 >= 0 ? 0.0 : -b1 / a11);
        }
    } else if (t < 0) {  // region 6
        const double tmp0 = a01 + b1, tmp1 = a00 + b0;
        if (tmp1 > tmp0) {
            const double numer = tmp1 - tmp0;
            if (numer >= denom_h) {
                t = 1;
                s = 0;
            } else {
                t = numer / denom_h;
                s = 1 - t;
            }
        } else {
            t = 0;
            s = tmp1 <= 0 ? 1.0 : (b0 >= 0 ? 0.0 : -b0 / a00);
        }
    } else {  // region 1
        const double numer = a11 + b1 - a01 - b0;
        if (numer <= 0) {
            s = 0;
            t = 1;
        } else if (numer >= denom_h) {
            s = 1;
            t = 0;
        } else {
            s = numer / denom_h;
            t = 1 - s;
        }
    }
    return {p0.x + s * e0.x + t * e1.x, p0.y + s * e0.y + t * e1.y, p0.z + s * e0.z + t * e1.z};
}

// Collisions::DistanceToTriangle (:215-218) = |closest - x|
inline double distance_to_triangle(const V3& x, const V3& p0, const V3& p1, const V3& p2) {
    return len3(sub3(closest_on_triangle(x, p0, p1, p2), x));
}

// TriangleMesh::IsInside (src/3d/triangle_mesh.cpp:120-152): the summed
// half solid angles atan2(det[A B C], |A||B||C| + |C| A.B + |A| B.C + |B| C.A)
// of the triangles seen from x (A, B, C = vertices - x; Van Oosterom and
// Strackee) reach 2 pi - 0.001 in magnitude.  det in ADet's term order (:154-163).
constexpr double kPi = 3.14159265358979323846;  // src/common/math_utils.hpp:11
bool mesh_inside(const std::vector<V3>& verts, const std::vector<int32_t>& tris, const V3& x) {
    double ret = 0;
    const size_t nt = tris.size() / 3;
    for (size_t t = 0; t < nt; ++t) {
        const V3 A = sub3(verts[tris[3 * t]], x), B = sub3(verts[tris[3 * t + 1]], x),
                 C = sub3(verts[tris[3 * t + 2]], x);
        double omega = A.x * B.y * C.z;
        omega += B.x * C.y * A.z;
        omega += C.x * A.y * B.z;
        omega -= C.x * B.y * A.z;
        omega -= B.x * A.y * C.z;
        omega -= A.x * C.y * B.z;
        const double nA = len3(A), nB = len3(B), nC = len3(C);
        double k = nA * nB * nC;
        k += nC * dot3(A, B);  // (A * B).Sum() sums x, y, z in Dot's order
        k += nA * dot3(B, C);
        k += nB * dot3(C, A);
        ret += std::atan2(omega, k);
    }
    return std::abs(ret) >= 2.0 * kPi - 0.001;
}

// Mesh2SDF::Build = ComputeExactBandDistanceField(bandwidth 1) + ComputeSigns.
// Every node starts at max(h) * (nx + ny + nz); a triangle updates the nodes in
// its index box, whose bounds divide the vertex coordinates by h.x on every
// axis without subtracting the grid origin (:33-56, as the reference), widened
// by 1 below and 2 above and clamped to the grid; nodes inside the mesh are
// negated.
Arr3 mesh_to_sdf(const int size[3], const double origin[3], const double h[3], const std::vector<V3>& verts,
                 const std::vector<int32_t>& tris) {
    Arr3 sdf;
    const double hmax = (h[0] >= h[1] && h[0] >= h[2]) ? h[0] : ((h[1] >= h[0] && h[1] >= h[2]) ? h[1] : h[2]);
    sdf.resize(size[0], size[1], size[2], hmax * (double)((size_t)size[0] + (size_t)size[1] + (size_t)size[2]));
    auto node = [&](int i, int j, int k) {  // Grid3D::GridIndexToPosition
        return V3{origin[0] + i * h[0], origin[1] + j * h[1], origin[2] + k * h[2]};
    };
    auto clampi = [](int v, int hi, int lo) { return v < lo ? lo : (v > hi ? hi : v); };
    const size_t nt = tris.size() / 3;
    if (!verts.empty()) {
        const double inv = 1.0 / h[0];
        for (size_t t = 0; t < nt; ++t) {
            const V3 &p = verts[tris[3 * t]], &q = verts[tris[3 * t + 1]], &r = verts[tris[3 * t + 2]];
            const double fi[3] = {p.x * inv, q.x * inv, r.x * inv}, fj[3] = {p.y * inv, q.y * inv, r.y * inv},
                         fk[3] = {p.z * inv, q.z * inv, r.z * inv};
            const int i0 = clampi(int(std::fmin(fi[0], std::fmin(fi[1], fi[2]))) - 1, size[0] - 1, 0);
            const int j0 = clampi(int(std::fmin(fj[0], std::fmin(fj[1], fj[2]))) - 1, size[1] - 1, 0);
            const int k0 = clampi(int(std::fmin(fk[0], std::fmin(fk[1], fk[2]))) - 1, size[2] - 1, 0);
            const int i1 = clampi(int(std::fmax(fi[0], std::fmax(fi[1], fi[2]))) + 1 + 1, size[0] - 1, 0);
            const int j1 = clampi(int(std::fmax(fj[0], std::fmax(fj[1], fj[2]))) + 1 + 1, size[1] - 1, 0);
            const int k1 = clampi(int(std::fmax(fk[0], std::fmax(fk[1], fk[2]))) + 1 + 1, size[2] - 1, 0);
            for (int k = k0; k <= k1; ++k)
                for (int j = j0; j <= j1; ++j)
                    for (int i = i0; i <= i1; ++i) {
                        const double d = distance_to_triangle(node(i, j, k), p, q, r);
                        if (d < sdf(i, j, k)) sdf(i, j, k) = d;
                    }
        }
    }
    for (int k = 0; k < size[2]; ++k)
        for (int j = 0; j < size[1]; ++j)
            for (int i = 0; i < size[0]; ++i)
                if (mesh_inside(verts, tris, node(i, j, k))) sdf(i, j, k) = sdf(i, j, k) * -1.0;
    return sdf;
}

// ---------------------------------------------------------------------------
// Scene sources: PICSimulator::InitializeFromTriangleMesh + InitializeParticles
// (src/pic_simulator.cpp:33-39,521-548) and VolumeEmitter::Emitt
// (src/emmiter/volume_emmiter.cpp:58-85).  Both put particlesPerBlock particles
// at pos + h * u (u uniform per axis) into each picked cell, pos the cell's
// GridIndexToPosition.  The reference draws u from std::rand() / RAND_MAX on
// racing threads, which no run reproduces; here u comes from SplitMix64
// (G. Steele, D. Lea, C. Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014) keyed by (seed, call, particle, axis), in [0, 1),
// and the particles are appended in ForEachIndex order (i outer, k inner).
// ---------------------------------------------------------------------------
inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline double uniform01(uint64_t seed, uint64_t call, uint64_t n, int axis) {
    const uint64_t r = splitmix64(splitmix64(splitmix64(seed) ^ call) ^ (3 * n + (uint64_t)axis));
    return (double)(r >> 11) * 0x1.0p-53;
}

template <class Pick>
int64_t seed_cells(orc_scene* s, Pick pick, const double o[3], const double h[3], int ppb, uint64_t seed,
                   uint64_t call, const double vel[3]) {
    uint64_t n = 0;
    for (int i = 0; i < s->n[0]; ++i)
        for (int j = 0; j < s->n[1]; ++j)
            for (int k = 0; k < s->n[2]; ++k) {
                if (!pick(i, j, k)) continue;
                const double pos[3] = {o[0] + i * h[0], o[1] + j * h[1], o[2] + k * h[2]};
                for (int p = 0; p < ppb; ++p, ++n) {
                    for (int a = 0; a < 3; ++a) s->x.push_back(h[a] * uniform01(seed, call, n, a) + pos[a]);
                    for (int a = 0; a < 3; ++a) s->vel.push_back(vel[a]);
                    for (int a = 0; a < 9; ++a) s->C.push_back(0.0);
                }
            }
    s->np += (int64_t)n;
    return (int64_t)n;
}

// VolumeEmitter::Emitt of every emitter, at the start of a substep
// (OnBeginAdvanceTimeStep, src/pic_simulator.cpp:156-163): a node of the
// emitter grid emits when emitterSdf <= 0, fluidSdf >= 0 and colliderSdf > 0
// there, the last two read at the same indices.  Emitt runs before
// BuildCollider, so in the first substep the collider SDF still holds its
// initial 0 and nothing is emitted.
void emit(orc_scene* s) {
    for (auto& e : s->emitters) {
        seed_cells(
            s,
            [&](int i, int j, int k) {
                const double c = s->substeps_run == 0 ? 0.0 : s->csdf(i, j, k);
                return e.sdf(i, j, k) <= 0 && s->sdf(i, j, k) >= 0 && c > 0;
            },
            e.origin, e.h, e.ppb, e.seed, e.calls, e.vel);
        ++e.calls;
    }
}

// InitializeFromTriangleMesh: Mesh2SDF of the mesh into the fluid SDF (cell
// centred grid, origin o + h/2), then InitializeParticles: a cell is picked
// when that SDF sampled at the cell's corner o + h*(i,j,k) is negative; the
// new particles have zero velocity.
void initialize_from_mesh(orc_scene* s, const std::vector<V3>& verts, const std::vector<int32_t>& tris, int ppb,
                          uint64_t seed) {
    const double so[3] = {s->origin[0] + s->h[0] / 2.0, s->origin[1] + s->h[1] / 2.0, s->origin[2] + s->h[2] / 2.0};
    s->sdf = mesh_to_sdf(s->n, so, s->h, verts, tris);
    const double zero[3] = {0.0, 0.0, 0.0};
    seed_cells(
        s,
        [&](int i, int j, int k) {
            const double p[3] = {s->origin[0] + i * s->h[0], s->origin[1] + j * s->h[1], s->origin[2] + k * s->h[2]};
            return sample_array(s->sdf, so, s->h, p) < 0;
        },
        s->origin, s->h, ppb, seed, 0, zero);
}

// ColliderCollection::BuildSdf with one more collider: the running minimum of
// the collider SDFs sampled at the scene's collider nodes origin + h*(i,j,k).
void add_mesh_collider(orc_scene* s, const int size[3], const double origin[3], const double h[3],
                       const std::vector<V3>& verts, const std::vector<int32_t>& tris) {
    const Arr3 c = mesh_to_sdf(size, origin, h, verts, tris);
    for (int k = 0; k < s->csdf.sz; ++k)
        for (int j = 0; j < s->csdf.sy; ++j)
            for (int i = 0; i < s->csdf.sx; ++i) {
                const double p[3] = {s->origin[0] + i * s->h[0], s->origin[1] + j * s->h[1], s->origin[2] + k * s->h[2]};
                const double d = sample_array(c, origin, h, p);
                if (d < s->csdf(i, j, k)) s->csdf(i, j, k) = d;
            }
    s->has_collider = 1;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

orc_scene* orc_create(const int n[3], const double origin[3], const double size[3], int method, double blend,
                      double max_cfl, const double gravity[3]) {
    auto* s = new orc_scene();
    for (int a = 0; a < 3; ++a) {
        s->n[a] = n[a];
        s->origin[a] = origin[a];
        s->size[a] = size[a];
        s->h[a] = size[a] / n[a];  // src/hybrid_simulator.cpp:7
        s->gravity[a] = gravity[a];
    }
    s->method = method;
    s->blend = blend;
    s->maxCfl = max_cfl;
    face_origins(s);
    s->u.resize(n[0] + 1, n[1], n[2]); s->v.resize(n[0], n[1] + 1, n[2]); s->w.resize(n[0], n[1], n[2] + 1);
    s->mu.resize(n[0] + 1, n[1], n[2]); s->mv.resize(n[0], n[1] + 1, n[2]); s->mw.resize(n[0], n[1], n[2] + 1);
    s->ru = s->u; s->rv = s->v; s->rw = s->w;
    s->sdf.resize(n[0], n[1], n[2], 0.0);
    s->csdf.resize(n[0], n[1], n[2], kMaxDistance);  // after BuildSdf with no colliders
    s->p.resize(n[0], n[1], n[2], 0.0);
    s->b.resize(n[0], n[1], n[2], 0.0);
    s->flags.assign((size_t)n[0] * n[1] * n[2], ORC_AIR);
    return s;
}

void orc_destroy(orc_scene* s) { delete s; }
void orc_set_threads(orc_scene* s, int nthreads) { s->nthreads = std::max(1, nthreads); }

void orc_plant_mistake(orc_scene* s, int which) { s->planted = which; }

void orc_set_fast_dots(orc_scene* s, int on) { s->fast_dots = on ? 1 : 0; }

void orc_set_surface_mode(orc_scene* s, int mode) { s->surface_mode = mode; }

void orc_set_particles(orc_scene* s, const double* x, const double* v, const double* C, int64_t np) {
    s->np = np;
    s->x.assign(x, x + 3 * np);
    s->vel.assign(v, v + 3 * np);
    if (C)
        s->C.assign(C, C + 9 * np);
    else
        s->C.assign(9 * np, 0.0);
}

void orc_get_particles(const orc_scene* s, double* x, double* v, double* C) {
    if (x) std::memcpy(x, s->x.data(), sizeof(double) * 3 * s->np);
    if (v) std::memcpy(v, s->vel.data(), sizeof(double) * 3 * s->np);
    if (C) {
        if ((int64_t)s->C.size() == 9 * s->np)
            std::memcpy(C, s->C.data(), sizeof(double) * 9 * s->np);
        else
            std::memset(C, 0, sizeof(double) * 9 * s->np);
    }
}

int64_t orc_num_particles(const orc_scene* s) { return s->np; }

static Arr3* grid_field(orc_scene* s, int what) {
    switch (what) {
        case ORC_U: return &s->u;
        case ORC_V: return &s->v;
        case ORC_W: return &s->w;
        case ORC_RU: return &s->ru;
        case ORC_RV: return &s->rv;
        case ORC_RW: return &s->rw;
        case ORC_SDF: return &s->sdf;
        case ORC_PRESSURE: return &s->p;
        case ORC_RHS: return &s->b;
        case ORC_COLLIDER_SDF: return &s->csdf;
        default: return nullptr;
    }
}

static Arr3i* grid_marker(orc_scene* s, int what) {
    switch (what) {
        case ORC_MU: return &s->mu;
        case ORC_MV: return &s->mv;
        case ORC_MW: return &s->mw;
        default: return nullptr;
    }
}

int64_t orc_grid_count(orc_scene* s, int what) {
    if (Arr3* a = grid_field(s, what)) return (int64_t)a->size();
    if (Arr3i* m = grid_marker(s, what)) return (int64_t)m->d.size();
    if (what == ORC_FLAGS) return (int64_t)s->flags.size();
    return -1;
}

// doubles for fields, int32 for markers, uint8 for cell flags
int orc_get_grid(orc_scene* s, int what, void* out) {
    if (Arr3* a = grid_field(s, what)) {
        std::memcpy(out, a->d.data(), sizeof(double) * a->size());
        return 0;
    }
    if (Arr3i* m = grid_marker(s, what)) {
        std::memcpy(out, m->d.data(), sizeof(int) * m->d.size());
        return 0;
    }
    if (what == ORC_FLAGS) {
        std::memcpy(out, s->flags.data(), s->flags.size());
        return 0;
    }
    return -1;
}

int orc_set_grid(orc_scene* s, int what, const void* in) {
    if (Arr3* a = grid_field(s, what)) {
        std::memcpy(a->d.data(), in, sizeof(double) * a->size());
        if (what == ORC_COLLIDER_SDF) s->has_collider = 1;  // ColliderCollection::BuildSdf output
        return 0;
    }
    if (Arr3i* m = grid_marker(s, what)) {
        std::memcpy(m->d.data(), in, sizeof(int) * m->d.size());
        return 0;
    }
    if (what == ORC_FLAGS) {
        std::memcpy(s->flags.data(), in, s->flags.size());
        return 0;
    }
    return -1;
}

void orc_p2g(orc_scene* s) { p2g(s); }
void orc_build_sdf(orc_scene* s) { build_sdf(s); }
void orc_build_markers(orc_scene* s) { build_markers(s); }
void orc_extrapolate_to_air(orc_scene* s) { extrapolate_to_air(s); }
void orc_apply_bc(orc_scene* s) { apply_bc(s); }
void orc_apply_gravity(orc_scene* s, double dt) { apply_gravity(s, dt); }
void orc_build_system(orc_scene* s) { build_system(s); }
void orc_apply_pressure(orc_scene* s) { apply_pressure(s); }
void orc_g2p(orc_scene* s) { g2p(s); }
void orc_move_particles(orc_scene* s, double dt) { move_particles(s, dt); }

void orc_pressure(orc_scene* s, const orc_solver_desc* d, orc_solve_stats* st) { pressure(s, d, st); }

int orc_surface(orc_scene* s, double** vertices, double** normals, int32_t** triangles, int64_t* nv, int64_t* nt) {
    return surface(s, vertices, normals, triangles, nv, nt);
}
void orc_free(void* p) { std::free(p); }

static std::vector<V3> to_v3(const double* v, int64_t nv) {
    std::vector<V3> out((size_t)nv);
    for (int64_t i = 0; i < nv; ++i) out[(size_t)i] = V3{v[3 * i], v[3 * i + 1], v[3 * i + 2]};
    return out;
}

int orc_mesh_to_sdf(const int size[3], const double origin[3], const double h[3], const double* vertices, int64_t nv,
                    const int32_t* triangles, int64_t nt, double* out) {
    const std::vector<int32_t> tris(triangles, triangles + 3 * nt);
    for (int32_t t : tris)
        if (t < 0 || t >= nv) return -1;
    const Arr3 a = mesh_to_sdf(size, origin, h, to_v3(vertices, nv), tris);
    std::copy(a.d.begin(), a.d.end(), out);
    return 0;
}

int orc_add_mesh_collider(orc_scene* s, const int size[3], const double origin[3], const double h[3],
                          const double* vertices, int64_t nv, const int32_t* triangles, int64_t nt) {
    const std::vector<int32_t> tris(triangles, triangles + 3 * nt);
    for (int32_t t : tris)
        if (t < 0 || t >= nv) return -1;
    add_mesh_collider(s, size, origin, h, to_v3(vertices, nv), tris);
    return 0;
}

void orc_closest_on_triangle(const double x[3], const double p0[3], const double p1[3], const double p2[3],
                             double out[3]) {
    const V3 c = closest_on_triangle(V3{x[0], x[1], x[2]}, V3{p0[0], p0[1], p0[2]}, V3{p1[0], p1[1], p1[2]},
                                     V3{p2[0], p2[1], p2[2]});
    out[0] = c.x;
    out[1] = c.y;
    out[2] = c.z;
}

int orc_add_volume_emitter(orc_scene* s, const int size[3], const double origin[3], const double h[3],
                           const double* vertices, int64_t nv, const int32_t* triangles, int64_t nt, int ppb,
                           const double velocity[3], uint64_t seed) {
    const std::vector<int32_t> tris(triangles, triangles + 3 * nt);
    for (int32_t t : tris)
        if (t < 0 || t >= nv) return -1;
    for (int a = 0; a < 3; ++a)
        if (size[a] != s->n[a]) return -2;
    orc_scene::Emitter e;
    e.sdf = mesh_to_sdf(size, origin, h, to_v3(vertices, nv), tris);
    for (int a = 0; a < 3; ++a) {
        e.origin[a] = origin[a];
        e.h[a] = h[a];
        e.vel[a] = velocity[a];
    }
    e.ppb = ppb;
    e.seed = seed;
    e.calls = 0;
    s->emitters.push_back(std::move(e));
    return 0;
}

int orc_initialize_from_mesh(orc_scene* s, const double* vertices, int64_t nv, const int32_t* triangles, int64_t nt,
                             int ppb, uint64_t seed) {
    const std::vector<int32_t> tris(triangles, triangles + 3 * nt);
    for (int32_t t : tris)
        if (t < 0 || t >= nv) return -1;
    initialize_from_mesh(s, to_v3(vertices, nv), tris, ppb, seed);
    return 0;
}

double orc_uniform01(uint64_t seed, uint64_t call, uint64_t n, int axis) { return uniform01(seed, call, n, axis); }

void orc_emit(orc_scene* s) { emit(s); }

int orc_mesh_inside(const double* vertices, int64_t nv, const int32_t* triangles, int64_t nt, const double x[3]) {
    const std::vector<int32_t> tris(triangles, triangles + 3 * nt);
    return mesh_inside(to_v3(vertices, nv), tris, V3{x[0], x[1], x[2]}) ? 1 : 0;
}

void orc_set_viscosity(orc_scene* s, double viscosity) { s->viscosity = viscosity; }
void orc_set_diffusion_solver(orc_scene* s, const orc_solver_desc* d) { s->diff = *d; }
void orc_diffusion(orc_scene* s, double dt, orc_solve_stats st[3]) {
    diffusion(s, dt);
    if (st)
        for (int c = 0; c < 3; ++c) st[c] = s->diff_last[c];
}
int orc_diffusion_markers(const orc_scene* s, int c, uint8_t* out) {
    if (c < 0 || c > 2) return -1;
    std::memcpy(out, s->dmark[c].data(), s->dmark[c].size());
    return (int)s->dmark[c].size();
}

// The substep: PICSimulator::OnAdvanceTimeStep (src/pic_simulator.cpp:111-154) with
// OnBeginAdvanceTimeStep (:156-196), ComputeExternalForces (:203-212), ComputeDiffusion
// (:214-222, skipped at viscosity 0), ComputePressure (:224-229), ComputeAdvection (:231-237).
void orc_substep(orc_scene* s, double dt, const orc_solver_desc* d, orc_solve_stats* st) {
    emit(s);
    ++s->substeps_run;
    p2g(s);
    build_sdf(s);
    extrapolate_to_air(s);
    apply_bc(s);
    apply_gravity(s, dt);
    apply_bc(s);
    if (s->viscosity > 0.0) {
        diffusion(s, dt);
        apply_bc(s);
    }
    pressure(s, d, st);
    apply_bc(s);
    extrapolate_to_air(s);
    apply_bc(s);
    g2p(s);
    move_particles(s, dt);
}

// PICSimulator::Cfl (src/pic_simulator.cpp:80-97) with ValueAtCellCenter
// (src/grid_systems/face_centered_grid3d.cpp:146-155)
double orc_cfl(const orc_scene* s, double dt) {
    double maxV = 0.0;
    for (int i = 0; i < s->n[0]; ++i)
        for (int j = 0; j < s->n[1]; ++j)
            for (int k = 0; k < s->n[2]; ++k) {
                double cx = 0.5 * (s->u(i, j, k) + s->u(i + 1, j, k));
                double cy = 0.5 * (s->v(i, j, k) + s->v(i, j + 1, k));
                double cz = 0.5 * (s->w(i, j, k) + s->w(i, j, k + 1));
                double vx = std::abs(cx) + dt * 0.0;
                double vy = std::abs(cy) + dt * -9.8;
                double vz = std::abs(cz) + dt * 0.0;
                maxV = std::max(maxV, vx);
                maxV = std::max(maxV, vy);
                maxV = std::max(maxV, vz);
            }
    double minH = std::min(s->h[0], std::min(s->h[1], s->h[2]));
    return maxV * dt / minH;
}

// PICSimulator::NumberOfSubTimeSteps (src/pic_simulator.cpp:349-354)
unsigned orc_num_substeps(const orc_scene* s, double dt) {
    double c = orc_cfl(s, dt);
    return static_cast<unsigned>(std::max(std::ceil(c / s->maxCfl), 1.0));
}

// PhysicsAnimation::AdvanceTimeStep (src/physics_animation.cpp:64-92)
int orc_advance_frame(orc_scene* s, double frame_dt, const orc_solver_desc* d) {
    double remaining = frame_dt;
    int count = 0;
    while (remaining > 0.0000000001) {
        unsigned n = orc_num_substeps(s, remaining);
        double sub = remaining / static_cast<double>(n);
        orc_substep(s, sub, d, nullptr);
        remaining -= sub;
        ++count;
    }
    return count;
}

void orc_last_stats(const orc_scene* s, orc_solve_stats* st) { *st = s->last; }

// ---------------------------------------------------------------------------
// Primitive restatements exported for the known-answer tests.
// ---------------------------------------------------------------------------
void orc_barycentric(double x, int iLow, int iHigh, int* i, double* f) { barycentric(x, iLow, iHigh, i, f); }

void orc_coords_weights(const int size[3], const double origin[3], const double h[3], const double x[3],
                        int idx_out[24], double w_out[8]) {
    Stencil st;
    coords_weights(size, origin, h, x, &st);
    std::memcpy(idx_out, st.idx, sizeof(int) * 24);
    std::memcpy(w_out, st.w, sizeof(double) * 8);
}

void orc_coords_grad_weights(const int size[3], const double origin[3], const double h[3], const double x[3],
                             int idx_out[24], double g_out[24]) {
    GradStencil st;
    coords_grad_weights(size, origin, h, x, &st);
    std::memcpy(idx_out, st.idx, sizeof(int) * 24);
    std::memcpy(g_out, st.g, sizeof(double) * 24);
}

double orc_trilerp(const double v[8], double fx, double fy, double fz) {
    return trilerp(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], fx, fy, fz);
}

double orc_sample_array(const double* data, const int dims[3], const double origin[3], const double h[3],
                        const double p[3]) {
    Arr3 a;
    a.sx = dims[0]; a.sy = dims[1]; a.sz = dims[2];
    a.d.assign(data, data + (size_t)dims[0] * dims[1] * dims[2]);
    return sample_array(a, origin, h, p);
}

void orc_extrapolate_to_region(const double* input, const int* valid, const int dims[3], int iterations,
                               double* output) {
    Arr3 in, out;
    in.sx = dims[0]; in.sy = dims[1]; in.sz = dims[2];
    in.d.assign(input, input + (size_t)dims[0] * dims[1] * dims[2]);
    Arr3i va;
    va.sx = dims[0]; va.sy = dims[1]; va.sz = dims[2];
    va.d.assign(valid, valid + va.sx * (size_t)va.sy * va.sz);
    extrapolate_to_region(in, va, iterations, out, 1);
    std::memcpy(output, out.d.data(), sizeof(double) * out.size());
}

// Nearby-point query of the restated hash grid (for the reference searcher tests).
int orc_hash_grid_query(const double* pts, int64_t npts, int sx, double spacing, const double q[3],
                        double radius, int64_t* out, int max_out) {
    // The reference's HashGridParticleSearcher with an arbitrary bucket count `sx`
    // per axis (its unit tests use 4, 6, 7).
    struct G {
        int S;
        double sp;
        std::vector<std::vector<int64_t>> b;
        int wrap(int v) const {
            int r = (int)((size_t)(int64_t)v % (size_t)S);
            if (r < 0) r += S;
            return r;
        }
        size_t key(int x, int y, int z) const { return ((size_t)wrap(z) * S + wrap(y)) * S + wrap(x); }
    } g{sx, spacing, {}};
    g.b.assign((size_t)sx * sx * sx, {});
    for (int64_t i = 0; i < npts; ++i) {
        int bx = (int)std::floor(pts[3 * i] / spacing), by = (int)std::floor(pts[3 * i + 1] / spacing),
            bz = (int)std::floor(pts[3 * i + 2] / spacing);
        g.b[g.key(bx, by, bz)].push_back(i);
    }
    int o[3] = {(int)std::floor(q[0] / spacing), (int)std::floor(q[1] / spacing), (int)std::floor(q[2] / spacing)};
    int nb[8][3];
    for (int i = 0; i < 8; ++i) { nb[i][0] = o[0]; nb[i][1] = o[1]; nb[i][2] = o[2]; }
    int dx = ((o[0] + 0.5f) * spacing <= q[0]) ? 1 : -1;
    int dy = ((o[1] + 0.5f) * spacing <= q[1]) ? 1 : -1;
    int dz = ((o[2] + 0.5f) * spacing <= q[2]) ? 1 : -1;
    nb[4][0] += dx; nb[5][0] += dx; nb[6][0] += dx; nb[7][0] += dx;
    nb[2][1] += dy; nb[3][1] += dy; nb[6][1] += dy; nb[7][1] += dy;
    nb[1][2] += dz; nb[3][2] += dz; nb[5][2] += dz; nb[7][2] += dz;
    int count = 0;
    const double r2 = radius * radius;
    for (int i = 0; i < 8; ++i) {
        for (int64_t pi : g.b[g.key(nb[i][0], nb[i][1], nb[i][2])]) {
            double ddx = pts[3 * pi] - q[0], ddy = pts[3 * pi + 1] - q[1], ddz = pts[3 * pi + 2] - q[2];
            if (ddx * ddx + ddy * ddy + ddz * ddz <= r2) {
                if (count < max_out) out[count] = pi;
                ++count;
            }
        }
    }
    return count;
}

// BLAS primitives on plain x-fastest arrays (src/linear_system/blas.cpp:4-83)
double orc_blas_dot(const double* a, const double* b, const int dims[3]) {
    Arr3 A, B;
    A.sx = B.sx = dims[0]; A.sy = B.sy = dims[1]; A.sz = B.sz = dims[2];
    size_t n = (size_t)dims[0] * dims[1] * dims[2];
    A.d.assign(a, a + n);
    B.d.assign(b, b + n);
    return dot(A, B);
}

// Solve a 7-point system given explicitly as {center,right,up,front} rows,
// with the Jacobi or CG restatement.  x is read as the initial guess.
void orc_solve_rows(const double* Ac, const double* Ar, const double* Au, const double* Af, const double* b,
                    double* x, const int dims[3], const orc_solver_desc* d, orc_solve_stats* st) {
    orc_scene s{};
    s.n[0] = dims[0]; s.n[1] = dims[1]; s.n[2] = dims[2];
    size_t n = (size_t)dims[0] * dims[1] * dims[2];
    auto fill = [&](Arr3& a, const double* src) {
        a.sx = dims[0]; a.sy = dims[1]; a.sz = dims[2];
        a.d.assign(src, src + n);
    };
    fill(s.Ac, Ac); fill(s.Ar, Ar); fill(s.Au, Au); fill(s.Af, Af); fill(s.b, b); fill(s.p, x);
    orc_solve_stats local{};
    if (d->kind == ORC_SOLVER_JACOBI)
        jacobi(&s, d, &local);
    else
        cg(&s, d, &local);
    std::memcpy(x, s.p.d.data(), sizeof(double) * n);
    if (st) *st = local;
}

void orc_residual_rows(const double* Ac, const double* Ar, const double* Au, const double* Af, const double* b,
                       const double* x, const int dims[3], double* out) {
    orc_scene s{};
    size_t n = (size_t)dims[0] * dims[1] * dims[2];
    auto fill = [&](Arr3& a, const double* src) {
        a.sx = dims[0]; a.sy = dims[1]; a.sz = dims[2];
        a.d.assign(src, src + n);
    };
    Arr3 X, B, R;
    fill(s.Ac, Ac); fill(s.Ar, Ar); fill(s.Au, Au); fill(s.Af, Af); fill(B, b); fill(X, x);
    R = X;
    residual(&s, X, B, R, 1);
    std::memcpy(out, R.d.data(), sizeof(double) * n);
}

void orc_spmv_rows(const double* Ac, const double* Ar, const double* Au, const double* Af, const double* v,
                   const int dims[3], double* out) {
    orc_scene s{};
    size_t n = (size_t)dims[0] * dims[1] * dims[2];
    auto fill = [&](Arr3& a, const double* src) {
        a.sx = dims[0]; a.sy = dims[1]; a.sz = dims[2];
        a.d.assign(src, src + n);
    };
    Arr3 V, O;
    fill(s.Ac, Ac); fill(s.Ar, Ar); fill(s.Au, Au); fill(s.Af, Af); fill(V, v);
    O = V;
    spmv(&s, V, O, 1);
    std::memcpy(out, O.d.data(), sizeof(double) * n);
}

}  // extern "C"
