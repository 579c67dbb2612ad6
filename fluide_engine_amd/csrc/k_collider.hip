// k_collider.hip -- colliders on the device: the collider branches of
// BlockedBoundryConditionSolver::ConstrainVelocity
// (src/fluid_solvers/blocked_boundry_condition_solver.cpp:18-291) and the
// collision response of moved particles.
//
// The collider is the ColliderCollection SDF (src/3d/collider_collection.cpp:33-50):
// n^3 FP64 values at the nodes o + h*(i,j,k), uploaded by the caller.  With a
// collider the boundary condition is
//   (1) faces whose collider fraction FractionInsideSdf(phi(x - h/2), phi(x + h/2))
//       is 1 are zeroed and marked invalid;
//   (2) ExtrapolateToRegion from the valid faces (depth = ceil(maxCfl));
//   (3) faces inside the collider take the tangential part of the sampled
//       velocity, v - (v.n) n with n the normalised collider gradient (0 where
//       the gradient vanishes); collider velocity and friction are 0 as in
//       the reference;
//   (4) the wall layers are zeroed;
//   (5) faces between a collider node and a free node are zeroed.
// Without a collider (1)-(3) and (5) are identities and only (4) runs.
#include <algorithm>
#include <cmath>
#include <vector>

#include "fe_kernels.hpp"

namespace fe {

namespace {

__device__ inline void face_of(const Geom& g, int64_t t, int& c, int64_t& f) {
    const int64_t n0 = face_count(g, 0), n1 = face_count(g, 1);
    c = t < n0 ? 0 : (t < n0 + n1 ? 1 : 2);
    f = t - (c == 0 ? 0 : (c == 1 ? n0 : n0 + n1));
}

__device__ inline void face_pos(const Geom& g, int c, int64_t f, double p[3]) {
#pragma clang fp contract(off)
    int i, j, k;
    split_index(f, face_dim(g, c, 0), face_dim(g, c, 1), i, j, k);
    const int idx[3] = {i, j, k};
    for (int a = 0; a < 3; ++a) p[a] = g.cod[c][a] + g.hd[a] * (double)idx[a];  // GetX/Y/ZPos
}

// FractionInsideSdf (:300-318)
__device__ inline double fraction_inside(double phi0, double phi1) {
    if (phi0 < 0.0 && phi1 < 0.0) return 1;
    if (phi0 < 0.0 && !(phi1 < 0.0)) return phi0 / (phi0 - phi1);
    if (!(phi0 < 0.0) && phi1 < 0.0) return phi1 / (phi1 - phi0);
    return 0;
}

// (1): blocked faces -> 0 and invalid; validity bytes over the three
// components back to back.
__global__ __launch_bounds__(kBlock) void k_bc_block(Geom g, FaceSet a, const double* __restrict__ csdf,
                                                     uint8_t* __restrict__ valid, int64_t total) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= total) return;
    int c;
    int64_t f;
    face_of(g, t, c, f);
    double p[3];
    face_pos(g, c, f, p);
    double p0[3] = {p[0], p[1], p[2]}, p1[3] = {p[0], p[1], p[2]};
    p0[c] = p[c] - 0.5 * g.hd[c];
    p1[c] = p[c] + 0.5 * g.hd[c];
    double frac = fraction_inside(sample_collider64(g, csdf, p0), sample_collider64(g, csdf, p1));
    frac = 1 - (frac > 1.0 ? 1.0 : (frac < 0.0 ? 0.0 : frac));  // 1 - Clamp(frac, 1, 0)
    if (frac > 0.0) {
        valid[t] = 1;
    } else {
        a.a[c][f] = 0.0f;
        valid[t] = 0;
    }
}

// (3): faces inside the collider get v - (v.n) n into tmp and a flag; the
// velocity is sampled from the grid before any face changes.
__global__ __launch_bounds__(kBlock) void k_bc_tangent(Geom g, FaceSet a, const double* __restrict__ csdf,
                                                       FaceSet tmp, uint8_t* __restrict__ inside, int64_t total) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= total) return;
    int c;
    int64_t f;
    face_of(g, t, c, f);
    double p[3];
    face_pos(g, c, f, p);
    if (!(sample_collider64(g, csdf, p) < 0)) {
        inside[t] = 0;
        return;
    }
    inside[t] = 1;
    double vel[3], gr[3];
    for (int q = 0; q < 3; ++q) vel[q] = (double)sample_face(a.a[q], g, q, (float)p[0], (float)p[1], (float)p[2]);
    collider_gradient64(g, csdf, p, gr);
    const double len = sqrt(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
    double out = 0.0;
    if (len * len > 0.0) {
        const double n[3] = {gr[0] / len, gr[1] / len, gr[2] / len};
        const double d = vel[0] * n[0] + vel[1] * n[1] + vel[2] * n[2];
        out = vel[c] - n[c] * d;
    }
    tmp.a[c][f] = (float)out;
}

__global__ __launch_bounds__(kBlock) void k_bc_copy(Geom g, FaceSet a, FaceSet tmp,
                                                    const uint8_t* __restrict__ inside, int64_t total) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= total || !inside[t]) return;
    int c;
    int64_t f;
    face_of(g, t, c, f);
    a.a[c][f] = tmp.a[c][f];
}

// (5): a collider node (colliderSdf(i,j,k) < 0, read at the index) zeroes its
// faces towards free nodes (:246-290).  Writes are zeros only, so the order of
// the reference's serial loop does not matter.
__global__ __launch_bounds__(kBlock) void k_bc_solid_faces(Geom g, FaceSet a, const double* __restrict__ csdf) {
    const int64_t cidx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (cidx >= cell_count(g)) return;
    if (!(csdf[cidx] < 0)) return;
    int i, j, k;
    split_index(cidx, g.n[0], g.n[1], i, j, k);
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    auto free_node = [&](int64_t q) { return !(csdf[q] < 0); };
    const int64_t iu = (int64_t)i + (int64_t)(nx + 1) * ((int64_t)j + (int64_t)ny * k);
    const int64_t iv = (int64_t)i + (int64_t)nx * ((int64_t)j + (int64_t)(ny + 1) * k);
    const int64_t iw = cidx;
    if (i > 0 && free_node(cidx - 1)) a.a[0][iu] = 0.0f;
    if (i < nx - 1 && free_node(cidx + 1)) a.a[0][iu + 1] = 0.0f;
    if (j > 0 && free_node(cidx - sy)) a.a[1][iv] = 0.0f;
    if (j < ny - 1 && free_node(cidx + sy)) a.a[1][iv + nx] = 0.0f;
    if (k > 0 && free_node(cidx - sz)) a.a[2][iw] = 0.0f;
    if (k < nz - 1 && free_node(cidx + sz)) a.a[2][iw + sz] = 0.0f;
}

// Particles that end inside the collider (MoveParticles, src/pic_simulator.cpp:297-304,
// which resolves them against the collider mesh but indexes the particle by
// the collider's loop index).  Here, on the particle itself and against the
// SDF: x -= phi n along the normalised gradient, clamped to the domain, and
// an inward normal velocity is dropped (restitution 0, friction 0).
__global__ __launch_bounds__(kBlock) void k_collide(Geom g, Particles P, int64_t cap, int64_t np,
                                                    const double* __restrict__ csdf) {
#pragma clang fp contract(off)
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= np) return;
    double p[3] = {(double)P.x[s], (double)P.x[cap + s], (double)P.x[2 * cap + s]};
    const double phi = sample_collider64(g, csdf, p);
    if (!(phi < 0)) return;
    double gr[3];
    collider_gradient64(g, csdf, p, gr);
    const double len = sqrt(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
    if (!(len * len > 0.0)) return;
    const double n[3] = {gr[0] / len, gr[1] / len, gr[2] / len};
    double v[3] = {(double)P.v[s], (double)P.v[cap + s], (double)P.v[2 * cap + s]};
    for (int a = 0; a < 3; ++a) {
        const double lo = g.od[a], hi = g.od[a] + g.Ld[a];
        p[a] = p[a] - phi * n[a];
        p[a] = p[a] < lo ? lo : (p[a] > hi ? hi : p[a]);
        P.x[a * cap + s] = (float)p[a];
    }
    const double vn = v[0] * n[0] + v[1] * n[1] + v[2] * n[2];
    if (vn < 0.0)
        for (int a = 0; a < 3; ++a) P.v[a * cap + s] = (float)(v[a] - vn * n[a]);
}

// ExtrapolateIntoCollider (src/pic_simulator.cpp:497-519) on the fluid SDF:
// a cell is valid iff the collider SDF at its centre is >= 0.
__global__ __launch_bounds__(kBlock) void k_sdf_valid(Geom g, const double* __restrict__ csdf,
                                                      uint8_t* __restrict__ valid) {
#pragma clang fp contract(off)
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= cell_count(g)) return;
    int idx[3];
    split_index(c, g.n[0], g.n[1], idx[0], idx[1], idx[2]);
    double cc[3];
    for (int a = 0; a < 3; ++a) cc[a] = (g.od[a] + 0.5 * g.hd[a]) + g.hd[a] * (double)idx[a];
    valid[c] = sample_collider64(g, csdf, cc) < 0 ? 0 : 1;
}

// One ExtrapolateToRegion sweep (src/common/math_utils.hpp:186-253) over the
// FP64 cell array, in place (invalid cells are written, only valid ones are
// read), sums in the reference's neighbour order.
__global__ __launch_bounds__(kBlock) void k_sdf_extrapolate(Geom g, double* __restrict__ sdf,
                                                            const uint8_t* __restrict__ vin,
                                                            uint8_t* __restrict__ vout) {
#pragma clang fp contract(off)
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= cell_count(g)) return;
    if (vin[c]) {
        vout[c] = 1;
        return;
    }
    int i, j, k;
    split_index(c, g.n[0], g.n[1], i, j, k);
    const int64_t st[3] = {1, g.n[0], (int64_t)g.n[0] * g.n[1]};
    const bool in[6] = {i + 1 < g.n[0], i > 0, j + 1 < g.n[1], j > 0, k + 1 < g.n[2], k > 0};
    double sum = 0;
    unsigned count = 0;
    for (int d = 0; d < 6; ++d) {
        if (!in[d]) continue;
        const int64_t q = (d & 1) ? c - st[d >> 1] : c + st[d >> 1];
        if (!vin[q]) continue;
        sum += sdf[q];
        ++count;
    }
    if (count > 0) {
        sdf[c] = sum / count;
        vout[c] = 1;
    } else {
        vout[c] = 0;
    }
}

// ---------------------------------------------------------------------------
// Mesh colliders: Mesh2SDF::Build (src/3d/mesh_2_sdf.cpp:14-95) of a
// TriangleMeshCollider (src/3d/triangle_mesh_collider.cpp:5-9) on the
// collider's own node grid, one thread per node looping over all triangles
// (the loop index is uniform, so each triangle's vertices and index box are
// scalar loads shared by the wave):
//   * distance: the minimum of DistanceToTriangle over the triangles whose
//     index box holds the node, starting from max(h) * (nx + ny + nz);
//   * sign: negated where TriangleMesh::IsInside (src/3d/triangle_mesh.cpp:120-163)
//     holds, i.e. the summed half solid angles reach 2 pi - 0.001 in magnitude.
// FP64 in the reference's operation order with contraction off, so the
// distances are the reference's bit for bit; the angle sum goes through the
// device atan2, which only feeds the threshold.
// ---------------------------------------------------------------------------
struct D3 {
    double x, y, z;
};
__device__ inline D3 dsub(const D3& a, const D3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline double ddot(const D3& a, const D3& b) {
#pragma clang fp contract(off)
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ inline double dlen(const D3& a) {
#pragma clang fp contract(off)
    return sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
}

// Collisions::ClossestPointOnTriangle (src/3d/collisions.cpp:21-213): D. Eberly's
// point-triangle distance (Geometric Tools, "Distance Between Point and
// Triangle in 3D"): the minimiser (s, t) of |p0 + s e0 + t e1 - x|^2 over the
// triangle, by the region of the (s, t) plane holding the free minimum.
__device__ inline double triangle_distance(const D3& x, const D3& p0, const D3& p1, const D3& p2) {
#pragma clang fp contract(off)
    const D3 d = dsub(p0, x), e0 = dsub(p1, p0), e1 = dsub(p2, p0);
    const double a00 = ddot(e0, e0), a01 = ddot(e0, e1), a11 = ddot(e1, e1);
    const double b0 = ddot(d, e0), b1 = ddot(d, e1);
    const double det = fmax(a00 * a11 - a01 * a01, 0.0);
    double s = a01 * b1 - a11 * b0;
    double t = a01 * b0 - a00 * b1;
    // on an edge from the (s, t) origin: 0 for a non-negative slope b, 1 once
    // -b reaches the curvature a, else -b / a
    auto along = [](double b, double a) { return b >= 0 ? 0.0 : (-b >= a ? 1.0 : -b / a); };
    if (s + t <= det) {
        if (s < 0) {
            if (t < 0 && b0 < 0) {  // region 4, towards the s edge
                t = 0;
                s = along(b0, a00);
            } else {  // region 4 towards the t edge, and region 3
                s = 0;
                t = along(b1, a11);
            }
        } else if (t < 0) {  // region 5
            t = 0;
            s = along(b0, a00);
        } else {  // region 0
            s /= det;
            t /= det;
        }
    } else {
        const double denom = a00 - 2 * a01 + a11;  // curvature along the edge s + t = 1
        if (s < 0 || t < 0) {  // regions 2 and 6: the roles of (s, b0, a00) and (t, b1, a11) swap
            const bool r2 = s < 0;
            const double bu = r2 ? b0 : b1, bw = r2 ? b1 : b0, aw = r2 ? a11 : a00;
            const double tmp0 = a01 + bu, tmp1 = aw + bw;
            double u, w;  // u: the parameter that is negative, w: the other
            if (tmp1 > tmp0) {
                const double numer = tmp1 - tmp0;
                if (numer >= denom) {
                    u = 1;
                    w = 0;
                } else {
                    u = numer / denom;
                    w = 1 - u;
                }
            } else {
                u = 0;
                w = tmp1 <= 0 ? 1.0 : (bw >= 0 ? 0.0 : -bw / aw);
            }
            s = r2 ? u : w;
            t = r2 ? w : u;
        } else {  // region 1
            const double numer = a11 + b1 - a01 - b0;
            if (numer <= 0) {
                s = 0;
                t = 1;
            } else if (numer >= denom) {
                s = 1;
                t = 0;
            } else {
                s = numer / denom;
                t = 1 - s;
            }
        }
    }
    const D3 c{p0.x + s * e0.x + t * e1.x, p0.y + s * e0.y + t * e1.y, p0.z + s * e0.z + t * e1.z};
    return dlen(dsub(c, x));
}

// One triangle's term of TriangleMesh::IsInside: atan2(det[A B C],
// |A||B||C| + |C| A.B + |A| B.C + |B| C.A), A, B, C = vertices - x, the
// determinant in ADet's term order (:154-163).
__device__ inline double half_solid_angle(const D3& x, const D3& p0, const D3& p1, const D3& p2) {
#pragma clang fp contract(off)
    const D3 A = dsub(p0, x), B = dsub(p1, x), C = dsub(p2, x);
    double omega = A.x * B.y * C.z;
    omega += B.x * C.y * A.z;
    omega += C.x * A.y * B.z;
    omega -= C.x * B.y * A.z;
    omega -= B.x * A.y * C.z;
    omega -= A.x * C.y * B.z;
    const double nA = dlen(A), nB = dlen(B), nC = dlen(C);
    double k = nA * nB * nC;
    k += nC * ddot(A, B);
    k += nA * ddot(B, C);
    k += nB * ddot(C, A);
    return atan2(omega, k);
}

constexpr double kPi = 3.14159265358979323846;  // src/common/math_utils.hpp:11

// tv: 9 doubles per triangle (its three vertices); tb: 6 ints per triangle,
// the index box i0, j0, k0, i1, j1, k1 (inclusive).
__global__ __launch_bounds__(kBlock) void k_mesh_sdf(NodeGrid3 G, const double* __restrict__ tv,
                                                     const int32_t* __restrict__ tb, int64_t nt, double fill,
                                                     double* __restrict__ out) {
#pragma clang fp contract(off)
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t nn = (int64_t)G.n[0] * G.n[1] * G.n[2];
    if (c >= nn) return;
    int i, j, k;
    split_index(c, G.n[0], G.n[1], i, j, k);
    const D3 x{G.o[0] + (double)i * G.h[0], G.o[1] + (double)j * G.h[1], G.o[2] + (double)k * G.h[2]};
    double d = fill, ret = 0;
    for (int64_t t = 0; t < nt; ++t) {
        const double* P = tv + 9 * t;
        const int32_t* B = tb + 6 * t;
        const D3 p0{P[0], P[1], P[2]}, p1{P[3], P[4], P[5]}, p2{P[6], P[7], P[8]};
        if (i >= B[0] && j >= B[1] && k >= B[2] && i <= B[3] && j <= B[4] && k <= B[5]) {
            const double dist = triangle_distance(x, p0, p1, p2);
            if (dist < d) d = dist;
        }
        ret += half_solid_angle(x, p0, p1, p2);
    }
    if (fabs(ret) >= 2.0 * kPi - 0.001) d = d * -1.0;
    out[c] = d;
}

__global__ __launch_bounds__(kBlock) void k_fill_f64(double* __restrict__ a, double v, int64_t n) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c < n) a[c] = v;
}

// ColliderCollection::BuildSdf with one more collider: the running minimum of
// the collider's SDF sampled at the scene's collider nodes origin + h*(i,j,k).
__global__ __launch_bounds__(kBlock) void k_collider_merge(Geom g, NodeGrid3 G, const double* __restrict__ c,
                                                           double* __restrict__ csdf) {
#pragma clang fp contract(off)
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= cell_count(g)) return;
    int idx[3];
    split_index(q, g.n[0], g.n[1], idx[0], idx[1], idx[2]);
    double p[3];
    for (int a = 0; a < 3; ++a) p[a] = g.od[a] + (double)idx[a] * g.hd[a];  // GridIndexToPosition
    const double d = sample_grid64(G, c, p);
    if (d < csdf[q]) csdf[q] = d;
}

inline dim3 grid_n(int64_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

}  // namespace

// Host part of Mesh2SDF::ComputeExactBandDistanceField (:20-80): each
// triangle's index box from its vertices divided by h.x on every axis, no
// origin subtracted, widened by 1 below and 2 above, clamped to the grid.
// Returns false when a quotient does not fit an int (the reference's int()
// conversion is undefined there) or a vertex index is out of range.
bool mesh_prepare(const NodeGrid3& G, const double* verts, int64_t nv, const int32_t* tris, int64_t nt,
                  std::vector<double>& tv, std::vector<int32_t>& tb, double& fill) {
    const double hx = G.h[0], hy = G.h[1], hz = G.h[2];
    const double hmax = (hx >= hy && hx >= hz) ? hx : ((hy >= hx && hy >= hz) ? hy : hz);  // Vector3::Max
    fill = hmax * (double)((size_t)G.n[0] + (size_t)G.n[1] + (size_t)G.n[2]);
    tv.assign((size_t)(9 * nt), 0.0);
    tb.assign((size_t)(6 * nt), 0);
    const double inv = 1.0 / hx;
    auto clampi = [](int v, int hi, int lo) { return v < lo ? lo : (v > hi ? hi : v); };
    for (int64_t t = 0; t < nt; ++t) {
        double f[3][3];  // f[axis][vertex]
        for (int q = 0; q < 3; ++q) {
            const int32_t vi = tris[3 * t + q];
            if (vi < 0 || vi >= nv) return false;
            for (int a = 0; a < 3; ++a) {
                tv[(size_t)(9 * t + 3 * q + a)] = verts[3 * (int64_t)vi + a];
                f[a][q] = verts[3 * (int64_t)vi + a] * inv;
                if (!(std::fabs(f[a][q]) < 2.0e9)) return false;
            }
        }
        for (int a = 0; a < 3; ++a) {
            const double lo = std::fmin(f[a][0], std::fmin(f[a][1], f[a][2]));
            const double hi = std::fmax(f[a][0], std::fmax(f[a][1], f[a][2]));
            tb[(size_t)(6 * t + a)] = clampi(int(lo) - 1, G.n[a] - 1, 0);
            tb[(size_t)(6 * t + 3 + a)] = clampi(int(hi) + 1 + 1, G.n[a] - 1, 0);
        }
    }
    return true;
}

hipError_t launch_mesh_sdf(hipStream_t st, const NodeGrid3& G, const double* tv_dev, const int32_t* tb_dev,
                           int64_t nt, double fill, double* out) {
    const int64_t nn = (int64_t)G.n[0] * G.n[1] * G.n[2];
    hipLaunchKernelGGL(k_mesh_sdf, grid_n(nn), dim3(kBlock), 0, st, G, tv_dev, tb_dev, nt, fill, out);
    return hipGetLastError();
}

hipError_t collider_merge(fe_scene* s, const NodeGrid3& G, const double* c_dev) {
    const int64_t nc = cell_count(s->g);
    hipError_t e;
    if (!s->csdf) {  // ColliderCollection's SDF before any collider: MAX_DISTANCE (collider_collection.cpp:3)
        if ((e = hipMalloc((void**)&s->csdf, sizeof(double) * nc)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_fill_f64, grid_n(nc), dim3(kBlock), 0, s->stream, s->csdf, 1.7976931348623157E+30, nc);
    }
    hipLaunchKernelGGL(k_collider_merge, grid_n(nc), dim3(kBlock), 0, s->stream, s->g, G, c_dev, s->csdf);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    s->has_collider = true;
    return hipSuccess;
}

hipError_t launch_sdf_into_collider(fe_scene* s) {
    if (!s->has_collider) return hipSuccess;
    const int64_t nc = cell_count(s->g);
    hipLaunchKernelGGL(k_sdf_valid, grid_n(nc), dim3(kBlock), 0, s->stream, s->g, (const double*)s->csdf, s->valid[0]);
    for (int it = 0; it < s->depth; ++it)
        hipLaunchKernelGGL(k_sdf_extrapolate, grid_n(nc), dim3(kBlock), 0, s->stream, s->g, s->sdf,
                           (const uint8_t*)s->valid[it & 1], s->valid[(it + 1) & 1]);
    return hipGetLastError();
}

hipError_t launch_bc(fe_scene* s) {
    if (!s->has_collider) return launch_walls(s);
    const Geom& g = s->g;
    const int64_t total = face_count(g, 0) + face_count(g, 1) + face_count(g, 2);
    hipError_t e;
    if (!s->bc_tmp[0])
        for (int c = 0; c < 3; ++c)
            if ((e = hipMalloc((void**)&s->bc_tmp[c], sizeof(float) * face_count(g, c))) != hipSuccess) return e;
    FaceSet a{{s->u[0], s->u[1], s->u[2]}};
    FaceSet tmp{{s->bc_tmp[0], s->bc_tmp[1], s->bc_tmp[2]}};
    // (1) + (2): the blocked-face markers go to the second validity buffer,
    // which the extrapolation reads first
    uint8_t* blocked = s->valid[1];
    hipLaunchKernelGGL(k_bc_block, grid_n(total), dim3(kBlock), 0, s->stream, g, a, (const double*)s->csdf, blocked,
                       total);
    const int64_t off[3] = {0, face_count(g, 0), face_count(g, 0) + face_count(g, 1)};
    const uint8_t* init[3] = {blocked + off[0], blocked + off[1], blocked + off[2]};
    if ((e = launch_extrapolate_from(s, init)) != hipSuccess) return e;
    // (3)
    uint8_t* inside = s->valid[0];
    hipLaunchKernelGGL(k_bc_tangent, grid_n(total), dim3(kBlock), 0, s->stream, g, a, (const double*)s->csdf, tmp,
                       inside, total);
    hipLaunchKernelGGL(k_bc_copy, grid_n(total), dim3(kBlock), 0, s->stream, g, a, tmp, (const uint8_t*)inside, total);
    // (4) + (5)
    if ((e = launch_walls(s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_bc_solid_faces, grid_n(cell_count(g)), dim3(kBlock), 0, s->stream, g, a,
                       (const double*)s->csdf);
    return hipGetLastError();
}

hipError_t launch_collide(fe_scene* s) {
    if (!s->has_collider || s->np == 0) return hipSuccess;
    hipLaunchKernelGGL(k_collide, grid_n(s->np), dim3(kBlock), 0, s->stream, s->g, s->P[s->cur], s->cap, s->np,
                       (const double*)s->csdf);
    return hipGetLastError();
}

void collider_free(fe_scene* s) {
    if (s->csdf) (void)hipFree(s->csdf);
    s->csdf = nullptr;
    s->has_collider = false;
    for (int c = 0; c < 3; ++c) {
        if (s->bc_tmp[c]) (void)hipFree(s->bc_tmp[c]);
        s->bc_tmp[c] = nullptr;
    }
}

}  // namespace fe
