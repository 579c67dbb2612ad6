// k_surface.hip -- fluid surface: MarchingCubesSolver::BuildSurface
// (src/3d/marching_cubes_solver.cpp:25-51) on the fluid SDF.
//
// Nodes are the SDF samples (cell centres, origin o + h/2); a node is inside
// when its value is <= 0.  The reference builds the mesh in four parts: the
// cubes between nodes (CalculateXYZMeshPart) and caps on the six boundary
// node planes (CalculateXY/YZ/XZMeshPart).  Here:
//   * cubes, on the device: every crossed node edge carries one vertex (the
//     reference's edge-keyed vertex map); vertices are numbered by a scan over
//     the crossed-edge flags (x edges, then y, then z), triangles by a scan
//     over per-cube counts, so the output is deterministic and needs no map;
//   * caps, on the host from the six boundary planes (O(n^2) work), one
//     vertex set per plane as in the reference (its separate vertex maps).
// Vertex position: linear zero of the SDF along the edge, alpha =
// |phi0| / (|phi0| + |phi1|) clamped to [1e-6, 0.999999]; normal: the same
// interpolation of the central-difference gradients at the two nodes,
// normalised (MarchingCubesSolver::Gradient and SafeNormalize).  The case
// tables come from fe_mc_rule.hpp (one face rule for cubes and caps, so the
// surface is closed); the reference's own table can leave holes on faces with
// diagonally opposite inside nodes.
#include <algorithm>
#include <cmath>
#include <vector>

#include "fe_kernels.hpp"
#include "fe_mc_rule.hpp"

namespace fe {

namespace {

struct NodeGrid {
    int n[3];
    double o[3], h[3], ih[3];  // node origin o + h/2, spacing, 1/h
    int64_t ne[3];             // edges per axis
};

__device__ __host__ inline int64_t node_index(const NodeGrid& G, int i, int j, int k) {
    return (int64_t)i + (int64_t)G.n[0] * ((int64_t)j + (int64_t)G.n[1] * k);
}

// Edge e of the concatenated x, y, z edge lists -> axis and low node.
__device__ inline void edge_node(const NodeGrid& G, int64_t e, int& axis, int& i, int& j, int& k) {
    axis = e < G.ne[0] ? 0 : (e < G.ne[0] + G.ne[1] ? 1 : 2);
    const int64_t l = e - (axis == 0 ? 0 : (axis == 1 ? G.ne[0] : G.ne[0] + G.ne[1]));
    const int dx = G.n[0] - (axis == 0), dy = G.n[1] - (axis == 1);
    split_index(l, dx, dy, i, j, k);
}

__device__ inline int64_t edge_id(const NodeGrid& G, int axis, int i, int j, int k) {
    if (axis == 0) return (int64_t)i + (int64_t)(G.n[0] - 1) * ((int64_t)j + (int64_t)G.n[1] * k);
    if (axis == 1) return G.ne[0] + (int64_t)i + (int64_t)G.n[0] * ((int64_t)j + (int64_t)(G.n[1] - 1) * k);
    return G.ne[0] + G.ne[1] + (int64_t)i + (int64_t)G.n[0] * ((int64_t)j + (int64_t)G.n[1] * k);
}

// MarchingCubesSolver::Gradient (src/3d/marching_cubes_solver.cpp:116-160).
__device__ inline void node_gradient(const NodeGrid& G, const double* __restrict__ sdf, int i, int j, int k,
                                     double g[3]) {
#pragma clang fp contract(off)
    const int idx[3] = {i, j, k};
    for (int a = 0; a < 3; ++a) {
        int p[3] = {i, j, k}, m[3] = {i, j, k};
        if (idx[a] > G.n[a] - 2) {
            p[a] = idx[a];
            m[a] = idx[a] - 1;
        } else if (idx[a] == 0) {
            p[a] = 1;
            m[a] = 0;
        } else {
            p[a] = idx[a] + 1;
            m[a] = idx[a] - 1;
        }
        g[a] = 0.5 * G.ih[a] * (sdf[node_index(G, p[0], p[1], p[2])] - sdf[node_index(G, m[0], m[1], m[2])]);
    }
}

__global__ __launch_bounds__(kBlock) void k_mc_edges(NodeGrid G, const double* __restrict__ sdf,
                                                     int32_t* __restrict__ flag, int64_t nedges) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= nedges) return;
    int axis, i, j, k;
    edge_node(G, e, axis, i, j, k);
    const double a = sdf[node_index(G, i, j, k)];
    const double b = sdf[node_index(G, i + (axis == 0), j + (axis == 1), k + (axis == 2))];
    flag[e] = (a <= 0.0) != (b <= 0.0) ? 1 : 0;
}

__device__ inline int cube_case(const NodeGrid& G, const double* __restrict__ sdf, int i, int j, int k) {
    int cs = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c)
        if (sdf[node_index(G, i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1))] <= 0.0) cs |= 1 << c;
    return cs;
}

__global__ __launch_bounds__(kBlock) void k_mc_count(NodeGrid G, const double* __restrict__ sdf,
                                                     const McTables* __restrict__ T, int32_t* __restrict__ cnt,
                                                     int64_t ncubes) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= ncubes) return;
    int i, j, k;
    split_index(c, G.n[0] - 1, G.n[1] - 1, i, j, k);
    cnt[c] = T->ntri[cube_case(G, sdf, i, j, k)];
}

__global__ __launch_bounds__(kBlock) void k_mc_vertices(NodeGrid G, const double* __restrict__ sdf,
                                                        const int32_t* __restrict__ vid, double* __restrict__ pos,
                                                        double* __restrict__ nrm, int64_t nedges) {
#pragma clang fp contract(off)
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= nedges) return;
    const int32_t v = vid[e];
    if (vid[e + 1] == v) return;  // not crossed
    int axis, i, j, k;
    edge_node(G, e, axis, i, j, k);
    const int i1 = i + (axis == 0), j1 = j + (axis == 1), k1 = k + (axis == 2);
    const double phi0 = sdf[node_index(G, i, j, k)], phi1 = sdf[node_index(G, i1, j1, k1)];
    // DistanceToZeroLevelSet (:591-601) and the clamps of SolveSingleCube (:263-268)
    double alpha = fabs(phi0) + fabs(phi1) > 0.000000001 ? fabs(phi0) / (fabs(phi0) + fabs(phi1)) : 0.5;
    alpha = alpha < 0.000001 ? 0.000001 : (alpha > 0.999999 ? 0.999999 : alpha);
    // The reference's edges run from the lower to the higher node
    // (_edgeConnection with its corner order) and its end points are
    // BoundingBox3D::Corner of the first cube in its i, j, k loop order that
    // holds the edge (the vertex map keeps that cube's vertex): the cube at
    // node - 1 on the two other axes, clamped at 0.
    const int a0[3] = {i, j, k};
    double g0[3], g1[3], nv[3];
    node_gradient(G, sdf, i, j, k, g0);
    node_gradient(G, sdf, i1, j1, k1, g1);
    for (int a = 0; a < 3; ++a) {
        const int cube = a == axis ? a0[a] : (a0[a] > 0 ? a0[a] - 1 : 0);
        const int off0 = a0[a] - cube, off1 = off0 + (a == axis);
        const double oc = G.o[a] + cube * G.h[a];  // GridIndexToPosition of the cube
        const double p0 = (G.h[a] * (off0 ? 0.5 : -0.5) + oc) + 0.5 * G.h[a];
        const double p1 = (G.h[a] * (off1 ? 0.5 : -0.5) + oc) + 0.5 * G.h[a];
        pos[3 * (int64_t)v + a] = (1.0 - alpha) * p0 + alpha * p1;
        nv[a] = (1.0 - alpha) * g0[a] + alpha * g1[a];
    }
    const double l2 = nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2];
    if (l2 > 0.0) {  // SafeNormalize: Vector3::Normalize divides by the length
        const double l = sqrt(l2);
        for (int a = 0; a < 3; ++a) nv[a] = nv[a] / l;
    }
    for (int a = 0; a < 3; ++a) nrm[3 * (int64_t)v + a] = nv[a];
}

__global__ __launch_bounds__(kBlock) void k_mc_emit(NodeGrid G, const double* __restrict__ sdf,
                                                    const McTables* __restrict__ T, const int32_t* __restrict__ toff,
                                                    const int32_t* __restrict__ vid, int32_t* __restrict__ tri,
                                                    int64_t ncubes) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= ncubes) return;
    const int32_t t0 = toff[c], nt = toff[c + 1] - t0;
    if (nt == 0) return;
    int i, j, k;
    split_index(c, G.n[0] - 1, G.n[1] - 1, i, j, k);
    const int cs = cube_case(G, sdf, i, j, k);
    for (int t = 0; t < nt; ++t)
        for (int q = 0; q < 3; ++q) {
            const int e = T->tri[cs][3 * t + q];
            const int axis = e >> 2, u = e & 1, w = (e >> 1) & 1;
            int64_t ge;
            if (axis == 0) ge = edge_id(G, 0, i, j + u, k + w);
            else if (axis == 1) ge = edge_id(G, 1, i + u, j, k + w);
            else ge = edge_id(G, 2, i + u, j + w, k);
            tri[3 * (int64_t)(t0 + t) + q] = vid[ge];
        }
}

// Boundary node planes, gathered for the host caps: plane p = 2*axis + side,
// each n[u] * n[v] values (u, v the other axes, lower first, u fastest).
__global__ __launch_bounds__(kBlock) void k_mc_planes(NodeGrid G, const double* __restrict__ sdf,
                                                      double* __restrict__ out, int64_t stride) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int p = (int)(t / stride);
    const int64_t r = t - (int64_t)p * stride;
    if (p >= 6) return;
    const int axis = p >> 1, side = p & 1;
    const int u = axis == 0 ? 1 : 0, v = axis == 2 ? 1 : 2;
    const int64_t nu = G.n[u];
    if (r >= nu * G.n[v]) return;
    int idx[3];
    idx[axis] = side ? G.n[axis] - 1 : 0;
    idx[u] = (int)(r % nu);
    idx[v] = (int)(r / nu);
    out[t] = sdf[node_index(G, idx[0], idx[1], idx[2])];
}

inline dim3 grid_of(int64_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

NodeGrid node_grid(const Geom& g) {
    NodeGrid G;
    for (int a = 0; a < 3; ++a) {
        G.n[a] = g.n[a];
        G.h[a] = g.hd[a];
        G.ih[a] = 1.0 / g.hd[a];
        G.o[a] = g.od[a] + g.hd[a] / 2.0;  // the fluid SDF grid origin (src/hybrid_simulator.cpp:13-15)
    }
    G.ne[0] = (int64_t)(G.n[0] - 1) * G.n[1] * G.n[2];
    G.ne[1] = (int64_t)G.n[0] * (G.n[1] - 1) * G.n[2];
    G.ne[2] = (int64_t)G.n[0] * G.n[1] * (G.n[2] - 1);
    return G;
}

template <typename T>
hipError_t grow(T** p, int64_t* cap, int64_t need) {
    if (*cap >= need) return hipSuccess;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    const int64_t n = std::max<int64_t>(need + need / 4, 1024);
    hipError_t e = hipMalloc((void**)p, sizeof(T) * n);
    if (e == hipSuccess) *cap = n;
    return e;
}

// Caps of one boundary plane (SolveXY / SolveYZ / SolveXZ with
// SolveSingleSquare, :272-589): squares of the plane's nodes, inside nodes and
// crossing points as vertices (one vertex set per plane), the plane's outward
// axis normal for every vertex.
void plane_caps(const NodeGrid& G, int p, const double* val, std::vector<double>& pos, std::vector<double>& nrm,
                std::vector<int32_t>& tri, int64_t vbase) {
    const int axis = p >> 1, side = p & 1;
    const int u = axis == 0 ? 1 : 0, v = axis == 2 ? 1 : 2;
    const int nu = G.n[u], nv = G.n[v];
    const double outward = side ? 1.0 : -1.0;
    const int fixed = side ? G.n[axis] - 1 : 0;
    // vertex ids: nodes, then u-edges, then v-edges of the plane
    const int64_t nnode = (int64_t)nu * nv, nue = (int64_t)(nu - 1) * nv;
    std::vector<int32_t> id(nnode + nue + (int64_t)nu * (nv - 1), -1);
    auto node_pos = [&](int a_u, int a_v, double out[3]) {
        int idx[3];
        idx[axis] = fixed;
        idx[u] = a_u;
        idx[v] = a_v;
        for (int a = 0; a < 3; ++a) out[a] = G.o[a] + idx[a] * G.h[a];
    };
    auto add_vertex = [&](const double q[3]) {  // caps follow the nv cube vertices
        const int64_t vi = vbase + (int64_t)pos.size() / 3;
        for (int a = 0; a < 3; ++a) {
            pos.push_back(q[a]);
            nrm.push_back(a == axis ? outward : 0.0);
        }
        return (int32_t)vi;
    };
    auto value = [&](int a_u, int a_v) { return val[(int64_t)a_u + (int64_t)nu * a_v]; };
    for (int jv = 0; jv + 1 < nv; ++jv)
        for (int iu = 0; iu + 1 < nu; ++iu) {
            // square corners in the (u, v) cyclic order of fe_mc_rule.hpp
            const int cu[4] = {iu, iu + 1, iu + 1, iu}, cv[4] = {jv, jv, jv + 1, jv + 1};
            bool in[4];
            double phi[4];
            for (int t = 0; t < 4; ++t) {
                phi[t] = value(cu[t], cv[t]);
                in[t] = phi[t] <= 0.0;
            }
            int poly[2][5], len[2];
            const int np = mc_square_caps(in, poly, len);
            for (int q = 0; q < np; ++q) {
                int32_t vid[5];
                for (int s = 0; s < len[q]; ++s) {
                    const int slot = poly[q][s];
                    if (slot < 4) {  // node
                        const int64_t key = (int64_t)cu[slot] + (int64_t)nu * cv[slot];
                        if (id[key] < 0) {
                            double pp[3];
                            node_pos(cu[slot], cv[slot], pp);
                            id[key] = add_vertex(pp);
                        }
                        vid[s] = id[key];
                    } else {  // crossing point on square edge slot-4 between corners t and t+1
                        const int t = slot - 4, t1 = (t + 1) & 3;
                        const int lo = (cu[t] + cv[t] <= cu[t1] + cv[t1]) ? t : t1, hi = lo == t ? t1 : t;
                        const bool along_u = cv[lo] == cv[hi];
                        const int64_t key = along_u ? nnode + (int64_t)cu[lo] + (int64_t)(nu - 1) * cv[lo]
                                                    : nnode + nue + (int64_t)cu[lo] + (int64_t)nu * cv[lo];
                        if (id[key] < 0) {
                            // SolveSingleSquare's alpha (:536-549)
                            const double f0 = std::fabs(phi[lo]), f1 = std::fabs(phi[hi]);
                            double alpha = f0 + f1 > 1e-12 ? f0 / (f0 + f1) : 0.5;
                            if (alpha < 0.000001f) alpha = 0.000001f;
                            if (alpha > 0.999999f) alpha = 0.999999f;
                            double p0[3], p1[3], pp[3];
                            node_pos(cu[lo], cv[lo], p0);
                            node_pos(cu[hi], cv[hi], p1);
                            for (int a = 0; a < 3; ++a) pp[a] = (1.0 - alpha) * p0[a] + alpha * p1[a];
                            id[key] = add_vertex(pp);
                        }
                        vid[s] = id[key];
                    }
                }
                // the (u, v) cyclic order turns counter-clockwise about u x v:
                // +x for x planes, -y for y planes, +z for z planes
                const double natural = axis == 1 ? -1.0 : 1.0;
                const bool flip = natural != outward;
                for (int s = 1; s + 1 < len[q]; ++s) {
                    tri.push_back(vid[0]);
                    tri.push_back(flip ? vid[s + 1] : vid[s]);
                    tri.push_back(flip ? vid[s] : vid[s + 1]);
                }
            }
        }
}

// Caps of one boundary plane as the reference builds them (SolveXY / SolveYZ /
// SolveXZ with SolveSingleSquare, :272-589, FE_SURFACE_REFERENCE): the
// reference's square loop order and corner order per plane, its square table
// (kMcRefSquare), a crossing placed from the square's own edge direction when
// the plane's vertex map first meets it, the plane's outward axis normal for
// every vertex.  Corner k of a square is node (cu, cv) + kRefSq[p][k] in the
// plane's (u, v) coordinates, u the lower other axis.
void plane_caps_reference(const NodeGrid& G, int p, const double* val, std::vector<double>& pos,
                          std::vector<double>& nrm, std::vector<int32_t>& tri, int64_t vbase) {
    // (du, dv) per corner; planes 2*axis + side: x (left, right), y (down, up), z (back, front)
    static const int kRefSq[6][4][2] = {
        {{0, 0}, {0, 1}, {1, 1}, {1, 0}},  // left  (u = y, v = z): (j,k) (j,k+1) (j+1,k+1) (j+1,k)
        {{0, 1}, {0, 0}, {1, 0}, {1, 1}},  // right: (j,k+1) (j,k) (j+1,k) (j+1,k+1)
        {{0, 0}, {1, 0}, {1, 1}, {0, 1}},  // down  (u = x, v = z): (i,k) (i+1,k) (i+1,k+1) (i,k+1)
        {{1, 0}, {0, 0}, {0, 1}, {1, 1}},  // up:    (i+1,k) (i,k) (i,k+1) (i+1,k+1)
        {{1, 0}, {0, 0}, {0, 1}, {1, 1}},  // back  (u = x, v = y): (i+1,j) (i,j) (i,j+1) (i+1,j+1)
        {{0, 0}, {1, 0}, {1, 1}, {0, 1}},  // front: (i,j) (i+1,j) (i+1,j+1) (i,j+1)
    };
    const int axis = p >> 1, side = p & 1;
    const int u = axis == 0 ? 1 : 0, v = axis == 2 ? 1 : 2;
    const int nu = G.n[u], nv = G.n[v];
    const double outward = side ? 1.0 : -1.0;
    const int fixed = side ? G.n[axis] - 1 : 0;
    const int64_t nnode = (int64_t)nu * nv, nue = (int64_t)(nu - 1) * nv;
    std::vector<int32_t> id(nnode + nue + (int64_t)nu * (nv - 1), -1);
    auto node_pos = [&](int a_u, int a_v, double out[3]) {
        int idx[3];
        idx[axis] = fixed;
        idx[u] = a_u;
        idx[v] = a_v;
        for (int a = 0; a < 3; ++a) out[a] = G.o[a] + idx[a] * G.h[a];
    };
    auto add_vertex = [&](const double q[3]) {
        const int64_t vi = vbase + (int64_t)pos.size() / 3;
        for (int a = 0; a < 3; ++a) {
            pos.push_back(q[a]);
            nrm.push_back(a == axis ? outward : 0.0);
        }
        return (int32_t)vi;
    };
    auto value = [&](int a_u, int a_v) { return val[(int64_t)a_u + (int64_t)nu * a_v]; };
    // the reference's loops: z planes j outer, i inner; x planes k outer, j
    // inner; y planes k outer, i inner -- always v outer, u inner here
    for (int jv = 0; jv + 1 < nv; ++jv)
        for (int iu = 0; iu + 1 < nu; ++iu) {
            int cu[4], cv[4];
            double phi[4];
            int cs = 0;
            for (int t = 0; t < 4; ++t) {
                cu[t] = iu + kRefSq[p][t][0];
                cv[t] = jv + kRefSq[p][t][1];
                phi[t] = value(cu[t], cv[t]);
                if (phi[t] <= 0.0) cs |= 1 << t;
            }
            if (cs == 0) continue;
            const char* tt = kMcRefSquare[cs];
            const int n = (int)std::strlen(tt);
            for (int q = 0; q + 2 < n; q += 3) {
                int32_t f[3];
                for (int r = 0; r < 3; ++r) {
                    const int slot = mc_hex(tt[q + r]);
                    if (slot < 4) {  // a corner node
                        const int64_t key = (int64_t)cu[slot] + (int64_t)nu * cv[slot];
                        if (id[key] < 0) {
                            double pp[3];
                            node_pos(cu[slot], cv[slot], pp);
                            id[key] = add_vertex(pp);
                        }
                        f[r] = id[key];
                    } else {  // the crossing on square edge slot-4, from corner t to t+1
                        const int t = slot - 4, t1 = (t + 1) & 3;
                        const int lo = (cu[t] + cv[t] <= cu[t1] + cv[t1]) ? t : t1;
                        const bool along_u = cv[t] == cv[t1];
                        const int64_t key = along_u ? nnode + (int64_t)cu[lo] + (int64_t)(nu - 1) * cv[lo]
                                                    : nnode + nue + (int64_t)cu[lo] + (int64_t)nu * cv[lo];
                        if (id[key] < 0) {
                            // SolveSingleSquare's alpha and clamps (:536-549)
                            const double f0 = std::fabs(phi[t]), f1 = std::fabs(phi[t1]);
                            double alpha = f0 + f1 > 1e-12 ? f0 / (f0 + f1) : 0.5f;
                            if (alpha < 0.000001f) alpha = 0.000001f;
                            if (alpha > 0.999999f) alpha = 0.999999f;
                            double p0[3], p1[3], pp[3];
                            node_pos(cu[t], cv[t], p0);
                            node_pos(cu[t1], cv[t1], p1);
                            for (int a = 0; a < 3; ++a) pp[a] = (1.f - alpha) * p0[a] + alpha * p1[a];
                            id[key] = add_vertex(pp);
                        }
                        f[r] = id[key];
                    }
                }
                tri.insert(tri.end(), f, f + 3);
            }
        }
}

// The surface SDF with a collider: the fluid SDF, replaced by minus the
// collider SDF where both are negative (GetSurface, src/hybrid_simulator.cpp:29-37).
__global__ __launch_bounds__(kBlock) void k_mc_combine(const double* __restrict__ sdf,
                                                       const double* __restrict__ csdf, double* __restrict__ out,
                                                       int64_t n) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const double a = sdf[c], b = csdf[c];
    out[c] = (a < 0 && b < 0) ? -b : a;
}

}  // namespace

void surface_free(fe_scene* s) {
    McState& m = s->mc;
    auto F = [](auto*& p) {
        if (p) (void)hipFree(p);
        p = nullptr;
    };
    F(m.flag);
    F(m.vid);
    F(m.cnt);
    F(m.toff);
    F(m.sums);
    F(m.tables);
    F(m.tables_ref);
    F(m.planes);
    F(m.pos);
    F(m.nrm);
    F(m.tri);
    F(m.comb);
    m = McState{};
}

hipError_t surface_build(fe_scene* s) {
    McState& m = s->mc;
    const NodeGrid G = node_grid(s->g);
    const int64_t nedges = G.ne[0] + G.ne[1] + G.ne[2];
    const int64_t ncubes = (int64_t)(G.n[0] - 1) * (G.n[1] - 1) * (G.n[2] - 1);
    hipError_t e;
    const bool ref = s->surface_mode == FE_SURFACE_REFERENCE;
    if (!m.tables) {
        const McTables host = make_mc_tables();
        if ((e = hipMalloc((void**)&m.tables, sizeof(McTables))) != hipSuccess) return e;
        if ((e = hipMemcpy(m.tables, &host, sizeof(McTables), hipMemcpyHostToDevice)) != hipSuccess) return e;
        const McTables href = make_mc_tables_reference();
        if ((e = hipMalloc((void**)&m.tables_ref, sizeof(McTables))) != hipSuccess) return e;
        if ((e = hipMemcpy(m.tables_ref, &href, sizeof(McTables), hipMemcpyHostToDevice)) != hipSuccess) return e;
        const int64_t tiles = scan_tiles(std::max(nedges, ncubes));
        if ((e = hipMalloc((void**)&m.flag, sizeof(int32_t) * (nedges + 1))) != hipSuccess) return e;
        if ((e = hipMalloc((void**)&m.vid, sizeof(int32_t) * (nedges + 1))) != hipSuccess) return e;
        if ((e = hipMalloc((void**)&m.cnt, sizeof(int32_t) * (ncubes + 1))) != hipSuccess) return e;
        if ((e = hipMalloc((void**)&m.toff, sizeof(int32_t) * (ncubes + 1))) != hipSuccess) return e;
        if ((e = hipMalloc((void**)&m.sums, sizeof(int32_t) * (tiles + 1))) != hipSuccess) return e;
        m.plane_stride = (int64_t)std::max({G.n[0], G.n[1], G.n[2]}) * std::max({G.n[0], G.n[1], G.n[2]});
        if ((e = hipMalloc((void**)&m.planes, sizeof(double) * 6 * m.plane_stride)) != hipSuccess) return e;
    }
    hipStream_t st = s->stream;
    // the reference's SDF values: the 8-bucket search over the same particles
    // where the last SDF came from the cell scan (the negative values, and so
    // the markers, are the same either way)
    if (s->sdf_from_cells && (e = launch_sdf(s, true)) != hipSuccess) return e;
    const McTables* tables = (const McTables*)(ref ? m.tables_ref : m.tables);
    const double* sdf = s->sdf;
    if (s->has_collider) {
        const int64_t nc = cell_count(s->g);
        if (!m.comb && (e = hipMalloc((void**)&m.comb, sizeof(double) * nc)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_mc_combine, grid_of(nc), dim3(kBlock), 0, st, sdf,
                           (const double*)s->csdf, m.comb, nc);
        sdf = m.comb;
    }
    // cubes: crossed edges -> vertex ids; per-cube triangle counts -> offsets
    hipLaunchKernelGGL(k_mc_edges, grid_of(nedges), dim3(kBlock), 0, st, G, sdf, m.flag, nedges);
    if ((e = launch_scan_with(st, m.flag, m.vid, nedges, m.sums)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_mc_count, grid_of(ncubes), dim3(kBlock), 0, st, G, sdf, tables, m.cnt, ncubes);
    if ((e = launch_scan_with(st, m.cnt, m.toff, ncubes, m.sums)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_mc_planes, grid_of(6 * m.plane_stride), dim3(kBlock), 0, st, G, sdf,
                       m.planes, m.plane_stride);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    int32_t nv = 0, nt = 0;
    if ((e = hipMemcpyAsync(&nv, m.vid + nedges, sizeof(int32_t), hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(&nt, m.toff + ncubes, sizeof(int32_t), hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    std::vector<double> planes((size_t)(6 * m.plane_stride));
    if ((e = hipMemcpyAsync(planes.data(), m.planes, sizeof(double) * planes.size(), hipMemcpyDeviceToHost, st)) !=
        hipSuccess)
        return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    if ((e = grow(&m.pos, &m.vcap, 3 * (int64_t)nv)) != hipSuccess) return e;
    if ((e = grow(&m.nrm, &m.ncap, 3 * (int64_t)nv)) != hipSuccess) return e;
    if ((e = grow(&m.tri, &m.tcap, 3 * (int64_t)nt)) != hipSuccess) return e;
    if (nv > 0)
        hipLaunchKernelGGL(k_mc_vertices, grid_of(nedges), dim3(kBlock), 0, st, G, sdf,
                           (const int32_t*)m.vid, m.pos, m.nrm, nedges);
    if (nt > 0)
        hipLaunchKernelGGL(k_mc_emit, grid_of(ncubes), dim3(kBlock), 0, st, G, sdf, tables,
                           (const int32_t*)m.toff, (const int32_t*)m.vid, m.tri, ncubes);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // caps on the host while the device writes the cube part
    m.cap_pos.clear();
    m.cap_nrm.clear();
    m.cap_tri.clear();
    const int order[6] = {4, 5, 0, 1, 2, 3};  // XY (z planes), YZ (x planes), XZ (y planes), as the reference
    for (int q = 0; q < 6; ++q) {
        const int p = order[q];
        if (ref)
            plane_caps_reference(G, p, planes.data() + (size_t)p * m.plane_stride, m.cap_pos, m.cap_nrm, m.cap_tri,
                                 nv);
        else
            plane_caps(G, p, planes.data() + (size_t)p * m.plane_stride, m.cap_pos, m.cap_nrm, m.cap_tri, nv);
    }
    m.nv3 = nv;
    m.nt3 = nt;
    m.built = true;
    return hipStreamSynchronize(st);
}

hipError_t surface_download(fe_scene* s, double* pos, double* nrm, int32_t* tri) {
    McState& m = s->mc;
    hipError_t e;
    const size_t v3 = 3 * (size_t)m.nv3, t3 = 3 * (size_t)m.nt3;
    if (pos && v3 && (e = hipMemcpy(pos, m.pos, sizeof(double) * v3, hipMemcpyDeviceToHost)) != hipSuccess) return e;
    if (nrm && v3 && (e = hipMemcpy(nrm, m.nrm, sizeof(double) * v3, hipMemcpyDeviceToHost)) != hipSuccess) return e;
    if (tri && t3 && (e = hipMemcpy(tri, m.tri, sizeof(int32_t) * t3, hipMemcpyDeviceToHost)) != hipSuccess) return e;
    if (pos) std::copy(m.cap_pos.begin(), m.cap_pos.end(), pos + v3);
    if (nrm) std::copy(m.cap_nrm.begin(), m.cap_nrm.end(), nrm + v3);
    if (tri) std::copy(m.cap_tri.begin(), m.cap_tri.end(), tri + t3);
    return hipSuccess;
}

}  // namespace fe
