// fe_mc_rule.hpp -- marching-cubes case tables generated from one rule.
//
// Cube corners are numbered by their offset bits (x = 1, y = 2, z = 4); a
// corner is inside when its SDF value is <= the iso value (0).  Edges:
//   0-3  along x, the other-axis bits (y, z) = (0,0), (1,0), (0,1), (1,1)
//   4-7  along y, (x, z) likewise
//   8-11 along z, (x, y) likewise
// On every cube face the crossed edges are joined in pairs; a face with four
// crossed edges (diagonally opposite inside corners) cuts each inside corner
// off on its own.  Neighbouring cubes see the same face corners and take the
// same decision, so the surface is closed.  The joined edges form loops; each
// loop is wound counter-clockwise seen from outside (its normal points from
// the inside corners to the outside ones) and fanned from the first loop
// position none of whose triangles lies in a cube face (such a triangle would
// overlap the neighbouring cube's surface).  Every case gives at most five
// triangles.  The domain-boundary caps use the same face rule on the
// boundary squares.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#include "fe_mc_table.hpp"

namespace fe {

constexpr int kMcMaxTri = 5;  // triangles per case (a loop of n edges gives n - 2)

struct McTables {
    uint8_t ntri[256];
    int8_t tri[256][kMcMaxTri * 3];  // cube edge indices, three per triangle
};

namespace mcr {

inline void edge_corners(int e, int& a, int& b) {
    const int axis = e >> 2, q = e & 3;
    const int u = q & 1, v = (q >> 1) & 1;  // bits of the two other axes, lower axis first
    int bits[3];
    bits[axis] = 0;
    const int o0 = axis == 0 ? 1 : 0, o1 = axis == 2 ? 1 : 2;
    bits[o0] = u;
    bits[o1] = v;
    a = bits[0] | (bits[1] << 1) | (bits[2] << 2);
    b = a | (1 << axis);
}

inline int edge_of(int c0, int c1) {  // corners differing in one bit
    const int d = c0 ^ c1, lo = c0 & c1;
    const int axis = d == 1 ? 0 : (d == 2 ? 1 : 2);
    const int o0 = axis == 0 ? 1 : 0, o1 = axis == 2 ? 1 : 2;
    const int u = (lo >> o0) & 1, v = (lo >> o1) & 1;
    return axis * 4 + u + 2 * v;
}

// Face f = 2 * axis + side: its four corners in cyclic order in the plane of
// the other two axes (lower axis first): (0,0), (1,0), (1,1), (0,1).
inline void face_corners(int f, int q[4]) {
    const int axis = f >> 1, side = f & 1;
    const int o0 = axis == 0 ? 1 : 0, o1 = axis == 2 ? 1 : 2;
    const int uv[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
    for (int t = 0; t < 4; ++t) q[t] = (side << axis) | (uv[t][0] << o0) | (uv[t][1] << o1);
}

// A cube face containing all three edges, or -1.
inline int face_shared(int e0, int e1, int e2) {
    for (int f = 0; f < 6; ++f) {
        int q[4];
        face_corners(f, q);
        int hit = 0;
        for (int t = 0; t < 4; ++t) {
            const int e = edge_of(q[t], q[(t + 1) & 3]);
            hit += (e == e0) + (e == e1) + (e == e2);
        }
        if (hit == 3) return f;
    }
    return -1;
}

}  // namespace mcr

// Pairs of crossed edges joined on one square with cyclic corners q[0..3]
// (inside flags in[0..3]); edge t joins q[t] and q[t+1].  Returns the number
// of pairs (0, 1 or 2) in seg[][2] as square-edge slots 0..3.
inline int mc_square_segments(const bool in[4], int seg[2][2]) {
    bool cr[4];
    int n = 0;
    for (int t = 0; t < 4; ++t) {
        cr[t] = in[t] != in[(t + 1) & 3];
        n += cr[t];
    }
    if (n == 2) {
        int a = -1, b = -1;
        for (int t = 0; t < 4; ++t)
            if (cr[t]) (a < 0 ? a : b) = t;
        seg[0][0] = a;
        seg[0][1] = b;
        return 1;
    }
    if (n == 4) {  // cut each inside corner off: corner m joins edges m-1 and m
        int k = 0;
        for (int m = 0; m < 4; ++m)
            if (in[m]) {
                seg[k][0] = (m + 3) & 3;
                seg[k][1] = m;
                ++k;
            }
        return k;
    }
    return 0;
}

// Cap polygons of a boundary square (cyclic corners q[0..3], inside flags
// in[0..3]) with the same face rule: slots 0-3 are the corners, 4-7 the
// crossing points on square edges 0-3 (edge t joins q[t] and q[t+1]).  The
// polygons follow the square's cyclic order.  Returns their number.
inline int mc_square_caps(const bool in[4], int poly[2][5], int len[2]) {
    bool cr[4];
    int n = 0;
    for (int t = 0; t < 4; ++t) {
        cr[t] = in[t] != in[(t + 1) & 3];
        n += cr[t];
    }
    if (n == 4) {  // each inside corner on its own
        int k = 0;
        for (int m = 0; m < 4; ++m)
            if (in[m]) {
                poly[k][0] = 4 + ((m + 3) & 3);
                poly[k][1] = m;
                poly[k][2] = 4 + m;
                len[k] = 3;
                ++k;
            }
        return k;
    }
    if (!in[0] && !in[1] && !in[2] && !in[3]) return 0;
    // one inside run: every inside corner and crossing point in cyclic order
    // from corner 0
    len[0] = 0;
    for (int t = 0; t < 4; ++t) {
        if (in[t]) poly[0][len[0]++] = t;
        if (cr[t]) poly[0][len[0]++] = 4 + t;
    }
    return 1;  // convex (the square cut by one segment): any fan triangulates it
}

inline McTables make_mc_tables() {
    McTables T;
    std::memset(T.ntri, 0, sizeof(T.ntri));
    std::memset(T.tri, -1, sizeof(T.tri));
    for (int cs = 0; cs < 256; ++cs) {
        auto inside = [&](int c) { return ((cs >> c) & 1) != 0; };
        // partner edges through every face
        int nbr[12][2];
        int nn[12] = {0};
        for (int f = 0; f < 6; ++f) {
            int q[4];
            mcr::face_corners(f, q);
            bool in[4];
            for (int t = 0; t < 4; ++t) in[t] = inside(q[t]);
            int seg[2][2];
            const int ns = mc_square_segments(in, seg);
            for (int s = 0; s < ns; ++s) {
                const int ea = mcr::edge_of(q[seg[s][0]], q[(seg[s][0] + 1) & 3]);
                const int eb = mcr::edge_of(q[seg[s][1]], q[(seg[s][1] + 1) & 3]);
                nbr[ea][nn[ea]++] = eb;
                nbr[eb][nn[eb]++] = ea;
            }
        }
        bool used[12] = {false};
        int nt = 0;
        for (int e0 = 0; e0 < 12; ++e0) {
            if (nn[e0] != 2 || used[e0]) continue;
            int loop[12], len = 0, prev = -1, cur = e0;
            do {
                loop[len++] = cur;
                used[cur] = true;
                const int nx = nbr[cur][0] != prev ? nbr[cur][0] : nbr[cur][1];
                prev = cur;
                cur = nx;
            } while (cur != e0 && len < 12);
            // orientation: Newell normal of the loop through edge midpoints
            // against the summed inside -> outside edge directions
            double P[12][3], N[3] = {0, 0, 0}, D[3] = {0, 0, 0};
            for (int t = 0; t < len; ++t) {
                int a, b;
                mcr::edge_corners(loop[t], a, b);
                for (int ax = 0; ax < 3; ++ax) {
                    P[t][ax] = 0.5 * (((a >> ax) & 1) + ((b >> ax) & 1));
                    const int in_c = inside(a) ? a : b, out_c = inside(a) ? b : a;
                    D[ax] += ((out_c >> ax) & 1) - ((in_c >> ax) & 1);
                }
            }
            for (int t = 0; t < len; ++t) {
                const double* p = P[t];
                const double* r = P[(t + 1) % len];
                N[0] += (p[1] - r[1]) * (p[2] + r[2]);
                N[1] += (p[2] - r[2]) * (p[0] + r[0]);
                N[2] += (p[0] - r[0]) * (p[1] + r[1]);
            }
            if (N[0] * D[0] + N[1] * D[1] + N[2] * D[2] < 0)
                for (int t = 1, u = len - 1; t < u; ++t, --u) {
                    const int tmp = loop[t];
                    loop[t] = loop[u];
                    loop[u] = tmp;
                }
            // fan from the first loop position whose triangles do not lie in
            // a cube face (three edges of one face): such a triangle would
            // overlap the neighbouring cube's surface on that face
            int s0 = 0;
            for (int s = 0; s < len; ++s) {
                bool ok = true;
                for (int t = 1; t + 1 < len && ok; ++t)
                    ok = mcr::face_shared(loop[s], loop[(s + t) % len], loop[(s + t + 1) % len]) < 0;
                if (ok) {
                    s0 = s;
                    break;
                }
            }
            for (int t = 1; t + 1 < len && nt < kMcMaxTri; ++t) {
                T.tri[cs][3 * nt] = (int8_t)loop[s0];
                T.tri[cs][3 * nt + 1] = (int8_t)loop[(s0 + t) % len];
                T.tri[cs][3 * nt + 2] = (int8_t)loop[(s0 + t + 1) % len];
                ++nt;
            }
        }
        T.ntri[cs] = (uint8_t)nt;
    }
    return T;
}

// Corner bits (x | y << 1 | z << 2) of the reference's cube corners 0-7
// (fe_mc_table.hpp).
constexpr int kMcRefCorner[8] = {0, 1, 5, 4, 2, 3, 7, 6};

inline int mc_hex(char c) { return c <= '9' ? c - '0' : c - 'a' + 10; }

// The reference's cube table (kMcRefCube) in this file's case and edge
// encoding, triangles in the table's order and winding.
inline McTables make_mc_tables_reference() {
    McTables T;
    std::memset(T.ntri, 0, sizeof(T.ntri));
    std::memset(T.tri, -1, sizeof(T.tri));
    for (int cs = 0; cs < 256; ++cs) {
        int rc = 0;
        for (int v = 0; v < 8; ++v)
            if ((cs >> kMcRefCorner[v]) & 1) rc |= 1 << v;
        const char* t = kMcRefCube[rc];
        const int n = (int)std::strlen(t);
        for (int q = 0; q < n; ++q) {
            const int e = mc_hex(t[q]);
            T.tri[cs][q] = (int8_t)mcr::edge_of(kMcRefCorner[kMcRefEdge[e][0]], kMcRefCorner[kMcRefEdge[e][1]]);
        }
        T.ntri[cs] = (uint8_t)(n / 3);
    }
    return T;
}

}  // namespace fe
