// fe_seed.cpp -- deterministic analytic scene seeding (host only).
//
// The reference seeds particles in PICSimulator::InitializeParticles
// (src/pic_simulator.cpp:521-548) with std::rand() from several threads, which
// is not reproducible.  This replacement fills every cell whose centre lies in
// the union of the given boxes and spheres with 2x2x2 particles, one per
// sub-cell, at cell_lo + (s + U) * h/2 where U in [0,1) comes from a
// counter-based splitmix64 keyed on (seed, cell, sub-cell, axis), so any two
// callers produce the same FP32 positions.
#include <cmath>
#include <cstdint>

#include "../../include/fluide_hip.h"

namespace {

inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline double uniform01(uint64_t seed, uint64_t cell, int sub, int axis) {
    const uint64_t ctr = ((cell * 8u + (uint64_t)sub) * 4u + (uint64_t)axis);
    const uint64_t r = splitmix64(splitmix64(seed) ^ ctr);
    return (double)(r >> 11) * (1.0 / 9007199254740992.0);  // 53 bits -> [0, 1)
}

bool inside(const fe_seed_shape& s, const double c[3]) {
    if (s.kind == 0) {
        for (int a = 0; a < 3; ++a)
            if (!(c[a] >= s.lo[a] && c[a] < s.hi[a])) return false;
        return true;
    }
    if (s.kind == 1) {
        const double dx = c[0] - s.lo[0], dy = c[1] - s.lo[1], dz = c[2] - s.lo[2];
        return dx * dx + dy * dy + dz * dz < s.hi[0] * s.hi[0];
    }
    return false;
}

}  // namespace

extern "C" int fe_seed_particles(const fe_scene_desc* desc, const fe_seed_shape* shapes, int32_t nshapes,
                                 uint64_t seed, float* x, int64_t capacity, int64_t* count) {
    if (!desc || (!shapes && nshapes > 0) || !count) return FE_E_INVALID;
    for (int a = 0; a < 3; ++a)
        if (desc->n[a] < 1 || !(desc->size[a] > 0)) return FE_E_INVALID;
    double h[3];
    for (int a = 0; a < 3; ++a) h[a] = desc->size[a] / desc->n[a];
    int64_t m = 0;
    for (int k = 0; k < desc->n[2]; ++k)
        for (int j = 0; j < desc->n[1]; ++j)
            for (int i = 0; i < desc->n[0]; ++i) {
                const int idx[3] = {i, j, k};
                double c[3];
                for (int a = 0; a < 3; ++a) c[a] = desc->origin[a] + h[a] * (idx[a] + 0.5);
                bool in = false;
                for (int q = 0; q < nshapes && !in; ++q) in = inside(shapes[q], c);
                if (!in) continue;
                const uint64_t cell = (uint64_t)i + (uint64_t)desc->n[0] * ((uint64_t)j + (uint64_t)desc->n[1] * k);
                for (int sub = 0; sub < 8; ++sub) {
                    if (x) {
                        if (m >= capacity) return FE_E_INVALID;
                        for (int a = 0; a < 3; ++a) {
                            const int sa = (sub >> a) & 1;
                            const double u = uniform01(seed, cell, sub, a);
                            x[3 * m + a] = (float)(desc->origin[a] + h[a] * (idx[a] + (sa + u) * 0.5));
                        }
                    }
                    ++m;
                }
            }
    *count = m;
    return FE_OK;
}
