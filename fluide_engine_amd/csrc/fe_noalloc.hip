// fe_noalloc.hip -- the device memory of fe_mesh_sdf_noalloc (Mesh2SDF of a
// triangle mesh on a node grid with no device allocation, fe_capi.hip).  Every
// device buffer of such a call (the raw mesh, the packed triangles with their
// index boxes, an error word and the SDF) is a slice of one file-scope device
// array, `scratch`, which each GPU holds from library load to process end.
// Calls that overlap on one GPU share that array and race; calls on different
// GPUs do not.
#include <mutex>

#include "fe_internal.hpp"
#include "fe_kernels.hpp"

namespace fe {

// The scratch layout (byte offsets, 256-byte aligned) for the limits of one
// call (kNaMaxVerts, kNaMaxTris, kNaMaxNodes in fe_kernels.hpp).
constexpr size_t kNaVerts = 0;                                           // 3 doubles per vertex
constexpr size_t kNaTris = kNaVerts + sizeof(double) * 3 * kNaMaxVerts;  // 3 int32 per triangle
constexpr size_t kNaTv = kNaTris + sizeof(int32_t) * 3 * kNaMaxTris;     // 9 doubles per triangle
constexpr size_t kNaTb = kNaTv + sizeof(double) * 9 * kNaMaxTris;        // 6 int32 per triangle
constexpr size_t kNaFlag = kNaTb + sizeof(int32_t) * 6 * kNaMaxTris;     // one int32 (256 bytes reserved)
constexpr size_t kNaSdf = kNaFlag + 256;                                 // 1 double per node
constexpr size_t kNaBytes = kNaSdf + sizeof(double) * kNaMaxNodes;
static_assert(kNaTris % 256 == 0 && kNaTv % 256 == 0 && kNaTb % 256 == 0 && kNaFlag % 256 == 0 &&
                  kNaSdf % 256 == 0,
              "scratch slices are 256-byte aligned");

// In host code this name is a host-side stand-in, not a device address: host
// code uses only the per-device pointer from hipGetSymbolAddress below.
static __device__ __align__(256) unsigned char scratch[kNaBytes];

constexpr int kNaMaxDevices = 64;
static std::mutex g_na_mu;
static unsigned char* g_na_ptr[kNaMaxDevices] = {nullptr};

// Once per device, with that device current: the device address of `scratch`.
hipError_t noalloc_setup(int device) {
    if (device < 0 || device >= kNaMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (g_na_ptr[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_na_ptr[device] = static_cast<unsigned char*>(p);
    return hipSuccess;
}

// The device form of mesh_prepare (k_collider.hip): the packed vertices and
// the index box of each triangle (vertex / h.x, widened by 1 below and 2
// above, clamped to the grid), with the same operations; a vertex index out
// of range or a quotient beyond the int range sets *bad.
__global__ __launch_bounds__(kBlock) void k_mesh_prepare(NodeGrid3 G, const double* __restrict__ verts, int64_t nv,
                                                         const int32_t* __restrict__ tris, int64_t nt, double inv,
                                                         double* __restrict__ tv, int32_t* __restrict__ tb,
                                                         int32_t* __restrict__ bad) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= nt) return;
    double f[3][3];  // f[axis][vertex]
    bool ok = true;
    for (int q = 0; q < 3; ++q) {
        const int32_t vi = tris[3 * t + q];
        const bool in = vi >= 0 && vi < nv;
        ok = ok && in;
        for (int a = 0; a < 3; ++a) {
            const double x = in ? verts[3 * (int64_t)vi + a] : 0.0;
            tv[9 * t + 3 * q + a] = x;
            f[a][q] = x * inv;
            ok = ok && fabs(f[a][q]) < 2.0e9;
        }
    }
    if (!ok) {
        atomicOr(bad, 1);
        return;
    }
    for (int a = 0; a < 3; ++a) {
        const double lo = fmin(f[a][0], fmin(f[a][1], f[a][2]));
        const double hi = fmax(f[a][0], fmax(f[a][1], f[a][2]));
        const int l = (int)lo - 1, h = (int)hi + 1 + 1;
        tb[6 * t + a] = l < 0 ? 0 : (l > G.n[a] - 1 ? G.n[a] - 1 : l);
        tb[6 * t + 3 + a] = h < 0 ? 0 : (h > G.n[a] - 1 ? G.n[a] - 1 : h);
    }
}

hipError_t noalloc_slices(int device, NoallocSlices* out) {
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (device < 0 || device >= kNaMaxDevices || !g_na_ptr[device]) return hipErrorNotInitialized;
    unsigned char* base = g_na_ptr[device];
    out->verts = reinterpret_cast<double*>(base + kNaVerts);
    out->tris = reinterpret_cast<int32_t*>(base + kNaTris);
    out->tv = reinterpret_cast<double*>(base + kNaTv);
    out->tb = reinterpret_cast<int32_t*>(base + kNaTb);
    out->bad = reinterpret_cast<int32_t*>(base + kNaFlag);
    out->sdf = reinterpret_cast<double*>(base + kNaSdf);
    return hipSuccess;
}

hipError_t launch_mesh_prepare(hipStream_t st, const NodeGrid3& G, const double* verts, int64_t nv,
                               const int32_t* tris, int64_t nt, double* tv, int32_t* tb, int32_t* bad) {
    if (nt <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mesh_prepare, dim3((unsigned)((nt + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, G, verts,
                       nv, tris, nt, 1.0 / G.h[0], tv, tb, bad);
    return hipGetLastError();
}

}  // namespace fe
