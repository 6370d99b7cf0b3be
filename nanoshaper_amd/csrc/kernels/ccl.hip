// Cavity detection as connected-component labelling of the TEMP_OUT nodes
// (6-connectivity) with a lock-free union-find: links always go from the
// larger root to the smaller, so each component's root is its smallest flat
// index, i.e. the first node the forward scan of
// getCavitiesWithBilevelStatusMap (Surface.cpp:1994-2126) meets.  The first
// node of the backward-z scan is recorded per component as a second key; the
// host replays the alternating scan order over the components to give each
// one the id the reference's flood fills (Surface.cpp:4284-4554) assign.
#include "launch.hpp"

namespace nsg {

__device__ __forceinline__ int ld(const int *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void st(int *p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// root with path halving; only non-roots are rewritten, always to an ancestor
__device__ int find_root(int *P, int x) {
    int p = ld(P + x);
    while (p != x) {
        const int gp = ld(P + p);
        if (gp != p) st(P + x, gp);
        x = p;
        p = gp;
    }
    return x;
}

__device__ void unite(int *P, int a, int b) {
    for (;;) {
        a = find_root(P, a);
        b = find_root(P, b);
        if (a == b) return;
        if (a < b) {
            const int t = a;
            a = b;
            b = t;
        }
        const int old = atomicCAS(P + a, a, b);
        if (old == a) return;
        a = old;
    }
}

__global__ void __launch_bounds__(256) k_ccl_init(const int *status, int *P, int64_t nvox) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nvox) P[f] = status[f] == NSGPU_STATUS_TEMP_OUT ? (int)f : -1;
}

__global__ void __launch_bounds__(256) k_ccl_merge(DevGrid g, int *P) {
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nvox || ld(P + f) < 0) return;
    const int i = (int)(f % g.nx);
    const int j = (int)((f / g.nx) % g.ny);
    const int k = (int)(f / ((int64_t)g.nx * g.ny));
    const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
    // skip the link when the neighbour already has the same parent (the common case)
    if (i + 1 < g.nx && ld(P + f + 1) >= 0 && ld(P + f + 1) != ld(P + f)) unite(P, (int)f, (int)(f + 1));
    if (j + 1 < g.ny && ld(P + f + sy) >= 0) unite(P, (int)f, (int)(f + sy));
    if (k + 1 < g.nz && ld(P + f + sz) >= 0) unite(P, (int)f, (int)(f + sz));
}

__global__ void __launch_bounds__(256) k_ccl_flatten(int *P, int64_t nvox) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nvox && ld(P + f) >= 0) st(P + f, find_root(P, (int)f));
}

// roots get a compact slot, stored in place as -(slot+2)
__global__ void __launch_bounds__(256) k_ccl_roots(int *P, int64_t nvox, int *nroots, int *rootlist) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nvox && P[f] == (int)f) {
        const int s = atomicAdd(nroots, 1);
        rootlist[s] = (int)f;
        P[f] = -(s + 2);
    }
}

__device__ __forceinline__ int slot_of(const int *P, int64_t f) {
    const int p = P[f];
    if (p == -1) return -1;
    return p < -1 ? -(p + 2) : -(P[p] + 2);
}

__global__ void __launch_bounds__(256) k_ccl_keys(DevGrid g, const int *P, unsigned int *bkey, int *count) {
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nvox) return;
    const int s = slot_of(P, f);
    if (s < 0) return;
    const int64_t plane = (int64_t)g.nx * g.ny;
    const int64_t k = f / plane;
    const unsigned int rank = (unsigned int)((g.nz - 1 - k) * plane + (f - k * plane));
    atomicMin(bkey + s, rank);
    atomicAdd(count + s, 1);
}

// final labels and fillCavities (Surface.cpp:2355-2487) for the filled ones
__global__ void __launch_bounds__(256) k_ccl_label(DevGrid g, const int *P, const int *label, const uint8_t *filled,
                                                   int build_eps, int accurate, int inside, int *status, int *epsmap,
                                                   uint8_t *idebmap, uint32_t *bits) {
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nvox) return;
    const int s = slot_of(P, f);
    if (s < 0) return;
    status[f] = label[s];
    if (!filled[s]) return;
    const int i = (int)(f % g.nx);
    const int j = (int)((f / g.nx) % g.ny);
    const int k = (int)(f / ((int64_t)g.nx * g.ny));
    const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
    if (build_eps) {
        epsmap[3 * f + 0] = inside;
        epsmap[3 * f + 1] = inside;
        epsmap[3 * f + 2] = inside;
        if (i > 0) epsmap[3 * (f - 1) + 0] = inside;
        if (j > 0) epsmap[3 * (f - sy) + 1] = inside;
        if (k > 0) epsmap[3 * (f - sz) + 2] = inside;
        idebmap[f] = 0;
    }
    if (accurate) {
        for (int dk = 0; dk < 2; ++dk)
            for (int dj = 0; dj < 2; ++dj)
                for (int di = 0; di < 2; ++di) {
                    if (i + di >= g.nx || j + dj >= g.ny || k + dk >= g.nz) continue;
                    const int64_t q = f + di + dj * sy + dk * sz;
                    atomicAnd(bits + (q >> 5), ~(1u << (uint32_t)(q & 31)));
                }
    }
}

static inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

void launch_ccl(hipStream_t s, const DevGrid &g, const int *status, int *P, int *nroots, int *rootlist) {
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_ccl_init, dim3(nblk(nvox)), dim3(256), 0, s, status, P, nvox);
    hipLaunchKernelGGL(k_ccl_merge, dim3(nblk(nvox)), dim3(256), 0, s, g, P);
    hipLaunchKernelGGL(k_ccl_flatten, dim3(nblk(nvox)), dim3(256), 0, s, P, nvox);
    hipLaunchKernelGGL(k_ccl_roots, dim3(nblk(nvox)), dim3(256), 0, s, P, nvox, nroots, rootlist);
}

void launch_ccl_keys(hipStream_t s, const DevGrid &g, const int *P, unsigned int *bkey, int *count) {
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_ccl_keys, dim3(nblk(nvox)), dim3(256), 0, s, g, P, bkey, count);
}

void launch_ccl_label(hipStream_t s, const DevGrid &g, const int *P, const int *label, const uint8_t *filled,
                      int build_eps, int accurate, int inside, int *status, int *eps, uint8_t *ideb,
                      uint32_t *bits) {
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_ccl_label, dim3(nblk(nvox)), dim3(256), 0, s, g, P, label, filled, build_eps, accurate,
                       inside, status, eps, ideb, bits);
}

}  // namespace nsg
