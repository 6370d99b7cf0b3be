// Cavity detection as connected-component labelling of the TEMP_OUT nodes
// (6-connectivity) with a lock-free union-find over the x-runs of TEMP_OUT
// nodes: links always go from the larger root to the smaller, so each
// component's root is its smallest flat index, i.e. the first node the forward scan of
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

// Runs of TEMP_OUT nodes along x are the union-find nodes: one wave walks a
// row in 64-node chunks, finds run starts with a ballot and points every node
// at the first node of its run (the run's smallest flat index).
__global__ void __launch_bounds__(256) k_ccl_rows(DevGrid g, const int *status, int *P) {
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= (int64_t)g.ny * g.nz) return;
    const int lane = threadIdx.x & 63;
    const int64_t base = row * g.nx;
    int carry = 0;  // run start (x index) of a run entering the chunk from the left
    for (int c0 = 0; c0 < g.nx; c0 += 64) {
        const int i = c0 + lane;
        const bool in = i < g.nx;
        const bool t = in && status[base + i] == NSGPU_STATUS_TEMP_OUT;
        const unsigned long long stop = __ballot(in && !t);
        const unsigned long long below = stop & ((lane == 0) ? 0ull : (~0ull >> (64 - lane)));
        const int start = below ? c0 + (63 - __clzll(below)) + 1 : carry;
        if (in) P[base + i] = t ? (int)(base + start) : -1;
        // the run reaching the end of this chunk continues into the next one
        const bool last_t = __shfl(t ? 1 : 0, 63, 64) != 0;
        const int last_start = __shfl(start, 63, 64);
        carry = last_t ? last_start : c0 + 64;
    }
}

// Links between runs of neighbouring rows (+y) and planes (+z), made once per
// overlap interval: at a node whose left neighbour does not also overlap.
__global__ void __launch_bounds__(256) k_ccl_link(DevGrid g, const int *status, int *P) {
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nvox || status[f] != NSGPU_STATUS_TEMP_OUT) return;
    const int i = (int)(f % g.nx);
    const int j = (int)((f / g.nx) % g.ny);
    const int k = (int)(f / ((int64_t)g.nx * g.ny));
    const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
    const bool left = i > 0 && status[f - 1] == NSGPU_STATUS_TEMP_OUT;
    if (j + 1 < g.ny && status[f + sy] == NSGPU_STATUS_TEMP_OUT &&
        !(left && status[f - 1 + sy] == NSGPU_STATUS_TEMP_OUT))
        unite(P, ld(P + f), ld(P + f + sy));
    if (k + 1 < g.nz && status[f + sz] == NSGPU_STATUS_TEMP_OUT &&
        !(left && status[f - 1 + sz] == NSGPU_STATUS_TEMP_OUT))
        unite(P, ld(P + f), ld(P + f + sz));
}

// Every node is pointed at its root.  The walk only reads: a compressing walk
// could overwrite a node another lane has just flattened with a stale
// ancestor.  Each lane writes its own entry only, so after the kernel every
// entry is the root.
__global__ void __launch_bounds__(256) k_ccl_flatten(int *P, int64_t nvox) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nvox) return;
    int x = ld(P + f);
    if (x < 0) return;
    for (int p = ld(P + x); p != x; p = ld(P + x)) x = p;
    st(P + f, x);
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

// Size and backward-scan key of every component.  The component of node 0
// (when node 0 is TEMP_OUT) is the first one any scan meets, so it is always
// labelled first and neither key is needed for it: it is skipped, which
// removes the only heavily contended address (the outside region).  Waves
// whose lanes all belong to one component combine before the atomics.
__global__ void __launch_bounds__(256) k_ccl_keys(DevGrid g, const int *P, unsigned int *bkey, int *count) {
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int skip = slot_of(P, 0);
    int s = -1;
    unsigned int rank = 0xffffffffu;
    if (f < nvox) {
        s = slot_of(P, f);
        if (s == skip) s = -1;
        const int64_t plane = (int64_t)g.nx * g.ny;
        const int64_t k = f / plane;
        rank = (unsigned int)((g.nz - 1 - k) * plane + (f - k * plane));
    }
    const bool valid = s >= 0;
    const unsigned long long act = __ballot(valid);
    if (act == 0) return;
    const int leader = __ffsll((long long)act) - 1;
    const int s0 = __shfl(s, leader, 64);
    if (__ballot(valid && s == s0) == act) {
        unsigned int r = valid ? rank : 0xffffffffu;
        for (int off = 32; off > 0; off >>= 1) r = min(r, (unsigned int)__shfl_down((int)r, off, 64));
        r = (unsigned int)__shfl((int)r, 0, 64);
        if ((int)(threadIdx.x & 63) == leader) {
            atomicMin(bkey + s0, r);
            atomicAdd(count + s0, __popcll(act));
        }
        return;
    }
    if (valid) {
        atomicMin(bkey + s, rank);
        atomicAdd(count + s, 1);
    }
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
    const int64_t rows = (int64_t)g.ny * g.nz;
    hipLaunchKernelGGL(k_ccl_rows, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, s, g, status, P);
    hipLaunchKernelGGL(k_ccl_link, dim3(nblk(nvox)), dim3(256), 0, s, g, status, P);
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
