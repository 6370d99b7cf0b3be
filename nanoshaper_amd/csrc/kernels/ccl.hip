// Cavity detection as connected-component labelling of the TEMP_OUT nodes
// (6-connectivity).  The union-find nodes are the x-runs of TEMP_OUT nodes,
// not the nodes: a 325^3 grid has ~34 M nodes but only a few hundred thousand
// runs, so only the row scan and the final label pass touch the status map.
//
// Run storage: run q of row r = k*ny + j (q < rowcnt[r], in x order) is slot
// r + q*rows of rx0/rx1 (its first and last x), rows = ny*nz; a row holds at
// most ceil(nx/2) runs.  Run-major slots put the q-th runs of consecutive
// rows in consecutive words, so the kernels' threads (one per row) read them
// coalesced.  The union-find runs on compact run ids
// c = roff[r] + q (roff: exclusive scan of rowcnt), which grow with the flat
// index of the run's first node.  Unions are made in two levels: inside each
// z-plane in LDS (one workgroup per plane, +y links), then across planes (+z
// links) on the plane roots in global memory, where the outside region is one
// component per plane and the cross-plane links are few distinct pairs.
//
// Links always go from the larger root to the smaller, so each component's
// root is its run of smallest id, whose first node is the component's
// smallest flat index: the first node the forward scan of
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

// One thread per row: the row's TEMP_OUT nodes are the zero bits of the x
// grid ray's status plane (the fill wrote status = inside exactly where that
// plane is set), so run starts and ends come from 64-node words with bit
// operations and the status map itself is not read.  The k-th start and the
// k-th end of a row belong to the same run.
// Also zeroes *nroots for k_ccl_roots and *nlist for k_ccl_flatten.
__global__ void __launch_bounds__(256) k_ccl_runs(DevGrid g, const unsigned long long *inside, int N, int64_t rs,
                                                  int *rx0, int *rx1, int *rowcnt, int *nroots, int *nlist) {
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row == 0) {
        *nroots = 0;
        *nlist = 0;
    }
    if (row >= (int64_t)g.ny * g.nz) return;
    const int j = (int)(row % g.ny), k = (int)(row / g.ny);
    const int64_t NN = (int64_t)N * N, ray = (int64_t)k * N + j;
    const int nw = (g.nx + 63) / 64;
    auto temp_out = [&](int w) -> unsigned long long {  // TEMP_OUT bits of word w (0 past the row)
        if (w >= nw) return 0ull;
        unsigned long long t = ~inside[(int64_t)w * NN + ray];
        const int valid = g.nx - 64 * w;
        if (valid < 64) t &= (1ull << valid) - 1ull;
        return t;
    };
    int ns = 0, ne = 0;
    unsigned long long prev = 0ull, t = temp_out(0);
    for (int w = 0; w < nw; ++w) {
        const unsigned long long next = temp_out(w + 1);
        unsigned long long starts = t & ~((t << 1) | prev);
        unsigned long long ends = t & ~((t >> 1) | (next << 63));
        while (starts) {
            rx0[row + rs * ns++] = 64 * w + __builtin_ctzll(starts);
            starts &= starts - 1ull;
        }
        while (ends) {
            rx1[row + rs * ne++] = 64 * w + __builtin_ctzll(ends);
            ends &= ends - 1ull;
        }
        prev = t >> 63;
        t = next;
    }
    rowcnt[row] = ns;
}

// root in LDS with path halving (the LDS twin of find_root)
__device__ int find_root_lds(int *P, int x) {
    int p = __hip_atomic_load(P + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    while (p != x) {
        const int gp = __hip_atomic_load(P + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (gp != p) __hip_atomic_store(P + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        x = p;
        p = gp;
    }
    return x;
}

__device__ void unite_lds(int *P, int a, int b) {
    for (;;) {
        a = find_root_lds(P, a);
        b = find_root_lds(P, b);
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

// merge walk of the sorted run lists of rows ra and rb; f(a, b) for every
// overlapping pair
template <class F>
__device__ __forceinline__ void overlaps(const int *rx0, const int *rx1, int64_t rs, int64_t ra, int na, int64_t rb,
                                         int nb, F f) {
    int a = 0, b = 0;
    while (a < na && b < nb) {
        const int64_t sa = ra + rs * a, sb = rb + rs * b;
        const int a0 = rx0[sa], a1 = rx1[sa], b0 = rx0[sb], b1 = rx1[sb];
        if (a1 < b0) {
            ++a;
        } else if (b1 < a0) {
            ++b;
        } else {
            f(a, b);
            if (a1 < b1)
                ++a;
            else
                ++b;
        }
    }
}

constexpr int kPlaneRuns = 16384;  // runs of one z-plane united in LDS

// One workgroup per z-plane: +y unions of the plane's runs, then every run
// of the plane points at its plane-level root.  A plane with more runs than
// the LDS holds unites in global memory instead.
__global__ void __launch_bounds__(1024) k_ccl_plane(DevGrid g, int64_t rs, const int *rx0, const int *rx1,
                                                    const int *rowcnt, const int *roff, int *R) {
    __shared__ int lp[kPlaneRuns];
    const int64_t row0 = (int64_t)blockIdx.x * g.ny, rowl = row0 + g.ny - 1;
    const int base = roff[row0];
    const int T = roff[rowl] + rowcnt[rowl] - base;
    if (T == 0) return;
    const bool in_lds = T <= kPlaneRuns;
    for (int l = threadIdx.x; l < T; l += blockDim.x) {
        if (in_lds)
            lp[l] = l;
        else
            st(R + base + l, base + l);
    }
    __syncthreads();
    for (int j = threadIdx.x; j + 1 < g.ny; j += blockDim.x) {
        const int64_t ra = row0 + j, rb = ra + 1;
        const int la = roff[ra] - base, lb = roff[rb] - base;
        overlaps(rx0, rx1, rs, ra, rowcnt[ra], rb, rowcnt[rb], [&](int a, int b) {
            if (in_lds)
                unite_lds(lp, la + a, lb + b);
            else
                unite(R, base + la + a, base + lb + b);
        });
    }
    __syncthreads();
    // read-only walks; each thread then writes its own entries, always a root
    for (int l = threadIdx.x; l < T; l += blockDim.x) {
        if (in_lds) {
            int x = lp[l];
            for (int p = lp[x]; p != x; p = lp[x]) x = p;
            R[base + l] = base + x;
        } else {
            int x = ld(R + base + l);
            for (int p = ld(R + x); p != x; p = ld(R + x)) x = p;
            st(R + base + l, x);
        }
    }
}

// +z links, one thread per row.  After the plane pass every run's entry is
// its plane root, so the overlaps of a wave's rows mostly repeat one root
// pair (the outside region of plane k and of plane k+1): the wave's first
// lane with an overlap unites its first pair for the whole wave, and every
// lane skips that pair and repeats of its own previous pair.
__global__ void __launch_bounds__(256) k_ccl_zlink(DevGrid g, int64_t rs, const int *rx0, const int *rx1,
                                                   const int *rowcnt, const int *roff, int *R) {
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = row < (int64_t)g.ny * (g.nz - 1);
    const int64_t other = row + g.ny;
    const int ca = valid ? roff[row] : 0, cb = valid ? roff[other] : 0;
    const int na = valid ? rowcnt[row] : 0, nb = valid ? rowcnt[other] : 0;
    // first overlapping pair of this row
    int fa = -1, fb = -1;
    overlaps(rx0, rx1, rs, row, na, other, nb, [&](int a, int b) {
        if (fa < 0) {
            fa = ld(R + ca + a);
            fb = ld(R + cb + b);
        }
    });
    const unsigned long long act = __ballot(fa >= 0);
    if (act == 0) return;
    const int leader = __ffsll((long long)act) - 1;
    const int wa = __shfl(fa, leader, 64), wb = __shfl(fb, leader, 64);
    if ((int)(threadIdx.x & 63) == leader) unite(R, wa, wb);
    if (fa < 0) return;
    int last_a = wa, last_b = wb;
    overlaps(rx0, rx1, rs, row, na, other, nb, [&](int a, int b) {
        const int pa = ld(R + ca + a), pb = ld(R + cb + b);
        if ((pa == last_a && pb == last_b) || (pa == wa && pb == wb)) return;
        last_a = pa;
        last_b = pb;
        unite(R, pa, pb);
    });
}

// Every run is pointed at its root: read-only walks, each thread writes its
// own row's entries, always with a root.  The runs outside the component of
// run 0 go to a compact list: that component has the smallest root, so the
// replay always numbers it first (NSGPU_STATUS_OUT, never a cavity), and the
// key and label passes need only the others, a small fraction of the runs.
__global__ void __launch_bounds__(256) k_ccl_flatten(int64_t rows, const int *rowcnt, const int *roff, int *R,
                                                     int *list_c, int *list_row, int *nlist) {
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    const int n = rowcnt[row], c0 = roff[row];
    for (int q = 0; q < n; ++q) {
        int x = ld(R + c0 + q);
        for (int p = ld(R + x); p != x; p = ld(R + x)) x = p;
        st(R + c0 + q, x);
        if (x != 0) {
            const int at = atomicAdd(nlist, 1);
            list_c[at] = c0 + q;
            list_row[at] = (int)row;
        }
    }
}

// roots get a compact slot, stored in place as -(slot+2); rootlist holds the
// root's first node (the component's smallest flat index)
// and the keys of slots below keycap start cleared (bkey all ones, count 0)
__global__ void __launch_bounds__(256) k_ccl_roots(DevGrid g, int64_t rs, const int *rx0, const int *rowcnt,
                                                   const int *roff, int *R, int *nroots, int *rootlist,
                                                   unsigned int *bkey, int *count, int keycap) {
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= (int64_t)g.ny * g.nz) return;
    const int n = rowcnt[row], c0 = roff[row];
    for (int q = 0; q < n; ++q) {
        if (R[c0 + q] != c0 + q) continue;
        const int s = atomicAdd(nroots, 1);
        rootlist[s] = (int)(row * g.nx + rx0[row + rs * q]);
        R[c0 + q] = -(s + 2);
        if (s < keycap) {
            bkey[s] = 0xffffffffu;
            count[s] = 0;
        }
    }
}

__device__ __forceinline__ int slot_of(const int *R, int c) {
    const int p = R[c];
    return p < -1 ? -(p + 2) : -(R[p] + 2);
}

// Size and backward-scan key of every component, per run: the run's first
// node has its smallest backward rank (k reversed, then j, then i).  The
// component of run 0 is numbered first by every replay, so neither key is
// needed for it: only the listed runs are visited (grid-stride over the list).
// A cavity's hundreds of runs would all hit its two key words, so a wave
// first combines its lanes per component (min of the ranks, sum of the
// sizes) and one lane per component does the two atomics.
__global__ void __launch_bounds__(256) k_ccl_keys(DevGrid g, int64_t rs, const int *rx0, const int *rx1,
                                                  const int *roff, const int *R, const int *list_c,
                                                  const int *list_row, const int *nlist, unsigned int *bkey,
                                                  int *count, int capacity) {
    const int nl = *nlist;
    const int lane = threadIdx.x & 63;
    const int64_t plane = (int64_t)g.nx * g.ny;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < nl; base += stride) {
        const int64_t e = base + lane;
        int s = -1, size = 0;
        unsigned int rank = 0xffffffffu;
        if (e < nl) {
            const int c = list_c[e], row = list_row[e];
            s = slot_of(R, c);
            if (s < capacity) {  // beyond capacity: the replay falls back to the host
                const int j = row % g.ny, k = row / g.ny;
                const int64_t id = row + rs * (c - roff[row]);
                rank = (unsigned int)((g.nz - 1 - k) * plane + (int64_t)j * g.nx + rx0[id]);
                size = rx1[id] - rx0[id] + 1;
            } else {
                s = -1;
            }
        }
        unsigned long long pending = __ballot(s >= 0);
        while (pending) {
            const int leader = __ffsll((unsigned long long)pending) - 1;
            const int key = __shfl(s, leader, 64);
            const bool mine = s == key;
            unsigned int r = mine ? rank : 0xffffffffu;
            int z = mine ? size : 0;
            for (int off = 32; off > 0; off >>= 1) {
                r = min(r, (unsigned int)__shfl_xor((int)r, off, 64));
                z += __shfl_xor(z, off, 64);
            }
            if (lane == leader) {
                atomicMin(bkey + key, r);
                atomicAdd(count + key, z);
            }
            pending &= ~__ballot(mine);
        }
    }
}

// Final labels and fillCavities (Surface.cpp:2355-2487) for the filled
// components: one wave per listed run (grid-stride), lanes stride over the
// run's nodes.  The map fill already wrote NSGPU_STATUS_OUT at every TEMP_OUT
// node, the label of the component of run 0 (the first one numbered), so only
// the listed runs of the other components are written here.
__global__ void __launch_bounds__(256) k_ccl_label(DevGrid g, int64_t rs, const int *rx0, const int *rx1,
                                                   const int *roff, const int *R, const int *list_c,
                                                   const int *list_row, const int *nlist, const int *label,
                                                   const uint8_t *filled, const int *done, int build_eps, int accurate,
                                                   int inside, int *status, int *epsmap, uint8_t *idebmap,
                                                   uint32_t *bits) {
    if (*done == 0) return;
    const int nl = *nlist;
    const int lane = threadIdx.x & 63;
    const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
    for (int e = blockIdx.x * 4 + (threadIdx.x >> 6); e < nl; e += gridDim.x * 4) {
        const int c = list_c[e], row = list_row[e];
        const int j = row % g.ny, k = row / g.ny;
        const int64_t base = (int64_t)row * g.nx;
        const int s = slot_of(R, c);
        const int lab = label[s];
        if (lab == NSGPU_STATUS_OUT) continue;  // never filled: not a cavity
        const bool fill = filled[s] != 0;
        const int64_t id = row + rs * (c - roff[row]);
        const int x1 = rx1[id];
        for (int i = rx0[id] + lane; i <= x1; i += 64) {
            const int64_t f = base + i;
            status[f] = lab;
            if (!fill) continue;
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
                            const int64_t qn = f + di + dj * sy + dk * sz;
                            atomicAnd(bits + (qn >> 5), ~(1u << (uint32_t)(qn & 31)));
                        }
            }
        }
    }
}

static inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

// grid of the passes over the run list (its length is known on the device only)
constexpr unsigned kListBlocks = 1024;

int ccl_row_capacity(const DevGrid &g) { return (g.nx + 1) / 2; }

void launch_ccl(hipStream_t s, const DevGrid &g, const DevRays &R, const CclRuns &c, int *scan_scratch, int *nroots,
                int *rootlist, unsigned int *bkey, int *count, int keycap) {
    const int64_t rows = (int64_t)g.ny * g.nz;
    hipLaunchKernelGGL(k_ccl_runs, dim3(nblk(rows)), dim3(256), 0, s, g, R.cov_st, R.N, rows, c.rx0, c.rx1, c.rowcnt,
                       nroots, c.nlist);
    launch_scan(s, c.rowcnt, rows, c.roff, scan_scratch);
    hipLaunchKernelGGL(k_ccl_plane, dim3((unsigned)g.nz), dim3(1024), 0, s, g, rows, c.rx0, c.rx1, c.rowcnt, c.roff,
                       c.R);
    if (g.nz > 1)
        hipLaunchKernelGGL(k_ccl_zlink, dim3(nblk((int64_t)g.ny * (g.nz - 1))), dim3(256), 0, s, g, rows, c.rx0, c.rx1,
                           c.rowcnt, c.roff, c.R);
    hipLaunchKernelGGL(k_ccl_flatten, dim3(nblk(rows)), dim3(256), 0, s, rows, c.rowcnt, c.roff, c.R, c.list_c,
                       c.list_row, c.nlist);
    hipLaunchKernelGGL(k_ccl_roots, dim3(nblk(rows)), dim3(256), 0, s, g, rows, c.rx0, c.rowcnt, c.roff, c.R, nroots,
                       rootlist, bkey, count, keycap);
}

void launch_ccl_keys(hipStream_t s, const DevGrid &g, const CclRuns &c, unsigned int *bkey, int *count,
                     int capacity) {
    const int64_t rows = (int64_t)g.ny * g.nz;
    hipLaunchKernelGGL(k_ccl_keys, dim3(kListBlocks), dim3(256), 0, s, g, rows, c.rx0, c.rx1, c.roff, c.R,
                       c.list_c, c.list_row, c.nlist, bkey, count, capacity);
}

void launch_ccl_label(hipStream_t s, const DevGrid &g, const CclRuns &c, const int *label, const uint8_t *filled,
                      const int *done, int build_eps, int accurate, int inside, int *status, int *eps, uint8_t *ideb,
                      uint32_t *bits) {
    const int64_t rows = (int64_t)g.ny * g.nz;
    hipLaunchKernelGGL(k_ccl_label, dim3(kListBlocks), dim3(256), 0, s, g, rows, c.rx0, c.rx1, c.roff, c.R,
                       c.list_c, c.list_row, c.nlist, label, filled, done, build_eps, accurate, inside, status, eps,
                       ideb, bits);
}

// keys of the *nroots components start at (no key, size 0)
__global__ void __launch_bounds__(256) k_ccl_clear_keys(const int *nroots, unsigned int *bkey, int *count,
                                                        int64_t capacity) {
    const int64_t n = min((int64_t)*nroots, capacity);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        bkey[i] = 0xffffffffu;
        count[i] = 0;
    }
}

void launch_ccl_clear_keys(hipStream_t s, const int *nroots, unsigned int *bkey, int *count, int64_t capacity) {
    hipLaunchKernelGGL(k_ccl_clear_keys, dim3(256), dim3(256), 0, s, nroots, bkey, count, capacity);
}

// bitonic sort of n (power of two, <= kReplayMax) keys in LDS, ascending
__device__ void bitonic_sort(unsigned long long *k, int n) {
    for (int size = 2; size <= n; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < n; i += blockDim.x) {
                const int j = i ^ stride;
                if (j > i) {
                    const bool up = (i & size) == 0;
                    const unsigned long long a = k[i], b = k[j];
                    if ((a > b) == up) {
                        k[i] = b;
                        k[j] = a;
                    }
                }
            }
            __syncthreads();
        }
}

// The alternating scan of getCavitiesWithBilevelStatusMap (Surface.cpp:
// 1994-2126) over the components: even passes take the not yet numbered
// component met first scanning k upwards (smallest root = smallest flat
// index), odd passes the one met first scanning k downwards (smallest
// backward key).  Ids count up from 3 (the first component is the outside);
// then fillCavities' sizes and filled flags (Surface.cpp:2355-2487).  One
// workgroup; more than kReplayMax components leave info[0] = 0 for the host.
__global__ void __launch_bounds__(1024) k_ccl_replay(CavityReplay r) {
    __shared__ unsigned long long keys[kReplayMax];
    __shared__ unsigned short orderF[kReplayMax], orderB[kReplayMax];
    __shared__ unsigned char done[kReplayMax];
    const int n = *r.nroots;
    if (n > kReplayMax) {
        if (threadIdx.x == 0) {
            r.info[0] = 0;
            r.info[1] = 0;
        }
        return;
    }
    int np = 1;
    while (np < n) np <<= 1;
    for (int pass = 0; pass < 2; ++pass) {
        for (int i = threadIdx.x; i < np; i += blockDim.x) {
            unsigned long long key = ~0ull;
            if (i < n) key = ((unsigned long long)(pass == 0 ? (unsigned int)r.rootlist[i] : r.bkey[i]) << 32) | i;
            keys[i] = key;
        }
        __syncthreads();
        bitonic_sort(keys, np);
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const unsigned short slot = (unsigned short)(keys[i] & 0xffffu);
            if (pass == 0)
                orderF[i] = slot;
            else
                orderB[i] = slot;
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) done[i] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        int pf = 0, pb = 0, id = NSGPU_STATUS_TEMP_OUT;
        for (int t = 0; t < n; ++t) {
            int pick;
            if ((t & 1) == 0) {
                while (done[orderF[pf]]) ++pf;
                pick = orderF[pf];
            } else {
                while (done[orderB[pb]]) ++pb;
                pick = orderB[pb];
            }
            done[pick] = 1;
            r.label[pick] = ++id;
        }
        r.info[1] = id - NSGPU_STATUS_OUT > 0 ? id - NSGPU_STATUS_OUT : 0;
    }
    __syncthreads();
    for (int s = threadIdx.x; s < n; s += blockDim.x) {
        const int lab = r.label[s];
        uint8_t f = 0;
        if (lab >= NSGPU_STATUS_FIRST_CAV) {
            const double v = r.count[s] * r.cube;
            f = v <= r.vol;
            r.cav_size[lab - NSGPU_STATUS_FIRST_CAV] = v;
            r.cav_filled[lab - NSGPU_STATUS_FIRST_CAV] = f;
        }
        r.filled[s] = f;
    }
    if (threadIdx.x == 0) r.info[0] = 1;
}

void launch_ccl_replay(hipStream_t s, const CavityReplay &r) {
    hipLaunchKernelGGL(k_ccl_replay, dim3(1), dim3(1024), 0, s, r);
}

}  // namespace nsg
