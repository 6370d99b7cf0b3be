// Pocket detection, the grid part of pocket mode (main.cpp:1411-1660), on
// bit planes of the grid: one bit per node, rows of W = ceil(nx / 64) words,
// word (row, w) holding nodes x = 64 w .. 64 w + 63 of row k * ny + j.
//
//   k_pk_diff     the difference map (Surface::difference,
//                 Surface.cpp:11471-11718): not out in the fat-probe status,
//                 out in the small-probe one
//   k_pk_morph    the reference's shape filter (filterCavities,
//                 Surface.cpp:2751-3010) is a morphological opening by the
//                 probe's digital ball: a support point is a node whose ball
//                 holds only set nodes (erosion; nodes beyond the grid are
//                 not tested), a kept node one within the ball of a support
//                 point (dilation).  The ball is a table of (dy, dz, half
//                 width in x); a word of the result combines the words of
//                 the table's rows, each widened or narrowed along x by word
//                 shifts
//   k_pk_uf_*     6-connected components of the kept nodes by union-find over
//                 node indices (each root the component's smallest flat
//                 index), the component keys of the reference's alternating
//                 scan, counts; the host numbers the components
//   k_pk_label    the final status map and per-component tallies
#include "launch.hpp"

namespace nsg {

namespace {

__device__ __forceinline__ int pk_ld(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void pk_st(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ int pk_find(int *P, int x) {
    int p = pk_ld(P + x);
    while (p != x) {
        const int gp = pk_ld(P + p);
        if (gp != p) pk_st(P + x, gp);
        x = p;
        p = gp;
    }
    return x;
}

__device__ void pk_unite(int *P, int a, int b) {
    for (;;) {
        a = pk_find(P, a);
        b = pk_find(P, b);
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

__device__ __forceinline__ bool bit_at(const unsigned long long *m, int W, int64_t row, int x) {
    return (m[row * W + (x >> 6)] >> (x & 63)) & 1ull;
}

// one wave per word: lane = node
__global__ void __launch_bounds__(256) k_pk_diff(DevGrid g, PocketArgs A) {
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave >= A.nwords) return;
    const int64_t row = wave / A.W;
    const int x = (int)(wave - row * A.W) * 64 + lane;
    bool d = false;
    if (x < g.nx) {
        const int64_t n = row * g.nx + x;
        const int s1 = A.st1[n], s2 = A.st2[n];
        const bool out2 = s2 == NSGPU_STATUS_TEMP_OUT || s2 == NSGPU_STATUS_OUT ||
                          (A.cav_and_pockets && s2 >= NSGPU_STATUS_FIRST_CAV);
        d = s1 != NSGPU_STATUS_TEMP_OUT && s1 != NSGPU_STATUS_OUT && out2;
    }
    const unsigned long long m = __ballot(d);
    if (lane == 0) A.D[wave] = m;
}

// erosion (dilate = false) or dilation of `in` by the ball; one thread per
// word.  A row's words outside the grid read as all ones for the erosion and
// zeros for the dilation, as do the bits past the row's last node.
__global__ void __launch_bounds__(256) k_pk_morph(DevGrid g, PocketArgs A, const unsigned long long *in,
                                                  const unsigned long long *and_mask, unsigned long long *out,
                                                  bool dilate) {
    const int64_t wq = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (wq >= A.nwords) return;
    const int64_t row = wq / A.W;
    const int w = (int)(wq - row * A.W);
    const int j = (int)(row % g.ny), k = (int)(row / g.ny);
    const unsigned long long fill = dilate ? 0ull : ~0ull;
    // bits of the last word past nx
    const int tail = g.nx - (A.W - 1) * 64;
    const unsigned long long past = tail >= 64 ? 0ull : (~0ull << tail);
    unsigned long long acc = dilate ? 0ull : ~0ull;
    for (int e = 0; e < A.nball; ++e) {
        const int dj = A.ball[3 * e], dk = A.ball[3 * e + 1], hw = A.ball[3 * e + 2];
        const int jj = j + dj, kk = k + dk;
        if (jj < 0 || jj >= g.ny || kk < 0 || kk >= g.nz) continue;  // erosion: ones; dilation: zeros
        const int64_t r2 = (int64_t)kk * g.ny + jj;
        auto word = [&](int ww) -> unsigned long long {
            if (ww < 0 || ww >= A.W) return fill;
            unsigned long long v = in[r2 * A.W + ww];
            if (ww == A.W - 1) v = dilate ? (v & ~past) : (v | past);
            return v;
        };
        const unsigned long long cur = word(w), lo = word(w - 1), hi = word(w + 1);
        unsigned long long r = cur;
        for (int d = 1; d <= hw; ++d) {
            const unsigned long long up = (cur >> d) | (hi << (64 - d));
            const unsigned long long dn = (cur << d) | (lo >> (64 - d));
            r = dilate ? (r | up | dn) : (r & up & dn);
        }
        acc = dilate ? (acc | r) : (acc & r);
    }
    if (w == A.W - 1) acc &= ~past;
    if (and_mask) acc &= and_mask[wq];
    out[wq] = acc;
}

// union-find over the set nodes of `m`, one wave per word.  A node's parent
// starts as the first node of its x-run inside the word (no atomics); runs
// continuing from the previous word are joined at bit 0, and each maximal
// stretch where a row and the row below (or the plane below) are both set is
// joined once, at its first node.
__global__ void __launch_bounds__(256) k_pk_uf_init(DevGrid g, PocketArgs A, const unsigned long long *m) {
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave >= A.nwords) return;
    const int64_t row = wave / A.W;
    const int x0 = (int)(wave - row * A.W) * 64, x = x0 + lane;
    if (x >= g.nx) return;
    const unsigned long long w = m[wave];
    const int64_t n = row * g.nx + x;
    if (!((w >> lane) & 1ull)) {
        A.parent[n] = -1;
        return;
    }
    // the highest clear bit below this one, + 1
    const unsigned long long below_clear = ~w & ((lane ? (~0ull >> (64 - lane)) : 0ull));
    const int start = below_clear ? 64 - __clzll((long long)below_clear) : 0;
    A.parent[n] = (int)(row * g.nx + x0 + start);
}

__global__ void __launch_bounds__(256) k_pk_uf_union(DevGrid g, PocketArgs A, const unsigned long long *m) {
    const int64_t wq = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (wq >= A.nwords) return;
    const int64_t row = wq / A.W;
    const int wi = (int)(wq - row * A.W);
    const int j = (int)(row % g.ny), k = (int)(row / g.ny);
    const unsigned long long w = m[wq];
    if (!w) return;
    const int64_t base = row * g.nx + wi * 64;
    // x: a run that continues from the previous word
    if (wi > 0 && (w & 1ull) && (m[wq - 1] >> 63)) pk_unite(A.parent, (int)base, (int)base - 1);
    // y and z: the first node of each stretch set in both rows
    for (int d = 0; d < 2; ++d) {
        if (d == 0 ? j == 0 : k == 0) continue;
        const int64_t oq = d == 0 ? wq - A.W : wq - (int64_t)g.ny * A.W;
        const int64_t step = d == 0 ? g.nx : (int64_t)g.nx * g.ny;
        const unsigned long long both = w & m[oq];
        const unsigned long long prev = wi > 0 ? ((m[wq - 1] & m[oq - 1]) >> 63) : 0ull;
        unsigned long long starts = both & ~((both << 1) | prev);
        while (starts) {
            const int b = __ffsll((long long)starts) - 1;
            starts &= starts - 1;
            pk_unite(A.parent, (int)(base + b), (int)(base + b - step));
        }
    }
}

// flatten; roots as a bit plane (their rank = the component's index)
__global__ void __launch_bounds__(256) k_pk_uf_roots(DevGrid g, PocketArgs A) {
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave >= A.nwords) return;
    const int64_t row = wave / A.W;
    const int x = (int)(wave - row * A.W) * 64 + lane;
    bool root = false;
    if (x < g.nx) {
        const int64_t n = row * g.nx + x;
        if (pk_ld(A.parent + n) >= 0) {
            // a find that writes only this node's own entry: another thread
            // reading it sees its old parent or its root, both ancestors
            int r = (int)n, p = pk_ld(A.parent + n);
            while (p != r) {
                r = p;
                p = pk_ld(A.parent + p);
            }
            pk_st(A.parent + n, r);
            root = r == (int)n;
        }
    }
    const unsigned long long m = __ballot(root);
    if (lane == 0) {
        A.roots[wave] = m;
        A.rcount[wave] = __popcll(m);
    }
}

// the component index of root flat index r: its rank among the roots
__device__ __forceinline__ int root_rank(const DevGrid &g, const PocketArgs &A, int r) {
    const int64_t row = r / g.nx;
    const int x = (int)(r - row * g.nx);
    const int64_t wq = row * A.W + (x >> 6);
    const int b = x & 63;
    return A.roff[wq] + __popcll(b ? (A.roots[wq] & (~0ull >> (64 - b))) : 0ull);
}

// Lanes of one wave holding nodes of one component combine their updates:
// the first pending lane leads the lanes with its component (whose mask is
// returned) and the pending set shrinks by them.  A wave's nodes lie in one
// grid row, usually in one or two components.
__device__ __forceinline__ unsigned long long next_group(unsigned long long &pending, int c, int &lead, int &cl) {
    lead = __ffsll((long long)pending) - 1;
    cl = __shfl(c, lead, 64);
    const unsigned long long m = __ballot(c == cl) & pending;
    pending &= ~m;
    return m;
}

// Per-component tallies of a wave's run of kWordsPerWave consecutive words,
// held by lane 0 while the groups keep naming the same component (they mostly
// do: consecutive words are consecutive nodes of a row) and added to the
// component's entries when it changes.
constexpr int kWordsPerWave = 16;
struct Tally {
    int c = -1, n = 0;
    long long key = 0x7fffffffffffffffll;
    int flag = 0;
    __device__ void flush(int *count, long long *keymin, int *flagv) {
        if (c < 0) return;
        if (count && n) atomicAdd(count + c, n);
        if (keymin) atomicMin(keymin + c, key);
        if (flagv && flag) flagv[c] = 1;
    }
    __device__ void add(int cl, int cnt, long long k, int f, int *count, long long *keymin, int *flagv) {
        if (cl != c) {
            flush(count, keymin, flagv);
            c = cl;
            n = 0;
            key = 0x7fffffffffffffffll;
            flag = 0;
        }
        n += cnt;
        key = min(key, k);
        flag |= f;
    }
};

// per component: node count and the backward-scan key (k descending, then j
// and i ascending: ((nz - 1 - k) ny + j) nx + i; along a row it grows with i,
// so a group's smallest is its leader's); the forward key is the root
__global__ void __launch_bounds__(256) k_pk_stats(DevGrid g, PocketArgs A) {
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t q0 = wave * kWordsPerWave;
    if (q0 >= A.nwords) return;
    const int64_t plane = (int64_t)g.nx * g.ny;
    Tally t;
    for (int64_t wq = q0; wq < min(q0 + kWordsPerWave, A.nwords); ++wq) {
        const int64_t row = wq / A.W;
        const int x = (int)(wq - row * A.W) * 64 + lane;
        const int64_t n = row * g.nx + x;
        int c = -1;
        if (x < g.nx) {
            const int r = A.parent[n];
            if (r >= 0) {
                c = root_rank(g, A, r);
                if (r == (int)n) A.croot[c] = r;
            }
        }
        unsigned long long pending = __ballot(c >= 0);
        const int64_t k = row / g.ny;
        const long long bkey = (long long)(g.nz - 1 - k) * plane + (n - k * plane);
        while (pending) {
            int lead, cl;
            const unsigned long long m = next_group(pending, c, lead, cl);
            const long long lk = __shfl(bkey, lead, 64);
            if (lane == 0) t.add(cl, __popcll(m), lk, 0, A.ccount, A.cbkey, nullptr);
        }
    }
    if (lane == 0) t.flush(A.ccount, A.cbkey, nullptr);
}

// the kept components' nodes (components not filled by the volume threshold)
__global__ void __launch_bounds__(256) k_pk_union_kept(DevGrid g, PocketArgs A) {
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave >= A.nwords) return;
    const int64_t row = wave / A.W;
    const int x = (int)(wave - row * A.W) * 64 + lane;
    bool u = false;
    if (x < g.nx) {
        const int r = A.parent[row * g.nx + x];
        if (r >= 0) u = !A.cfilled[root_rank(g, A, r)];
    }
    const unsigned long long m = __ballot(u);
    if (lane == 0) A.U[wave] = m;
}

// final status map (support -id, kept id, a filled component id, else
// inside), per component the kept nodes and whether a node lies in the small
// surface's outside; support points as a bit plane for the ordered lists
__global__ void __launch_bounds__(256) k_pk_label(DevGrid g, PocketArgs A) {
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t q0 = wave * kWordsPerWave;
    if (q0 >= A.nwords) return;
    Tally t;
    for (int64_t wq = q0; wq < min(q0 + kWordsPerWave, A.nwords); ++wq) {
        const int64_t row = wq / A.W;
        const int x = (int)(wq - row * A.W) * 64 + lane;
        bool sup = false, kept = false, out2 = false;
        int c = -1;
        if (x < g.nx) {
            const int64_t n = row * g.nx + x;
            const int r = A.parent[n];
            int st = NSGPU_STATUS_INSIDE;
            if (r >= 0) {
                c = root_rank(g, A, r);
                const int id = A.cid[c];
                const unsigned long long bit = 1ull << (x & 63);
                if (A.cfilled[c]) {
                    st = id;
                } else if (A.S2[wq] & bit) {
                    st = -id;
                    sup = kept = true;
                } else if (A.K2[wq] & bit) {
                    st = id;
                    kept = true;
                }
                out2 = A.st2[n] == NSGPU_STATUS_TEMP_OUT || A.st2[n] == NSGPU_STATUS_OUT;
            }
            A.st1[n] = st;
        }
        unsigned long long pending = __ballot(c >= 0);
        const unsigned long long km = __ballot(kept), om = __ballot(out2);
        while (pending) {
            int lead, cl;
            const unsigned long long m = next_group(pending, c, lead, cl);
            if (lane == 0) t.add(cl, __popcll(m & km), 0, (m & om) != 0, A.ckept, nullptr, A.cpocket);
        }
        const unsigned long long sm = __ballot(sup);
        if (lane == 0) A.scount[wq] = __popcll(sm);
    }
    if (lane == 0) t.flush(A.ckept, nullptr, A.cpocket);
}

// support nodes in flat order with their component index
__global__ void __launch_bounds__(256) k_pk_support_list(DevGrid g, PocketArgs A) {
    const int64_t wq = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (wq >= A.nwords) return;
    unsigned long long m = A.S2[wq];
    if (!m) return;
    const int64_t row = wq / A.W;
    const int x0 = (int)(wq - row * A.W) * 64;
    int at = A.soff[wq];
    while (m) {
        const int b = __ffsll((long long)m) - 1;
        m &= m - 1;
        const int64_t n = row * g.nx + x0 + b;
        A.sup_node[at] = (long long)n;
        A.sup_comp[at] = root_rank(g, A, A.parent[n]);
        ++at;
    }
}

// the components' ids before the second filter (the status map that
// saveCavities and getCavitiesAtoms read, main.cpp:1590-1610), -1 elsewhere
__global__ void __launch_bounds__(256) k_pk_comp_labels(DevGrid g, PocketArgs A, int *out) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (int64_t)g.nx * g.ny * g.nz) return;
    const int r = A.parent[n];
    out[n] = r >= 0 ? A.cid[root_rank(g, A, r)] : NSGPU_STATUS_INSIDE;
}

}  // namespace

void launch_pk_comp_labels(hipStream_t s, const DevGrid &g, const PocketArgs &A, int *out) {
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_pk_comp_labels, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, s, g, A, out);
}

static unsigned blocks_of(int64_t threads) { return (unsigned)((threads + 255) / 256); }

void launch_pk_diff(hipStream_t s, const DevGrid &g, const PocketArgs &A) {
    hipLaunchKernelGGL(k_pk_diff, dim3(blocks_of(A.nwords * 64)), dim3(256), 0, s, g, A);
}

void launch_pk_morph(hipStream_t s, const DevGrid &g, const PocketArgs &A, const unsigned long long *in,
                     const unsigned long long *and_mask, unsigned long long *out, bool dilate) {
    hipLaunchKernelGGL(k_pk_morph, dim3(blocks_of(A.nwords)), dim3(256), 0, s, g, A, in, and_mask, out, dilate);
}

void launch_pk_components(hipStream_t s, const DevGrid &g, const PocketArgs &A, const unsigned long long *m) {
    hipLaunchKernelGGL(k_pk_uf_init, dim3(blocks_of(A.nwords * 64)), dim3(256), 0, s, g, A, m);
    hipLaunchKernelGGL(k_pk_uf_union, dim3(blocks_of(A.nwords)), dim3(256), 0, s, g, A, m);
    hipLaunchKernelGGL(k_pk_uf_roots, dim3(blocks_of(A.nwords * 64)), dim3(256), 0, s, g, A);
}

void launch_pk_stats(hipStream_t s, const DevGrid &g, const PocketArgs &A) {
    const int64_t waves = (A.nwords + kWordsPerWave - 1) / kWordsPerWave;
    hipLaunchKernelGGL(k_pk_stats, dim3(blocks_of(waves * 64)), dim3(256), 0, s, g, A);
}

void launch_pk_union_kept(hipStream_t s, const DevGrid &g, const PocketArgs &A) {
    hipLaunchKernelGGL(k_pk_union_kept, dim3(blocks_of(A.nwords * 64)), dim3(256), 0, s, g, A);
}

void launch_pk_label(hipStream_t s, const DevGrid &g, const PocketArgs &A) {
    const int64_t waves = (A.nwords + kWordsPerWave - 1) / kWordsPerWave;
    hipLaunchKernelGGL(k_pk_label, dim3(blocks_of(waves * 64)), dim3(256), 0, s, g, A);
}

void launch_pk_support_list(hipStream_t s, const DevGrid &g, const PocketArgs &A) {
    hipLaunchKernelGGL(k_pk_support_list, dim3(blocks_of(A.nwords)), dim3(256), 0, s, g, A);
}

}  // namespace nsg
