// Analytic vertices and marching cubes.
//
// Vertex assembly (assembleVerticesList, Surface.cpp:7415-7660): every edge
// ray segment records a vertex on the edge it enters and the edge it leaves;
// a record is dropped when the two corners of its edge agree (after cavity
// filling), and the first record of an edge wins.  Records of one edge can
// only come from one ray, so "first" is the smaller depth on that ray.  The
// kernels count per ray, prefix-sum, and write in ray order (panel, row,
// column), which is the order a single-threaded reference run produces.
//
// Marching cubes (triangulateSurface, Surface.cpp:9909-10300): cubes
// i,j,k in [1, N-2]; a corner is inside when its bit is 0; edges 5, 6 and 10
// of an active cube get a mid-point vertex when no analytic one exists
// (getVertices :8002-8327); triangles come from the case table with missing
// vertices skipped and are written reversed (triangulationKernel :9075-9318).
// Vertex and triangle order is cube order, as a single-threaded run.
#include "launch.hpp"
#include "../common/mc_cases.h"

namespace nsg {

// No node-indexed table of vertex ids: an edge's analytic vertex is found on
// its edge ray (the ray's scanned vertex offset plus the vertex's rank among
// the ray's vertices, which are in node order, from a bitmap of the ray's
// nodes and per-word prefix counts: three independent loads), and a mid point
// through the active index of the cube that adds it (segment mask and scan).
// A table of three ids per node (12 bytes a node) was written and read at
// random in HBM; the bitmaps take 3 N^2 (N + 63) / 64 words of 12 bytes.

__device__ __forceinline__ int bit_at(const uint32_t *bits, int64_t f, int64_t nvox) {
    if (f < 0 || f >= nvox) return 1;
    return (bits[f >> 5] >> (uint32_t)(f & 31)) & 1u;
}

struct EdgeSlots {
    int s[3];  // edge-ray slots of panels 0 (x), 1 (z), 2 (y)
};

// Walk the vertex records of one edge ray.  emit(q, dir, node, px, py, pz,
// depth, ia) is called for every kept record in order (ia: the node's index
// along the ray, increasing); returns the count.  One instance
// per panel (0: x rays, 1: z rays, 2: y rays), so the grid arrays and extents
// it reads are fixed at compile time.
template <int PANEL, class Emit>
__device__ int walk_edge_ray_p(const DevGrid &g, const DevRays &R, const uint32_t *bits, int slot, int64_t ray,
                               Emit emit) {
    const int c = R.cnt[(int64_t)slot * R.N * R.N + ray];
    if (c == 0 || (c & 1) || c > R.K) return 0;
    const double *T = R.t + ((int64_t)slot * R.N * R.N + ray) * R.K;
    const int n = (int)(ray / R.N), m = (int)(ray % R.N);
    const double delta = kDelta - g.hside;
    const double *ga = PANEL == 0 ? g.x : (PANEL == 1 ? g.z : g.y);
    const int na = PANEL == 0 ? g.nx : (PANEL == 1 ? g.nz : g.ny);
    const double gmin = PANEL == 0 ? g.xmin : (PANEL == 1 ? g.zmin : g.ymin);
    const double pa = ga[0] + delta;
    const double dir = (ga[na - 1] + delta) - pa;
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    // fixed coordinates of the ray and the flat-index stride along it
    double p0, p1;
    int64_t stride;
    if (PANEL == 0) {
        if (m >= g.ny || n >= g.nz) return 0;
        p0 = g.y[m] + delta;
        p1 = g.z[n] + delta;
        stride = 1;
    } else if (PANEL == 1) {
        if (m >= g.nx || n >= g.ny) return 0;
        p0 = g.x[m] + delta;
        p1 = g.y[n] + delta;
        stride = (int64_t)g.nx * g.ny;
    } else {
        if (m >= g.nx || n >= g.nz) return 0;
        p0 = g.x[m] + delta;
        p1 = g.z[n] + delta;
        stride = g.nx;
    }
    // The depths are sorted and dir > 0, so the segment ends visited below come
    // in non-decreasing node order: a node already given a vertex can only be
    // the last one given one.
    int last = -1, count = 0;
    for (int q = 0; q < c; q += 2) {
        const double t0 = T[q], t1 = T[q + 1];
        if (!(fabs(t1 - t0) >= kEpsInt)) continue;
        const double lim0 = pa + dir * t0, lim1 = pa + dir * t1;
        const int a = (int)rintp((lim0 - gmin) * g.scale);
        const int b = (int)rintp((lim1 - gmin) * g.scale);
        if (b <= a) continue;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int ia = h == 0 ? a : b;
            const double lim = h == 0 ? lim0 : lim1;
            if (ia < 0 || ia >= na) continue;
            int ix, iy, iz, d;
            if (PANEL == 0) { ix = ia; iy = m; iz = n; d = 0; }
            else if (PANEL == 1) { ix = m; iy = n; iz = ia; d = 2; }
            else { ix = m; iy = ia; iz = n; d = 1; }
            const int64_t f = ((int64_t)iz * g.ny + iy) * g.nx + ix;
            if (bit_at(bits, f, nvox) == bit_at(bits, f + stride, nvox)) continue;
            if (ia == last) continue;
            last = ia;
            double px, py, pz;
            if (PANEL == 0) { px = lim; py = p0; pz = p1; }
            else if (PANEL == 1) { px = p0; py = p1; pz = lim; }
            else { px = p0; py = lim; pz = p1; }
            emit(count, d, f, px, py, pz, q + h, ia);
            ++count;
        }
    }
    return count;
}

template <class Emit>
__device__ __forceinline__ int walk_edge_ray(const DevGrid &g, const DevRays &R, const uint32_t *bits, int panel,
                                             int slot, int64_t ray, Emit emit) {
    if (panel == 0) return walk_edge_ray_p<0>(g, R, bits, slot, ray, emit);
    if (panel == 1) return walk_edge_ray_p<1>(g, R, bits, slot, ray, emit);
    return walk_edge_ray_p<2>(g, R, bits, slot, ray, emit);
}

__global__ void __launch_bounds__(256) k_vtx_count(DevGrid g, DevRays R, EdgeSlots es, const uint32_t *bits,
                                                   int *counts) {
    const int64_t NN = (int64_t)R.N * R.N;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= 3 * NN) return;
    const int panel = (int)(r / NN);
    counts[r] = walk_edge_ray(g, R, bits, panel, es.s[panel], r - panel * NN,
                              [](int, int, int64_t, double, double, double, int, int) {});
}

// The vertices of a wave's rays are one contiguous range from lane 0's
// offset (the offsets are the scanned counts).  Its first kVtxStage vertices
// are staged in LDS and stored by consecutive lanes after the walk; any beyond
// go straight out.  Stored per lane at its own offset, every store of a
// vertex's 24 bytes, slot and depth id left partial lines.
constexpr int kVtxStage = 192;

__global__ void __launch_bounds__(256) k_vtx_write(DevGrid g, DevRays R, EdgeSlots es, const uint32_t *bits,
                                                   const int *offs, double *verts, uint64_t *vbits, int *vpre,
                                                   int64_t *vslot, int64_t *vhit, int64_t vcap) {
    __shared__ double sv[4][3 * kVtxStage];
    __shared__ int64_t ss[4][kVtxStage];
    __shared__ int64_t sh[4][kVtxStage];
    const int64_t NN = (int64_t)R.N * R.N;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const bool valid = r < 3 * NN;
    const unsigned long long vm = __ballot(valid);
    if (vm == 0) return;
    const int base = valid ? offs[r] : 0;
    const int wbase = __shfl(base, 0, 64);
    int cnt = 0;
    // the ray's node bitmap (word-major, so that a wave's stores of one word
    // are contiguous): up to kRegWords words are kept in registers and stored
    // by the whole wave after the walk; longer rays store each word as the
    // walk leaves it
    constexpr int kRegWords = 8;
    const int W = (R.N + 63) >> 6;
    const bool in_regs = W <= kRegWords;
    const int64_t NR = 3 * NN;
    uint64_t wb[kRegWords];
#pragma unroll
    for (int u = 0; u < kRegWords; ++u) wb[u] = 0;
    int cw = 0, cpre = 0;
    uint64_t cbits = 0;
    auto flush_to = [&](int w) {
        for (; cw < w; ++cw) {
            vbits[(int64_t)cw * NR + r] = cbits;
            vpre[(int64_t)cw * NR + r] = cpre;
            cpre += __popcll(cbits);
            cbits = 0;
        }
    };
    if (valid) {
        const int panel = (int)(r / NN);
        cnt = walk_edge_ray(g, R, bits, panel, es.s[panel], r - panel * NN,
                            [&](int q, int d, int64_t f, double px, double py, double pz, int hit, int ia) {
                                const int v = base + q;
                                if (in_regs) {
#pragma unroll
                                    for (int u = 0; u < kRegWords; ++u)
                                        if (u == (ia >> 6)) wb[u] |= 1ull << (ia & 63);
                                } else {
                                    flush_to(ia >> 6);
                                    cbits |= 1ull << (ia & 63);
                                }
                                if (v >= vcap) return;
                                // the depth this vertex comes from, for its analytic normal
                                const int64_t h = ((int64_t)es.s[panel] * NN + (r - panel * NN)) * R.K + hit;
                                const int so = v - wbase;
                                if (so < kVtxStage) {
                                    sv[wv][3 * so + 0] = px;
                                    sv[wv][3 * so + 1] = py;
                                    sv[wv][3 * so + 2] = pz;
                                    ss[wv][so] = 3 * f + d;
                                    sh[wv][so] = h;
                                } else {
                                    verts[3 * (int64_t)v + 0] = px;
                                    verts[3 * (int64_t)v + 1] = py;
                                    verts[3 * (int64_t)v + 2] = pz;
                                    vslot[v] = 3 * f + d;
                                    vhit[v] = h;
                                }
                            });
        if (!in_regs) flush_to(W);  // the ray's last words
    }
    if (in_regs && valid) {
        int pre = 0;
#pragma unroll
        for (int u = 0; u < kRegWords; ++u) {
            if (u >= W) break;
            vbits[(int64_t)u * NR + r] = wb[u];
            vpre[(int64_t)u * NR + r] = pre;
            pre += __popcll(wb[u]);
        }
    }
    // staged: v in [wbase, wbase + n), below vcap and the stage's size
    const int64_t wend = (int64_t)__shfl(base + cnt, 63 - __builtin_clzll(vm), 64);
    const int n = (int)max((int64_t)0, min(min(wend, vcap) - wbase, (int64_t)kVtxStage));
    __builtin_amdgcn_wave_barrier();
    for (int x = lane; x < 3 * n; x += 64) verts[3 * (int64_t)wbase + x] = sv[wv][x];
    for (int x = lane; x < n; x += 64) {
        vslot[wbase + x] = ss[wv][x];
        vhit[wbase + x] = sh[wv][x];
    }
}

// ---------------------------------------------------------------- marching cubes
// Cubes are classified 32 at a time: a lane owns a row segment (32 cubes
// along x at fixed j,k), reads a 64-bit window of corner bits from each of the
// four corner rows (j,k), (j+1,k), (j,k+1), (j+1,k+1), and finds the active
// cubes (neither all inside nor all outside) with bitwise operations.  Active
// cubes are compacted in (k, j, i) order; the vertex and triangle passes then
// visit only them.

__constant__ uint64_t c_tri_cases[256];

__device__ __forceinline__ bool edge_on(int c, int e) {
    const int a = (e < 8) ? (e & 3) + (e & 4) : e - 8;
    int b;
    if (e < 8) b = ((e & 3) == 3 ? (e & 4) : (e & 3) + 1 + (e & 4));
    else b = e - 4;
    return ((c >> a) & 1) != ((c >> b) & 1);
}

// edge e of the cube at node f -> the edge's key 3 * node + direction
// (getTriangles' edge numbering), the vslot of a vertex on it
__device__ __forceinline__ int64_t edge_slot(int e, int64_t f, int64_t sy, int64_t sz, int64_t) {
    switch (e) {
        case 0: return 3 * f + 0;
        case 1: return 3 * (f + 1) + 1;
        case 2: return 3 * (f + sy) + 0;
        case 3: return 3 * f + 1;
        case 4: return 3 * (f + sz) + 0;
        case 5: return 3 * (f + 1 + sz) + 1;
        case 6: return 3 * (f + sy + sz) + 0;
        case 7: return 3 * (f + sz) + 1;
        case 8: return 3 * f + 2;
        case 9: return 3 * (f + 1) + 2;
        case 10: return 3 * (f + 1 + sy) + 2;
        default: return 3 * (f + sy) + 2;
    }
}

// Whether the cube that owns edge e (the cube where it is edge 5, 6 or 10)
// lies in the cube range, i.e. whether getVertices adds a mid point for it
// when no analytic vertex exists.  Owners sit at -1 in some of i, j, k.
__device__ __forceinline__ bool owner_in_range(int e, int i, int j, int k) {
    switch (e) {
        case 5: case 6: case 10: return true;
        case 7: return i > 1;             // Y(i, j, k+1)   = edge 5 of (i-1, j, k)
        case 1: return k > 1;             // Y(i+1, j, k)   = edge 5 of (i, j, k-1)
        case 3: return i > 1 && k > 1;    // Y(i, j, k)     = edge 5 of (i-1, j, k-1)
        case 4: return j > 1;             // X(i, j, k+1)   = edge 6 of (i, j-1, k)
        case 2: return k > 1;             // X(i, j+1, k)   = edge 6 of (i, j, k-1)
        case 0: return j > 1 && k > 1;    // X(i, j, k)     = edge 6 of (i, j-1, k-1)
        case 9: return j > 1;             // Z(i+1, j, k)   = edge 10 of (i, j-1, k)
        case 11: return i > 1;            // Z(i, j+1, k)   = edge 10 of (i-1, j, k)
        default: return i > 1 && j > 1;  // 8: Z(i, j, k) = edge 10 of (i-1, j-1, k)
    }
}

// 64 corner bits starting at flat node f (bits past the grid read as outside)
__device__ __forceinline__ uint64_t bit_window(const uint32_t *bits, int64_t f, int64_t nwords) {
    const int64_t w = f >> 5;
    const uint32_t s = (uint32_t)(f & 31);
    const uint64_t lo = w < nwords ? bits[w] : 0xffffffffu;
    const uint64_t hi = w + 1 < nwords ? bits[w + 1] : 0xffffffffu;
    return ((hi << 32) | lo) >> s | (s ? (~0ull << (64 - s)) : 0ull);
}

struct SegIter {
    int nseg_row;  // segments per row
    int cx, cy, cz;
    int64_t nseg;
};

__device__ __forceinline__ void seg_decode(const SegIter &si, int64_t s, int &i0, int &j, int &k, int &len) {
    const int sr = (int)(s % si.nseg_row);
    const int64_t row = s / si.nseg_row;
    j = 1 + (int)(row % si.cy);
    k = 1 + (int)(row / si.cy);
    i0 = 1 + 32 * sr;
    len = min(32, si.cx - 32 * sr);
}

// active-cube mask of a segment; rows A=(j,k) B=(j+1,k) C=(j,k+1) D=(j+1,k+1)
__device__ __forceinline__ uint32_t seg_active(const DevGrid &g, const uint32_t *bits, int i0, int j, int k, int len,
                                               uint64_t rows[4]) {
    const int64_t nwords = ((int64_t)g.nx * g.ny * g.nz + 31) >> 5;
    const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
    const int64_t f = ((int64_t)k * g.ny + j) * g.nx + i0;
    rows[0] = bit_window(bits, f, nwords);
    rows[1] = bit_window(bits, f + sy, nwords);
    rows[2] = bit_window(bits, f + sz, nwords);
    rows[3] = bit_window(bits, f + sy + sz, nwords);
    uint64_t out = ~0ull, in = ~0ull;  // bit b: all eight corners of cube i0+b outside / inside
    for (int r = 0; r < 4; ++r) {
        out &= rows[r] & (rows[r] >> 1);
        in &= ~rows[r] & (~rows[r] >> 1);
    }
    const uint32_t valid = len >= 32 ? 0xffffffffu : ((1u << len) - 1u);
    return (uint32_t)(~(out | in)) & valid;
}

__device__ __forceinline__ int seg_case(const uint64_t rows[4], int b) {
    // corner q inside (bit 0) -> case bit q; corners 0..7 as in classifyCube
    const uint64_t A = rows[0] >> b, B = rows[1] >> b, C = rows[2] >> b, D = rows[3] >> b;
    return (int)(((~A) & 1) | (((~A >> 1) & 1) << 1) | (((~B >> 1) & 1) << 2) | (((~B) & 1) << 3) |
                 (((~C) & 1) << 4) | (((~C >> 1) & 1) << 5) | (((~D >> 1) & 1) << 6) | (((~D) & 1) << 7));
}

__global__ void __launch_bounds__(256) k_mc_classify(DevGrid g, SegIter si, const uint32_t *bits, uint32_t *segmask,
                                                     int *segcnt) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= si.nseg) return;
    int i0, j, k, len;
    seg_decode(si, s, i0, j, k, len);
    uint64_t rows[4];
    const uint32_t m = seg_active(g, bits, i0, j, k, len, rows);
    segmask[s] = m;
    segcnt[s] = __popc(m);
}

// active list in cube order: node index and case of each active cube
__global__ void __launch_bounds__(256) k_mc_compact(DevGrid g, SegIter si, const uint32_t *bits,
                                                    const uint32_t *segmask, const int *segoff, int *anode,
                                                    uint8_t *acase) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= si.nseg) return;
    uint32_t m = segmask[s];
    if (!m) return;
    int i0, j, k, len;
    seg_decode(si, s, i0, j, k, len);
    const int64_t nwords = ((int64_t)g.nx * g.ny * g.nz + 31) >> 5;
    const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
    const int64_t f = ((int64_t)k * g.ny + j) * g.nx + i0;
    const uint64_t rows[4] = {bit_window(bits, f, nwords), bit_window(bits, f + sy, nwords),
                              bit_window(bits, f + sz, nwords), bit_window(bits, f + sy + sz, nwords)};
    int o = segoff[s];
    while (m) {
        const int b = __ffs(m) - 1;
        m &= m - 1;
        anode[o] = (int)(f + b);
        acase[o] = (uint8_t)seg_case(rows, b);
        ++o;
    }
}

// The analytic vertex of edge e of cube (i, j, k), or -1: the edge's ray
// (panels 0, 1, 2 hold the x, z and y edge rays, ray n * N + m) and the
// node's index ia along it; its rank on the ray from the ray's bitmap.
__device__ __forceinline__ int analytic_vertex(const McVertexMap &vm, int e, int i, int j, int k) {
    if (!vm.vbits) return -1;
    int panel, m, n, ia;
    switch (e) {
        case 0: panel = 0; m = j; n = k; ia = i; break;
        case 1: panel = 2; m = i + 1; n = k; ia = j; break;
        case 2: panel = 0; m = j + 1; n = k; ia = i; break;
        case 3: panel = 2; m = i; n = k; ia = j; break;
        case 4: panel = 0; m = j; n = k + 1; ia = i; break;
        case 5: panel = 2; m = i + 1; n = k + 1; ia = j; break;
        case 6: panel = 0; m = j + 1; n = k + 1; ia = i; break;
        case 7: panel = 2; m = i; n = k + 1; ia = j; break;
        case 8: panel = 1; m = i; n = j; ia = k; break;
        case 9: panel = 1; m = i + 1; n = j; ia = k; break;
        case 10: panel = 1; m = i + 1; n = j + 1; ia = k; break;
        default: panel = 1; m = i; n = j + 1; ia = k; break;
    }
    const int64_t r = ((int64_t)panel * vm.N + n) * vm.N + m;
    const int64_t wi = (int64_t)(ia >> 6) * (3 * (int64_t)vm.N * vm.N) + r;
    const uint64_t word = vm.vbits[wi];
    const int pre = vm.vpre[wi], off = vm.offs[r];
    const uint64_t below = (1ull << (ia & 63)) - 1ull;
    return ((word >> (ia & 63)) & 1ull) ? off + pre + __popcll(word & below) : -1;
}

// node (i, j, k) of a flat index below 2^31 (the active list holds int32
// nodes): 32-bit divisions, a fraction of the cost of 64-bit ones
__device__ __forceinline__ void cube_ijk(const DevGrid &g, int f, int &i, int &j, int &k) {
    const int row = f / g.nx;
    i = f - row * g.nx;
    k = row / g.ny;
    j = row - k * g.ny;
}

// index of cube (i, j, k) in the active list, or -1 when it is not active
__device__ __forceinline__ int active_index(const DevGrid &g, const McVertexMap &vm, int i, int j, int k) {
    const int nseg_row = (g.nx - 2 + 31) / 32;
    const int64_t s = ((int64_t)(k - 1) * (g.ny - 2) + (j - 1)) * nseg_row + ((i - 1) >> 5);
    const uint32_t m = vm.segmask[s];
    const int b = (i - 1) & 31;
    if (!((m >> b) & 1u)) return -1;
    return vm.segoff[s] + __popc(m & ((1u << b) - 1u));
}

// The mid point on edge e of cube (i, j, k), or -1: it is added by the cube
// that owns the edge (where it is edge 5, 6 or 10, see owner_in_range), in
// the order 5, 6, 10 from that cube's offset.
__device__ __forceinline__ int midpoint_vertex(const DevGrid &g, const McVertexMap &vm, int e, int i, int j, int k) {
    int io = i, jo = j, ko = k, bit;
    switch (e) {
        case 5: bit = 0; break;
        case 6: bit = 1; break;
        case 10: bit = 2; break;
        case 7: io = i - 1; bit = 0; break;
        case 1: ko = k - 1; bit = 0; break;
        case 3: io = i - 1; ko = k - 1; bit = 0; break;
        case 4: jo = j - 1; bit = 1; break;
        case 2: ko = k - 1; bit = 1; break;
        case 0: jo = j - 1; ko = k - 1; bit = 1; break;
        case 9: jo = j - 1; bit = 2; break;
        case 11: io = i - 1; bit = 2; break;
        default: io = i - 1; jo = j - 1; bit = 2; break;
    }
    if (io < 1 || jo < 1 || ko < 1) return -1;
    const int q = active_index(g, vm, io, jo, ko);
    if (q < 0) return -1;
    const int mk = vm.amask[q];
    if (!((mk >> bit) & 1)) return -1;
    return vm.base + vm.avoff[q] + __popc((unsigned)mk & ((1u << bit) - 1u));
}

// per active cube: the mid points it adds (edges 5, 6, 10 that are active
// and have no analytic vertex) as a mask and a count, and the triangles it
// emits (all three vertices present after the mid points are added)
__global__ void __launch_bounds__(256) k_mc_counts(DevGrid g, int nact, const int *anode, const uint8_t *acase,
                                                   McVertexMap vm, uint8_t *amask, int *nv, int *nt) {
    const int q = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (q >= nact) return;
    const int64_t sz = (int64_t)g.nx * g.ny;
    const int64_t f = anode[q];
    const int c = acase[q];
    int i, j, k;
    cube_ijk(g, (int)f, i, j, k);
    // only active edges carry vertices, and only they appear in triangles
    unsigned have = 0;  // bit e: edge e has an analytic vertex
#pragma unroll
    for (int e = 0; e < 12; ++e)
        if (edge_on(c, e) && analytic_vertex(vm, e, i, j, k) != -1) have |= 1u << e;
    const int mk = (edge_on(c, 5) && !((have >> 5) & 1u) ? 1 : 0) | (edge_on(c, 6) && !((have >> 6) & 1u) ? 2 : 0) |
                   (edge_on(c, 10) && !((have >> 10) & 1u) ? 4 : 0);
    amask[q] = (uint8_t)mk;
    nv[q] = __popc((unsigned)mk);
    const uint64_t w = c_tri_cases[c];
    int t = 0;
    for (int tr = 0; tr < 5; ++tr) {
        bool ok = true;
        int e0 = (int)((w >> (12 * tr)) & 15);
        if (e0 == 15) break;
        for (int u = 0; u < 3; ++u) {
            const int e = (int)((w >> (12 * tr + 4 * u)) & 15);
            ok &= ((have >> e) & 1u) || owner_in_range(e, i, j, k);
        }
        t += ok;
    }
    nt[q] = t;
}

// vertexInterp (Surface.cpp:14516-14523): p1 + mu (p2 - p1), mu = (iso - v1) / (v2 - v1)
__device__ __forceinline__ void vertex_interp(double iso, const double p1[3], const double p2[3], double v1, double v2,
                                              double *out) {
    const double mu = (iso - v1) / (v2 - v1);
    out[0] = p1[0] + mu * (p2[0] - p1[0]);
    out[1] = p1[1] + mu * (p2[1] - p1[1]);
    out[2] = p1[2] + mu * (p2[2] - p1[2]);
}

__global__ void __launch_bounds__(256) k_mc_vwrite(DevGrid g, int nact, const int *anode, const uint8_t *acase,
                                                   McVertexMap vm, double *verts, int64_t *vslot, int *vorig,
                                                   const double *field, double iso) {
    const int q = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    // cubes that add no mid point read nothing more
    if (q >= nact) return;
    const int mk = vm.amask[q];
    if (mk == 0) return;
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz, sy = g.nx, sz = (int64_t)g.nx * g.ny;
    const int64_t f = anode[q];
    int i, j, k;
    cube_ijk(g, (int)f, i, j, k);
    const int base = vm.base;
    int v = base + vm.avoff[q];
    // the edges 5, 6, 10 of a cube belong to it only: no other lane writes them
    const int64_t s5 = edge_slot(5, f, sy, sz, nvox), s6 = edge_slot(6, f, sy, sz, nvox),
                  s10 = edge_slot(10, f, sy, sz, nvox);
    const bool m5 = mk & 1, m6 = mk & 2, m10 = mk & 4;
    if (field) {
        // corner positions x_i +- hside (the mulTable corners, tools.h:154-162)
        // and field values of the cube (getVertices :8143-8169)
        const double xm = g.x[i] - g.hside, xp = g.x[i] + g.hside, ym = g.y[j] - g.hside, yp = g.y[j] + g.hside,
                     zm = g.z[k] - g.hside, zp = g.z[k] + g.hside;
        const double c2[3] = {xp, yp, zm}, c5[3] = {xp, ym, zp}, c6[3] = {xp, yp, zp}, c7[3] = {xm, yp, zp};
        const double f2 = field[f + 1 + sy], f5 = field[f + 1 + sz], f6 = field[f + 1 + sy + sz],
                     f7 = field[f + sy + sz];
        if (m5) {
            vertex_interp(iso, c5, c6, f5, f6, verts + 3 * (int64_t)v);
            vslot[v] = s5;
            vorig[v++ - base] = 4 * q + 0;
        }
        if (m6) {
            vertex_interp(iso, c6, c7, f6, f7, verts + 3 * (int64_t)v);
            vslot[v] = s6;
            vorig[v++ - base] = 4 * q + 1;
        }
        if (m10) {
            vertex_interp(iso, c2, c6, f2, f6, verts + 3 * (int64_t)v);
            vslot[v] = s10;
            vorig[v++ - base] = 4 * q + 2;
        }
        return;
    }
    if (m5) {
        verts[3 * (int64_t)v] = g.x[i + 1] - g.hside;
        verts[3 * (int64_t)v + 1] = g.y[j];
        verts[3 * (int64_t)v + 2] = g.z[k + 1] - g.hside;
        vslot[v] = s5;
        vorig[v++ - base] = 4 * q + 0;
    }
    if (m6) {
        verts[3 * (int64_t)v] = g.x[i];
        verts[3 * (int64_t)v + 1] = g.y[j + 1] - g.hside;
        verts[3 * (int64_t)v + 2] = g.z[k + 1] - g.hside;
        vslot[v] = s6;
        vorig[v++ - base] = 4 * q + 1;
    }
    if (m10) {
        verts[3 * (int64_t)v] = g.x[i + 1] - g.hside;
        verts[3 * (int64_t)v + 1] = g.y[j + 1] - g.hside;
        verts[3 * (int64_t)v + 2] = g.z[k];
        vslot[v] = s10;
        vorig[v++ - base] = 4 * q + 2;
    }
}

// The vertex of edge e: the analytic one, else the mid point its owner adds.
__device__ __forceinline__ int edge_vertex(const DevGrid &g, const McVertexMap &vm, int e, int i, int j, int k) {
    const int v = analytic_vertex(vm, e, i, j, k);
    return v != -1 ? v : midpoint_vertex(g, vm, e, i, j, k);
}

__global__ void __launch_bounds__(256) k_mc_twrite(DevGrid g, int nact, const int *anode, const uint8_t *acase,
                                                   McVertexMap vm, const int *toff, const double *verts, int *tris,
                                                   double *tarea, int reverse) {
    const int q = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    // The valid lanes of a wave are a prefix and their triangles one
    // contiguous range from lane 0's offset (at most 5 per cube): they are
    // staged in LDS and the wave stores the range with consecutive lanes.
    // Stored per lane at its own offset, every store left partial lines.
    __shared__ int stri[4][3 * 5 * 64];
    __shared__ double sarea[4][5 * 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const bool valid = q < nact;
    const unsigned long long wmask = __ballot(valid);
    if (wmask == 0) return;
    const int base = __shfl(valid ? toff[q] : 0, 0, 64);
    const int64_t sz = (int64_t)g.nx * g.ny;
    const int64_t f = valid ? anode[q] : 0;
    int i, j, k;
    cube_ijk(g, (int)f, i, j, k);
    const uint64_t w = valid ? c_tri_cases[acase[q]] : ~0ull;
    int o = valid ? toff[q] : base;
    for (int t = 0; t < 5; ++t) {
        const int e0 = (int)((w >> (12 * t)) & 15), e1 = (int)((w >> (12 * t + 4)) & 15),
                  e2 = (int)((w >> (12 * t + 8)) & 15);
        if (e0 == 15) break;
        const int v0 = edge_vertex(g, vm, e0, i, j, k);
        const int v1 = edge_vertex(g, vm, e1, i, j, k);
        const int v2 = edge_vertex(g, vm, e2, i, j, k);
        if (v0 == -1 || v1 == -1 || v2 == -1) continue;
        const int so = o - base;  // < 5 * 64: at most five triangles per lane before it
        stri[wv][3 * so + 0] = reverse ? v2 : v0;
        stri[wv][3 * so + 1] = v1;
        stri[wv][3 * so + 2] = reverse ? v0 : v2;
        const double *a0 = verts + 3 * (int64_t)v0, *a1 = verts + 3 * (int64_t)v1, *a2 = verts + 3 * (int64_t)v2;
        const double a = sqrt((a0[0] - a1[0]) * (a0[0] - a1[0]) + (a0[1] - a1[1]) * (a0[1] - a1[1]) +
                              (a0[2] - a1[2]) * (a0[2] - a1[2]));
        const double b = sqrt((a0[0] - a2[0]) * (a0[0] - a2[0]) + (a0[1] - a2[1]) * (a0[1] - a2[1]) +
                              (a0[2] - a2[2]) * (a0[2] - a2[2]));
        const double cc = sqrt((a1[0] - a2[0]) * (a1[0] - a2[0]) + (a1[1] - a2[1]) * (a1[1] - a2[1]) +
                               (a1[2] - a2[2]) * (a1[2] - a2[2]));
        const double ttt = (a + b + cc) * (b + cc - a) * (cc + a - b) * (a + b - cc);
        sarea[wv][so] = ttt > 0 ? 0.25 * sqrt(ttt) : 0.0;
        ++o;
    }
    // the wave's end: the last valid lane's
    const int n = __shfl(o, 63 - __builtin_clzll(wmask), 64) - base;
    __builtin_amdgcn_wave_barrier();
    for (int x = lane; x < 3 * n; x += 64) tris[3 * (int64_t)base + x] = stri[wv][x];
    for (int x = lane; x < n; x += 64) tarea[base + x] = sarea[wv][x];
}

static inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

void mc_upload_tables(hipStream_t s) {
    hipMemcpyToSymbolAsync(HIP_SYMBOL(c_tri_cases), kMcTriCases, sizeof(kMcTriCases), 0, hipMemcpyHostToDevice, s);
}

void launch_vtx_count(hipStream_t s, const DevGrid &g, const DevRays &R, const int eslots[3], const uint32_t *bits,
                      int *counts) {
    EdgeSlots es{{eslots[0], eslots[1], eslots[2]}};
    const int64_t n = 3 * (int64_t)R.N * R.N;
    hipLaunchKernelGGL(k_vtx_count, dim3(nblk(n)), dim3(256), 0, s, g, R, es, bits, counts);
}

void launch_vtx_write(hipStream_t s, const DevGrid &g, const DevRays &R, const int eslots[3], const uint32_t *bits,
                      const int *offs, double *verts, uint64_t *vbits, int *vpre, int64_t *vslot, int64_t *vhit,
                      int64_t vcap) {
    EdgeSlots es{{eslots[0], eslots[1], eslots[2]}};
    const int64_t n = 3 * (int64_t)R.N * R.N;
    hipLaunchKernelGGL(k_vtx_write, dim3(nblk(n)), dim3(256), 0, s, g, R, es, bits, offs, verts, vbits, vpre, vslot,
                       vhit, vcap);
}

static SegIter seg_iter(const DevGrid &g) {
    SegIter si;
    si.cx = g.nx > 2 ? g.nx - 2 : 0;
    si.cy = g.ny > 2 ? g.ny - 2 : 0;
    si.cz = g.nz > 2 ? g.nz - 2 : 0;
    si.nseg_row = (si.cx + 31) / 32;
    si.nseg = (int64_t)si.nseg_row * si.cy * si.cz;
    return si;
}

int64_t mc_num_segments(const DevGrid &g) { return seg_iter(g).nseg; }

void launch_mc_classify(hipStream_t s, const DevGrid &g, const uint32_t *bits, uint32_t *segmask, int *segcnt) {
    const SegIter si = seg_iter(g);
    if (si.nseg) hipLaunchKernelGGL(k_mc_classify, dim3(nblk(si.nseg)), dim3(256), 0, s, g, si, bits, segmask, segcnt);
}

void launch_mc_compact(hipStream_t s, const DevGrid &g, const uint32_t *bits, const uint32_t *segmask,
                       const int *segoff, int *anode, uint8_t *acase) {
    const SegIter si = seg_iter(g);
    if (si.nseg)
        hipLaunchKernelGGL(k_mc_compact, dim3(nblk(si.nseg)), dim3(256), 0, s, g, si, bits, segmask, segoff, anode,
                           acase);
}

void launch_mc_counts(hipStream_t s, const DevGrid &g, int nact, const int *anode, const uint8_t *acase,
                      const McVertexMap &vm, uint8_t *amask, int *nv, int *nt) {
    if (nact)
        hipLaunchKernelGGL(k_mc_counts, dim3(nblk(nact)), dim3(256), 0, s, g, nact, anode, acase, vm, amask, nv, nt);
}

void launch_mc_vwrite(hipStream_t s, const DevGrid &g, int nact, const int *anode, const uint8_t *acase,
                      const McVertexMap &vm, double *verts, int64_t *vslot, int *vorig, const double *field,
                      double iso) {
    if (nact)
        hipLaunchKernelGGL(k_mc_vwrite, dim3(nblk(nact)), dim3(256), 0, s, g, nact, anode, acase, vm, verts, vslot,
                           vorig, field, iso);
}

void launch_mc_twrite(hipStream_t s, const DevGrid &g, int nact, const int *anode, const uint8_t *acase,
                      const McVertexMap &vm, const int *toff, const double *verts, int *tris, double *tarea,
                      int reverse) {
    if (nact)
        hipLaunchKernelGGL(k_mc_twrite, dim3(nblk(nact)), dim3(256), 0, s, g, nact, anode, acase, vm, toff, verts,
                           tris, tarea, reverse);
}

}  // namespace nsg
