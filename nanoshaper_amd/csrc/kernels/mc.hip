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
#include "launch.hpp"
#include "../common/mc_cases.h"

namespace nsg {

__device__ __forceinline__ int bit_at(const uint32_t *bits, int64_t f, int64_t nvox) {
    if (f < 0 || f >= nvox) return 1;
    return (bits[f >> 5] >> (uint32_t)(f & 31)) & 1u;
}

struct EdgeSlots {
    int s[3];  // edge-ray slots of panels 0 (x), 1 (z), 2 (y)
};

// Walk the vertex records of one edge ray.  emit(q, dir, ix, iy, iz, px, py, pz)
// is called for every kept record in order; returns the count.
template <class Emit>
__device__ int walk_edge_ray(const DevGrid &g, const DevRays &R, const uint32_t *bits, int panel, int slot,
                             int64_t ray, Emit emit) {
    const int c = R.cnt[(int64_t)slot * R.N * R.N + ray];
    if (c == 0 || (c & 1) || c > R.K) return 0;
    const double *T = R.t + ((int64_t)slot * R.N * R.N + ray) * R.K;
    const int n = (int)(ray / R.N), m = (int)(ray % R.N);
    const double delta = kDelta - g.hside;
    const double *ga = panel == 0 ? g.x : (panel == 1 ? g.z : g.y);
    const int na = panel == 0 ? g.nx : (panel == 1 ? g.nz : g.ny);
    const double gmin = panel == 0 ? g.xmin : (panel == 1 ? g.zmin : g.ymin);
    const double pa = ga[0] + delta;
    const double dir = (ga[na - 1] + delta) - pa;
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    // fixed coordinates of the ray and the flat-index stride along it
    double p0, p1;
    int64_t stride;
    if (panel == 0) {
        if (m >= g.ny || n >= g.nz) return 0;
        p0 = g.y[m] + delta;
        p1 = g.z[n] + delta;
        stride = 1;
    } else if (panel == 1) {
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
    int claimed[2 * 16];
    int nclaimed = 0, count = 0;
    for (int q = 0; q < c; q += 2) {
        const double t0 = T[q], t1 = T[q + 1];
        if (!(fabs(t1 - t0) >= kEpsInt)) continue;
        const double lim[2] = {pa + dir * t0, pa + dir * t1};
        const int a = (int)rintp((lim[0] - gmin) * g.scale);
        const int b = (int)rintp((lim[1] - gmin) * g.scale);
        if (b <= a) continue;
        const int e[2] = {a, b};
        for (int h = 0; h < 2; ++h) {
            const int ia = e[h];
            if (ia < 0 || ia >= na) continue;
            int ix, iy, iz, d;
            if (panel == 0) { ix = ia; iy = m; iz = n; d = 0; }
            else if (panel == 1) { ix = m; iy = n; iz = ia; d = 2; }
            else { ix = m; iy = ia; iz = n; d = 1; }
            const int64_t f = ((int64_t)iz * g.ny + iy) * g.nx + ix;
            if (bit_at(bits, f, nvox) == bit_at(bits, f + stride, nvox)) continue;
            bool dup = false;
            for (int u = 0; u < nclaimed; ++u) dup |= claimed[u] == ia;
            if (dup) continue;
            if (nclaimed < 32) claimed[nclaimed++] = ia;
            double px, py, pz;
            if (panel == 0) { px = lim[h]; py = p0; pz = p1; }
            else if (panel == 1) { px = p0; py = p1; pz = lim[h]; }
            else { px = p0; py = lim[h]; pz = p1; }
            emit(count, d, f, px, py, pz);
            ++count;
        }
    }
    return count;
}

__global__ void __launch_bounds__(256) k_vtx_count(DevGrid g, DevRays R, EdgeSlots es, const uint32_t *bits,
                                                   int *counts) {
    const int64_t NN = (int64_t)R.N * R.N;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= 3 * NN) return;
    const int panel = (int)(r / NN);
    counts[r] = walk_edge_ray(g, R, bits, panel, es.s[panel], r - panel * NN,
                              [](int, int, int64_t, double, double, double) {});
}

__global__ void __launch_bounds__(256) k_vtx_write(DevGrid g, DevRays R, EdgeSlots es, const uint32_t *bits,
                                                   const int *offs, double *verts, int *idx) {
    const int64_t NN = (int64_t)R.N * R.N;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= 3 * NN) return;
    const int panel = (int)(r / NN);
    const int base = offs[r];
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    walk_edge_ray(g, R, bits, panel, es.s[panel], r - panel * NN,
                  [&](int q, int d, int64_t f, double px, double py, double pz) {
                      const int v = base + q;
                      verts[3 * (int64_t)v + 0] = px;
                      verts[3 * (int64_t)v + 1] = py;
                      verts[3 * (int64_t)v + 2] = pz;
                      idx[d * nvox + f] = v;
                  });
}

// ---------------------------------------------------------------- marching cubes
__device__ __forceinline__ int cube_case(const uint32_t *bits, int64_t f, int64_t sy, int64_t sz, int64_t nvox) {
    int c = 0;
    c |= (!bit_at(bits, f, nvox)) << 0;
    c |= (!bit_at(bits, f + 1, nvox)) << 1;
    c |= (!bit_at(bits, f + 1 + sy, nvox)) << 2;
    c |= (!bit_at(bits, f + sy, nvox)) << 3;
    c |= (!bit_at(bits, f + sz, nvox)) << 4;
    c |= (!bit_at(bits, f + 1 + sz, nvox)) << 5;
    c |= (!bit_at(bits, f + 1 + sy + sz, nvox)) << 6;
    c |= (!bit_at(bits, f + sy + sz, nvox)) << 7;
    return c;
}

__constant__ uint64_t c_tri_cases[256];

__device__ __forceinline__ bool edge_on(int c, int e) {
    const int a = (e < 8) ? (e & 3) + (e & 4) : e - 8;
    int b;
    if (e < 8) b = ((e & 3) == 3 ? (e & 4) : (e & 3) + 1 + (e & 4));
    else b = e - 4;
    return ((c >> a) & 1) != ((c >> b) & 1);
}

// edge e of cube (i,j,k) -> vertex index slot: direction and node offset (getTriangles)
__device__ __forceinline__ int64_t edge_slot(int e, int64_t f, int64_t sy, int64_t sz, int64_t nvox) {
    switch (e) {
        case 0: return 0 * nvox + f;
        case 1: return 1 * nvox + f + 1;
        case 2: return 0 * nvox + f + sy;
        case 3: return 1 * nvox + f;
        case 4: return 0 * nvox + f + sz;
        case 5: return 1 * nvox + f + 1 + sz;
        case 6: return 0 * nvox + f + sy + sz;
        case 7: return 1 * nvox + f + sz;
        case 8: return 2 * nvox + f;
        case 9: return 2 * nvox + f + 1;
        case 10: return 2 * nvox + f + 1 + sy;
        default: return 2 * nvox + f + sy;
    }
}

struct CubeIter {
    int64_t ncubes;
    int cx, cy;  // cube counts along x and y
};

__device__ __forceinline__ int64_t cube_node(const DevGrid &g, const CubeIter &ci, int64_t q) {
    const int i = 1 + (int)(q % ci.cx);
    const int j = 1 + (int)((q / ci.cx) % ci.cy);
    const int k = 1 + (int)(q / ((int64_t)ci.cx * ci.cy));
    return ((int64_t)k * g.ny + j) * g.nx + i;
}

__global__ void __launch_bounds__(256) k_mc_vcount(DevGrid g, CubeIter ci, const uint32_t *bits, const int *idx,
                                                   int *counts) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ci.ncubes) return;
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz, sy = g.nx, sz = (int64_t)g.nx * g.ny;
    const int64_t f = cube_node(g, ci, q);
    const int c = cube_case(bits, f, sy, sz, nvox);
    int n = 0;
    if (c != 0 && c != 255) {
        n += edge_on(c, 5) && idx[edge_slot(5, f, sy, sz, nvox)] == -1;
        n += edge_on(c, 6) && idx[edge_slot(6, f, sy, sz, nvox)] == -1;
        n += edge_on(c, 10) && idx[edge_slot(10, f, sy, sz, nvox)] == -1;
    }
    counts[q] = n;
}

__global__ void __launch_bounds__(256) k_mc_vwrite(DevGrid g, CubeIter ci, const uint32_t *bits, int *idx,
                                                   const int *offs, int base, double *verts) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ci.ncubes) return;
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz, sy = g.nx, sz = (int64_t)g.nx * g.ny;
    const int64_t f = cube_node(g, ci, q);
    const int c = cube_case(bits, f, sy, sz, nvox);
    if (c == 0 || c == 255) return;
    const int i = (int)(f % g.nx), j = (int)((f / g.nx) % g.ny), k = (int)(f / sz);
    int v = base + offs[q];
    // the three slots are owned by this cube only, so reading them here is race free
    const int64_t s5 = edge_slot(5, f, sy, sz, nvox), s6 = edge_slot(6, f, sy, sz, nvox),
                  s10 = edge_slot(10, f, sy, sz, nvox);
    const bool m5 = edge_on(c, 5) && idx[s5] == -1;
    const bool m6 = edge_on(c, 6) && idx[s6] == -1;
    const bool m10 = edge_on(c, 10) && idx[s10] == -1;
    if (m5) {
        verts[3 * (int64_t)v] = g.x[i + 1] - g.hside;
        verts[3 * (int64_t)v + 1] = g.y[j];
        verts[3 * (int64_t)v + 2] = g.z[k + 1] - g.hside;
        idx[s5] = v++;
    }
    if (m6) {
        verts[3 * (int64_t)v] = g.x[i];
        verts[3 * (int64_t)v + 1] = g.y[j + 1] - g.hside;
        verts[3 * (int64_t)v + 2] = g.z[k + 1] - g.hside;
        idx[s6] = v++;
    }
    if (m10) {
        verts[3 * (int64_t)v] = g.x[i + 1] - g.hside;
        verts[3 * (int64_t)v + 1] = g.y[j + 1] - g.hside;
        verts[3 * (int64_t)v + 2] = g.z[k];
        idx[s10] = v++;
    }
}

template <bool kWrite>
__device__ int cube_tris(const DevGrid &g, const CubeIter &ci, const uint32_t *bits, const int *idx, int64_t q,
                         const double *verts, int *tris, double *tarea, int base) {
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz, sy = g.nx, sz = (int64_t)g.nx * g.ny;
    const int64_t f = cube_node(g, ci, q);
    const int c = cube_case(bits, f, sy, sz, nvox);
    if (c == 0 || c == 255) return 0;
    const uint64_t w = c_tri_cases[c];
    int n = 0;
    for (int t = 0; t < 5; ++t) {
        const int e0 = (int)((w >> (12 * t)) & 15), e1 = (int)((w >> (12 * t + 4)) & 15),
                  e2 = (int)((w >> (12 * t + 8)) & 15);
        if (e0 == 15) break;
        const int v0 = idx[edge_slot(e0, f, sy, sz, nvox)];
        const int v1 = idx[edge_slot(e1, f, sy, sz, nvox)];
        const int v2 = idx[edge_slot(e2, f, sy, sz, nvox)];
        if (v0 == -1 || v1 == -1 || v2 == -1) continue;
        if (kWrite) {
            const int o = base + n;
            tris[3 * (int64_t)o + 0] = v2;
            tris[3 * (int64_t)o + 1] = v1;
            tris[3 * (int64_t)o + 2] = v0;
            const double *a0 = verts + 3 * (int64_t)v0, *a1 = verts + 3 * (int64_t)v1, *a2 = verts + 3 * (int64_t)v2;
            const double a = sqrt((a0[0] - a1[0]) * (a0[0] - a1[0]) + (a0[1] - a1[1]) * (a0[1] - a1[1]) +
                                  (a0[2] - a1[2]) * (a0[2] - a1[2]));
            const double b = sqrt((a0[0] - a2[0]) * (a0[0] - a2[0]) + (a0[1] - a2[1]) * (a0[1] - a2[1]) +
                                  (a0[2] - a2[2]) * (a0[2] - a2[2]));
            const double cc = sqrt((a1[0] - a2[0]) * (a1[0] - a2[0]) + (a1[1] - a2[1]) * (a1[1] - a2[1]) +
                                   (a1[2] - a2[2]) * (a1[2] - a2[2]));
            const double ttt = (a + b + cc) * (b + cc - a) * (cc + a - b) * (a + b - cc);
            tarea[o] = ttt > 0 ? 0.25 * sqrt(ttt) : 0.0;
        }
        ++n;
    }
    return n;
}

__global__ void __launch_bounds__(256) k_mc_tcount(DevGrid g, CubeIter ci, const uint32_t *bits, const int *idx,
                                                   int *counts) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ci.ncubes) return;
    counts[q] = cube_tris<false>(g, ci, bits, idx, q, nullptr, nullptr, nullptr, 0);
}

__global__ void __launch_bounds__(256) k_mc_twrite(DevGrid g, CubeIter ci, const uint32_t *bits, const int *idx,
                                                   const int *offs, const double *verts, int *tris, double *tarea) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ci.ncubes) return;
    cube_tris<true>(g, ci, bits, idx, q, verts, tris, tarea, offs[q]);
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
                      const int *offs, double *verts, int *idx) {
    EdgeSlots es{{eslots[0], eslots[1], eslots[2]}};
    const int64_t n = 3 * (int64_t)R.N * R.N;
    hipLaunchKernelGGL(k_vtx_write, dim3(nblk(n)), dim3(256), 0, s, g, R, es, bits, offs, verts, idx);
}

static CubeIter cube_iter(const DevGrid &g) {
    CubeIter ci;
    ci.cx = g.nx > 2 ? g.nx - 2 : 0;
    ci.cy = g.ny > 2 ? g.ny - 2 : 0;
    const int cz = g.nz > 2 ? g.nz - 2 : 0;
    ci.ncubes = (int64_t)ci.cx * ci.cy * cz;
    return ci;
}

int64_t mc_num_cubes(const DevGrid &g) { return cube_iter(g).ncubes; }

void launch_mc_vcount(hipStream_t s, const DevGrid &g, const uint32_t *bits, const int *idx, int *counts) {
    const CubeIter ci = cube_iter(g);
    if (ci.ncubes) hipLaunchKernelGGL(k_mc_vcount, dim3(nblk(ci.ncubes)), dim3(256), 0, s, g, ci, bits, idx, counts);
}

void launch_mc_vwrite(hipStream_t s, const DevGrid &g, const uint32_t *bits, int *idx, const int *offs, int base,
                      double *verts) {
    const CubeIter ci = cube_iter(g);
    if (ci.ncubes)
        hipLaunchKernelGGL(k_mc_vwrite, dim3(nblk(ci.ncubes)), dim3(256), 0, s, g, ci, bits, idx, offs, base, verts);
}

void launch_mc_tcount(hipStream_t s, const DevGrid &g, const uint32_t *bits, const int *idx, int *counts) {
    const CubeIter ci = cube_iter(g);
    if (ci.ncubes) hipLaunchKernelGGL(k_mc_tcount, dim3(nblk(ci.ncubes)), dim3(256), 0, s, g, ci, bits, idx, counts);
}

void launch_mc_twrite(hipStream_t s, const DevGrid &g, const uint32_t *bits, const int *idx, const int *offs,
                      const double *verts, int *tris, double *tarea) {
    const CubeIter ci = cube_iter(g);
    if (ci.ncubes)
        hipLaunchKernelGGL(k_mc_twrite, dim3(nblk(ci.ncubes)), dim3(256), 0, s, g, ci, bits, idx, offs, verts, tris,
                           tarea);
}

}  // namespace nsg
