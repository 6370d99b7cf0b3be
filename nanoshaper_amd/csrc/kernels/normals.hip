// Vertex normals of the analytic marching-cubes mesh (Compute_Vertex_Normals).
//
// Analytic vertices come from edge-ray depths: their normal is the patch's
// analytic normal at the intersection point, as getPatchIntersectionData
// stores it with the depth (ConnollySurface::getNormal, ConnollySurface.cpp:
// 5736-5766; getNormalToTorus :5794-5819; the skin cells' quadric gradient,
// SkinSurface.cpp:3680-3696).  Mid-point vertices have no
// analytic normal: like approximateNormals (Surface.cpp:10588-10737) their
// normal is the normalised sum of plane3points (tools.cpp:258-274, not
// normalised) over the triangles that use them, in triangle order.
#include "launch.hpp"

namespace nsg {

__global__ void __launch_bounds__(256) k_normals_analytic(DevGrid g, DevPatches P, DevRays R, NormalArgs a) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.n_analytic) return;
    const int64_t NN = (int64_t)R.N * R.N;
    const int64_t h = a.vhit[v];
    const int slot = (int)(h / (NN * R.K));
    const int64_t ray = (h / R.K) % NN;
    const int panel = slot - a.eslot0;
    const int m = (int)(ray % R.N), nn = (int)(ray / R.N);
    int fd, ld, vc, na;
    const double *ga, *gf, *gl;
    if (panel == 0) { fd = 1; ld = 2; vc = 0; ga = g.x; gf = g.y; gl = g.z; na = g.nx; }
    else if (panel == 1) { fd = 0; ld = 1; vc = 2; ga = g.z; gf = g.x; gl = g.y; na = g.nz; }
    else { fd = 0; ld = 2; vc = 1; ga = g.y; gf = g.x; gl = g.z; na = g.ny; }
    // the intersection point as the ray/patch test formed it (edge rays)
    const double delta = kDelta - g.hside;
    double y[3];
    y[vc] = ga[0] + delta;
    y[fd] = gf[m] + delta;
    y[ld] = gl[nn] + delta;
    y[vc] += R.t[h] * (ga[na - 1] - ga[0]);
    double n[3];
    analytic_normal(P, R.pid[h], a.probe, y, n);
    for (int q = 0; q < 3; ++q) a.normals[3 * v + q] = n[q];
}

// first index of `key` in the sorted active-cube list, or -1
__device__ __forceinline__ int find_cube(const int *anode, int nact, int key) {
    int lo = 0, hi = nact;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (anode[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return (lo < nact && anode[lo] == key) ? lo : -1;
}

__global__ void __launch_bounds__(256) k_normals_midpoint(DevGrid g, NormalArgs a) {
    const int64_t v = a.n_analytic + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.nverts) return;
    const int o = a.vorig[v - a.n_analytic];
    const int f = a.anode[o >> 2];
    const int sy = g.nx, sz = g.nx * g.ny;
    // the four cubes around the mid point's edge, in increasing flat index:
    // edge 5 (y edge at (i+1, j, k+1)), edge 6 (x edge at (i, j+1, k+1)),
    // edge 10 (z edge at (i+1, j+1, k))
    int cubes[4];
    switch (o & 3) {
        case 0: cubes[0] = f; cubes[1] = f + 1; cubes[2] = f + sz; cubes[3] = f + 1 + sz; break;
        case 1: cubes[0] = f; cubes[1] = f + sy; cubes[2] = f + sz; cubes[3] = f + sy + sz; break;
        default: cubes[0] = f; cubes[1] = f + 1; cubes[2] = f + sy; cubes[3] = f + 1 + sy; break;
    }
    double s[3] = {0., 0., 0.};
    for (int c = 0; c < 4; ++c) {
        const int q = find_cube(a.anode, a.nact, cubes[c]);
        if (q < 0) continue;
        for (int t = a.atoff[q]; t < a.atoff[q] + a.ant[q]; ++t) {
            const int *tv = a.tris + 3 * (int64_t)t;
            if (tv[0] != v && tv[1] != v && tv[2] != v) continue;
            const double *p1 = a.verts + 3 * (int64_t)tv[0], *p2 = a.verts + 3 * (int64_t)tv[1],
                         *p3 = a.verts + 3 * (int64_t)tv[2];
            s[0] += p1[1] * (p2[2] - p3[2]) + p2[1] * (p3[2] - p1[2]) + p3[1] * (p1[2] - p2[2]);
            s[1] += p1[2] * (p2[0] - p3[0]) + p2[2] * (p3[0] - p1[0]) + p3[2] * (p1[0] - p2[0]);
            s[2] += p1[0] * (p2[1] - p3[1]) + p2[0] * (p3[1] - p1[1]) + p3[0] * (p1[1] - p2[1]);
        }
    }
    const double norm = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (int q = 0; q < 3; ++q) a.normals[3 * v + q] = s[q] / norm;
}

void launch_normals(hipStream_t s, const DevGrid &g, const DevPatches &P, const DevRays &R, const NormalArgs &a) {
    if (a.n_analytic > 0)
        hipLaunchKernelGGL(k_normals_analytic, dim3((unsigned)((a.n_analytic + 255) / 256)), dim3(256), 0, s, g, P, R,
                           a);
    const int64_t nm = a.nverts - a.n_analytic;
    if (nm > 0) hipLaunchKernelGGL(k_normals_midpoint, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, s, g, a);
}

}  // namespace nsg
