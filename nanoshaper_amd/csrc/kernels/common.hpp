// Device-side types and helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nsgpu.h"

namespace nsg {

constexpr double kEpsInt = 1e-8;  // Surface.h:286 EPS_INT
constexpr double kDelta = 1e-10;  // Surface.h:290 delta_accurate_triangulation

// The axes of a panel: rays along vc, pixels over (fd, ld) (Surface.cpp:652-659)
template <int PANEL>
struct PanelAxes {
    static constexpr int fd = PANEL == 0 ? 1 : 0;
    static constexpr int ld = PANEL == 1 ? 1 : 2;
    static constexpr int vc = PANEL == 0 ? 0 : (PANEL == 1 ? 2 : 1);
};

// Grid as the kernels see it; node coordinates come from host-computed
// arrays so that device and host agree bit for bit.
struct DevGrid {
    int nx, ny, nz, N;
    double xmin, ymin, zmin, xmax, ymax, zmax;
    double scale, side, hside;
    const double *x, *y, *z;
};

struct DevPatches {
    int n;
    const int *kind, *off, *cnt, *flags;
    const double *bsph;  // 4 per patch
    const double *data;
    // patch ids ordered by class: [0, n_plain) spheres and facets, then tori,
    // then skin cells; each class runs as its own launch (the torus path needs
    // far more registers than the plain one)
    const int *order;
    int n_plain;
    int n_skin;  // skin cells, last in order (after the tori)
    const int *atom;  // 3 per patch: the atoms the patch is built on (-1 unused)
};

// Ray hit storage: ray r of slot s keeps up to K depths at hits[(s*NN+r)*K].
struct DevRays {
    int N;            // rays per panel side (max(nx,ny,nz))
    int K;            // capacity per ray
    int nslots;       // panels[0] + panels[1]
    int panels0;      // grid-ray panels
    int *cnt;         // [nslots*N*N]
    double *t;        // [nslots*N*N*K]
    double *rvol;     // [nslots*N*N] segment length per ray
    // Coverage bit planes written by the sort: bit n of a ray's plane is set
    // when the ray's kept segments cover node n of its axis (the map-fill rule
    // of its panel and phase).  W = ceil(N/64) words per ray; word w of ray r
    // of slot s at cov[(s*W + w)*N*N + r], so consecutive rays are
    // consecutive words.  cov_st: the status/idebmap plane of the x grid rays.
    int W;
    unsigned long long *cov;     // [nslots*W*N*N]
    unsigned long long *cov_st;  // [W*N*N]
    int *pid;         // [nslots*N*N*K] patch of each depth (sorted with it), or null
};

// rintp, tools.cpp:43-46
__device__ __forceinline__ double rintp(double x) { return floor(x + 0.5); }

// appends depth t of patch `patch` to ray `ray` of slot `slot`; a full list
// only counts (the host regrows K from the longest count and traces again)
__device__ __forceinline__ void push_hit(const DevRays &R, int slot, int64_t ray, double t, int patch) {
    const int64_t r = (int64_t)slot * R.N * R.N + ray;
    const int pos = atomicAdd(&R.cnt[r], 1);
    if (pos < R.K) {
        R.t[r * R.K + pos] = t;
        if (R.pid) R.pid[r * R.K + pos] = patch;
    }
}

// Workgroups are dealt to the eight XCDs (each with its own L2) round robin:
// block b runs on XCD b % 8.  This maps the blocks of one XCD onto one
// contiguous range of [0, nb), so that neighbouring work items, which share
// cache lines, meet in the same L2.  A bijection on [0, nb).
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int q = nb >> 3, r = nb & 7, x = b & 7, y = b >> 3;
    return x < r ? x * (q + 1) + y : r * (q + 1) + (x - r) * q + y;
}

__device__ __forceinline__ int64_t flat(const DevGrid &g, int i, int j, int k) {
    return ((int64_t)k * g.ny + j) * g.nx + i;
}

}  // namespace nsg

#define NSG_CHECK_LAUNCH() (void)hipGetLastError()
