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

// Analytic normal of patch pid at its point y: the sphere's or torus's unit
// normal (ConnollySurface::getNormal, ConnollySurface.cpp:5736-5766;
// getNormalToTorus :5794-5819) or a skin cell's 2 grad Q, negated in the
// concave tetra cells and not normalised (SkinSurface.cpp:3680-3696).
__device__ __forceinline__ void analytic_normal(const DevPatches &P, int pid, double probe, const double y[3], double n[3]) {
    const int kind = P.kind[pid];
    const double *c = P.bsph + 4 * pid;
    if (kind >= NSGPU_PATCH_SKIN_POINT) {
        // skin cell: 2 grad Q / 2 = Q (x, 1), doubled and negated in the
        // concave tetra cells, not normalised (SkinSurface.cpp:3680-3696)
        const double *Q = P.data + P.off[pid] + 6;
        const double sg = kind == NSGPU_PATCH_SKIN_TETRA ? -2.0 : 2.0;
        n[0] = sg * (Q[0] * y[0] + Q[4] * y[1] + Q[7] * y[2] + Q[9]);
        n[1] = sg * (Q[4] * y[0] + Q[1] * y[1] + Q[5] * y[2] + Q[8]);
        n[2] = sg * (Q[7] * y[0] + Q[5] * y[1] + Q[2] * y[2] + Q[6]);
        return;
    }
    if (kind == NSGPU_PATCH_FACET_REG || kind == NSGPU_PATCH_FACET_SING) {
        for (int q = 0; q < 3; ++q) n[q] = (c[q] - y[q]) / probe;
        return;
    }
    if (kind == NSGPU_PATCH_POINT) {
        for (int q = 0; q < 3; ++q) n[q] = (y[q] - c[q]) / c[3];
        return;
    }
    // torus: into the torus frame, direction from the point to the circle,
    // back to the world frame (the frame is orthonormal: Rot = invrot^T)
    const double *td = P.data + P.off[pid];
    const double *iv = td + 3;
    const double t[3] = {y[0] - td[0], y[1] - td[1], y[2] - td[2]};
    double C[3];
    for (int i = 0; i < 3; ++i) C[i] = iv[3 * i + 0] * t[0] + iv[3 * i + 1] * t[1] + iv[3 * i + 2] * t[2];
    double d = C[0] * C[0] + C[1] * C[1];
    const double rm1 = td[12] / sqrt(d) - 1.;
    d = sqrt(d * rm1 * rm1 + C[2] * C[2]);
    C[0] *= rm1 / d;
    C[1] *= rm1 / d;
    C[2] /= -d;
    for (int i = 0; i < 3; ++i) n[i] = iv[0 + i] * C[0] + iv[3 + i] * C[1] + iv[6 + i] * C[2];
}

// Skin mixed cells (SkinSurface.cpp:4374-4708).
// unit-axis ray o + t e_axis against the quadric; the coefficient order and
// the expressions are the reference's
__device__ __forceinline__ bool ray_quadric(const double *Q, const double o[3], const int axis, double t[2]) {
    const double q44 = Q[3], q33 = Q[2], q22 = Q[1], q11 = Q[0], q34 = Q[6], q24 = Q[8], q23 = Q[5], q14 = Q[9],
                 q13 = Q[7], q12 = Q[4];
    double qa, b;
    if (axis == 0) {
        qa = q11;
        b = q14 + o[2] * q13 + o[1] * q12 + o[0] * q11;
    } else if (axis == 1) {
        qa = q22;
        b = q24 + o[2] * q23 + o[1] * q22 + o[0] * q12;
    } else {
        qa = q33;
        b = q34 + o[2] * q33 + o[1] * q23 + o[0] * q13;
    }
    if (qa == 0) return false;
    const double c = q44 + (o[2] + o[2]) * q34 + (o[2] * o[2]) * q33 + (o[1] + o[1]) * q24 + (2 * o[1] * o[2]) * q23 +
                     (o[1] * o[1]) * q22 + (o[0] + o[0]) * q14 + (2 * o[0] * o[2]) * q13 + (2 * o[0] * o[1]) * q12 +
                     (o[0] * o[0]) * q11;
    double delta = b * b - qa * c;
    if (delta < 0) return false;
    delta = sqrt(delta);
    t[0] = (-b - delta) / qa;
    t[1] = (-b + delta) / qa;
    return true;
}

// both roots of a ray against the cell's planes (isFeasible for each): one
// pass loads each plane once; a lane stops when both points have failed
__device__ __forceinline__ void skin_feasible2(const double *pl, int cnt, const double p[3], const double r[3],
                                               bool &ok_p, bool &ok_r) {
    bool out_p = false, out_r = false;
    for (int q = 0; q < cnt && !(out_p && out_r); ++q) {
        const double *w = pl + 4 * q;
        out_p |= w[0] * p[0] + w[1] * p[1] + w[2] * p[2] + w[3] > 0;
        out_r |= w[0] * r[0] + w[1] * r[1] + w[2] * r[2] + w[3] > 0;
    }
    ok_p = !out_p;
    ok_r = !out_r;
}

}  // namespace nsg

#define NSG_CHECK_LAUNCH() (void)hipGetLastError()
