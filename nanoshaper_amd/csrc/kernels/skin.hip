// Ray casting of the skin's mixed cells (SkinSurface::getPatchIntersectionData,
// SkinSurface.cpp:3422-4188, with RAY_VS_PATCH_TESTS_CULLING off).
//
// One 64-lane workgroup per (cell, panel slot), lanes striding over the
// pixel rectangle of the cell's bounding box, as for the SES patches.  Per
// pixel inside the box: the disc cull of a point cell's atom (EQ_CULLING),
// the unit-axis ray/quadric roots (rayQuadricIntersectionX/Y/Z,
// :4632-4708), then the cell's planes at each root (isFeasible, :4374-4447).
// The ray starts at xmin (+ delta) and a root t (in A) is stored as
// t / (xmax - xmin), the reference's t / ray_dir.
#include "launch.hpp"

namespace nsg {

// unit-axis ray o + t e_axis against the quadric; the coefficient order and
// the expressions are the reference's
__device__ __forceinline__ bool ray_quadric(const double *Q, const double o[3], int axis, double t[2]) {
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

__device__ __forceinline__ bool skin_feasible(const double *pl, int cnt, const double p[3]) {
    for (int q = 0; q < cnt; ++q) {
        const double *w = pl + 4 * q;
        if (w[0] * p[0] + w[1] * p[1] + w[2] * p[2] + w[3] > 0) return false;
    }
    return true;
}

__global__ void __launch_bounds__(64) k_intersect_skin(DevGrid g, DevPatches P, DevRays R, unsigned long long *part) {
    const int first = P.n - P.n_skin;
    const int64_t wg = (int64_t)blockIdx.y * P.n + first + blockIdx.x;
    const int it = P.order[first + blockIdx.x];
    const int slot = blockIdx.y;
    const int phase = slot < R.panels0 ? 0 : 1;
    const int panel = phase == 0 ? slot : slot - R.panels0;
    const int kind = P.kind[it];
    const double *pd = P.data + P.off[it];
    const double *Q = pd + 6;
    const double *pl = pd + NSGPU_SKIN_HDR;
    const int cnt = P.cnt[it];
    const double down[3] = {pd[0], pd[1], pd[2]};
    const double up[3] = {pd[3], pd[4], pd[5]};
    const double sc[3] = {P.bsph[4 * it], P.bsph[4 * it + 1], P.bsph[4 * it + 2]};
    const double radius = P.bsph[4 * it + 3];
    // CELL_CULLING (SkinSurface.cpp:3487-3500)
    if (down[0] > g.xmax + kEpsInt || down[1] > g.ymax + kEpsInt || down[2] > g.zmax + kEpsInt) return;
    const double lowpad = g.hside - kDelta + kEpsInt;
    if (up[0] < g.xmin - lowpad || up[1] < g.ymin - lowpad || up[2] < g.zmin - lowpad) return;
    const double gmin[3] = {g.xmin, g.ymin, g.zmin}, gmax[3] = {g.xmax, g.ymax, g.zmax};
    const int gn[3] = {g.nx, g.ny, g.nz};
    int64_t is[3], ie[3];
    for (int d = 0; d < 3; ++d) {
        is[d] = (int64_t)rintp(fmax(0., down[d] - gmin[d]) * g.scale);
        ie[d] = (int64_t)rintp(fmax(0., up[d] - gmin[d]) * g.scale);
        if (ie[d] >= gn[d]) ie[d] = gn[d] - 1;
    }
    const int64_t pix[3] = {1 + ie[0] - is[0], 1 + ie[1] - is[1], 1 + ie[2] - is[2]};
    const double delta = (phase == 0) ? 0. : kDelta - g.hside;
    int fd, ld, vc;
    const double *gf, *gl;
    if (panel == 0) { fd = 1; ld = 2; vc = 0; gf = g.y; gl = g.z; }
    else if (panel == 1) { fd = 0; ld = 1; vc = 2; gf = g.x; gl = g.y; }
    else { fd = 0; ld = 2; vc = 1; gf = g.x; gl = g.z; }
    const double ray_dir = gmax[vc] - gmin[vc];
    const int64_t total = pix[fd] * pix[ld];
    unsigned int cand = 0;
    for (int64_t rp = threadIdx.x; rp < total; rp += 64) {
        const int64_t nl = rp / pix[fd];
        const int64_t m = rp - nl * pix[fd] + is[fd];
        const int64_t n = nl + is[ld];
        double pa[3];
        pa[vc] = gmin[vc] + delta;
        pa[fd] = gf[m] + delta;
        if (pa[fd] < down[fd] || pa[fd] > up[fd]) continue;
        pa[ld] = gl[n] + delta;
        if (pa[ld] < down[ld] || pa[ld] > up[ld]) continue;
        ++cand;
        if (kind == NSGPU_PATCH_SKIN_POINT) {  // EQ_CULLING (:3602-3608)
            const double dx = sc[fd] - pa[fd], dy = sc[ld] - pa[ld];
            if (dx * dx + dy * dy > radius * radius) continue;
        }
        double t[2];
        if (!ray_quadric(Q, pa, vc, t)) continue;
        for (int i = 0; i < 2; ++i) {
            double ip[3] = {pa[0], pa[1], pa[2]};
            ip[vc] += t[i];
            if (!skin_feasible(pl, cnt, ip)) continue;
            push_hit(R, slot, n * R.N + m, t[i] / ray_dir, it);
        }
    }
    for (int off = 32; off > 0; off >>= 1) cand += __shfl_down(cand, off, 64);
    if (threadIdx.x == 0 && cand) part[wg] = cand;
}

void launch_intersect_skin(hipStream_t s, const DevGrid &g, const DevPatches &P, const DevRays &R,
                           unsigned long long *part) {
    if (P.n_skin <= 0 || R.nslots <= 0) return;
    hipLaunchKernelGGL(k_intersect_skin, dim3((unsigned)P.n_skin, (unsigned)R.nslots), dim3(64), 0, s, g, P, R, part);
}

}  // namespace nsg
