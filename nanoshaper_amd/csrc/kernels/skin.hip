// Ray casting of the skin's mixed cells (SkinSurface::getPatchIntersectionData,
// SkinSurface.cpp:3422-4188, with RAY_VS_PATCH_TESTS_CULLING off).
//
// Per (cell, panel slot), the pixels of the rectangle of the cell's bounding
// box are shared out over sixteen lanes (or a wave, for large cells).  Per
// pixel inside the box: the disc cull of a point cell's atom (EQ_CULLING),
// the unit-axis ray/quadric roots (rayQuadricIntersectionX/Y/Z,
// :4632-4708), then the cell's planes at each root (isFeasible, :4374-4447).
// The ray starts at xmin (+ delta) and a root t (in A) is stored as
// t / (xmax - xmin), the reference's t / ray_dir.
#include "launch.hpp"

namespace nsg {

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

// A cell's rectangle for one panel slot.
struct SkinRect {
    int it, kind, cnt;
    int is0, is1, pf, total;
    double down[3], up[3];
};

template <int PANEL>
__device__ __forceinline__ SkinRect skin_rect(const DevGrid &g, const DevPatches &P, int it) {
    using A = PanelAxes<PANEL>;
    SkinRect c;
    c.it = it;
    c.kind = P.kind[it];
    c.cnt = P.cnt[it];
    const double *pd = P.data + P.off[it];
    for (int q = 0; q < 3; ++q) {
        c.down[q] = pd[q];
        c.up[q] = pd[3 + q];
    }
    // CELL_CULLING (SkinSurface.cpp:3487-3500)
    const double lowpad = g.hside - kDelta + kEpsInt;
    const bool culled = c.down[0] > g.xmax + kEpsInt || c.down[1] > g.ymax + kEpsInt || c.down[2] > g.zmax + kEpsInt ||
                        c.up[0] < g.xmin - lowpad || c.up[1] < g.ymin - lowpad || c.up[2] < g.zmin - lowpad;
    const double gmin[3] = {g.xmin, g.ymin, g.zmin};
    const int gn[3] = {g.nx, g.ny, g.nz};
    int64_t is0 = (int64_t)rintp(fmax(0., c.down[A::fd] - gmin[A::fd]) * g.scale);
    int64_t ie0 = (int64_t)rintp(fmax(0., c.up[A::fd] - gmin[A::fd]) * g.scale);
    int64_t is1 = (int64_t)rintp(fmax(0., c.down[A::ld] - gmin[A::ld]) * g.scale);
    int64_t ie1 = (int64_t)rintp(fmax(0., c.up[A::ld] - gmin[A::ld]) * g.scale);
    if (ie0 >= gn[A::fd]) ie0 = gn[A::fd] - 1;
    if (ie1 >= gn[A::ld]) ie1 = gn[A::ld] - 1;
    const int64_t pf = 1 + ie0 - is0, pl = 1 + ie1 - is1;
    c.is0 = (int)min(is0, (int64_t)gn[A::fd]);
    c.is1 = (int)min(is1, (int64_t)gn[A::ld]);
    c.pf = (int)max(pf, (int64_t)1);
    c.total = (!culled && pf > 0 && pl > 0) ? (int)(pf * pl) : 0;
    return c;
}

// pixels lane, lane + stride, ... of the rectangle; returns the candidates
template <int PANEL>
__device__ __forceinline__ unsigned int skin_rect_rays(const DevGrid &g, const DevPatches &P, const DevRays &R,
                                                       const SkinRect &c, int slot, double delta, int lane,
                                                       int stride) {
    using A = PanelAxes<PANEL>;
    const double *gf = A::fd == 0 ? g.x : g.y;
    const double *gl = A::ld == 1 ? g.y : g.z;
    const double gmin[3] = {g.xmin, g.ymin, g.zmin}, gmax[3] = {g.xmax, g.ymax, g.zmax};
    const double ray_dir = gmax[A::vc] - gmin[A::vc];
    const double *pd = P.data + P.off[c.it];
    const double *sp = P.bsph + 4 * c.it;
    const double sfd = sp[A::fd], sld = sp[A::ld], srad = sp[3];
    unsigned int cand = 0;
    for (int rp = lane; rp < c.total; rp += stride) {
        const int nl = rp / c.pf;
        const int m = rp - nl * c.pf + c.is0;
        const int n = nl + c.is1;
        double pa[3];
        pa[A::vc] = gmin[A::vc] + delta;
        pa[A::fd] = gf[m] + delta;
        if (pa[A::fd] < c.down[A::fd] || pa[A::fd] > c.up[A::fd]) continue;
        pa[A::ld] = gl[n] + delta;
        if (pa[A::ld] < c.down[A::ld] || pa[A::ld] > c.up[A::ld]) continue;
        ++cand;
        if (c.kind == NSGPU_PATCH_SKIN_POINT) {  // EQ_CULLING (:3602-3608)
            const double dx = sfd - pa[A::fd], dy = sld - pa[A::ld];
            if (dx * dx + dy * dy > srad * srad) continue;
        }
        double t[2];
        if (!ray_quadric(pd + 6, pa, A::vc, t)) continue;
        double ip0[3] = {pa[0], pa[1], pa[2]}, ip1[3] = {pa[0], pa[1], pa[2]};
        ip0[A::vc] += t[0];
        ip1[A::vc] += t[1];
        bool ok0, ok1;
        skin_feasible2(pd + NSGPU_SKIN_HDR, c.cnt, ip0, ip1, ok0, ok1);
        if (ok0) push_hit(R, slot, (int64_t)n * R.N + m, t[0] / ray_dir, c.it);
        if (ok1) push_hit(R, slot, (int64_t)n * R.N + m, t[1] / ray_dir, c.it);
    }
    return cand;
}

// grid: (ceil(n_skin / 4), nslots).  A wave takes four cells, sixteen lanes
// each, so the lanes reading a cell's quadric and planes share their
// addresses.  Rectangles above kWaveCell pixels (edge and facet cells next to
// the molecule's outside span tens of A) are walked by the whole wave after.
constexpr int kWaveCell = 256;

template <int PANEL>
__device__ __forceinline__ unsigned int skin_slot(const DevGrid &g, const DevPatches &P, const DevRays &R, int slot,
                                                  double delta) {
    const int first = P.n - P.n_skin;
    const int lane = threadIdx.x, grp = lane >> 4, gl16 = lane & 15;
    const int ci = blockIdx.x * 4 + grp;
    unsigned int cand = 0;
    bool big = false;
    if (ci < P.n_skin) {
        const SkinRect c = skin_rect<PANEL>(g, P, P.order[first + ci]);
        big = c.total > kWaveCell;
        if (!big) cand += skin_rect_rays<PANEL>(g, P, R, c, slot, delta, gl16, 16);
    }
    uint64_t mask = __ballot(big && gl16 == 0);
    while (mask) {
        const int b = __ffsll((unsigned long long)mask) - 1;
        mask &= mask - 1;
        const SkinRect c = skin_rect<PANEL>(g, P, P.order[first + blockIdx.x * 4 + (b >> 4)]);
        cand += skin_rect_rays<PANEL>(g, P, R, c, slot, delta, lane, 64);
    }
    return cand;
}

__global__ void __launch_bounds__(64) k_intersect_skin(DevGrid g, DevPatches P, DevRays R, unsigned long long *part) {
    const int first = P.n - P.n_skin;
    const int64_t wg = (int64_t)blockIdx.y * P.n + first + blockIdx.x;
    const int slot = blockIdx.y;
    const int phase = slot < R.panels0 ? 0 : 1;
    const int panel = phase == 0 ? slot : slot - R.panels0;
    const double delta = (phase == 0) ? 0. : kDelta - g.hside;
    unsigned int cand;
    if (panel == 0) cand = skin_slot<0>(g, P, R, slot, delta);
    else if (panel == 1) cand = skin_slot<1>(g, P, R, slot, delta);
    else cand = skin_slot<2>(g, P, R, slot, delta);
    for (int off = 32; off > 0; off >>= 1) cand += __shfl_down(cand, off, 64);
    if (threadIdx.x == 0 && cand) part[wg] = cand;
}

void launch_intersect_skin(hipStream_t s, const DevGrid &g, const DevPatches &P, const DevRays &R,
                           unsigned long long *part) {
    if (P.n_skin <= 0 || R.nslots <= 0) return;
    hipLaunchKernelGGL(k_intersect_skin, dim3((unsigned)((P.n_skin + 3) / 4), (unsigned)R.nslots), dim3(64), 0, s,
                       g, P, R, part);
}

}  // namespace nsg
