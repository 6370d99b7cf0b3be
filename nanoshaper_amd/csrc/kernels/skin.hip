// Ray casting of the skin's mixed cells (SkinSurface::getPatchIntersectionData,
// SkinSurface.cpp:3422-4188, with RAY_VS_PATCH_TESTS_CULLING off).
//
// Per (cell, panel slot), the pixels of the rectangle of the cell's bounding
// box (64 cells' rectangles per wave, shared out evenly).  Per pixel inside
// the box: the disc cull of a point cell's atom (EQ_CULLING),
// the unit-axis ray/quadric roots (rayQuadricIntersectionX/Y/Z,
// :4632-4708), then the cell's planes at each root (isFeasible, :4374-4447).
// The ray starts at xmin (+ delta) and a root t (in A) is stored as
// t / (xmax - xmin), the reference's t / ray_dir.
#include "launch.hpp"

namespace nsg {

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

// One workgroup (a wave) takes kSkinCells cells for one panel slot.  Lane q <
// kSkinCells sets up one cell's rectangle; a wave scan lays the rectangles'
// pixels end to end, and the lanes take consecutive pixels of that list, 64 at
// a time, finding their cell by a binary search over the scanned ends in LDS.
// Small and large cells share the lanes evenly, and a cell's set-up is done
// once (sixteen lanes per cell, walking rectangles of a few dozen pixels, left
// half the lanes idle and repeated the set-up per group).  The cells come
// sorted largest first and are dealt to the waves in turn, so every wave gets
// one cell of each size band and the waves' totals are alike.
constexpr int kSkinCells = 32;
struct SkinCells {
    int end[64], it[64], kind[64], cnt[64], off[64], is0[64], is1[64], pf[64];
    float rpf[64];
    double box[4][64], sph[3][64];  // down/up along fd, ld; bounding sphere fd, ld, radius
};

template <int PANEL>
__device__ __forceinline__ unsigned int skin_slot(const DevGrid &g, const DevPatches &P, const DevRays &R, int slot,
                                                  double delta, int group, SkinCells &L) {
    using A = PanelAxes<PANEL>;
    int *s_end = L.end, *s_it = L.it, *s_kind = L.kind, *s_cnt = L.cnt, *s_off = L.off, *s_is0 = L.is0,
        *s_is1 = L.is1, *s_pf = L.pf;
    float *s_rpf = L.rpf;
    double(*s_box)[64] = L.box;
    double(*s_sph)[64] = L.sph;
    const int lane = threadIdx.x;
    const int first = P.n - P.n_skin;
    // cells come largest first; wave w takes cells w, w + W, w + 2 W, ... (W
    // waves per slot), one from each size band, so the waves' totals are alike
    const int nw = (P.n_skin + kSkinCells - 1) / kSkinCells;
    const int ci = lane * nw + group;
    int total = 0;
    if (lane < kSkinCells && ci < P.n_skin) {
        const SkinRect c = skin_rect<PANEL>(g, P, P.order[first + ci]);
        total = c.total;
        s_it[lane] = c.it;
        s_kind[lane] = c.kind;
        s_cnt[lane] = c.cnt;
        s_off[lane] = P.off[c.it];
        s_is0[lane] = c.is0;
        s_is1[lane] = c.is1;
        s_pf[lane] = c.pf;
        s_rpf[lane] = 1.f / (float)c.pf;
        s_box[0][lane] = c.down[A::fd];
        s_box[1][lane] = c.up[A::fd];
        s_box[2][lane] = c.down[A::ld];
        s_box[3][lane] = c.up[A::ld];
        const double *sp = P.bsph + 4 * c.it;
        s_sph[0][lane] = sp[A::fd];
        s_sph[1][lane] = sp[A::ld];
        s_sph[2][lane] = sp[3];
    }
    int incl = total;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    s_end[lane] = incl;
    const int wtotal = __shfl(incl, 63, 64);
    __syncthreads();
    const double *gf = A::fd == 0 ? g.x : g.y;
    const double *gl = A::ld == 1 ? g.y : g.z;
    const double gmin[3] = {g.xmin, g.ymin, g.zmin}, gmax[3] = {g.xmax, g.ymax, g.zmax};
    const double ray_dir = gmax[A::vc] - gmin[A::vc];
    unsigned int cand = 0;
    for (int p = lane; p < wtotal; p += 64) {
        // cell j: the first whose scanned end exceeds p (ends never decrease)
        int j = 0;
#pragma unroll
        for (int st = kSkinCells / 2; st > 0; st >>= 1)
            if (s_end[j + st - 1] <= p) j += st;
        const int lp = p - (j ? s_end[j - 1] : 0);
        const int pf = s_pf[j];
        int nl = (int)((float)lp * s_rpf[j]);  // lp / pf within one, then exact
        while (nl > 0 && nl * pf > lp) --nl;
        while ((nl + 1) * pf <= lp) ++nl;
        const int m = lp - nl * pf + s_is0[j];
        const int n = nl + s_is1[j];
        double pa[3];
        pa[A::vc] = gmin[A::vc] + delta;
        pa[A::fd] = gf[m] + delta;
        if (pa[A::fd] < s_box[0][j] || pa[A::fd] > s_box[1][j]) continue;
        pa[A::ld] = gl[n] + delta;
        if (pa[A::ld] < s_box[2][j] || pa[A::ld] > s_box[3][j]) continue;
        ++cand;
        if (s_kind[j] == NSGPU_PATCH_SKIN_POINT) {  // EQ_CULLING (:3602-3608)
            const double dx = s_sph[0][j] - pa[A::fd], dy = s_sph[1][j] - pa[A::ld], srad = s_sph[2][j];
            if (dx * dx + dy * dy > srad * srad) continue;
        }
        const double *pd = P.data + s_off[j];
        double t[2];
        if (!ray_quadric(pd + 6, pa, A::vc, t)) continue;
        double ip0[3] = {pa[0], pa[1], pa[2]}, ip1[3] = {pa[0], pa[1], pa[2]};
        ip0[A::vc] += t[0];
        ip1[A::vc] += t[1];
        bool ok0, ok1;
        skin_feasible2(pd + NSGPU_SKIN_HDR, s_cnt[j], ip0, ip1, ok0, ok1);
        if (ok0) push_hit(R, slot, (int64_t)n * R.N + m, t[0] / ray_dir, s_it[j]);
        if (ok1) push_hit(R, slot, (int64_t)n * R.N + m, t[1] / ray_dir, s_it[j]);
    }
    return cand;
}

// grid: ceil(n_skin / kSkinCells) groups (padded to a multiple of 8) times
// nslots waves, one per workgroup: the cells come largest first, and so do
// their waves in every slot
__global__ void __launch_bounds__(64) k_intersect_skin(DevGrid g, DevPatches P, DevRays R, unsigned long long *part) {
    const int first = P.n - P.n_skin;
    // Block b runs on XCD b % 8.  XCD x takes groups x, x + 8, x + 16, ...,
    // each with all its slots in turn, so a group's cells are fetched into one
    // L2 (six times the HBM traffic when every slot of a group ran on another
    // XCD), while the groups still start largest first across the chip (one
    // contiguous range of groups per XCD left the heaviest range on one XCD
    // and doubled the time).  The grid is padded to a multiple of 8 groups.
    const int nw = (P.n_skin + kSkinCells - 1) / kSkinCells;
    const int xcd = (int)(blockIdx.x & 7), r = (int)(blockIdx.x >> 3);
    const int slot = r % R.nslots, group = (r / R.nslots) * 8 + xcd;
    if (group >= nw) return;
    const int64_t wg = (int64_t)slot * P.n + first + group;
    const int phase = slot < R.panels0 ? 0 : 1;
    const int panel = phase == 0 ? slot : slot - R.panels0;
    const double delta = (phase == 0) ? 0. : kDelta - g.hside;
    __shared__ SkinCells cells;
    unsigned int cand;
    if (panel == 0) cand = skin_slot<0>(g, P, R, slot, delta, group, cells);
    else if (panel == 1) cand = skin_slot<1>(g, P, R, slot, delta, group, cells);
    else cand = skin_slot<2>(g, P, R, slot, delta, group, cells);
    for (int off = 32; off > 0; off >>= 1) cand += __shfl_down(cand, off, 64);
    if (threadIdx.x == 0 && cand) part[wg] = cand;
}

void launch_intersect_skin(hipStream_t s, const DevGrid &g, const DevPatches &P, const DevRays &R,
                           unsigned long long *part) {
    if (P.n_skin <= 0 || R.nslots <= 0) return;
    const int64_t groups = ((P.n_skin + kSkinCells - 1) / kSkinCells + 7) / 8 * 8;
    hipLaunchKernelGGL(k_intersect_skin, dim3((unsigned)(groups * R.nslots)), dim3(64), 0, s, g, P, R, part);
}

}  // namespace nsg
