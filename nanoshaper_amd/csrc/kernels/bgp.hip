// Boundary grid points and their projection on the SES: phase 2 of
// Surface::getSurf (Surface.cpp:1476-1588) and Surface::projector
// (:4657-4680) over ConnollySurface's projection (buildAuxiliaryGrid
// ConnollySurface.cpp:3666-3909, getProjection :5822-6106).
//
// Three passes, all over HBM-resident maps:
//   k_bgp_flags    one wave per 64 nodes of a grid row: the node is a boundary
//                  grid point when its six epsmap faces differ; the wave's
//                  ballot is the row word of a bitmask, its popcount the
//                  word's count (scanned on the device into output offsets,
//                  so points keep the reference's z, y, x order)
//   k_bgp_bin      the auxiliary grid: every patch in the cells its bounding
//                  box covers (count, then fill)
//   k_bgp_project  one wave per auxiliary cell: the cell's boundary points are
//                  gathered from the bitmask into LDS, then each lane projects
//                  one point on every patch of the cell's list; the patch is
//                  the same for all lanes, so its kind, data and feasibility
//                  loop are uniform across the wave
#include <hip/hip_runtime.h>

#include "launch.hpp"

namespace nsg {

namespace {

constexpr int kBgpCap = 512;  // LDS entries per gathered batch of one cell

// the six epsmap faces around node n: +x, +y, +z faces of the node, the +x
// face of node - x, +y of node - y, +z of node - z (Surface.cpp:1490-1525)
struct Faces {
    int v[6];
};
__device__ __forceinline__ Faces faces_at(const DevGrid &g, const int *eps, int64_t n) {
    Faces f;
    const int *e = eps + 3 * n;
    f.v[0] = e[0];
    f.v[1] = e[1];
    f.v[2] = e[2];
    f.v[3] = e[-3];
    f.v[4] = eps[3 * (n - g.nx) + 1];
    f.v[5] = eps[3 * (n - (int64_t)g.nx * g.ny) + 2];
    return f;
}

__global__ void __launch_bounds__(256) k_bgp_flags(DevGrid g, const int *eps, int W, unsigned long long *bits,
                                                   int *wcount, int *n_internal) {
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t nwords = (int64_t)g.ny * g.nz * W;
    if (wave >= nwords) return;
    const int64_t row = wave / W;
    const int w = (int)(wave - row * W);
    const int iy = (int)(row % g.ny), iz = (int)(row / g.ny);
    const int ix = w * 64 + lane;
    bool is_bgp = false, internal = false;
    if (ix >= 1 && ix < g.nx && iy >= 1 && iz >= 1) {
        const Faces f = faces_at(g, eps, row * g.nx + ix);
        const int k = f.v[0];
        is_bgp = k != f.v[1] || k != f.v[2] || k != f.v[3] || k != f.v[4] || k != f.v[5];
        bool any_out = false;
        for (int q = 0; q < 6; ++q) any_out |= f.v[q] == 0;
        internal = is_bgp && !any_out;
    }
    const unsigned long long m = __ballot(is_bgp);
    const unsigned long long mi = __ballot(internal);
    if (lane == 0) {
        bits[wave] = m;
        wcount[wave] = __popcll(m);
        if (mi) atomicAdd(n_internal, __popcll(mi));
    }
}

// the auxiliary cells the patch's bounding box covers
// (ConnollySurface.cpp:3847-3874)
__device__ __forceinline__ void aux_box(const BgpArgs &A, const double *c, int lo[3], int hi[3]) {
    const double radius = c[3];
    for (int d = 0; d < 3; ++d) {
        int64_t a = (int64_t)rintp(fmax(0., (c[d] - radius) - A.amin[d]) * A.ascale);
        int64_t b = (int64_t)rintp(fmax(0., (c[d] + radius) - A.amin[d]) * A.ascale);
        if (a >= A.na) a = A.na - 1;
        if (b >= A.na) b = A.na - 1;
        lo[d] = (int)a;
        hi[d] = (int)b;
    }
}

__global__ void __launch_bounds__(256) k_bgp_bin(DevPatches P, BgpArgs A, bool fill) {
    const int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= P.n) return;
    int lo[3], hi[3];
    aux_box(A, P.bsph + 4 * it, lo, hi);
    for (int iz = lo[2]; iz <= hi[2]; ++iz)
        for (int iy = lo[1]; iy <= hi[1]; ++iy)
            for (int ix = lo[0]; ix <= hi[0]; ++ix) {
                const int64_t c = ((int64_t)iz * A.na + iy) * A.na + ix;
                if (!fill) {
                    atomicAdd(A.ccount + c, 1);
                } else {
                    const int at = atomicAdd(A.ccursor + c, 1);
                    A.clist[A.coff[c] + at] = it;
                }
            }
}

// tools.cpp:1112-1120
__device__ __forceinline__ double project_sphere(const double y[3], const double *c, double r, double proj[3]) {
    double d = sqrt((y[0] - c[0]) * (y[0] - c[0]) + (y[1] - c[1]) * (y[1] - c[1]) + (y[2] - c[2]) * (y[2] - c[2]));
    d = r / d;
    for (int q = 0; q < 3; ++q) proj[q] = c[q] + d * (y[q] - c[q]);
    return sqrt((proj[0] - y[0]) * (proj[0] - y[0]) + (proj[1] - y[1]) * (proj[1] - y[1]) +
                (proj[2] - y[2]) * (proj[2] - y[2]));
}

// ConnollySurface.cpp:5693-5702 (point2plane, tools.cpp:277-287)
__device__ __forceinline__ void project_circle(const double pt[3], double radius, const double c[3], const double w[4],
                                               double proj[3]) {
    const double den = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    const double val = (w[0] * pt[0] + w[1] * pt[1] + w[2] * pt[2]) + w[3];
    const double cc = val / den;
    for (int q = 0; q < 3; ++q) proj[q] = (pt[q] - w[q] * cc) - c[q];
    radius /= sqrt(proj[0] * proj[0] + proj[1] * proj[1] + proj[2] * proj[2]);
    for (int q = 0; q < 3; ++q) proj[q] = c[q] + radius * proj[q];
}

// ConnollySurface.cpp:5705-5734 with plane3points (tools.cpp:258-274); the
// torus frame Rot is taken as invrot transposed
__device__ __forceinline__ double project_torus(const double y[3], const double *td, double probe, double proj[3]) {
    const double *iv = td + 3;
    const double t[3] = {y[0] - td[0], y[1] - td[1], y[2] - td[2]};
    double pp[3], pm[3], mp[4];
    for (int i = 0; i < 3; ++i) pp[i] = iv[3 * i] * t[0] + iv[3 * i + 1] * t[1] + iv[3 * i + 2] * t[2];
    const double tplane[4] = {0, 0, +1, 0}, zero[3] = {0, 0, 0};
    project_circle(pp, td[12], zero, tplane, pm);
    const double p2[3] = {0, 0, +1}, p3[3] = {0, 0, -1};
    mp[0] = pm[1] * (p2[2] - p3[2]) + p2[1] * (p3[2] - pm[2]) + p3[1] * (pm[2] - p2[2]);
    mp[1] = pm[2] * (p2[0] - p3[0]) + p2[2] * (p3[0] - pm[0]) + p3[2] * (pm[0] - p2[0]);
    mp[2] = pm[0] * (p2[1] - p3[1]) + p2[0] * (p3[1] - pm[1]) + p3[0] * (pm[1] - p2[1]);
    mp[3] = -(mp[0] * pm[0] + mp[1] * pm[1] + mp[2] * pm[2]);
    const double nrm = sqrt(mp[0] * mp[0] + mp[1] * mp[1] + mp[2] * mp[2]);
    for (int q = 0; q < 4; ++q) mp[q] /= nrm;
    double pr[3];
    project_circle(pp, probe, pm, mp, pr);
    for (int i = 0; i < 3; ++i) proj[i] = (iv[i] * pr[0] + iv[3 + i] * pr[1] + iv[6 + i] * pr[2]) + td[i];
    return sqrt((proj[0] - y[0]) * (proj[0] - y[0]) + (proj[1] - y[1]) * (proj[1] - y[1]) +
                (proj[2] - y[2]) * (proj[2] - y[2]));
}

// ConnollySurface::isFeasible (ConnollySurface.cpp:4769-5021) on the patch
// SoA's normalised tests, for one point
__device__ __forceinline__ bool feasible1(const DevPatches &P, int it, const double p[3]) {
    const int kind = P.kind[it], cnt = P.cnt[it];
    const double *d = P.data + P.off[it];
    if (kind == NSGPU_PATCH_POINT) {
        for (int i = 0; i < cnt; ++i) {
            const double *s = d + 4 * i;
            double dd = (s[0] - p[0]) * (s[0] - p[0]) + (s[1] - p[1]) * (s[1] - p[1]) + (s[2] - p[2]) * (s[2] - p[2]);
            dd -= s[3];
            if (dd < 0) return false;
        }
        return true;
    }
    if (kind == NSGPU_PATCH_FACET_REG || kind == NSGPU_PATCH_FACET_SING) {
        for (int i = 0; i < cnt; ++i) {
            const double *w = d + 4 * i;
            if ((w[0] * p[0] + w[1] * p[1] + w[2] * p[2]) + w[3] > 0) return false;
        }
        return true;
    }
    const double r = d[13];
    if (r >= 0.) {
        const double dd = (d[0] - p[0]) * (d[0] - p[0]) + (d[1] - p[1]) * (d[1] - p[1]) + (d[2] - p[2]) * (d[2] - p[2]);
        if (dd < r * r) return false;
    }
    if (kind == NSGPU_PATCH_TORUS_SING) return true;
    const double *pl = d + NSGPU_TORUS_HDR;
    const int flags = P.flags[it];
    for (int i = 0; i < cnt; ++i, pl += 8) {
        const bool le1 = !((pl[0] * p[0] + pl[1] * p[1] + pl[2] * p[2]) + pl[3] > 0);
        const bool le2 = !((pl[4] * p[0] + pl[5] * p[1] + pl[6] * p[2]) + pl[7] > 0);
        if ((flags >> i) & 1 ? (le1 && le2) : (le1 || le2)) return true;
    }
    return false;
}

__device__ __forceinline__ bool lex_less(double d, int it, int kk, double bd, int bit, int bkk) {
    return d < bd || (d == bd && (it < bit || (it == bit && kk < bkk)));
}

// no patch projection of p is feasible: sample the probe positions along the
// circle of every self-intersecting regular torus of the cell and keep the
// nearest probe-sphere projection that a facet on that edge accepts
// (ConnollySurface.cpp:5916-6086); (distance, patch, sample) decide ties as
// the reference's ascending loops do
__device__ bool sampled_projection(const DevPatches &P, const BgpArgs &A, const int *cl, int nc, double probe,
                                   const double p[3], double proj[3]) {
    double best = INFINITY;
    int bit = 0x7fffffff, bkk = 0x7fffffff;
    const double step = (2.0 * 3.14159265358979323846264338327950288) / 500;
    for (int k = 0; k < nc; ++k) {
        const int it = cl[k];
        if (P.kind[it] != NSGPU_PATCH_TORUS_REG) continue;
        const double *td = P.data + P.off[it];
        if (!(td[13] >= 0.)) continue;
        const int a = P.atom[3 * it], b = P.atom[3 * it + 1];
        const long long key = ((long long)min(a, b) << 32) | (unsigned)max(a, b);
        int lo = 0, hi = A.nf;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (A.fkey[mid] < key) lo = mid + 1;
            else hi = mid;
        }
        if (lo >= A.nf || A.fkey[lo] != key) continue;
        const double *iv = td + 3;
        double tcoi = 0.;
        for (int kk = 0; kk < 500; ++kk) {
            double sp[3], loc[3];
            const double ct = cos(tcoi), st = sin(tcoi);
            for (int d = 0; d < 3; ++d) sp[d] = td[d] + td[12] * (ct * iv[3 + d] + st * iv[d]);
            tcoi += step;
            const double dist = project_sphere(p, sp, probe, loc);
            if (!lex_less(dist, it, kk, best, bit, bkk)) continue;
            for (int f = lo; f < A.nf && A.fkey[f] == key; ++f)
                if (feasible1(P, A.fval[f], loc)) {
                    best = dist;
                    bit = it;
                    bkk = kk;
                    proj[0] = loc[0];
                    proj[1] = loc[1];
                    proj[2] = loc[2];
                    break;
                }
        }
    }
    return best != INFINITY;
}

__global__ void __launch_bounds__(64) k_bgp_project(DevGrid g, DevPatches P, BgpArgs A, double probe) {
    __shared__ int sx[kBgpCap], sy[kBgpCap], sz[kBgpCap], spos[kBgpCap];
    __shared__ int s_total;
    const int lane = threadIdx.x;
    const int64_t cell = blockIdx.x;
    const int ax = (int)(cell % A.na), ay = (int)((cell / A.na) % A.na), az = (int)(cell / ((int64_t)A.na * A.na));
    const int xlo = A.range[2 * ax], xhi = A.range[2 * ax + 1];
    const int ylo = A.range[2 * (A.na + ay)], yhi = A.range[2 * (A.na + ay) + 1];
    const int zlo = A.range[2 * (2 * A.na + az)], zhi = A.range[2 * (2 * A.na + az) + 1];
    if (xhi < xlo || yhi < ylo || zhi < zlo) return;
    const int ny_c = yhi - ylo + 1, nrows = ny_c * (zhi - zlo + 1);
    const int xspan = xhi - xlo + 1;
    const int rows_per_batch = min(64, max(1, kBgpCap / xspan));
    const int nc = A.ccount[cell];
    const int *cl = A.clist + A.coff[cell];
    for (int r0 = 0; r0 < nrows; r0 += rows_per_batch) {
        // gather: lane l < rows_per_batch reads its row's words over [xlo, xhi]
        int cnt = 0;
        unsigned long long wv[2] = {0, 0};
        int iy = 0, iz = 0, w0 = 0;
        int64_t wbase = 0;
        const int r = r0 + lane;
        if (lane < rows_per_batch && r < nrows) {
            iy = ylo + r % ny_c;
            iz = zlo + r / ny_c;
            w0 = xlo >> 6;
            const int w1 = xhi >> 6;
            wbase = ((int64_t)iz * g.ny + iy) * A.W;
            for (int w = w0; w <= w1; ++w) {
                unsigned long long m = A.bits[wbase + w];
                const int b0 = w * 64;
                if (xlo > b0) m &= ~0ull << (xlo - b0);
                if (xhi < b0 + 63) m &= ~0ull >> (63 - (xhi - b0));
                wv[w - w0] = m;
                cnt += __popcll(m);
            }
        }
        // wave exclusive prefix of cnt
        int pre = cnt;
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(pre, o, 64);
            if (lane >= o) pre += v;
        }
        const int total = __shfl(pre, 63, 64);
        pre -= cnt;
        if (cnt) {
            int at = pre;
            for (int q = 0; q < 2; ++q) {
                unsigned long long m = wv[q];
                const int w = w0 + q;
                while (m) {
                    const int b = __ffsll((long long)m) - 1;
                    m &= m - 1;
                    // output position: offset of the word plus the points before
                    // bit b in it
                    const unsigned long long full = A.bits[wbase + w];
                    const int pos = A.woff[wbase + w] + __popcll(b ? (full & (~0ull >> (64 - b))) : 0ull);
                    sx[at] = w * 64 + b;
                    sy[at] = iy;
                    sz[at] = iz;
                    spos[at] = pos;
                    ++at;
                }
            }
        }
        __syncthreads();
        // project the gathered points, one per lane
        for (int b0 = 0; b0 < total; b0 += 64) {
            const int e = b0 + lane;
            const bool active = e < total;
            int ix = 0, jy = 0, kz = 0, pos = 0;
            double p[3] = {0, 0, 0}, proj[3] = {0, 0, 0};
            bool internal = false;
            if (active) {
                ix = sx[e];
                jy = sy[e];
                kz = sz[e];
                pos = spos[e];
                p[0] = g.x[ix];
                p[1] = g.y[jy];
                p[2] = g.z[kz];
                proj[0] = p[0];
                proj[1] = p[1];
                proj[2] = p[2];
                const Faces f = faces_at(g, A.eps, ((int64_t)kz * g.ny + jy) * g.nx + ix);
                bool any_out = false;
                for (int q = 0; q < 6; ++q) any_out |= f.v[q] == 0;
                internal = !any_out;
            }
            double best = INFINITY;
            int bit = 0x7fffffff;
            const bool work = active && !internal;
            for (int k = 0; k < nc; ++k) {
                const int it = cl[k];
                const int kind = P.kind[it];
                double loc[3], dist;
                if (kind == NSGPU_PATCH_TORUS_REG || kind == NSGPU_PATCH_TORUS_SING)
                    dist = project_torus(p, P.data + P.off[it], probe, loc);
                else
                    dist = project_sphere(p, P.bsph + 4 * it, kind == NSGPU_PATCH_POINT ? P.bsph[4 * it + 3] : probe,
                                          loc);
                if (!work || !lex_less(dist, it, 0, best, bit, 0)) continue;
                if (!feasible1(P, it, loc)) continue;
                best = dist;
                bit = it;
                proj[0] = loc[0];
                proj[1] = loc[1];
                proj[2] = loc[2];
            }
            if (work && best == INFINITY) {
                if (nc > 0 && sampled_projection(P, A, cl, nc, probe, p, proj))
                    atomicAdd(A.counters + 0, 1);
                else
                    atomicAdd(A.counters + 1, 1);
            }
            if (active) {
                A.ibgp[3 * (int64_t)pos] = ix;
                A.ibgp[3 * (int64_t)pos + 1] = jy;
                A.ibgp[3 * (int64_t)pos + 2] = kz;
                for (int q = 0; q < 3; ++q) {
                    A.pos[3 * (int64_t)pos + q] = proj[q];
                    A.nor[3 * (int64_t)pos + q] = 0.;
                }
            }
        }
        __syncthreads();
    }
}

}  // namespace

void launch_bgp_flags(hipStream_t s, const DevGrid &g, const int *eps, int W, unsigned long long *bits, int *wcount,
                      int *n_internal) {
    const int64_t waves = (int64_t)g.ny * g.nz * W;
    if (waves <= 0) return;
    hipLaunchKernelGGL(k_bgp_flags, dim3((unsigned)((waves * 64 + 255) / 256)), dim3(256), 0, s, g, eps, W, bits,
                       wcount, n_internal);
}

void launch_bgp_bin(hipStream_t s, const DevPatches &P, const BgpArgs &A, bool fill) {
    const int64_t cells = (int64_t)A.na * A.na * A.na;
    launch_fill32(s, fill ? A.ccursor : A.ccount, 0u, cells);
    if (P.n > 0) hipLaunchKernelGGL(k_bgp_bin, dim3((P.n + 255) / 256), dim3(256), 0, s, P, A, fill);
}

void launch_bgp_project(hipStream_t s, const DevGrid &g, const DevPatches &P, const BgpArgs &A, double probe) {
    const int64_t cells = (int64_t)A.na * A.na * A.na;
    if (cells <= 0) return;
    hipLaunchKernelGGL(k_bgp_project, dim3((unsigned)cells), dim3(64), 0, s, g, P, A, probe);
}

}  // namespace nsg
