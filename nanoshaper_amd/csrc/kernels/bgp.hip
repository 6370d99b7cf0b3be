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
//   k_bgp_project  one wave per auxiliary cell that holds boundary points
//                  (flagged by k_bgp_flags, listed by k_bgp_cells): the
//                  cell's points are gathered from the bitmask into LDS, then
//                  each lane projects one point on every patch of the cell's
//                  list; the patch is the same for all lanes, so its kind,
//                  data and feasibility loop are uniform across the wave
//   k_bgp_fallback one workgroup per point with no feasible patch projection,
//                  sampling probe positions along self-intersecting tori
#include <hip/hip_runtime.h>

#include "launch.hpp"

namespace nsg {

namespace {

constexpr int kBgpCap = 256;  // LDS entries per gathered batch of one cell
constexpr int kFallbackThreads = 512;  // one thread per sample angle (kCoiSamples)
static_assert(kFallbackThreads >= kCoiSamples, "a sample per thread");

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

__global__ void __launch_bounds__(256) k_bgp_flags(DevGrid g, BgpArgs A, unsigned long long *bits, int *wcount,
                                                   int *n_internal) {
    const int *eps = A.eps;
    const int W = A.W;
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
    // the auxiliary cell of a boundary point has work for k_bgp_project
    if (is_bgp) A.cflag[((int64_t)A.node_aux[2 * g.N + iz] * A.na + A.node_aux[g.N + iy]) * A.na + A.node_aux[ix]] = 1;
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
// SoA's normalised tests, for one point; d = the patch's data
__device__ __forceinline__ bool feasible_hdr(int kind, int cnt, int flags, const double *d, const double p[3]) {
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
    for (int i = 0; i < cnt; ++i, pl += 8) {
        const bool le1 = !((pl[0] * p[0] + pl[1] * p[1] + pl[2] * p[2]) + pl[3] > 0);
        const bool le2 = !((pl[4] * p[0] + pl[5] * p[1] + pl[6] * p[2]) + pl[7] > 0);
        if ((flags >> i) & 1 ? (le1 && le2) : (le1 || le2)) return true;
    }
    return false;
}
__device__ __forceinline__ bool feasible1(const DevPatches &P, int it, const double p[3]) {
    return feasible_hdr(P.kind[it], P.cnt[it], P.flags[it], P.data + P.off[it], p);
}

__device__ __forceinline__ bool lex_less(double d, int it, int kk, double bd, int bit, int bkk) {
    return d < bd || (d == bd && (it < bit || (it == bit && kk < bkk)));
}

// No patch projection of p is feasible: sample the probe positions along the
// circle of every self-intersecting regular torus of the cell and keep the
// nearest probe-sphere projection that a facet on that edge accepts
// (ConnollySurface.cpp:5916-6086).  The whole wave searches for one point:
// lane l takes samples l, l + 64, ... of each torus, whose cos and sin come
// from the host's table (the reference's accumulated angle, glibc's cos and
// sin); the lanes' bests are reduced in (distance, patch, sample) order,
// which decides ties as the reference's ascending loops do.  Called by the
// whole wave with p uniform; returns whether a sample was kept.
__device__ bool sampled_projection_block(const DevPatches &P, const BgpArgs &A, const int *cl, int nc, double probe,
                                         const double p[3], double proj[3]) {
    const int lane = threadIdx.x & 63, tid = threadIdx.x, nt = blockDim.x;
    double best = INFINITY;
    int bit = 0x7fffffff, bkk = 0x7fffffff;
    double bp[3] = {0, 0, 0};
    // the cell's eligible patches (self-intersecting regular tori with facets
    // on their edge), found by the lanes in parallel, 64 patches at a time
    for (int k0 = 0; k0 < nc; k0 += 64) {
        bool elig = false;
        int my_it = 0;
        if (k0 + lane < nc) {
            my_it = cl[k0 + lane];
            elig = P.kind[my_it] == NSGPU_PATCH_TORUS_REG && P.data[P.off[my_it] + 13] >= 0. &&
                   A.efac[2 * my_it + 1] > 0;
        }
        unsigned long long em = __ballot(elig);
        while (em) {
            const int L = __ffsll((long long)em) - 1;
            em &= em - 1;
            const int it = __shfl(my_it, L, 64);
            const double *td = P.data + P.off[it];
            // the edge's facets (usually two): entries [lo, lo + n_all) of
            // fval, the headers of the first kMaxEdgeFacets loaded once
            const int lo = A.efac[2 * it], n_all = A.efac[2 * it + 1];
            constexpr int kMaxEdgeFacets = 8;
            int fkind[kMaxEdgeFacets], fcnt[kMaxEdgeFacets];
            const double *fdat[kMaxEdgeFacets];
            const int nfc = min(n_all, kMaxEdgeFacets);
            for (int f = 0; f < nfc; ++f) {
                const int fi = A.fval[lo + f];
                fkind[f] = P.kind[fi];
                fcnt[f] = P.cnt[fi];
                fdat[f] = P.data + P.off[fi];
            }
            const double *iv = td + 3;
            for (int kk = tid; kk < kCoiSamples; kk += nt) {
                double sp[3], loc[3];
                const double ct = A.coi[2 * kk], st = A.coi[2 * kk + 1];
                for (int d = 0; d < 3; ++d) sp[d] = td[d] + td[12] * (ct * iv[3 + d] + st * iv[d]);
                const double dist = project_sphere(p, sp, probe, loc);
                if (!lex_less(dist, it, kk, best, bit, bkk)) continue;
                bool ok = false;
                for (int f = 0; f < nfc && !ok; ++f) ok = feasible_hdr(fkind[f], fcnt[f], 0, fdat[f], loc);
                // facets past the first kMaxEdgeFacets (a degenerate edge)
                for (int f = kMaxEdgeFacets; !ok && f < n_all; ++f) ok = feasible1(P, A.fval[lo + f], loc);
                if (ok) {
                    best = dist;
                    bit = it;
                    bkk = kk;
                    bp[0] = loc[0];
                    bp[1] = loc[1];
                    bp[2] = loc[2];
                }
            }
        }
    }
    // wave reduction in (distance, patch, sample) order
    for (int o = 32; o >= 1; o >>= 1) {
        const double ob = __shfl_xor(best, o, 64);
        const int oit = __shfl_xor(bit, o, 64), okk = __shfl_xor(bkk, o, 64);
        const double o0 = __shfl_xor(bp[0], o, 64), o1 = __shfl_xor(bp[1], o, 64), o2 = __shfl_xor(bp[2], o, 64);
        if (lex_less(ob, oit, okk, best, bit, bkk)) {
            best = ob;
            bit = oit;
            bkk = okk;
            bp[0] = o0;
            bp[1] = o1;
            bp[2] = o2;
        }
    }
    // then across the block's waves
    __shared__ double r_best[kFallbackThreads / 64], r_p[kFallbackThreads / 64][3];
    __shared__ int r_it[kFallbackThreads / 64], r_kk[kFallbackThreads / 64];
    if (lane == 0) {
        r_best[tid >> 6] = best;
        r_it[tid >> 6] = bit;
        r_kk[tid >> 6] = bkk;
        for (int d = 0; d < 3; ++d) r_p[tid >> 6][d] = bp[d];
    }
    __syncthreads();
    best = r_best[0];
    bit = r_it[0];
    bkk = r_kk[0];
    for (int d = 0; d < 3; ++d) bp[d] = r_p[0][d];
    for (int w = 1; w < nt / 64; ++w)
        if (lex_less(r_best[w], r_it[w], r_kk[w], best, bit, bkk)) {
            best = r_best[w];
            bit = r_it[w];
            bkk = r_kk[w];
            for (int d = 0; d < 3; ++d) bp[d] = r_p[w][d];
        }
    __syncthreads();
    if (best == INFINITY) return false;
    proj[0] = bp[0];
    proj[1] = bp[1];
    proj[2] = bp[2];
    return true;
}

__global__ void __launch_bounds__(64) k_bgp_project(DevGrid g, DevPatches P, BgpArgs A, double probe) {
    // gathered points: node offsets within the cell's ranges (6 bits each,
    // a cell spans at most 64 nodes per axis) and output positions
    __shared__ int sloc[kBgpCap], spos[kBgpCap];
    __shared__ int h_it[64], h_kind[64], h_cnt[64], h_flags[64], h_off[64];
    __shared__ double h_bs[64][4];
    __shared__ double h_td[64][NSGPU_TORUS_HDR];
    __shared__ int s_total;
    const int lane = threadIdx.x;
    const int64_t cell = A.active[blockIdx.x];
    const int ax = (int)(cell % A.na), ay = (int)((cell / A.na) % A.na), az = (int)(cell / ((int64_t)A.na * A.na));
    const int xlo = A.range[2 * ax], xhi = A.range[2 * ax + 1];
    const int ylo = A.range[2 * (A.na + ay)], yhi = A.range[2 * (A.na + ay) + 1];
    const int zlo = A.range[2 * (2 * A.na + az)], zhi = A.range[2 * (2 * A.na + az) + 1];
    if (xhi < xlo || yhi < ylo || zhi < zlo) return;
    const int ny_c = yhi - ylo + 1, nrows = ny_c * (zhi - zlo + 1);
    const int nc = A.ccount[cell];
    const int *cl = A.clist + A.coff[cell];
    for (int r0 = 0; r0 < nrows;) {
        // gather: lane l reads row r0 + l's words over [xlo, xhi]; the batch
        // takes the leading rows whose points fit the LDS lists (a row holds
        // at most 64 points, so at least kBgpCap / 64 rows are taken)
        int cnt = 0;
        unsigned long long wv[2] = {0, 0};
        int iy = 0, iz = 0, w0 = 0;
        int64_t wbase = 0;
        const int r = r0 + lane;
        if (r < nrows) {
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
        // wave inclusive prefix of cnt; rows up to the last lane whose prefix
        // fits are taken
        int pre = cnt;
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(pre, o, 64);
            if (lane >= o) pre += v;
        }
        const unsigned long long fits = __ballot(r < nrows && pre <= kBgpCap);
        const int take = __popcll(fits);  // fits is a prefix of the lanes
        const int total = take ? __shfl(pre, take - 1, 64) : 0;
        pre -= cnt;
        r0 += take;
        if (cnt && lane < take) {
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
                    sloc[at] = (w * 64 + b - xlo) | ((iy - ylo) << 6) | ((iz - zlo) << 12);
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
                const int lc = sloc[e];
                ix = xlo + (lc & 63);
                jy = ylo + ((lc >> 6) & 63);
                kz = zlo + (lc >> 12);
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
            // the cell's patches, 64 at a time: their headers staged in LDS by
            // one load per lane, then every lane tests its point on each
            for (int c0 = 0; c0 < nc; c0 += 64) {
                const int nk = min(64, nc - c0);
                __syncthreads();
                if (lane < nk) {
                    const int it = cl[c0 + lane];
                    h_it[lane] = it;
                    h_kind[lane] = P.kind[it];
                    h_cnt[lane] = P.cnt[it];
                    h_flags[lane] = P.flags[it];
                    h_off[lane] = P.off[it];
                    for (int q = 0; q < 4; ++q) h_bs[lane][q] = P.bsph[4 * it + q];
                    // a torus's centre, frame, radii: read by every lane
                    if (h_kind[lane] == NSGPU_PATCH_TORUS_REG || h_kind[lane] == NSGPU_PATCH_TORUS_SING) {
                        const double *td = P.data + h_off[lane];
                        for (int q = 0; q < NSGPU_TORUS_HDR; ++q) h_td[lane][q] = td[q];
                    }
                }
                __syncthreads();
                for (int k = 0; k < nk; ++k) {
                    const int it = h_it[k], kind = h_kind[k];
                    const double *d = P.data + h_off[k];
                    double loc[3], dist;
                    if (kind == NSGPU_PATCH_TORUS_REG || kind == NSGPU_PATCH_TORUS_SING)
                        dist = project_torus(p, h_td[k], probe, loc);
                    else
                        dist = project_sphere(p, h_bs[k], kind == NSGPU_PATCH_POINT ? h_bs[k][3] : probe, loc);
                    if (!work || !lex_less(dist, it, 0, best, bit, 0)) continue;
                    if (!feasible_hdr(kind, h_cnt[k], h_flags[k], d, loc)) continue;
                    best = dist;
                    bit = it;
                    proj[0] = loc[0];
                    proj[1] = loc[1];
                    proj[2] = loc[2];
                }
            }
            // points with no feasible patch projection: queued for
            // k_bgp_fallback, which gives each its own wave
            if (work && best == INFINITY) {
                if (nc > 0) {
                    const int q = atomicAdd(A.counters + 3, 1);
                    A.fq[2 * q] = pos;
                    A.fq[2 * q + 1] = (int)cell;
                } else {
                    atomicAdd(A.counters + 1, 1);
                }
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

// one workgroup per queued point (grid-stride over the queue): the sampled
// projection of the point's cell
__global__ void __launch_bounds__(kFallbackThreads) k_bgp_fallback(DevGrid g, DevPatches P, BgpArgs A, double probe) {
    const int nq = A.counters[3];
    for (int q = blockIdx.x; q < nq; q += gridDim.x) {
        const int pos = A.fq[2 * q];
        const int64_t cell = A.fq[2 * q + 1];
        const int *ib = A.ibgp + 3 * (int64_t)pos;
        const double p[3] = {g.x[ib[0]], g.y[ib[1]], g.z[ib[2]]};
        double fp[3];
        const bool kept = sampled_projection_block(P, A, A.clist + A.coff[cell], A.ccount[cell], probe, p, fp);
        if (threadIdx.x == 0) {
            if (kept)
                for (int d = 0; d < 3; ++d) A.pos[3 * (int64_t)pos + d] = fp[d];
            atomicAdd(A.counters + (kept ? 0 : 1), 1);
        }
    }
}

// list of the auxiliary cells flagged by k_bgp_flags
__global__ void __launch_bounds__(256) k_bgp_cells(BgpArgs A, int64_t ncell) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncell && A.cflag[c]) A.active[A.coff_active[c]] = (int)c;
}

}  // namespace

void launch_bgp_flags(hipStream_t s, const DevGrid &g, const BgpArgs &A, unsigned long long *bits, int *wcount,
                      int *n_internal) {
    const int64_t waves = (int64_t)g.ny * g.nz * A.W;
    if (waves <= 0) return;
    hipLaunchKernelGGL(k_bgp_flags, dim3((unsigned)((waves * 64 + 255) / 256)), dim3(256), 0, s, g, A, bits, wcount,
                       n_internal);
}

void launch_bgp_cells(hipStream_t s, const BgpArgs &A) {
    const int64_t cells = (int64_t)A.na * A.na * A.na;
    hipLaunchKernelGGL(k_bgp_cells, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s, A, cells);
}

void launch_bgp_bin(hipStream_t s, const DevPatches &P, const BgpArgs &A, bool fill) {
    const int64_t cells = (int64_t)A.na * A.na * A.na;
    launch_fill32(s, fill ? A.ccursor : A.ccount, 0u, cells);
    if (P.n > 0) hipLaunchKernelGGL(k_bgp_bin, dim3((P.n + 255) / 256), dim3(256), 0, s, P, A, fill);
}

void launch_bgp_project(hipStream_t s, const DevGrid &g, const DevPatches &P, const BgpArgs &A, double probe,
                        int64_t n_active) {
    if (n_active <= 0) return;
    hipLaunchKernelGGL(k_bgp_project, dim3((unsigned)n_active), dim3(64), 0, s, g, P, A, probe);
    hipLaunchKernelGGL(k_bgp_fallback, dim3(1024), dim3(kFallbackThreads), 0, s, g, P, A, probe);
}

}  // namespace nsg
