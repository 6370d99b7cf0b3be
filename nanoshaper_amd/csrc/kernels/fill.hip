// Map fill from the sorted ray lists, as a gather: one lane per grid node
// reads the three rays through it and writes epsmap (3 faces), idebmap and
// status once, so every map byte is written exactly once and coalesced.  The
// decisions are those of setVerticesAndGridsWithIntersectionData
// (Surface.cpp:5101-6250) with NEW_INTERSECTION_FILTERING: a ray with an even
// number of hits pairs (2q, 2q+1) when the hits are >= EPS_INT apart, a ray
// with an odd count is left at background (the failed-ray copy of the
// reference is unreachable under that filter).
#include "launch.hpp"

namespace nsg {

// Does one segment of a grid ray cover node 'i' along the ray axis?
// panel 0 (x rays): eps face if i in [i1,i2] and x_i <= lim2-hside, inside if x_i in [lim1,lim2]
// panels 1,2: eps face if i in [i1,i2) and c_i <= lim2-hside (Surface.cpp:5506-5520, :5643-5650, :5743-5750)
template <bool kClosedEnd, bool kStatus>
__device__ __forceinline__ void grid_ray(const DevRays &R, int64_t ray, const double *ga, int na, double gmin,
                                         double scale, double hside, int i, bool &eps, bool &in) {
    const int c = R.cnt[ray];
    if (c == 0 || (c & 1) || c > R.K) return;
    const double *T = R.t + ray * R.K;
    const double pa = ga[0];
    const double dir = ga[na - 1] - pa;
    const double ci = ga[i];
    for (int q = 0; q < c; q += 2) {
        const double t0 = T[q], t1 = T[q + 1];
        if (!(fabs(t1 - t0) >= kEpsInt)) continue;
        const double lim1 = pa + dir * t0;
        const double lim2 = pa + dir * t1;
        int i1 = (int)rintp((lim1 - gmin) * scale);
        int i2 = (int)rintp((lim2 - gmin) * scale);
        i1 = i1 < 0 ? 0 : (i1 >= na ? na - 1 : i1);
        i2 = i2 < 0 ? 0 : (i2 >= na ? na - 1 : i2);
        if (i2 < i1) continue;
        const bool span = kClosedEnd ? (i >= i1 && i <= i2) : (i >= i1 && i < i2);
        if (!span) continue;
        if (ci <= lim2 - hside) eps = true;
        if (kStatus && ci <= lim2 && ci >= lim1) in = true;
    }
}

// Edge ray: node i is inside when some segment has xa < i <= xb (Surface.cpp:5999-6008).
__device__ __forceinline__ bool edge_ray(const DevRays &R, int64_t ray, const double *ga, int na, double gmin,
                                         double scale, double hside, int i) {
    const int c = R.cnt[ray];
    if (c == 0 || (c & 1) || c > R.K) return false;
    const double *T = R.t + ray * R.K;
    const double delta = kDelta - hside;
    const double pa = ga[0] + delta;
    const double dir = (ga[na - 1] + delta) - pa;
    for (int q = 0; q < c; q += 2) {
        const double t0 = T[q], t1 = T[q + 1];
        if (!(fabs(t1 - t0) >= kEpsInt)) continue;
        const double lim1 = pa + dir * t0;
        const double lim2 = pa + dir * t1;
        const int a = (int)rintp((lim1 - gmin) * scale);
        const int b = (int)rintp((lim2 - gmin) * scale);
        if (b <= a) continue;
        if (i > a && i <= b) return true;
    }
    return false;
}

__global__ void __launch_bounds__(256) k_fill_maps(DevGrid g, DevRays R, FillOpts o, int *epsmap,
                                                   uint8_t *idebmap, int *status) {
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nvox) return;
    const int i = (int)(f % g.nx);
    const int j = (int)((f / g.nx) % g.ny);
    const int k = (int)(f / ((int64_t)g.nx * g.ny));
    const int64_t NN = (int64_t)R.N * R.N;
    bool e0 = false, e1 = false, e2 = false, in = false;
    if (o.slot_x >= 0) {
        const int64_t ray = o.slot_x * NN + (int64_t)k * R.N + j;
        grid_ray<true, true>(R, ray, g.x, g.nx, g.xmin, g.scale, g.hside, i, e0, in);
    }
    if (o.build_eps) {
        bool dummy = false;
        if (o.slot_z >= 0)
            grid_ray<false, false>(R, o.slot_z * NN + (int64_t)j * R.N + i, g.z, g.nz, g.zmin, g.scale, g.hside, k,
                                   e2, dummy);
        if (o.slot_y >= 0)
            grid_ray<false, false>(R, o.slot_y * NN + (int64_t)k * R.N + i, g.y, g.ny, g.ymin, g.scale, g.hside, j,
                                   e1, dummy);
        int *e = epsmap + 3 * f;
        e[0] = e0 ? o.inside : 0;
        e[1] = e1 ? o.inside : 0;
        e[2] = e2 ? o.inside : 0;
        idebmap[f] = in ? 0 : 1;
    }
    if (o.build_status) status[f] = in ? NSGPU_STATUS_INSIDE : NSGPU_STATUS_TEMP_OUT;
}

// Corner bitmap (bit = 1 outside), 32 nodes per word in flat order; each wave
// packs its 64 decisions with a ballot.
__global__ void __launch_bounds__(256) k_fill_corners(DevGrid g, DevRays R, FillOpts o, uint32_t *bits) {
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool inside = false;
    if (f < nvox) {
        const int i = (int)(f % g.nx);
        const int j = (int)((f / g.nx) % g.ny);
        const int k = (int)(f / ((int64_t)g.nx * g.ny));
        const int64_t NN = (int64_t)R.N * R.N;
        inside = edge_ray(R, o.eslot_x * NN + (int64_t)k * R.N + j, g.x, g.nx, g.xmin, g.scale, g.hside, i) ||
                 edge_ray(R, o.eslot_z * NN + (int64_t)j * R.N + i, g.z, g.nz, g.zmin, g.scale, g.hside, k) ||
                 edge_ray(R, o.eslot_y * NN + (int64_t)k * R.N + i, g.y, g.ny, g.ymin, g.scale, g.hside, j);
    }
    // nodes past the end of the grid read as outside, as the initial fill
    const unsigned long long outside = __ballot(f >= nvox || !inside);
    const int lane = threadIdx.x & 63;
    const int64_t w0 = f >> 5;  // f is a multiple of 64 at lane 0
    const int64_t nwords = (nvox + 31) >> 5;
    if (lane == 0 && w0 < nwords) bits[w0] = (uint32_t)(outside & 0xffffffffull);
    if (lane == 0 && w0 + 1 < nwords) bits[w0 + 1] = (uint32_t)(outside >> 32);
}

// ------------------------------------------------------------- tiled fill
// A workgroup owns a 64 x 8 x 8 (i, j, k) tile of nodes and stages the interval
// records of the rays through the tile in LDS once: 8 x 64 z rays (i, j),
// 8 x 64 y rays (i, k) and 8 x 8 x rays (j, k).  Each wave then walks (j, k)
// pairs; its 64 lanes take 64 consecutive i, test integer intervals and write
// epsmap, idebmap and status for those nodes once, coalesced.
constexpr int TI = 64, TJ = 8, TK = 8;

// word q of a record at rec[q * stride]
template <int kStride>
__device__ __forceinline__ bool rec_has(const int *rec, int i) {
    const int n = rec[0];
    for (int q = 0; q < n; ++q) {
        const unsigned int w = (unsigned int)rec[(1 + q) * kStride];
        if (i >= (int)(w & 0xffffu) && i <= (int)(w >> 16)) return true;
    }
    return false;
}

// LDS copy of the records of the rays (outer0+o)*N + inner0+i of one slot
// (zeros outside the grid), read with two 16-byte loads per record.
// kTransposed: word q of ray (o, i) at dst[(o*8 + q)*ninner + i], so that 64
// lanes reading 64 consecutive rays hit 64 consecutive words; otherwise
// dst[(o*ninner + i)*8 + q] (records read wave-uniformly).
template <bool kTransposed>
__device__ __forceinline__ void stage_records(int *dst, const int *src, int slot, int64_t NN, int N, int outer0,
                                              int nouter, int inner0, int ninner, int outer_lim, int inner_lim) {
    static_assert(kSegWords == 8, "records are two int4");
    const int total = nouter * ninner * 2;
    const int4 *s4 = reinterpret_cast<const int4 *>(src);
    for (int w = threadIdx.x; w < total; w += blockDim.x) {
        int o, ii, h;
        if (kTransposed) {  // consecutive threads take consecutive rays
            o = w / (2 * ninner);
            const int rem = w - o * 2 * ninner;
            h = rem / ninner;
            ii = rem - h * ninner;
        } else {
            o = w / (ninner * 2);
            const int rem = w - o * (ninner * 2);
            ii = rem >> 1;
            h = rem & 1;
        }
        const int oo = outer0 + o, ix = inner0 + ii;
        int4 v = make_int4(0, 0, 0, 0);
        if (slot >= 0 && oo < outer_lim && ix < inner_lim) v = s4[((int64_t)slot * NN + (int64_t)oo * N + ix) * 2 + h];
        if (kTransposed) {
            int *d = dst + (o * 8 + 4 * h) * ninner + ii;
            d[0] = v.x;
            d[ninner] = v.y;
            d[2 * ninner] = v.z;
            d[3 * ninner] = v.w;
        } else {
            reinterpret_cast<int4 *>(dst)[(o * ninner + ii) * 2 + h] = v;
        }
    }
}

__global__ void __launch_bounds__(256) k_fill_tiled(DevGrid g, DevRays R, FillOpts o, int *epsmap, uint8_t *idebmap,
                                                    int *status) {
    __shared__ __align__(16) int zrec[TJ * TI * kSegWords];
    __shared__ __align__(16) int yrec[TK * TI * kSegWords];
    __shared__ __align__(16) int xrec[TK * TJ * kSegWords];
    __shared__ __align__(16) int xst[TK * TJ * kSegWords];
    const int i0 = blockIdx.x * TI, j0 = blockIdx.y * TJ, k0 = blockIdx.z * TK;
    const int64_t NN = (int64_t)R.N * R.N;
    stage_records<true>(zrec, R.seg, o.build_eps ? o.slot_z : -1, NN, R.N, j0, TJ, i0, TI, g.ny, g.nx);
    stage_records<true>(yrec, R.seg, o.build_eps ? o.slot_y : -1, NN, R.N, k0, TK, i0, TI, g.nz, g.nx);
    stage_records<false>(xrec, R.seg, o.slot_x, NN, R.N, k0, TK, j0, TJ, g.nz, g.ny);
    stage_records<false>(xst, R.seg_st, o.slot_x >= 0 ? 0 : -1, 0, R.N, k0, TK, j0, TJ, g.nz, g.ny);
    __syncthreads();
    const int ii = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = i0 + ii;
    if (i >= g.nx) return;
    for (int p = wv; p < TJ * TK; p += 4) {
        const int jj = p % TJ, kk = p / TJ;
        const int j = j0 + jj, k = k0 + kk;
        if (j >= g.ny || k >= g.nz) continue;
        bool e0 = false, e1 = false, e2 = false, in = false;
        if (o.slot_x >= 0) {
            const int *xe = xrec + (kk * TJ + jj) * kSegWords, *xs = xst + (kk * TJ + jj) * kSegWords;
            if (xe[0] >= 0 && xs[0] >= 0) {
                e0 = rec_has<1>(xe, i);
                in = rec_has<1>(xs, i);
            } else {
                grid_ray<true, true>(R, o.slot_x * NN + (int64_t)k * R.N + j, g.x, g.nx, g.xmin, g.scale, g.hside, i,
                                     e0, in);
            }
        }
        const int64_t f = ((int64_t)k * g.ny + j) * g.nx + i;
        if (o.build_eps) {
            bool dummy = false;
            if (o.slot_z >= 0) {
                const int *zr = zrec + jj * 8 * TI + ii;
                if (zr[0] >= 0) e2 = rec_has<TI>(zr, k);
                else grid_ray<false, false>(R, o.slot_z * NN + (int64_t)j * R.N + i, g.z, g.nz, g.zmin, g.scale,
                                            g.hside, k, e2, dummy);
            }
            if (o.slot_y >= 0) {
                const int *yr = yrec + kk * 8 * TI + ii;
                if (yr[0] >= 0) e1 = rec_has<TI>(yr, j);
                else grid_ray<false, false>(R, o.slot_y * NN + (int64_t)k * R.N + i, g.y, g.ny, g.ymin, g.scale,
                                            g.hside, j, e1, dummy);
            }
            int *e = epsmap + 3 * f;
            e[0] = e0 ? o.inside : 0;
            e[1] = e1 ? o.inside : 0;
            e[2] = e2 ? o.inside : 0;
            idebmap[f] = in ? 0 : 1;
        }
        if (o.build_status) status[f] = in ? NSGPU_STATUS_INSIDE : NSGPU_STATUS_TEMP_OUT;
    }
}

// Corner bitmap from the edge rays, same tiling; bits start at 1 (outside)
// and each 64-node row segment clears its inside bits with at most three
// atomicAnd on the words it overlaps.
__global__ void __launch_bounds__(256) k_corners_tiled(DevGrid g, DevRays R, FillOpts o, uint32_t *bits) {
    __shared__ __align__(16) int zrec[TJ * TI * kSegWords];
    __shared__ __align__(16) int yrec[TK * TI * kSegWords];
    __shared__ __align__(16) int xrec[TK * TJ * kSegWords];
    const int i0 = blockIdx.x * TI, j0 = blockIdx.y * TJ, k0 = blockIdx.z * TK;
    const int64_t NN = (int64_t)R.N * R.N;
    stage_records<true>(zrec, R.seg, o.eslot_z, NN, R.N, j0, TJ, i0, TI, g.ny, g.nx);
    stage_records<true>(yrec, R.seg, o.eslot_y, NN, R.N, k0, TK, i0, TI, g.nz, g.nx);
    stage_records<false>(xrec, R.seg, o.eslot_x, NN, R.N, k0, TK, j0, TJ, g.nz, g.ny);
    __syncthreads();
    const int ii = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = i0 + ii;
    for (int p = wv; p < TJ * TK; p += 4) {
        const int jj = p % TJ, kk = p / TJ;
        const int j = j0 + jj, k = k0 + kk;
        if (j >= g.ny || k >= g.nz) continue;
        bool inside = false;
        if (i < g.nx) {
            const int *xr = xrec + (kk * TJ + jj) * kSegWords;
            const int *zr = zrec + jj * 8 * TI + ii;
            const int *yr = yrec + kk * 8 * TI + ii;
            inside = (xr[0] >= 0 ? rec_has<1>(xr, i)
                                 : edge_ray(R, o.eslot_x * NN + (int64_t)k * R.N + j, g.x, g.nx, g.xmin, g.scale,
                                            g.hside, i)) ||
                     (zr[0] >= 0 ? rec_has<TI>(zr, k)
                                 : edge_ray(R, o.eslot_z * NN + (int64_t)j * R.N + i, g.z, g.nz, g.zmin, g.scale,
                                            g.hside, k)) ||
                     (yr[0] >= 0 ? rec_has<TI>(yr, j)
                                 : edge_ray(R, o.eslot_y * NN + (int64_t)k * R.N + i, g.y, g.ny, g.ymin, g.scale,
                                            g.hside, j));
        }
        const unsigned long long m = __ballot(inside);
        if (ii == 0 && m) {
            const int64_t f0 = ((int64_t)k * g.ny + j) * g.nx + i0;
            const int64_t w0 = f0 >> 5;
            const unsigned int s = (unsigned int)(f0 & 31);
            const uint32_t a = (uint32_t)(m << s);
            const uint32_t b = (uint32_t)((s ? (m >> (32 - s)) : (m >> 32)) & 0xffffffffull);
            const uint32_t c = s ? (uint32_t)(m >> (64 - s)) : 0u;
            if (a) atomicAnd(bits + w0, ~a);
            if (b) atomicAnd(bits + w0 + 1, ~b);
            if (c) atomicAnd(bits + w0 + 2, ~c);
        }
    }
}

void launch_fill_tiled(hipStream_t s, const DevGrid &g, const DevRays &R, const FillOpts &o, int *eps, uint8_t *ideb,
                       int *status) {
    dim3 grid((unsigned)((g.nx + TI - 1) / TI), (unsigned)((g.ny + TJ - 1) / TJ), (unsigned)((g.nz + TK - 1) / TK));
    hipLaunchKernelGGL(k_fill_tiled, grid, dim3(256), 0, s, g, R, o, eps, ideb, status);
}

void launch_corners_tiled(hipStream_t s, const DevGrid &g, const DevRays &R, const FillOpts &o, uint32_t *bits) {
    dim3 grid((unsigned)((g.nx + TI - 1) / TI), (unsigned)((g.ny + TJ - 1) / TJ), (unsigned)((g.nz + TK - 1) / TK));
    hipLaunchKernelGGL(k_corners_tiled, grid, dim3(256), 0, s, g, R, o, bits);
}

// Deterministic two-level sum of the per-ray segment lengths of each panel.
__global__ void __launch_bounds__(256) k_sum_blocks(const double *v, int64_t n, double *partial) {
    __shared__ double sh[256];
    double acc = 0;
    const int64_t base = (int64_t)blockIdx.x * 4096;
    for (int q = 0; q < 16; ++q) {
        const int64_t idx = base + q * 256 + threadIdx.x;
        if (idx < n) acc += v[idx];
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

void launch_fill_maps(hipStream_t s, const DevGrid &g, const DevRays &R, const FillOpts &o, int *eps, uint8_t *ideb,
                      int *status) {
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_fill_maps, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, s, g, R, o, eps, ideb,
                       status);
}

void launch_fill_corners(hipStream_t s, const DevGrid &g, const DevRays &R, const FillOpts &o, uint32_t *bits) {
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_fill_corners, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, s, g, R, o, bits);
}

int64_t sum_partials_count(int64_t n) { return (n + 4095) / 4096; }

void launch_sum_blocks(hipStream_t s, const double *v, int64_t n, double *partial) {
    const int64_t nb = sum_partials_count(n);
    if (nb > 0) hipLaunchKernelGGL(k_sum_blocks, dim3((unsigned)nb), dim3(256), 0, s, v, n, partial);
}

}  // namespace nsg
