// Map fill from the coverage bit planes of the sorted rays, as a gather: one
// lane per grid node reads the bits of the three rays through it and writes
// epsmap (3 faces), idebmap and status once, so every map byte is written
// exactly once and coalesced.  The per-node decisions were made by the sort
// with the rules of setVerticesAndGridsWithIntersectionData (Surface.cpp:
// 5101-6250) under NEW_INTERSECTION_FILTERING: a ray with an even number of
// hits pairs (2q, 2q+1) when the hits are >= EPS_INT apart, a ray with an odd
// count is left at background (the failed-ray copy of the reference is
// unreachable under that filter).
#include "launch.hpp"

namespace nsg {

// coverage bit of node n on ray `ray` of one slot's plane
__device__ __forceinline__ bool cov_bit(const unsigned long long *plane, int64_t NN, int64_t ray, int n) {
    return (plane[(int64_t)(n >> 6) * NN + ray] >> (n & 63)) & 1ull;
}

// A lane takes kFillNodes consecutive nodes of the flat order, so its epsmap
// faces (12 bytes a node), idebmap bytes and status words are aligned 16-byte
// stores; the x-ray words are shared by the lanes of a row and the z- and
// y-ray words of consecutive i are consecutive words of their planes.
constexpr int kFillNodes = 4;  // 8 was slower (0.25 vs 0.22 ms at 325^3)

// The maps are written once and not read again by this kernel: nontemporal
// stores (streamed past the caches' write allocation) took the fill from 236
// to 127 us at 335^3 (639 MB, 5.0 TB/s).
typedef int v4i __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void store_nt(int *p, int a, int b, int c, int d) {
    const v4i v = {a, b, c, d};
    __builtin_nontemporal_store(v, reinterpret_cast<v4i *>(p));
}

__global__ void __launch_bounds__(256) k_fill_bits(DevGrid g, DevRays R, FillOpts o, int *epsmap, uint8_t *idebmap,
                                                   int *status) {
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    const int64_t f0 = kFillNodes * ((int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x);
    if (f0 >= nvox) return;
    if (o.init_bits && (f0 & 31) == 0) o.init_bits[f0 >> 5] = 0xffffffffu;  // word f0 / 32, < ceil(nvox / 32)
    const int64_t NN = (int64_t)R.N * R.N, SW = (int64_t)R.W * NN;
    const int64_t plane = (int64_t)g.nx * g.ny;
    int k = (int)(f0 / plane);
    const int64_t rem = f0 - (int64_t)k * plane;
    int j = (int)(rem / g.nx);
    int i = (int)(rem - (int64_t)j * g.nx);
    int eps[3 * kFillNodes], st[kFillNodes];
    uint8_t ideb[kFillNodes];
    const int nq = (int)min((int64_t)kFillNodes, nvox - f0);
    for (int q = 0; q < kFillNodes; ++q) {
        bool e0 = false, in = false, e1 = false, e2 = false;
        if (q < nq) {
            if (o.slot_x >= 0) {
                const int64_t ray = (int64_t)k * R.N + j;
                e0 = cov_bit(R.cov + o.slot_x * SW, NN, ray, i);
                in = cov_bit(R.cov_st, NN, ray, i);
            }
            if (o.build_eps) {
                e2 = o.slot_z >= 0 && cov_bit(R.cov + o.slot_z * SW, NN, (int64_t)j * R.N + i, k);
                e1 = o.slot_y >= 0 && cov_bit(R.cov + o.slot_y * SW, NN, (int64_t)k * R.N + i, j);
            }
        }
        eps[3 * q + 0] = e0 ? o.inside : 0;
        eps[3 * q + 1] = e1 ? o.inside : 0;
        eps[3 * q + 2] = e2 ? o.inside : 0;
        ideb[q] = in ? 0 : 1;
        st[q] = in ? NSGPU_STATUS_INSIDE : o.out_code;
        if (++i == g.nx) {
            i = 0;
            if (++j == g.ny) {
                j = 0;
                ++k;
            }
        }
    }
    // A full wave stages its 256 nodes' epsmap (3 KB) in LDS, so that each of
    // its three 16-byte stores writes one contiguous 1 KB: stored from the
    // lanes' own 48-byte runs, every store instruction left partial lines at a
    // 48-byte stride and the map fill wrote at half the bandwidth of a plain
    // fill.  A wave's LDS accesses are performed in order, so the reads see
    // the other lanes' writes; the wave barrier keeps the compiler from
    // moving them.
    __shared__ int4 stage[4][3 * 64];
    const bool full = __ballot(nq == kFillNodes) == ~0ull;
    if (nq == kFillNodes) {
        if (o.build_eps && full) {
            const int lane = threadIdx.x & 63;
            int4 *w = stage[threadIdx.x >> 6];
            for (int v = 0; v < 3; ++v) w[3 * lane + v] = make_int4(eps[4 * v], eps[4 * v + 1], eps[4 * v + 2], eps[4 * v + 3]);
            __builtin_amdgcn_wave_barrier();
            int4 *dst = reinterpret_cast<int4 *>(epsmap + 3 * (f0 - kFillNodes * lane));
            for (int v = 0; v < 3; ++v) {
                const int4 x = w[lane + 64 * v];
                store_nt(reinterpret_cast<int *>(dst + lane + 64 * v), x.x, x.y, x.z, x.w);
            }
        } else if (o.build_eps) {
            for (int v = 0; v < 3 * kFillNodes / 4; ++v)
                store_nt(epsmap + 3 * f0 + 4 * v, eps[4 * v], eps[4 * v + 1], eps[4 * v + 2], eps[4 * v + 3]);
        }
        if (o.build_eps) {
            uint32_t *iw = reinterpret_cast<uint32_t *>(idebmap + f0);
            for (int v = 0; v < kFillNodes / 4; ++v)
                __builtin_nontemporal_store((uint32_t)ideb[4 * v] | ((uint32_t)ideb[4 * v + 1] << 8) |
                                                ((uint32_t)ideb[4 * v + 2] << 16) | ((uint32_t)ideb[4 * v + 3] << 24),
                                            iw + v);
        }
        if (o.build_status) {
            for (int v = 0; v < kFillNodes / 4; ++v)
                store_nt(status + f0 + 4 * v, st[4 * v], st[4 * v + 1], st[4 * v + 2], st[4 * v + 3]);
        }
    } else {
        for (int q = 0; q < nq; ++q) {
            if (o.build_eps) {
                for (int l = 0; l < 3; ++l) epsmap[3 * (f0 + q) + l] = eps[3 * q + l];
                idebmap[f0 + q] = ideb[q];
            }
            if (o.build_status) status[f0 + q] = st[q];
        }
    }
}

// Corner bitmap (1 = outside) from the edge-ray planes: bits start at 1 and
// each 64-node row segment clears its inside bits with at most three
// atomicAnd on the 32-bit words it overlaps.
//
// A block takes a tile of 64 (i) x 16 (j) x 64 (k) nodes.  Node (i, j, k) is
// bit k & 63 of z-ray word (j, i) and bit j & 63 of y-ray word (k, i), so
// the tile's 16 x 64 z words and 64 x 64 y words are read once into LDS and
// each serves every row of the tile that needs it; read per row, they were
// fetched 64 times each.
constexpr int kCornJ = 16;

__global__ void __launch_bounds__(256) k_corners_bits(DevGrid g, DevRays R, FillOpts o, uint32_t *bits) {
    __shared__ unsigned long long zw[kCornJ][64], yw[64][64], xw[64][kCornJ];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i0 = blockIdx.x * 64, j0 = blockIdx.y * kCornJ, k0 = blockIdx.z * 64;
    const int64_t NN = (int64_t)R.N * R.N, SW = (int64_t)R.W * NN;
    const unsigned long long *px = R.cov + o.eslot_x * SW + (int64_t)(i0 >> 6) * NN;  // ray k N + j, bit i & 63
    const unsigned long long *pz = R.cov + o.eslot_z * SW + (int64_t)(k0 >> 6) * NN;  // ray j N + i, bit k & 63
    const unsigned long long *py = R.cov + o.eslot_y * SW + (int64_t)(j0 >> 6) * NN;  // ray k N + i, bit j & 63
    const int i = i0 + lane;
    for (int jj = wave; jj < kCornJ; jj += 4) {
        const int j = j0 + jj;
        zw[jj][lane] = (i < g.nx && j < g.ny) ? pz[(int64_t)j * R.N + i] : 0ull;
    }
    for (int kk = wave; kk < 64; kk += 4) {
        const int k = k0 + kk;
        yw[kk][lane] = (i < g.nx && k < g.nz) ? py[(int64_t)k * R.N + i] : 0ull;
    }
    for (int q = threadIdx.x; q < 64 * kCornJ; q += 256) {
        const int kk = q / kCornJ, jj = q % kCornJ, k = k0 + kk, j = j0 + jj;
        xw[kk][jj] = (k < g.nz && j < g.ny) ? px[(int64_t)k * R.N + j] : 0ull;
    }
    __syncthreads();
    const int jbit = j0 & 63;
    // eight rows at a time: eight independent ballots, then lane q clears the
    // words of row q, so the rows' LDS reads and atomics overlap
    for (int r0 = wave * 8; r0 < 64 * kCornJ; r0 += 32) {
        unsigned long long mine = 0ull;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int r = r0 + q, kk = r / kCornJ, jj = r % kCornJ;
            const bool inside = i < g.nx && k0 + kk < g.nz && j0 + jj < g.ny &&
                                (((xw[kk][jj] >> lane) | (zw[jj][lane] >> kk) | (yw[kk][lane] >> (jbit + jj))) & 1ull);
            const unsigned long long m = __ballot(inside);
            if (lane == q) mine = m;
        }
        if (lane < 8 && mine) {
            const int r = r0 + lane, k = k0 + r / kCornJ, j = j0 + r % kCornJ;
            const unsigned long long m = mine;
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

void launch_fill_bits(hipStream_t s, const DevGrid &g, const DevRays &R, const FillOpts &o, int *eps, uint8_t *ideb,
                      int *status) {
    const int64_t groups = ((int64_t)g.nx * g.ny * g.nz + kFillNodes - 1) / kFillNodes;
    hipLaunchKernelGGL(k_fill_bits, dim3((unsigned)((groups + 255) / 256)), dim3(256), 0, s, g, R, o, eps, ideb,
                       status);
}

void launch_corners_bits(hipStream_t s, const DevGrid &g, const DevRays &R, const FillOpts &o, uint32_t *bits) {
    dim3 grid((unsigned)((g.nx + 63) / 64), (unsigned)((g.ny + kCornJ - 1) / kCornJ), (unsigned)((g.nz + 63) / 64));
    hipLaunchKernelGGL(k_corners_bits, grid, dim3(256), 0, s, g, R, o, bits);
}

// Deterministic two-level sum: block b adds elements [4096 b, 4096 b + 4096)
// of v in a fixed order and writes one partial.
__device__ __forceinline__ void sum_block(const double *v, int64_t n, double *partial, int64_t b) {
    __shared__ double sh[256];
    double acc = 0;
    const int64_t base = b * 4096;
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
    if (threadIdx.x == 0) partial[b] = sh[0];
}

__global__ void __launch_bounds__(256) k_sum_blocks(const double *v, int64_t n, double *partial) {
    sum_block(v, n, partial, blockIdx.x);
}

// the per-ray segment lengths of every traced slot in one launch:
// grid (blocks per slot, slots)
__global__ void __launch_bounds__(256) k_sum_slots(const double *v, int64_t n, VolumeSlots vs, double *partial,
                                                   int64_t per) {
    const int64_t slot = vs.slot[blockIdx.y];
    sum_block(v + slot * n, n, partial + slot * per, blockIdx.x);
}

// the same sums with n = *cnt + *off (an exclusive scan's total, on the device)
__global__ void __launch_bounds__(256) k_sum_blocks_scan(const double *v, const int *cnt, const int *off,
                                                         double *partial) {
    sum_block(v, (int64_t)*cnt + *off, partial, blockIdx.x);
}

int64_t sum_partials_count(int64_t n) { return (n + 4095) / 4096; }

void launch_sum_blocks(hipStream_t s, const double *v, int64_t n, double *partial) {
    const int64_t nb = sum_partials_count(n);
    if (nb > 0) hipLaunchKernelGGL(k_sum_blocks, dim3((unsigned)nb), dim3(256), 0, s, v, n, partial);
}

void launch_sum_blocks_scan(hipStream_t s, const double *v, int64_t ncap, const int *cnt, const int *off,
                            double *partial) {
    const int64_t nb = sum_partials_count(ncap);
    if (nb > 0) hipLaunchKernelGGL(k_sum_blocks_scan, dim3((unsigned)nb), dim3(256), 0, s, v, cnt, off, partial);
}

void launch_sum_slots(hipStream_t s, const double *v, int64_t n, const VolumeSlots &vs, double *partial) {
    const int64_t nb = sum_partials_count(n);
    if (nb > 0 && vs.n > 0)
        hipLaunchKernelGGL(k_sum_slots, dim3((unsigned)nb, (unsigned)vs.n), dim3(256), 0, s, v, n, vs, partial, nb);
}

}  // namespace nsg
