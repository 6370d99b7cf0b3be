// Exclusive prefix sum of int32 counts (block sums, then each block adds the
// sums before it), used for the ordered compaction of vertices, cubes and
// triangles.  Blocks of 256 lanes cover 1024 items; beyond 4096 blocks the
// block sums are scanned recursively.
#include "launch.hpp"

namespace nsg {

constexpr int kScanItems = 1024;

__device__ __forceinline__ int block_excl_scan(int v, int *sh, int &total) {
    // inclusive scan within waves, then across the 4 waves
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) sh[w] = x;
    __syncthreads();
    int base = 0;
    for (int q = 0; q < w; ++q) base += sh[q];
    total = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return base + x - v;
}

__global__ void __launch_bounds__(256) k_scan_reduce(const int *in, int64_t n, int *bsum) {
    __shared__ int sh[4];
    const int64_t base = (int64_t)blockIdx.x * kScanItems + threadIdx.x * 4;
    int v = 0;
    for (int q = 0; q < 4; ++q)
        if (base + q < n) v += in[base + q];
    int total;
    block_excl_scan(v, sh, total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// total (optional): total[0] = the sum of all n inputs, total[1] = 0, written
// by the thread of the last element
__global__ void __launch_bounds__(256) k_scan_down(const int *in, int64_t n, const int *boff, int *out, int *total) {
    __shared__ int sh[4];
    const int64_t base = (int64_t)blockIdx.x * kScanItems + threadIdx.x * 4;
    int v[4], s = 0;
    for (int q = 0; q < 4; ++q) {
        v[q] = (base + q < n) ? in[base + q] : 0;
        s += v[q];
    }
    int btotal;
    int run = block_excl_scan(s, sh, btotal) + (boff ? boff[blockIdx.x] : 0);
    for (int q = 0; q < 4; ++q) {
        if (base + q < n) out[base + q] = run;
        run += v[q];
        if (total && base + q == n - 1) {
            total[0] = run;
            total[1] = 0;
        }
    }
}

// As k_scan_down, with the block's offset summed from the block sums of the
// blocks before it (nb <= kFusedBlocks), which saves the scan of the block
// sums as a launch of its own.
constexpr int64_t kFusedBlocks = 4096;

__global__ void __launch_bounds__(256) k_scan_down_fused(const int *in, int64_t n, const int *bsum, int *out,
                                                         int *total) {
    __shared__ int sh[4], red[4];
    int pre = 0;
    for (int q = threadIdx.x; q < (int)blockIdx.x; q += 256) pre += bsum[q];
    for (int off = 32; off > 0; off >>= 1) pre += __shfl_down(pre, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = pre;
    __syncthreads();
    const int boff = red[0] + red[1] + red[2] + red[3];
    const int64_t base = (int64_t)blockIdx.x * kScanItems + threadIdx.x * 4;
    int v[4], sum = 0;
    for (int q = 0; q < 4; ++q) {
        v[q] = (base + q < n) ? in[base + q] : 0;
        sum += v[q];
    }
    int btotal;
    int run = block_excl_scan(sum, sh, btotal) + boff;
    for (int q = 0; q < 4; ++q) {
        if (base + q < n) out[base + q] = run;
        run += v[q];
        if (total && base + q == n - 1) {
            total[0] = run;
            total[1] = 0;
        }
    }
}

// scratch must hold scan_scratch_ints(n) ints
int64_t scan_scratch_ints(int64_t n) {
    int64_t s = 0;
    while (n > kScanItems) {
        n = (n + kScanItems - 1) / kScanItems;
        s += 2 * n;
    }
    return s + 2;
}

void launch_scan(hipStream_t st, const int *in, int64_t n, int *out, int *scratch, int *total) {
    if (n <= 0) {
        if (total) launch_fill32(st, total, 0u, 2);
        return;
    }
    const int64_t nb = (n + kScanItems - 1) / kScanItems;
    if (nb == 1) {
        hipLaunchKernelGGL(k_scan_down, dim3(1), dim3(256), 0, st, in, n, (const int *)nullptr, out, total);
        return;
    }
    int *bsum = scratch, *boff = scratch + nb;
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(256), 0, st, in, n, bsum);
    if (nb <= kFusedBlocks) {
        hipLaunchKernelGGL(k_scan_down_fused, dim3((unsigned)nb), dim3(256), 0, st, in, n, (const int *)bsum, out,
                           total);
        return;
    }
    launch_scan(st, bsum, nb, boff, scratch + 2 * nb);
    hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nb), dim3(256), 0, st, in, n, boff, out, total);
}

// dst[0..n) = *src[q] (the scalars the host reads between stages), then
// dst + 8 (as doubles) = dsrc[0..nd): one copy to the host takes them all
__global__ void __launch_bounds__(256) k_gather(GatherSrc g, int *dst) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < g.n) dst[t] = *g.src[t];
    double *dd = reinterpret_cast<double *>(dst + 8);
    for (int64_t q = t; q < g.nd; q += (int64_t)gridDim.x * blockDim.x) dd[q] = g.dsrc[q];
}

void launch_gather(hipStream_t st, const GatherSrc &g, int *dst) {
    const int64_t work = g.nd > 8 ? g.nd : 8;
    const unsigned nb = (unsigned)((work + 255) / 256 < 1024 ? (work + 255) / 256 : 1024);
    hipLaunchKernelGGL(k_gather, dim3(nb), dim3(256), 0, st, g, dst);
}

}  // namespace nsg
