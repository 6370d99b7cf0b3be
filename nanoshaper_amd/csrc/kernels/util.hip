// Small device utilities: fills (used instead of memset nodes, so that the
// enqueued work can be captured into a graph and replayed) and the
// deterministic volume reduction.
#include <algorithm>

#include "launch.hpp"

namespace nsg {

__global__ void __launch_bounds__(256) k_fill32(uint32_t *p, uint32_t v, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// up to four fills in one launch: the words of the segments in turn
__global__ void __launch_bounds__(256) k_fill_multi(FillList f) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int q = 0; q < f.k; ++q)
        for (int64_t i = t0; i < f.s[q].n; i += stride) f.s[q].p[i] = f.s[q].v;
}

void launch_fill_multi(hipStream_t s, const FillList &f) {
    int64_t n = 0;
    for (int q = 0; q < f.k; ++q) n = std::max(n, f.s[q].n);
    if (n <= 0) return;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_multi, dim3((unsigned)blocks), dim3(256), 0, s, f);
}

void launch_fill32(hipStream_t s, void *p, uint32_t v, int64_t n) {
    if (n <= 0) return;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill32, dim3((unsigned)blocks), dim3(256), 0, s, (uint32_t *)p, v, n);
}

// volume = A * sum of the listed slots' block partials / flag_sum, summed in
// slot then block order by one thread
__global__ void k_volume(const double *partial, int64_t per, VolumeSlots vs, double *out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double sum = 0;
    for (int k = 0; k < vs.n; ++k)
        for (int64_t q = 0; q < per; ++q) sum += partial[(int64_t)vs.slot[k] * per + q];
    *out = vs.flag_sum ? vs.area_px * sum / vs.flag_sum : 0.0;
}

void launch_volume(hipStream_t s, const double *partial, int64_t per, const VolumeSlots &vs, double *out) {
    hipLaunchKernelGGL(k_volume, dim3(1), dim3(64), 0, s, partial, per, vs, out);
}

// device wall clock (constant rate, hipDeviceAttributeWallClockRate kHz)
__global__ void k_stamp(unsigned long long *out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = wall_clock64();
}

void launch_stamp(hipStream_t s, unsigned long long *out) { hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, s, out); }

}  // namespace nsg
