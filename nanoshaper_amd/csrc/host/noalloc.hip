// getSurf for callers that cannot allocate device memory: every scratch
// buffer of the call (ray lists, coverage planes, partial sums, cavity runs)
// is carved from one file-scope device array per GPU, and the maps go to the
// caller's device arrays.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.hpp"
#include "nsgpu.h"

namespace {

constexpr int kMaxDevices = 64;

// The scratch array itself.  In host code `scratch` names a host-side
// stand-in, not device memory: its device address on each GPU comes from
// hipGetSymbolAddress (nsg_scratch_setup) and only that address is passed on.
static __device__ __align__(256) unsigned char scratch[NSGPU_NOALLOC_SCRATCH_BYTES];

std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};
bool g_looked_up[kMaxDevices] = {};

}  // namespace

// once per device, with that device current (nsgpu_create sets it)
void nsg_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_looked_up[device]) return;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) g_ptr[device] = p;
    else (void)hipGetLastError();  // a failed lookup must not surface in a later call
    g_looked_up[device] = true;
}

extern "C" int nsgpu_get_surf_noalloc(nsgpu_ctx *ctx, const nsgpu_surf_opts *opts, nsgpu_surf_result *res,
                                      int32_t *epsmap, uint8_t *idebmap, int32_t *status) {
    if (!ctx || !opts) return NSGPU_E_ARG;
    const int device = nsg_ctx_device(ctx);
    void *p = nullptr;
    if (device >= 0 && device < kMaxDevices) {
        std::lock_guard<std::mutex> lock(g_mu);
        p = g_ptr[device];
    }
    return nsg_get_surf_in(ctx, opts, res, epsmap, idebmap, status, p, NSGPU_NOALLOC_SCRATCH_BYTES);
}
