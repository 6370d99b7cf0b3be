// Library-internal entry points shared by the host sources (hidden: not part
// of the C ABI).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "nsgpu.h"

#define NSG_HIDDEN __attribute__((visibility("hidden")))

// getSurf with every scratch buffer carved from [scratch, scratch + bytes)
// and the maps written to the caller's device arrays (api.cpp)
NSG_HIDDEN int nsg_get_surf_in(nsgpu_ctx *c, const nsgpu_surf_opts *o, nsgpu_surf_result *res, int32_t *epsmap,
                               uint8_t *idebmap, int32_t *status, void *scratch, size_t bytes);
// the context's device (api.cpp)
NSG_HIDDEN int nsg_ctx_device(const nsgpu_ctx *c);
// looks up the device's scratch array once, with that device current (noalloc.hip)
NSG_HIDDEN void nsg_scratch_setup(int device);
