// Host patch set behind the opaque nsgpu_patch_set handle: the SoA of
// include/nsgpu.h, owned by the library (SES and skin builders).
#pragma once
#include <cstdint>
#include <vector>

struct nsgpu_patch_set {
    std::vector<int32_t> kind, off, cnt, flags, atom;
    std::vector<double> bsph, data;
    int64_t counts[5] = {0, 0, 0, 0, 0};  // SES kinds
    int errors = 0;
};
