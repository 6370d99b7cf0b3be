// FP64 cube root and cos(acos(x) / 3) for the closed-form cubic (Cardano's
// two cube roots, or the trigonometric form of three real roots) as an FP32
// seed from the hardware transcendentals, one Newton step in FP64 and a last
// Newton step on a residual computed with fma splits, so that the result is
// within about one unit in the last place of the exact value, as the FP64
// library functions are.  Inputs outside the seeds' well-conditioned ranges
// take the library functions.
//
// The functions compile for the host too (g++), so that the accuracy check
// runs against the host's libm on the same code.
#pragma once
#include <math.h>

#ifndef __HIPCC__
#define NSG_HD
#else
#include <hip/hip_runtime.h>
#define NSG_HD __host__ __device__
#endif

namespace nsg {

// about 1 / x to FP32 precision (the Newton steps only need ~1e-7)
NSG_HD inline float rcp_seed(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// cbrt(a).  |a| in [2^-120, 2^120]: the seed's relative error e0 ~ 1e-7
// (float rounding of a and cbrtf) drops to ~e0^2 after the first step; the
// second step's residual y^3 - a is exact to ~2^-100 relative (y^2 = h + l
// exactly, then one fma), so the step leaves ~e0^3 plus the final rounding.
NSG_HD inline double cbrt_fast(double a) {
    const double aa = fabs(a);
    if (!(aa >= 0x1p-120 && aa <= 0x1p120)) return cbrt(a);  // 0, tiny, huge, inf, nan
    const float yf = cbrtf((float)a);
    const double inv = (double)rcp_seed(3.0f * yf * yf);  // 1 / (3 y^2)
    double y = (double)yf;
    y -= (y * y * y - a) * inv;
    const double h = y * y;
    const double l = fma(y, y, -h);
    const double r = fma(h, y, -a) + l * y;
    return y - r * inv;
}

// cos(acos(x) / 3) for x in [-1, 1], the root in [1/2, 1] of
// g(y) = 4 y^3 - 3 y - x.  Below x = -0.99 the root approaches the double
// root of g at x = -1 (g'(y) = 12 y^2 - 3 -> 0) and the library composition
// is used; x > 1 and NaN also go there (acos returns NaN).
NSG_HD inline double cos_third_acos(double x) {
    if (!(x > -0.99 && x <= 1.0)) return cos((1. / 3.) * acos(x));
    const float yf = cosf(acosf((float)x) * (1.0f / 3.0f));
    const double inv = (double)rcp_seed(12.0f * yf * yf - 3.0f);  // 1 / g'(y), g' >= 0.5 here
    double y = (double)yf;
    y -= ((4. * y * y - 3.) * y - x) * inv;
    // g(y) with y^3 = ph + (pl + l y) exactly to ~2^-100 and the O(1) terms
    // summed by two-sums, then one step on that residual
    const double h = y * y;
    const double l = fma(y, y, -h);
    const double ph = h * y;
    const double pl = fma(h, y, -ph);
    const double t = 3. * y;
    const double tl = fma(3., y, -t);
    const double s1 = 4. * ph - t;  // two-sum of 4 ph and -t
    const double bb1 = s1 - 4. * ph;
    const double e1 = (4. * ph - (s1 - bb1)) + (-t - bb1);
    const double s2 = s1 - x;  // two-sum of s1 and -x
    const double bb2 = s2 - s1;
    const double e2 = (s1 - (s2 - bb2)) + (-x - bb2);
    const double g = s2 + ((e1 + e2) + (4. * (pl + l * y) - tl));
    return y - g * inv;
}

}  // namespace nsg
