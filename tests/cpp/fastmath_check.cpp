// Accuracy of the kernels' FP64 cbrt and cos(acos(x)/3) (fastmath.hpp, the
// same source the device build compiles) against long-double references.
// Prints: "<name> <samples> <max ulp error> <max ulp error of the libm FP64
// function on the same samples>" per function.
#include <math.h>
#include <stdio.h>

#include <random>

#include "../../nanoshaper_amd/csrc/kernels/fastmath.hpp"

static double ulp_err(double y, long double ref) {
    const double r = (double)ref;
    const double u = nextafter(fabs(r), INFINITY) - fabs(r);
    return (double)(fabsl((long double)y - ref) / (long double)u);
}

int main() {
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    double worst = 0, worst_lib = 0;
    long n = 0;
    for (int i = 0; i < 2000000; ++i) {
        const double mag = pow(10.0, -30.0 + 60.0 * U(rng));
        const double a = (i & 1) ? -mag : mag;
        const long double ref = cbrtl((long double)a);
        const double e = ulp_err(nsg::cbrt_fast(a), ref), el = ulp_err(cbrt(a), ref);
        if (e > worst) worst = e;
        if (el > worst_lib) worst_lib = el;
        ++n;
    }
    printf("cbrt %ld %.4f %.4f\n", n, worst, worst_lib);
    worst = worst_lib = 0;
    n = 0;
    for (int i = 0; i < 2000000; ++i) {
        double x;
        if (i % 4 == 0) x = -0.99 + 0.02 * U(rng);        // around the fallback threshold
        else if (i % 4 == 1) x = 1.0 - pow(10.0, -16.0 * U(rng));  // towards x = 1
        else x = -1.0 + 2.0 * U(rng);
        const long double ref = cosl(acosl((long double)x) / 3.0L);
        const double e = ulp_err(nsg::cos_third_acos(x), ref), el = ulp_err(cos((1. / 3.) * acos(x)), ref);
        if (e > worst) worst = e;
        if (el > worst_lib) worst_lib = el;
        ++n;
    }
    printf("cos_third_acos %ld %.4f %.4f\n", n, worst, worst_lib);
    return 0;
}
