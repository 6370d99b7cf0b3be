// Ray/patch intersection over the grid and edge-ray panels, and the per-ray
// sort.  One 64-lane workgroup owns one (patch, panel slot): the patch record
// is loaded once (wave-uniform, scalar path), the lanes stride over the pixel
// rectangle of the patch's bounding sphere, and every feasible depth is
// appended to its ray's slot list.  The FP64 path evaluates the reference's
// expressions in the reference's order (the file is built with
// -ffp-contract=off), so roots, feasibility and rounding decisions match the
// CPU code (ConnollySurface::getPatchIntersectionData, ConnollySurface.cpp:
// 4319-4591; raySphere/rayTorus/quarticEqSolutions, tools.cpp:350-1099).
#include <algorithm>
#include <type_traits>

#include "fastmath.hpp"
#include "launch.hpp"

namespace nsg {

// raySphere, tools.cpp:833-862, for an axis-aligned direction.  The products
// with the two zero direction components are exact zeros and drop out.
__device__ __forceinline__ bool ray_sphere(const double o[3], const double dir[3], const double c[3], double r,
                                           double &t1, double &t2) {
    const double A = dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2];
    const double tx = o[0] - c[0], ty = o[1] - c[1], tz = o[2] - c[2];
    const double B = tx * dir[0] + ty * dir[1] + tz * dir[2];
    const double C = tx * tx + ty * ty + tz * tz - r * r;
    double det = B * B - A * C;
    if (det < 0.) return false;
    det = sqrt(det);
    t1 = (-B - det) / A;
    t2 = (-B + det) / A;
    if (t2 < t1) {
        const double tt = t1;
        t1 = t2;
        t2 = tt;
    }
    return true;
}

// realCubicSolution, tools.cpp:350-393.  The cube roots and cos(acos(.)/3)
// come from fastmath.hpp (FP32 seeds refined in FP64 to within about one unit
// in the last place, as the FP64 library functions are).
__device__ double cubic_root(double b, double c, double d) {
    const double p = -(1.0 / 3.0) * b * b + c;
    const double q = (2. / 27.) * b * b * b - (1. / 3.) * b * c + d;
    const double D3p = (1. / 27.) * p * p * p;
    const double D3 = D3p + 0.25 * q * q;
    if (fabs(c) < b * b * 1.E-6 && fabs(d) < fabs(b) * 1.E-6) return c / b - b - d / (b * b);
    if (fabs(d / c) < 4.E-4) {
        const double ratio = d / c;
        if (D3 <= 0.) return 0.5 * (ratio - b + sqrt((b - ratio) * (b - ratio) - 4. * (c - ratio * (b - ratio))));
        const double eps = -ratio * ratio * (b - ratio) / (3. * ratio * ratio - 2. * b * ratio + c);
        return eps - ratio;
    }
    if (D3 >= 0.) {
        const double radD3 = sqrt(D3);
        const double A = cbrt_fast(radD3 - 0.5 * q);
        const double psu3A = cbrt_fast(radD3 + 0.5 * q);
        return A - psu3A - (1. / 3.) * b;
    }
    return 2. * sqrt(-(1. / 3.) * p) * cos_third_acos(-0.5 * q / sqrt(-D3p)) - (1. / 3.) * b;
}

// quarticEqSolutions, tools.cpp:525-641: Ferrari's depressed quartic with the
// resolvent cubic, the discriminant-based root count and the reciprocal
// variant for large coefficients.
__device__ int quartic_roots(double roots[4], double b, double c, double d, double e) {
    const double b_2 = 0.5 * b, b_4 = 0.25 * b;
    double p = c - 1.5 * (b_2 * b_2);
    double q = d + b_2 * (b_2 * b_2 - c);
    double r = e + b_4 * (b_4 * (c - 3.0 * b_4 * b_4) - d);
    double f = p * p - 4. * r;
    double delta = 4. * f * (4. * r * (p * p) - p * (q * q) - 16. * (r * r)) + (q * q) * (128. * p * r - 27. * (q * q));
    bool change_variant = false;
    if (e != 0. && (fabs(p) > 3.E+3 || fabs(r) > 3.E+3 || fabs(q) > 3.E+3 || fabs(delta) > 1.E+14)) {
        const double dsu2e = 0.5 * d / e;
        change_variant = true;
        p = -1.5 * dsu2e * dsu2e + c / e;
        q = dsu2e * (dsu2e * dsu2e - c / e) + b / e;
        r = dsu2e * (-3. / 16. * (dsu2e * dsu2e * dsu2e) + 0.25 * dsu2e * (c / e) - 0.5 * b / e) + 1. / e;
        f = p * p - 4. * r;
        delta = 4. * f * (4. * r * (p * p) - p * (q * q) - 16. * (r * r)) + (q * q) * (128. * p * r - 27. * (q * q));
    }
    int num;
    if (delta < 0. || (delta == 0. && (p >= 0. || f < 0.)))
        num = 2;
    else if ((p < 0. && f > 0.) || (delta == 0. && p == 0. && r == 0.))
        num = 4;
    else
        return 0;
    const double m = cubic_root(p, 0.25 * (p * p) - r, (-0.125) * (q * q));
    if (m <= 0) return 0;
    double coeff, p1, p2, p3 = 0, smax;
    if (m < 1.E-14) {
        p1 = p + p;
        p2 = (q * q) / f;
        p3 = 2. * sqrt(f);
        smax = p3 - p1 - p2;
        coeff = 2. * fabs(q) / f;
    } else {
        p1 = -2. * (p + m);
        p2 = sqrt(2. * (q * q) / m);
        smax = p1 + p2;
        coeff = sqrt(2. * m);
    }
    roots[0] = 0.5 * (((q >= 0.) ? -coeff : coeff) - sqrt(smax));
    roots[1] = 0.5 * (((q >= 0.) ? -coeff : coeff) + sqrt(smax));
    if (num == 4) {
        const double smin = (m < 1.E-14) ? -p1 - p2 - p3 : p1 - p2;
        roots[2] = 0.5 * (((q >= 0.) ? coeff : -coeff) - sqrt(smin));
        roots[3] = 0.5 * (((q >= 0.) ? coeff : -coeff) + sqrt(smin));
    }
    if (change_variant) {
        for (int i = 0; i < num; i++) roots[i] = 1. / (roots[i] - 0.25 * d / e);
    } else {
        for (int i = 0; i < num; i++) roots[i] -= 0.25 * b;
    }
    return num;
}

// rayTorus, tools.cpp:907-1099 (analytical).  t[0] is the clipping-sphere
// entry; the ray is moved there, brought into the torus frame and the quartic
// in the distance along the (unit, axis-aligned) direction is solved; roots
// behind the entry or outside the clipping sphere are dropped.
__device__ __forceinline__ int ray_torus(const double *td, const double sc[3], double probe, double radius, int vc,
                         const double orig[3], double ray_dir, double t[4]) {
    double o1[3] = {orig[0], orig[1], orig[2]};
    o1[vc] += t[0] * ray_dir;
    const double tx = o1[0] - td[0], ty = o1[1] - td[1], tz = o1[2] - td[2];
    const double *iv = td + 3;
    double pp[3];
    for (int i = 0; i < 3; i++) pp[i] = iv[3 * i + 0] * tx + iv[3 * i + 1] * ty + iv[3 * i + 2] * tz;
    const double r2 = probe * probe;
    const double R2 = td[12] * td[12];
    const double pd = pp[0] * iv[0 + vc] + pp[1] * iv[3 + vc] + pp[2] * iv[6 + vc];
    const double p2 = pp[0] * pp[0] + pp[1] * pp[1] + pp[2] * pp[2];
    const double coeff = p2 - r2 - R2;
    const double uz = iv[6 + vc];
    const double B = 4 * pd;
    const double C = 2 * coeff + 4 * pd * pd + 4 * R2 * uz * uz;
    const double D = 4 * pd * coeff + 8 * R2 * pp[2] * uz;
    const double E = coeff * coeff - 4 * R2 * (r2 - pp[2] * pp[2]);
    double roots[4];
    const int nr = quartic_roots(roots, B, C, D, E);
    const double t0 = t[0];
    int ni = 0;
    for (int i = 0; i < nr; i++) {
        if (roots[i] < 0) continue;
        double pt[3] = {o1[0], o1[1], o1[2]};
        pt[vc] += roots[i];
        const double dd = (sc[0] - pt[0]) * (sc[0] - pt[0]) + (sc[1] - pt[1]) * (sc[1] - pt[1]) +
                          (sc[2] - pt[2]) * (sc[2] - pt[2]);
        if (dd < radius * radius) t[ni++] = t0 + roots[i] / ray_dir;
    }
    return ni;
}

// A chord of the clipping sphere that stays farther than R + r from the torus
// axis, or (ring torus, R > r) nearer than R - r, lies outside the solid torus,
// so the torus quartic has no root on it and rayTorus finds nothing there; the
// walk drops such rays before the solve.  Along the ray (x = pa + s e_vc) the
// squared distance to the axis is the convex quadratic A s^2 + B s + C; the
// chord is widened and the radii padded by kTorusCullPad (A), far beyond the
// error of any computed root, and the comparisons are made without division.
constexpr double kTorusCullPad = 1e-3;

template <int VC>
__device__ __forceinline__ bool torus_far(const double *td, const double pa[3], double s0, double s1, double probe) {
    const double u[3] = {pa[0] - td[0], pa[1] - td[1], pa[2] - td[2]};
    const double w[3] = {td[9], td[10], td[11]};  // third row of the inverse frame: the axis
    const double aw = u[0] * w[0] + u[1] * w[1] + u[2] * w[2];
    const double A = 1. - w[VC] * w[VC];
    const double B = 2. * (u[VC] - aw * w[VC]);
    const double C = u[0] * u[0] + u[1] * u[1] + u[2] * u[2] - aw * aw;
    const double lo = s0 - kTorusCullPad, hi = s1 + kTorusCullPad;
    const double rlo = (A * lo + B) * lo + C, rhi = (A * hi + B) * hi + C;
    const double R = td[12];
    const double Ro = R + probe + kTorusCullPad;
    double rmin = fmin(rlo, rhi);
    bool far_out;
    if (A > 1e-9 && -B >= 2. * A * lo && -B <= 2. * A * hi)  // vertex inside the chord
        far_out = 4. * A * C - B * B > 4. * A * (Ro * Ro);
    else
        far_out = rmin > Ro * Ro;
    if (far_out) return true;
    const double Ri = R - probe - kTorusCullPad;
    return Ri > 0. && fmax(rlo, rhi) < Ri * Ri;
}

// ConnollySurface::isFeasible, ConnollySurface.cpp:4769-5021, on the SoA form
// (plane signs applied when the patches were built).  Every item is the
// reference's FP64 expression; items are taken four at a time (the tail
// repeats the last item), so the wave issues the uniform loads of a group
// together and branches once per group.  Each test is a pure function of the
// point, so evaluating a group's extra items does not change the result.
__device__ __forceinline__ bool plane_out(const double *q, const double p[3]) {
    return (q[0] * p[0] + q[1] * p[1] + q[2] * p[2]) + q[3] > 0;
}
__device__ __forceinline__ bool in_clip_sphere(const double *q, const double p[3]) {
    double dd = (q[0] - p[0]) * (q[0] - p[0]) + (q[1] - p[1]) * (q[1] - p[1]) + (q[2] - p[2]) * (q[2] - p[2]);
    dd -= q[3];
    return dd < 0;
}

// Point cells (outside every clipping sphere) and facets (inside every
// plane), for both depths of a sphere or facet ray at once: one pass over the
// items tests the two points, so each item is loaded once, and a lane stops
// only when both points have failed.  Each result is the AND of its own
// tests.
__device__ __forceinline__ void feasible_plain2(int kind, const double *d, int cnt, const double p[3],
                                                const double q[3], bool &ok_p, bool &ok_q) {
    bool out_p = false, out_q = false;
    if (kind == NSGPU_PATCH_POINT) {
        for (int i = 0; i < cnt; i += 4) {
            const int j1 = min(i + 1, cnt - 1), j2 = min(i + 2, cnt - 1), j3 = min(i + 3, cnt - 1);
            out_p |= in_clip_sphere(d + 4 * i, p) | in_clip_sphere(d + 4 * j1, p) | in_clip_sphere(d + 4 * j2, p) |
                     in_clip_sphere(d + 4 * j3, p);
            out_q |= in_clip_sphere(d + 4 * i, q) | in_clip_sphere(d + 4 * j1, q) | in_clip_sphere(d + 4 * j2, q) |
                     in_clip_sphere(d + 4 * j3, q);
            if (out_p & out_q) break;
        }
    } else {
        for (int i = 0; i < cnt; i += 4) {
            const int j1 = min(i + 1, cnt - 1), j2 = min(i + 2, cnt - 1), j3 = min(i + 3, cnt - 1);
            out_p |= plane_out(d + 4 * i, p) | plane_out(d + 4 * j1, p) | plane_out(d + 4 * j2, p) |
                     plane_out(d + 4 * j3, p);
            out_q |= plane_out(d + 4 * i, q) | plane_out(d + 4 * j1, q) | plane_out(d + 4 * j2, q) |
                     plane_out(d + 4 * j3, q);
            if (out_p & out_q) break;
        }
    }
    ok_p = !out_p;
    ok_q = !out_q;
}

// torus: outside the self-intersection sphere, then any visible plane pair
// (acute pair: both planes; obtuse pair: either plane)
__device__ __forceinline__ bool pair_ok(const double *pl, int acute, const double p[3]) {
    const bool le1 = !plane_out(pl, p), le2 = !plane_out(pl + 4, p);
    return acute ? (le1 & le2) : (le1 | le2);
}
__device__ bool feasible_torus(int kind, const double *d, int cnt, int flags, const double p[3]) {
    const double r = d[13];
    if (r >= 0.) {
        const double dd = (d[0] - p[0]) * (d[0] - p[0]) + (d[1] - p[1]) * (d[1] - p[1]) + (d[2] - p[2]) * (d[2] - p[2]);
        if (dd < r * r) return false;
    }
    if (kind == NSGPU_PATCH_TORUS_SING) return true;
    const double *pl = d + NSGPU_TORUS_HDR;
    for (int i = 0; i < cnt; i += 2) {
        const int j = min(i + 1, cnt - 1);
        if (pair_ok(pl + 8 * i, (flags >> i) & 1, p) | pair_ok(pl + 8 * j, (flags >> j) & 1, p)) return true;
    }
    return false;
}

// FP64 closed forms as out-of-line calls for the mixed solver's rare
// fallbacks, so that their registers do not count against the common path
__device__ __attribute__((noinline)) double cubic_root_fallback(double b, double c, double d) {
    return cubic_root(b, c, d);
}
__device__ __attribute__((noinline)) int quartic_roots_fallback(double roots[4], double b, double c, double d,
                                                                double e) {
    return quartic_roots(roots, b, c, d, e);
}

// Mixed precision, resolvent cubic of the quartic: the branch decisions of
// realCubicSolution are taken in FP64 exactly as in cubic_root; where the
// FP64 path would evaluate cbrt or acos/cos, the closed form is evaluated in
// FP32 (hardware transcendentals) as a seed and refined by Newton's method in
// FP64 on x^3 + b x^2 + c x + d.  A seed whose refinement does not settle
// within two steps to 1e-12 relative falls back to the FP64 closed form.
__device__ double cubic_root_mixed(double b, double c, double d) {
    const double p = -(1.0 / 3.0) * b * b + c;
    const double q = (2. / 27.) * b * b * b - (1. / 3.) * b * c + d;
    const double D3p = (1. / 27.) * p * p * p;
    const double D3 = D3p + 0.25 * q * q;
    if (fabs(c) < b * b * 1.E-6 && fabs(d) < fabs(b) * 1.E-6) return c / b - b - d / (b * b);
    if (fabs(d / c) < 4.E-4) return cubic_root_fallback(b, c, d);
    float x32;
    if (D3 >= 0.) {
        const float radD3 = sqrtf((float)D3), hq = (float)(0.5 * q);
        x32 = cbrtf(radD3 - hq) - cbrtf(radD3 + hq) - (float)((1. / 3.) * b);
    } else {
        const float Aarg = acosf((float)(-0.5 * q / sqrt(-D3p)));
        x32 = 2.f * sqrtf((float)(-(1. / 3.) * p)) * cosf((1.f / 3.f) * Aarg) - (float)((1. / 3.) * b);
    }
    double x = x32;
    for (int it = 0; it < 2; ++it) {
        const double g = ((x + b) * x + c) * x + d;
        const double gp = (3. * x + 2. * b) * x + c;
        if (gp == 0.) return cubic_root_fallback(b, c, d);
        x -= g / gp;
    }
    const double g = ((x + b) * x + c) * x + d;
    const double gp = (3. * x + 2. * b) * x + c;
    if (!(fabs(g) <= 1e-12 * fabs(gp) * (1. + fabs(x)))) return cubic_root_fallback(b, c, d);
    return x;
}

// relative size of the quartic's discriminant below which a root pair counts
// as near-double
constexpr double kNearTangent = 1e-8;

// quarticEqSolutions with the mixed resolvent root: the classification, the
// variant switch and the root formulas are those of quartic_roots in FP64;
// only the resolvent root m comes from cubic_root_mixed, except for
// near-double roots, which take quartic_roots whole.  A root m close to
// the m <= 0 or m < 1e-14 thresholds is recomputed by the FP64 closed form so
// that the root count and formula branch are the reference's.
__device__ __forceinline__ int quartic_roots_mixed(double roots[4], double b, double c, double d, double e) {
    const double b_2 = 0.5 * b, b_4 = 0.25 * b;
    double p = c - 1.5 * (b_2 * b_2);
    double q = d + b_2 * (b_2 * b_2 - c);
    double r = e + b_4 * (b_4 * (c - 3.0 * b_4 * b_4) - d);
    double f = p * p - 4. * r;
    double delta = 4. * f * (4. * r * (p * p) - p * (q * q) - 16. * (r * r)) + (q * q) * (128. * p * r - 27. * (q * q));
    bool change_variant = false;
    if (e != 0. && (fabs(p) > 3.E+3 || fabs(r) > 3.E+3 || fabs(q) > 3.E+3 || fabs(delta) > 1.E+14)) {
        const double dsu2e = 0.5 * d / e;
        change_variant = true;
        p = -1.5 * dsu2e * dsu2e + c / e;
        q = dsu2e * (dsu2e * dsu2e - c / e) + b / e;
        r = dsu2e * (-3. / 16. * (dsu2e * dsu2e * dsu2e) + 0.25 * dsu2e * (c / e) - 0.5 * b / e) + 1. / e;
        f = p * p - 4. * r;
        delta = 4. * f * (4. * r * (p * p) - p * (q * q) - 16. * (r * r)) + (q * q) * (128. * p * r - 27. * (q * q));
    }
    // near-multiple roots (grazing rays): the discriminant is small against
    // its own terms and the resolvent cubic shares the near-double root; the
    // reference's FP64 closed form decides these rays
    const double dmag = fabs(4. * f) * (fabs(4. * r * (p * p)) + fabs(p * (q * q)) + 16. * (r * r)) +
                        (q * q) * (fabs(128. * p * r) + 27. * (q * q));
    if (fabs(delta) <= kNearTangent * dmag) {
        // the out-of-line FP64 solver writes to an array of its own, so that
        // roots itself never needs an address and stays in registers
        double fb[4];
        const int nf = quartic_roots_fallback(fb, b, c, d, e);
#pragma unroll
        for (int i = 0; i < 4; ++i) roots[i] = fb[i];
        return nf;
    }
    int num;
    if (delta < 0. || (delta == 0. && (p >= 0. || f < 0.)))
        num = 2;
    else if ((p < 0. && f > 0.) || (delta == 0. && p == 0. && r == 0.))
        num = 4;
    else
        return 0;
    const double cb = p, cc = 0.25 * (p * p) - r, cd = (-0.125) * (q * q);
    double m = cubic_root_mixed(cb, cc, cd);
    if (m <= 1e-9 * (1. + fabs(cb))) m = cubic_root_fallback(cb, cc, cd);
    if (m <= 0) return 0;
    double coeff, p1, p2, p3 = 0, smax;
    if (m < 1.E-14) {
        p1 = p + p;
        p2 = (q * q) / f;
        p3 = 2. * sqrt(f);
        smax = p3 - p1 - p2;
        coeff = 2. * fabs(q) / f;
    } else {
        p1 = -2. * (p + m);
        p2 = sqrt(2. * (q * q) / m);
        smax = p1 + p2;
        coeff = sqrt(2. * m);
    }
    roots[0] = 0.5 * (((q >= 0.) ? -coeff : coeff) - sqrt(smax));
    roots[1] = 0.5 * (((q >= 0.) ? -coeff : coeff) + sqrt(smax));
    if (num == 4) {
        const double smin = (m < 1.E-14) ? -p1 - p2 - p3 : p1 - p2;
        roots[2] = 0.5 * (((q >= 0.) ? coeff : -coeff) - sqrt(smin));
        roots[3] = 0.5 * (((q >= 0.) ? coeff : -coeff) + sqrt(smin));
    }
    if (change_variant) {
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (i < num) roots[i] = 1. / (roots[i] - 0.25 * d / e);
    } else {
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (i < num) roots[i] -= 0.25 * b;
    }
    return num;
}

// Bracketed Newton polish of a root s of x^4 + B x^3 + C x^2 + D x + E (the
// closed form is off by up to ~1e-5 A on grazing rays of a 5,000-atom SES).
// A step is taken only when it is a refinement (at most kPolishStep, in A
// along the ray), the quartic changes sign across a bracket of twice the step
// around s, and it lowers |q|; clustered pairs closer than the bracket keep s.
constexpr double kPolishStep = 1e-4;

__device__ __forceinline__ double polish_quartic_root(double s, double B, double C, double D, double E) {
    auto q = [&](double x) { return (((x + B) * x + C) * x + D) * x + E; };
    auto dq = [&](double x) { return ((4 * x + 3 * B) * x + 2 * C) * x + D; };
    for (int it = 0; it < 3; ++it) {
        const double f = q(s), fp = dq(s);
        if (f == 0 || fp == 0) break;
        const double step = f / fp;
        if (fabs(step) <= 1e-12 * (1. + fabs(s))) {  // converged: a step this small cannot change roots
            s -= step;
            break;
        }
        if (!(fabs(step) <= kPolishStep)) break;
        const double h = 2 * fabs(step) + 1e-12;
        if (q(s - h) * q(s + h) > 0) break;
        const double sn = s - step;
        if (fabs(q(sn)) >= fabs(f)) break;
        s = sn;
    }
    return s;
}

// rayTorus, mixed precision: the quartic's resolvent root from an FP32 seed
// refined in FP64 and every quartic root Newton-polished in FP64; the frame
// change, coefficients and clipping are those of ray_torus.
__device__ __forceinline__ int ray_torus_mixed(const double *td, const double sc[3], double probe, double radius, int vc,
                               const double orig[3], double ray_dir, double t[4]) {
    double o1[3] = {orig[0], orig[1], orig[2]};
    o1[vc] += t[0] * ray_dir;
    const double tx = o1[0] - td[0], ty = o1[1] - td[1], tz = o1[2] - td[2];
    const double *iv = td + 3;
    double pp[3];
    for (int i = 0; i < 3; i++) pp[i] = iv[3 * i + 0] * tx + iv[3 * i + 1] * ty + iv[3 * i + 2] * tz;
    const double r2 = probe * probe;
    const double R2 = td[12] * td[12];
    const double pd = pp[0] * iv[0 + vc] + pp[1] * iv[3 + vc] + pp[2] * iv[6 + vc];
    const double p2 = pp[0] * pp[0] + pp[1] * pp[1] + pp[2] * pp[2];
    const double coeff = p2 - r2 - R2;
    const double uz = iv[6 + vc];
    const double B = 4 * pd;
    const double C = 2 * coeff + 4 * pd * pd + 4 * R2 * uz * uz;
    const double D = 4 * pd * coeff + 8 * R2 * pp[2] * uz;
    const double E = coeff * coeff - 4 * R2 * (r2 - pp[2] * pp[2]);
    double roots[4];
    const int nr = quartic_roots_mixed(roots, B, C, D, E);
    const double t0 = t[0];
    int ni = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (i >= nr) break;
        double rt = roots[i];
        if (rt < 0) continue;
        rt = polish_quartic_root(rt, B, C, D, E);
        if (rt < 0) continue;
        double pt[3] = {o1[0], o1[1], o1[2]};
        pt[vc] += rt;
        const double dd = (sc[0] - pt[0]) * (sc[0] - pt[0]) + (sc[1] - pt[1]) * (sc[1] - pt[1]) +
                          (sc[2] - pt[2]) * (sc[2] - pt[2]);
        if (dd < radius * radius) t[ni++] = t0 + rt / ray_dir;
    }
    return ni;
}

// ---------------------------------------------------------------------------
// Sturm-sequence roots of the torus quartic, the reference's path when
// Analytical_Ray_Vs_Torus_Intersection is false: getRealRootsSturm
// (tools.cpp:199-255) over sturm/sturm.cpp:20-281 and sturm/util.cpp:17-112
// (RELERROR 1e-7, MAXPOW 32, MAXIT 800, SMALL_ENOUGH 1e-12, solve.h:12-24).
// The recursive bisection runs from an explicit stack; its tasks write
// disjoint root ranges, so the order they run in does not change the roots.
namespace sturm {
constexpr double kRelErr = 1.0e-7, kSmall = 1.0e-12;
constexpr int kMaxPow = 32, kMaxIt = 800;

struct Poly {
    int ord;
    double c[5];
};

__device__ __forceinline__ double eval(const Poly &p, double x) {
    double f = p.c[p.ord];
    for (int i = p.ord - 1; i >= 0; --i) f = x * f + p.c[i];
    return f;
}

__device__ int modp(const Poly &u, const Poly &v, Poly &r) {
    for (int i = 0; i <= u.ord; ++i) r.c[i] = u.c[i];
    if (v.c[v.ord] < 0.0) {
        for (int k = u.ord - v.ord - 1; k >= 0; k -= 2) r.c[k] = -r.c[k];
        for (int k = u.ord - v.ord; k >= 0; --k)
            for (int j = v.ord + k - 1; j >= k; --j) r.c[j] = -r.c[j] - r.c[v.ord + k] * v.c[j - k];
    } else {
        for (int k = u.ord - v.ord; k >= 0; --k)
            for (int j = v.ord + k - 1; j >= k; --j) r.c[j] -= r.c[v.ord + k] * v.c[j - k];
    }
    int k = v.ord - 1;
    while (k >= 0 && fabs(r.c[k]) < kSmall) {
        r.c[k] = 0.0;
        --k;
    }
    r.ord = k < 0 ? 0 : k;
    return r.ord;
}

__device__ int build(Poly *s) {
    s[0].ord = 4;
    s[1].ord = 3;
    double f = fabs(s[0].c[4] * 4);
    for (int i = 1; i <= 4; ++i) s[1].c[i - 1] = s[0].c[i] * i / f;
    int sp = 2;
    for (; modp(s[sp - 2], s[sp - 1], s[sp]); ++sp) {
        f = -fabs(s[sp].c[s[sp].ord]);
        for (int i = s[sp].ord; i >= 0; --i) s[sp].c[i] /= f;
    }
    s[sp].c[0] = -s[sp].c[0];
    return sp;
}

__device__ int changes(int np, const Poly *s, double a) {
    int n = 0;
    double lf = eval(s[0], a);
    for (int q = 1; q <= np; ++q) {
        const double f = eval(s[q], a);
        if (lf == 0.0 || lf * f < 0) ++n;
        lf = f;
    }
    return n;
}

// modified regula falsi on [a, b]
__device__ bool modrf(const Poly &p, double a, double b, double &val) {
    double fa = p.c[p.ord], fb = fa;
    for (int i = p.ord - 1; i >= 0; --i) {
        fa = a * fa + p.c[i];
        fb = b * fb + p.c[i];
    }
    if (fa * fb > 0.0) return false;
    if (fabs(fa) < kRelErr) {
        val = a;
        return true;
    }
    if (fabs(fb) < kRelErr) {
        val = b;
        return true;
    }
    double lfx = fa;
    for (int its = 0; its < kMaxIt; ++its) {
        const double x = (fb * a - fa * b) / (fb - fa);
        const double fx = eval(p, x);
        if (fabs(x) > kRelErr) {
            if (fabs(fx / x) < kRelErr) {
                val = x;
                return true;
            }
        } else if (fabs(fx) < kRelErr) {
            val = x;
            return true;
        }
        if ((fa * fx) < 0) {
            b = x;
            fb = fx;
            if ((lfx * fx) > 0) fa /= 2;
        } else {
            a = x;
            fa = fx;
            if ((lfx * fx) > 0) fb /= 2;
        }
        lfx = fx;
    }
    return false;
}

struct Task {
    double min, max;
    int atmin, atmax, first;
};

__device__ void bisect(int np, const Poly *s, double min0, double max0, int atmin0, int atmax0, double *roots) {
    Task st[8];
    int top = 0;
    st[top++] = Task{min0, max0, atmin0, atmax0, 0};
    while (top > 0) {
        Task t = st[--top];
        double min = t.min, max = t.max, mid = 0;
        int its;
        if (t.atmin - t.atmax == 1) {
            if (modrf(s[0], min, max, roots[t.first])) continue;
            bool done = false;
            for (its = 0; its < kMaxIt; ++its) {
                mid = (min + max) / 2;
                const int atmid = changes(np, s, mid);
                if (atmid < t.atmax || atmid > t.atmin) {
                    done = true;
                    break;
                }
                if (fabs(mid) > kRelErr) {
                    if (fabs((max - min) / mid) < kRelErr) {
                        roots[t.first] = mid;
                        done = true;
                        break;
                    }
                } else if (fabs(max - min) < kRelErr) {
                    roots[t.first] = mid;
                    done = true;
                    break;
                }
                if (t.atmin - atmid == 0) min = mid;
                else max = mid;
            }
            if (!done) roots[t.first] = mid;
            continue;
        }
        for (its = 0; its < kMaxIt; ++its) {
            mid = (min + max) / 2;
            const int atmid = changes(np, s, mid);
            if (atmid < t.atmax || atmid > t.atmin) break;
            const int n1 = t.atmin - atmid, n2 = atmid - t.atmax;
            if (n1 != 0 && n2 != 0) {
                if (top < 7) {
                    st[top++] = Task{mid, max, atmid, t.atmax, t.first + n1};
                    st[top++] = Task{min, mid, t.atmin, atmid, t.first};
                }
                break;
            }
            if (n1 == 0) min = mid;
            else max = mid;
        }
        if (its == kMaxIt)
            for (int q = t.atmax; q < t.atmin; ++q) roots[t.first + q - t.atmax] = mid;
    }
}

// roots of x^4 + B x^3 + C x^2 + D x + E, ascending; a root the bisection
// gives up on stays NaN (the reference leaves it unwritten) and is dropped
__device__ int quartic(double roots[4], double B, double C, double D, double E) {
    Poly s[5];
    s[0].c[0] = E;
    s[0].c[1] = D;
    s[0].c[2] = C;
    s[0].c[3] = B;
    s[0].c[4] = 1.0;
    for (int i = 0; i < 4; ++i) roots[i] = __longlong_as_double(0x7ff8000000000000ll);
    const int np = build(s);
    // numroots: sign changes at -inf and +inf
    int atpos = 0, atneg = 0;
    double lf = s[0].c[s[0].ord];
    for (int q = 1; q <= np; ++q) {
        const double f = s[q].c[s[q].ord];
        if (lf == 0.0 || lf * f < 0) ++atpos;
        lf = f;
    }
    lf = (s[0].ord & 1) ? -s[0].c[s[0].ord] : s[0].c[s[0].ord];
    for (int q = 1; q <= np; ++q) {
        const double f = (s[q].ord & 1) ? -s[q].c[s[q].ord] : s[q].c[s[q].ord];
        if (lf == 0.0 || lf * f < 0) ++atneg;
        lf = f;
    }
    int atmin = atneg, atmax = atpos;
    if (atmin - atmax == 0) return 0;
    double min = -1.0;
    int nc = changes(np, s, min);
    for (int i = 0; nc != atmin && i != kMaxPow; ++i) {
        min *= 10.0;
        nc = changes(np, s, min);
    }
    if (nc != atmin) atmin = nc;
    double max = 1.0;
    nc = changes(np, s, max);
    for (int i = 0; nc != atmax && i != kMaxPow; ++i) {
        max *= 10.0;
        nc = changes(np, s, max);
    }
    if (nc != atmax) atmax = nc;
    const int n = atmin - atmax;
    if (n <= 0 || n > 4) return n <= 0 ? 0 : 4;
    bisect(np, s, min, max, atmin, atmax, roots);
    return n;
}
}  // namespace sturm

// rayTorus with the Sturm roots; frame change, coefficients and clipping as ray_torus
__device__ __forceinline__ int ray_torus_sturm(const double *td, const double sc[3], double probe, double radius, int vc,
                               const double orig[3], double ray_dir, double t[4]) {
    double o1[3] = {orig[0], orig[1], orig[2]};
    o1[vc] += t[0] * ray_dir;
    const double tx = o1[0] - td[0], ty = o1[1] - td[1], tz = o1[2] - td[2];
    const double *iv = td + 3;
    double pp[3];
    for (int i = 0; i < 3; i++) pp[i] = iv[3 * i + 0] * tx + iv[3 * i + 1] * ty + iv[3 * i + 2] * tz;
    const double r2 = probe * probe;
    const double R2 = td[12] * td[12];
    const double pd = pp[0] * iv[0 + vc] + pp[1] * iv[3 + vc] + pp[2] * iv[6 + vc];
    const double p2 = pp[0] * pp[0] + pp[1] * pp[1] + pp[2] * pp[2];
    const double coeff = p2 - r2 - R2;
    const double uz = iv[6 + vc];
    const double B = 4 * pd;
    const double C = 2 * coeff + 4 * pd * pd + 4 * R2 * uz * uz;
    const double D = 4 * pd * coeff + 8 * R2 * pp[2] * uz;
    const double E = coeff * coeff - 4 * R2 * (r2 - pp[2] * pp[2]);
    double roots[4];
    const int nr = sturm::quartic(roots, B, C, D, E);
    const double t0 = t[0];
    int ni = 0;
    for (int i = 0; i < nr; i++) {
        if (roots[i] < 0) continue;
        double pt[3] = {o1[0], o1[1], o1[2]};
        pt[vc] += roots[i];
        const double dd = (sc[0] - pt[0]) * (sc[0] - pt[0]) + (sc[1] - pt[1]) * (sc[1] - pt[1]) +
                          (sc[2] - pt[2]) * (sc[2] - pt[2]);
        if (dd < radius * radius) t[ni++] = t0 + roots[i] / ray_dir;
    }
    return ni;
}

// Path counters of the counting variant (see intersect_path_counts).
enum { PC_CAND, PC_DISC, PC_SPHERE, PC_TORUS, PC_ROOTS, PC_HITS, PC_FEAS_ITEMS, PC_FEAS_CALLS, PC_N };

// The tori of P.order (spheres and facets are k_intersect_plain's), one
// 64-lane wave per (torus, slot): as its own launch over one frame
// (k_intersect_torus, item = blockIdx.x) or as an item of a multi-frame launch
// (torus_item).  Per-item counts go to the item's own row of 'part' (zeroed
// before the launch) and are summed by k_reduce_cols: one atomic per item on
// a shared address would serialise 10^5 workgroups at the L2.  kRoots: torus
// roots from the closed form (kRootsClosed), the mixed-precision solver
// (kRootsMixed) or Sturm bisection (kRootsSturm).  kCount: instead of hits,
// per-item path counts.
// Five waves per SIMD for the closed-form roots: capped at 96 VGPRs that item
// needs no spills, and 5 waves hide a little more of its memory latency than
// the 4 that its unconstrained 102 VGPRs allow (0.424 -> 0.419 ms per launch
// at 325^3).  The mixed-precision and Sturm items would spill: no cap.
template <bool kCount, int kRoots>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kRoots == kRootsClosed ? 5 : 1)))
k_intersect_torus(DevGrid g, DevPatches P, DevRays R, double probe, unsigned long long *part) {
    // workgroup = patch rank * nslots + slot: the patches of a class come
    // largest first, and every slot starts on them together
    const int rank = (int)(blockIdx.x / R.nslots);
    const int slot = (int)(blockIdx.x % R.nslots);
#include "torus_item.inc"
}

template <bool kCount, int kRoots>
__device__ __forceinline__ void torus_item(const DevGrid &g, const DevPatches &P, const DevRays &R, double probe,
                                           unsigned long long *part, int bid) {
    const int rank = bid / R.nslots;
    const int slot = bid % R.nslots;
#include "torus_item.inc"
}



// Sphere and facet patches (P.order [0, n_plain)), one item (a 64-lane
// wave) per patch over every slot.  The walk over each slot's pixel rectangle (the culls of
// k_intersect) appends the pixels that reach the sphere test to a ring queue
// in LDS; a lane takes the next queued pixel as soon as its own is decided
// and tests its two depths against the patch's items in order, eight at a
// time (isFeasible, ConnollySurface.cpp:4769-5021, the same FP64
// expressions); the walk refills the queue whenever it holds less than a
// wave's worth.  So the lanes stay busy until the patch's last pixel: held
// until the slowest of a batch of 64 was decided (a pixel with a feasible
// depth tests every item and nearly every batch holds one), and drained at
// the end of every slot, the waves tested about 2.8 times the items needed.
// Items are staged in LDS when a patch has at most kStageItems of them.
constexpr int kRing = 128;  // a power of two, >= 2 * 64
constexpr int kStageItems = 64;

template <bool kCount>
__global__ void __launch_bounds__(64) k_intersect_plain(DevGrid g, DevPatches P, DevRays R,
                                                        unsigned long long *part) {
    const int rank = (int)blockIdx.x;
#include "plain_item.inc"
}

template <bool kCount>
__device__ __forceinline__ void plain_item(const DevGrid &g, const DevPatches &P, const DevRays &R,
                                           unsigned long long *part, int rank) {
#include "plain_item.inc"
}


// Writes one ray's coverage plane: node intervals arrive in ascending,
// disjoint order (the segments of a sorted ray), and every one of the W words
// is stored once, in order.
struct PlaneWriter {
    unsigned long long *base;  // word w at base[w * stride]
    int64_t stride;
    int W, cw;
    unsigned long long acc;
    __device__ PlaneWriter(unsigned long long *b, int64_t s, int w) : base(b), stride(s), W(w), cw(0), acc(0ull) {}
    __device__ void add(int lo, int hi) {
        if (lo < cw * 64) lo = cw * 64;  // never reached with disjoint ascending intervals
        while (lo <= hi) {
            const int w = lo >> 6;
            while (cw < w) {
                base[cw * stride] = acc;
                acc = 0ull;
                ++cw;
            }
            const int top = min(hi, (w << 6) + 63);
            const int n = top - lo + 1;
            acc |= (n >= 64 ? ~0ull : ((1ull << n) - 1ull)) << (lo & 63);
            lo = top + 1;
        }
    }
    __device__ void finish() {
        while (cw < W) {
            base[cw * stride] = acc;
            acc = 0ull;
            ++cw;
        }
    }
};

// Ascending sort of a ray's n <= NS depths (and patch ids) in registers: a
// bitonic network over NS slots padded with +inf, every index a constant.
// A pair is exchanged only when the first is strictly greater, as in the
// insertion sort, so the depths come out in the same order.
template <int NS, bool kPid>
__device__ __forceinline__ void sort_small(double *T, int *I, int n) {
    double v[NS];
    int id[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        v[q] = q < n ? T[q] : __longlong_as_double(0x7ff0000000000000ll);
        id[q] = (kPid && q < n) ? I[q] : 0;
    }
#pragma unroll
    for (int k = 2; k <= NS; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < NS; ++i) {
                const int l = i ^ j;
                if (l <= i) continue;
                const bool up = (i & k) == 0;
                const int lo = up ? i : l, hi = up ? l : i;  // after the exchange v[lo] <= v[hi]
                if (v[lo] > v[hi]) {
                    const double t = v[lo];
                    v[lo] = v[hi];
                    v[hi] = t;
                    if (kPid) {
                        const int u = id[lo];
                        id[lo] = id[hi];
                        id[hi] = u;
                    }
                }
            }
#pragma unroll
    for (int q = 0; q < NS; ++q)
        if (q < n) {
            T[q] = v[q];
            if (kPid) I[q] = id[q];
        }
}

template <bool kPid>
__device__ __forceinline__ void sort_depths(double *T, int *I, int n) {
    if (n <= 1) return;
    if (n <= 4) return sort_small<4, kPid>(T, I, n);
    if (n <= 8) return sort_small<8, kPid>(T, I, n);
    if (n <= 16) return sort_small<16, kPid>(T, I, n);
    for (int a = 1; a < n; ++a) {  // longer lists: insertion sort in place
        const double v = T[a];
        const int id = kPid ? I[a] : 0;
        int b = a - 1;
        while (b >= 0 && T[b] > v) {
            T[b + 1] = T[b];
            if (kPid) I[b + 1] = I[b];
            --b;
        }
        T[b + 1] = v;
        if (kPid) I[b + 1] = id;
    }
}

// Per-ray insertion sort of the depths (Surface::reorderPatchIntersections,
// Surface.cpp:5080-5095), plus the ray's segment-length integral of
// setVerticesAndGridsWithIntersectionData (Surface.cpp:5265-5290, :5425-5440)
// and the statistics of the ray lists.
// part: 3 values per block (hits, odd rays, longest list), reduced by k_reduce_cols
__global__ void __launch_bounds__(256) k_sort(DevGrid g, DevRays R, unsigned long long *part) {
    const int64_t nper = (int64_t)R.N * R.N;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = nper * R.nslots;
    int c = 0;
    unsigned int odd = 0;
    if (r < total) {
        c = R.cnt[r];
        const int n = c < R.K ? c : R.K;
        double *T = R.t + r * R.K;
        if (R.pid)
            sort_depths<true>(T, R.pid + r * R.K, n);
        else
            sort_depths<false>(T, nullptr, n);
        const int slot = (int)(r / nper);
        const int phase = slot < R.panels0 ? 0 : 1;
        const int panel = phase == 0 ? slot : slot - R.panels0;
        const double *ga = panel == 0 ? g.x : (panel == 1 ? g.z : g.y);
        const int na = panel == 0 ? g.nx : (panel == 1 ? g.nz : g.ny);
        double pa, dir;
        if (phase == 0) {
            pa = ga[0];
            dir = ga[na - 1] - pa;
        } else {
            const double delta = kDelta - g.hside;
            pa = ga[0] + delta;
            dir = (ga[na - 1] + delta) - pa;
        }
        double vol = 0.0;
        const double gmin = panel == 0 ? g.xmin : (panel == 1 ? g.zmin : g.ymin);
        // node intervals of the map fill, from the same double comparisons
        // setVerticesAndGridsWithIntersectionData makes per node
        // (Surface.cpp:5427-5520, :5636-5650, :5736-5750, :5990-6008),
        // written as coverage bits
        const int64_t ray = r - (int64_t)slot * nper;
        PlaneWriter E(R.cov + (int64_t)slot * R.W * nper + ray, nper, R.W);
        const bool st = phase == 0 && panel == 0;
        PlaneWriter S(st ? R.cov_st + ray : nullptr, nper, st ? R.W : 0);
        auto push = [](PlaneWriter &w, int lo, int hi) {
            if (lo <= hi) w.add(lo, hi);
        };
        if (c <= R.K && (c & 1) == 0) {
            for (int i = 0; i < c; i += 2)
                if (fabs(T[i + 1] - T[i]) >= kEpsInt) {
                    const double lim1 = pa + dir * T[i];
                    const double lim2 = pa + dir * T[i + 1];
                    vol += lim2 - lim1;
                    const int a = (int)rintp((lim1 - gmin) * g.scale);
                    const int b = (int)rintp((lim2 - gmin) * g.scale);
                    if (phase == 1) {
                        // corners a < n <= b
                        if (b > a) push(E, max(a + 1, 0), min(b, na - 1));
                        continue;
                    }
                    const int i1 = a < 0 ? 0 : (a >= na ? na - 1 : a);
                    const int i2 = b < 0 ? 0 : (b >= na ? na - 1 : b);
                    if (i2 < i1) continue;
                    // panel 0 covers [i1, i2], panels 1 and 2 the half-open [i1, i2)
                    const int last = panel == 0 ? i2 : i2 - 1;
                    int ie = last;
                    while (ie >= i1 && !(ga[ie] <= lim2 - g.hside)) --ie;
                    push(E, i1, ie);
                    if (st) {
                        int is = i1, it = i2;
                        while (is <= i2 && !(ga[is] >= lim1)) ++is;
                        while (it >= is && !(ga[it] <= lim2)) --it;
                        push(S, is, it);
                    }
                }
        }
        E.finish();
        S.finish();
        R.rvol[r] = vol;
        odd = (c & 1);
    }
    // block reduction of the statistics
    __shared__ unsigned long long sh[256 / 64][2];
    __shared__ int shm[256 / 64];
    unsigned long long h = (unsigned long long)c, o = odd;
    int mx = c;
    for (int off = 32; off > 0; off >>= 1) {
        h += __shfl_down(h, off, 64);
        o += __shfl_down(o, off, 64);
        mx = max(mx, __shfl_down(mx, off, 64));
    }
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) {
        sh[w][0] = h;
        sh[w][1] = o;
        shm[w] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        h = sh[0][0];
        o = sh[0][1];
        mx = shm[0];
        for (int q = 1; q < 256 / 64; ++q) {
            h += sh[q][0];
            o += sh[q][1];
            mx = max(mx, shm[q]);
        }
        part[3 * (int64_t)blockIdx.x + 0] = h;
        part[3 * (int64_t)blockIdx.x + 1] = o;
        part[3 * (int64_t)blockIdx.x + 2] = (unsigned long long)mx;
    }
}

// out[c] = sum (or max when bit c of max_mask is set) over rows of in[row*ncols + c]
// Column sums (or maxima, columns in max_mask) of a row-major table of
// per-workgroup partials: blockIdx.y splits the rows, each block combines its
// share and adds it to out[c] with one atomic (out is cleared first).
__global__ void __launch_bounds__(1024) k_reduce_cols(const unsigned long long *in, int64_t nrows, int ncols,
                                                      unsigned int max_mask, unsigned long long *out) {
    __shared__ unsigned long long sh[1024 / 64];
    const int c = blockIdx.x;
    const bool is_max = (max_mask >> c) & 1u;
    const int64_t per = (nrows + gridDim.y - 1) / gridDim.y;
    const int64_t r0 = (int64_t)blockIdx.y * per, r1 = min(nrows, r0 + per);
    unsigned long long acc = 0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += 1024) {
        const unsigned long long v = in[r * ncols + c];
        acc = is_max ? (v > acc ? v : acc) : acc + v;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_down(acc, off, 64);
        acc = is_max ? (o > acc ? o : acc) : acc + o;
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        acc = sh[0];
        for (int q = 1; q < 1024 / 64; ++q) acc = is_max ? (sh[q] > acc ? sh[q] : acc) : acc + sh[q];
        if (acc) {
            if (is_max)
                atomicMax(out + c, acc);
            else
                atomicAdd(out + c, acc);
        }
    }
}

// ---------------------------------------------------------------- cell grid
// A uniform grid of kCell x kCell pixel cells over each panel's plane; a cell
// lists the patches whose pixel rectangle (the rectangle the patch-centric
// cast walks: the bounding box rounded to pixels, ConnollySurface.cpp:
// 4378-4410) meets it.  It is the auxiliary 2D grid of the ray-centric path
// (ConnollySurface::preProcessPanel, ConnollySurface.cpp:3318-3560, which bins
// the same bounding boxes) laid on the grid's own pixels.  Built on the device:
// count, scan, scatter, then each cell's list in ascending patch id.

// the pixel rectangle [is0, ie0] x [is1, ie1] of patch it on panel PANEL;
// false when the patch is culled (CELL_CULLING) or the rectangle is empty
template <int PANEL>
__device__ __forceinline__ bool patch_rect(const DevGrid &g, const DevPatches &P, int it, int &is0, int &ie0, int &is1,
                                           int &ie1) {
    using A = PanelAxes<PANEL>;
    double down[3], up[3];
    if (P.kind[it] >= NSGPU_PATCH_SKIN_POINT) {
        const double *pd = P.data + P.off[it];
        for (int q = 0; q < 3; ++q) {
            down[q] = pd[q];
            up[q] = pd[3 + q];
        }
    } else {
        const double *c = P.bsph + 4 * it;
        for (int q = 0; q < 3; ++q) {
            down[q] = c[q] - c[3];
            up[q] = c[q] + c[3];
        }
    }
    if (down[0] > g.xmax + kEpsInt || down[1] > g.ymax + kEpsInt || down[2] > g.zmax + kEpsInt) return false;
    const double lowpad = g.hside - kDelta + kEpsInt;
    if (up[0] < g.xmin - lowpad || up[1] < g.ymin - lowpad || up[2] < g.zmin - lowpad) return false;
    const double gmin[3] = {g.xmin, g.ymin, g.zmin};
    const int gn[3] = {g.nx, g.ny, g.nz};
    int64_t a0 = (int64_t)rintp(fmax(0., down[A::fd] - gmin[A::fd]) * g.scale);
    int64_t b0 = (int64_t)rintp(fmax(0., up[A::fd] - gmin[A::fd]) * g.scale);
    int64_t a1 = (int64_t)rintp(fmax(0., down[A::ld] - gmin[A::ld]) * g.scale);
    int64_t b1 = (int64_t)rintp(fmax(0., up[A::ld] - gmin[A::ld]) * g.scale);
    if (b0 >= gn[A::fd]) b0 = gn[A::fd] - 1;
    if (b1 >= gn[A::ld]) b1 = gn[A::ld] - 1;
    if (b0 < a0 || b1 < a1) return false;
    is0 = (int)a0;
    ie0 = (int)b0;
    is1 = (int)a1;
    ie1 = (int)b1;
    return true;
}

// every (panel, cell) the patch's rectangles meet: count (fill = false) or list
__global__ void __launch_bounds__(256) k_cells_bin(DevGrid g, DevPatches P, CellGrid C, bool fill) {
    const int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= P.n) return;
    const int nc = C.nc, nn = nc * nc;
    for (int panel = 0; panel < 3; ++panel) {
        int is0, ie0, is1, ie1;
        bool ok;
        if (panel == 0) ok = patch_rect<0>(g, P, it, is0, ie0, is1, ie1);
        else if (panel == 1) ok = patch_rect<1>(g, P, it, is0, ie0, is1, ie1);
        else ok = patch_rect<2>(g, P, it, is0, ie0, is1, ie1);
        if (!ok) continue;
        for (int cv = is1 / kCell; cv <= ie1 / kCell; ++cv)
            for (int cu = is0 / kCell; cu <= ie0 / kCell; ++cu) {
                const int c = panel * nn + cv * nc + cu;
                if (!fill) {
                    atomicAdd(C.count + c, 1);
                } else {
                    const int at = atomicAdd(C.cursor + c, 1);
                    C.list[C.off[c] + at] = it;
                }
            }
    }
}

// each cell's list in ascending patch id (insertion sort: lists are short)
__global__ void __launch_bounds__(256) k_cells_sort(CellGrid C) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= 3 * C.nc * C.nc) return;
    int *l = C.list + C.off[c];
    const int n = C.count[c];
    for (int i = 1; i < n; ++i) {
        const int v = l[i];
        int j = i - 1;
        while (j >= 0 && l[j] > v) {
            l[j + 1] = l[j];
            --j;
        }
        l[j + 1] = v;
    }
}

int cell_grid_side(const DevRays &R) { return (R.N + kCell - 1) / kCell; }

void launch_cells_count(hipStream_t s, const DevGrid &g, const DevPatches &P, const CellGrid &C) {
    launch_fill32(s, C.count, 0u, 3 * (int64_t)C.nc * C.nc);
    if (P.n > 0) hipLaunchKernelGGL(k_cells_bin, dim3((P.n + 255) / 256), dim3(256), 0, s, g, P, C, false);
}

void launch_cells_fill(hipStream_t s, const DevGrid &g, const DevPatches &P, const CellGrid &C) {
    const int64_t cells = 3 * (int64_t)C.nc * C.nc;
    launch_fill32(s, C.cursor, 0u, cells);
    if (P.n > 0) hipLaunchKernelGGL(k_cells_bin, dim3((P.n + 255) / 256), dim3(256), 0, s, g, P, C, true);
    hipLaunchKernelGGL(k_cells_sort, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s, C);
}

// One patch of a cell against the lane's ray (pa, ray_dir along vc): the
// bounding box, the disc cull and the roots of getPatchIntersectionData
// (ConnollySurface.cpp:4496-4586; SkinSurface.cpp:3602-3700), appending
// feasible depths (ray parameter in [0, 1] units of ray_dir) via emit.
template <int PANEL, int kRoots, class Emit>
__device__ __forceinline__ bool tile_patch(const DevPatches &P, int it, double probe, const double pa[3],
                                           double ray_dir, double skin_dir, Emit emit) {
    using A = PanelAxes<PANEL>;
    constexpr int fd = A::fd, ld = A::ld, vc = A::vc;
    const int kind = P.kind[it];
    const double *pd = P.data + P.off[it];
    if (kind >= NSGPU_PATCH_SKIN_POINT) {
        if (pa[fd] < pd[fd] || pa[fd] > pd[3 + fd] || pa[ld] < pd[ld] || pa[ld] > pd[3 + ld]) return false;
        if (kind == NSGPU_PATCH_SKIN_POINT) {  // EQ_CULLING (SkinSurface.cpp:3602-3608)
            const double *sp = P.bsph + 4 * it;
            const double dx = sp[fd] - pa[fd], dy = sp[ld] - pa[ld];
            if (dx * dx + dy * dy > sp[3] * sp[3]) return true;
        }
        double t[2];
        if (!ray_quadric(pd + 6, pa, vc, t)) return true;
        double ip0[3] = {pa[0], pa[1], pa[2]}, ip1[3] = {pa[0], pa[1], pa[2]};
        ip0[vc] += t[0];
        ip1[vc] += t[1];
        bool ok0, ok1;
        skin_feasible2(pd + NSGPU_SKIN_HDR, P.cnt[it], ip0, ip1, ok0, ok1);
        if (ok0) emit(t[0] / skin_dir);
        if (ok1) emit(t[1] / skin_dir);
        return true;
    }
    const double sc[3] = {P.bsph[4 * it], P.bsph[4 * it + 1], P.bsph[4 * it + 2]};
    const double radius = P.bsph[4 * it + 3];
    if (pa[fd] < sc[fd] - radius || pa[fd] > sc[fd] + radius || pa[ld] < sc[ld] - radius || pa[ld] > sc[ld] + radius)
        return false;
    const double dx = sc[fd] - pa[fd], dy = sc[ld] - pa[ld];
    if (dx * dx + dy * dy > radius * radius) return true;
    double dir[3] = {0., 0., 0.};
    dir[vc] = ray_dir;
    double t[4];
    if (!ray_sphere(pa, dir, sc, radius, t[0], t[1])) return true;
    const int pcnt = P.cnt[it];
    if (kind == NSGPU_PATCH_TORUS_REG || kind == NSGPU_PATCH_TORUS_SING) {
        int ni;
        if (kRoots == kRootsMixed) ni = ray_torus_mixed(pd, sc, probe, radius, vc, pa, ray_dir, t);
        else if (kRoots == kRootsSturm) ni = ray_torus_sturm(pd, sc, probe, radius, vc, pa, ray_dir, t);
        else ni = ray_torus(pd, sc, probe, radius, vc, pa, ray_dir, t);
        const int pflags = P.flags[it];
        for (int i = 0; i < ni; i++) {
            double ip[3] = {pa[0], pa[1], pa[2]};
            ip[vc] += t[i] * ray_dir;
            if (feasible_torus(kind, pd, pcnt, pflags, ip)) emit(t[i]);
        }
        return true;
    }
    double ip0[3] = {pa[0], pa[1], pa[2]}, ip1[3] = {pa[0], pa[1], pa[2]};
    ip0[vc] += t[0] * ray_dir;
    ip1[vc] += t[1] * ray_dir;
    bool ok0, ok1;
    feasible_plain2(kind, pd, pcnt, ip0, ip1, ok0, ok1);
    if (ok0) emit(t[0]);
    if (ok1) emit(t[1]);
    return true;
}

// One wave per (cell, slot): lane l takes the ray of pixel (8 cu + l % 8,
// 8 cv + l / 8) and walks the cell's patch list, which every lane reads in
// the same order.  The lane owns its ray, so depths are written without
// atomics; the per-ray sort that follows (k_sort) orders them.
template <int PANEL, int kRoots>
__global__ void __launch_bounds__(64) k_cast_tiles(DevGrid g, DevPatches P, DevRays R, CellGrid C, double probe,
                                                   int slot, unsigned long long *part) {
    using A = PanelAxes<PANEL>;
    const int nc = C.nc, cell = blockIdx.x;
    const int cu = cell % nc, cv = cell / nc;
    const int lane = threadIdx.x;
    const int m = cu * kCell + (lane % kCell), n = cv * kCell + lane / kCell;
    const int gn[3] = {g.nx, g.ny, g.nz};
    const bool valid = m < gn[A::fd] && n < gn[A::ld];
    const double delta = slot < R.panels0 ? 0. : kDelta - g.hside;
    const double *ga = A::vc == 0 ? g.x : (A::vc == 1 ? g.y : g.z);
    const double *gf = A::fd == 0 ? g.x : g.y;
    const double *gl = A::ld == 1 ? g.y : g.z;
    const double gmin[3] = {g.xmin, g.ymin, g.zmin}, gmax[3] = {g.xmax, g.ymax, g.zmax};
    const double ray_dir = ga[gn[A::vc] - 1] - ga[0], skin_dir = gmax[A::vc] - gmin[A::vc];
    double pa[3];
    pa[A::vc] = ga[0] + delta;
    pa[A::fd] = (valid ? gf[m] : gf[0]) + delta;
    pa[A::ld] = (valid ? gl[n] : gl[0]) + delta;
    const int64_t NN = (int64_t)R.N * R.N, r = (int64_t)slot * NN + (int64_t)n * R.N + m;
    double *tout = R.t + r * R.K;
    int *pout = R.pid ? R.pid + r * R.K : nullptr;
    int cnt = 0;
    unsigned int cand = 0;
    const int pc = PANEL * nc * nc + cell;
    const int b = C.off[pc], e = b + C.count[pc];
    for (int q = b; q < e; ++q) {
        const int it = C.list[q];
        if (!valid) continue;
        const bool in_box = tile_patch<PANEL, kRoots>(P, it, probe, pa, ray_dir, skin_dir, [&](double t) {
            if (cnt < R.K) {
                tout[cnt] = t;
                if (pout) pout[cnt] = it;
            }
            ++cnt;
        });
        cand += in_box;
    }
    if (valid) R.cnt[r] = cnt;
    for (int off = 32; off > 0; off >>= 1) cand += __shfl_down(cand, off, 64);
    if (lane == 0) part[(int64_t)slot * nc * nc + cell] = cand;
}

int64_t tile_workgroups(const DevRays &R, int nc) { return (int64_t)nc * nc * R.nslots; }

void launch_cast_tiles(hipStream_t s, const DevGrid &g, const DevPatches &P, const DevRays &R, const CellGrid &C,
                       double probe, unsigned long long *part, int roots) {
    const dim3 grid((unsigned)((int64_t)C.nc * C.nc)), block(64);
    for (int slot = 0; slot < R.nslots; ++slot) {
        const int panel = slot < R.panels0 ? slot : slot - R.panels0;
#define NSG_TILES(PN, RT) hipLaunchKernelGGL((k_cast_tiles<PN, RT>), grid, block, 0, s, g, P, R, C, probe, slot, part)
#define NSG_TILES_P(RT)               \
    if (panel == 0) NSG_TILES(0, RT); \
    else if (panel == 1) NSG_TILES(1, RT); \
    else NSG_TILES(2, RT)
        if (roots == kRootsMixed) { NSG_TILES_P(kRootsMixed); }
        else if (roots == kRootsSturm) { NSG_TILES_P(kRootsSturm); }
        else { NSG_TILES_P(kRootsClosed); }
#undef NSG_TILES_P
#undef NSG_TILES
    }
}

// getRayIntersection for arbitrary axis-aligned rays (castAxisOrientedRay,
// nanoshaper.cpp:220-297; ConnollySurface::getRayIntersection,
// ConnollySurface.cpp:5494-5608): one thread per ray o -> o + len e_vc, the
// patches of the ray's cell, their feasible roots as coordinates along the
// axis, sorted, with depths closer than EPS_INT to the last kept one dropped.
// count[r] = hits kept (-1: len <= 0; -2: the ray misses the grid's cells);
// more than maxh hits keep the first maxh in depth order and report all.
template <int PANEL>
__global__ void __launch_bounds__(64) k_cast_rays(DevGrid g, DevPatches P, CellGrid C, double probe, int n,
                                                  const double *orig, const double *coord, int maxh, int *count,
                                                  double *hits, double *normals) {
    using A = PanelAxes<PANEL>;
    constexpr int vc = A::vc, fd = A::fd, ld = A::ld;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const double pa[3] = {orig[3 * r], orig[3 * r + 1], orig[3 * r + 2]};
    const double len = coord[r] - pa[vc];
    if (!(len > 0)) {
        count[r] = -1;
        return;
    }
    const double gmin[3] = {g.xmin, g.ymin, g.zmin};
    const int gn[3] = {g.nx, g.ny, g.nz};
    const double pf = rintp((pa[fd] - gmin[fd]) * g.scale), pl = rintp((pa[ld] - gmin[ld]) * g.scale);
    if (!(pf >= 0 && pl >= 0 && pf < gn[fd] && pl < gn[ld])) {
        count[r] = -2;
        return;
    }
    const int cell = PANEL * C.nc * C.nc + ((int)pl / kCell) * C.nc + (int)pf / kCell;
    double *H = hits + (int64_t)r * maxh;
    double *Nm = normals ? normals + (int64_t)r * maxh * 3 : nullptr;
    int total = 0;
    const int b = C.off[cell], e = b + C.count[cell];
    for (int q = b; q < e; ++q) {
        const int it = C.list[q];
        tile_patch<PANEL, kRootsClosed>(P, it, probe, pa, len, 1.0, [&](double t) {
            // the skin form returns a distance along the axis, the others a
            // fraction of len; both become a coordinate
            const double a = P.kind[it] >= NSGPU_PATCH_SKIN_POINT ? pa[vc] + t : pa[vc] + len * t;
            double nrm[3] = {0., 0., 0.};
            if (Nm) {
                double y[3] = {pa[0], pa[1], pa[2]};
                y[vc] = a;
                analytic_normal(P, it, probe, y, nrm);
            }
            // insertion into the ray's sorted list (ties keep arrival order)
            int pos = total < maxh ? total : maxh;
            while (pos > 0 && H[pos - 1] > a) {
                if (pos < maxh) {
                    H[pos] = H[pos - 1];
                    if (Nm)
                        for (int k = 0; k < 3; ++k) Nm[3 * pos + k] = Nm[3 * (pos - 1) + k];
                }
                --pos;
            }
            if (pos < maxh) {
                H[pos] = a;
                if (Nm)
                    for (int k = 0; k < 3; ++k) Nm[3 * pos + k] = nrm[k];
            }
            ++total;
        });
    }
    // near-coincident depths (castAxisOrientedRay's clean-up): keep a depth
    // only when it is at least EPS_INT from the last kept one
    const int kept_in = total < maxh ? total : maxh;
    int kept = 0;
    double last = INFINITY;
    for (int q = 0; q < kept_in; ++q) {
        if (fabs(H[q] - last) >= kEpsInt) {
            H[kept] = H[q];
            if (Nm)
                for (int k = 0; k < 3; ++k) Nm[3 * kept + k] = Nm[3 * q + k];
            ++kept;
            last = H[q];
        }
    }
    count[r] = kept + (total - kept_in);
}

void launch_cast_rays(hipStream_t s, const DevGrid &g, const DevPatches &P, const CellGrid &C, double probe,
                      int panel, int n, const double *orig, const double *coord, int maxh, int *count, double *hits,
                      double *normals) {
    if (n <= 0) return;
    const dim3 grid((unsigned)((n + 63) / 64)), block(64);
    if (panel == 0) hipLaunchKernelGGL(k_cast_rays<0>, grid, block, 0, s, g, P, C, probe, n, orig, coord, maxh, count, hits, normals);
    else if (panel == 1) hipLaunchKernelGGL(k_cast_rays<1>, grid, block, 0, s, g, P, C, probe, n, orig, coord, maxh, count, hits, normals);
    else hipLaunchKernelGGL(k_cast_rays<2>, grid, block, 0, s, g, P, C, probe, n, orig, coord, maxh, count, hits, normals);
}

// The intersection of one or many frames: every frame's items of one class
// (0: spheres and facets, one item per patch over every slot; 1: tori, one
// item per (torus, slot)).  A single getSurf is the one-frame case.  Frames
// come largest first and each frame's items largest patch first, so the long
// items are taken early and the launch ends on short ones.  The frames'
// descriptors are read from device memory through the constant address
// space, as wave-uniform scalar loads.

// the frame holding item i: the last whose first item is <= i
template <int kClass>
__device__ __forceinline__ int frame_of(const BatchCast &B, int64_t i) {
    int lo = 0, hi = B.nf - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (B.f[mid].first[kClass] <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

template <int kClass, int kRoots>
__device__ __forceinline__ void run_item(const BatchCast &B, int64_t i) {
    typedef const __attribute__((address_space(4))) FrameCast ConstFrame;
    ConstFrame *F = (ConstFrame *)(B.f + frame_of<kClass>(B, i));
    if (kClass == 0)
        plain_item<false>((const DevGrid &)F->g, (const DevPatches &)F->P, (const DevRays &)F->R, F->part,
                          (int)(i - F->first[0]));
    else
        torus_item<false, kRoots>((const DevGrid &)F->g, (const DevPatches &)F->P, (const DevRays &)F->R,
                                  F->probe, F->part, (int)(i - F->first[1]));
}

// workgroup b runs item b
template <int kClass, int kRoots>
__global__ void __launch_bounds__(64) k_intersect_frames(BatchCast B) {
    if ((int64_t)blockIdx.x < B.items[kClass]) run_item<kClass, kRoots>(B, blockIdx.x);
}

// Persistent: each workgroup takes the next B.chunk items from the class's
// queue until the list is used up (one atomic per chunk: with one per item
// the queue's address serialises millions of returns at the L2).  The loop
// keeps the item's invariant values (LDS addresses, lane masks, constants)
// live across iterations, which costs a wave per SIMD against the one-item
// kernel (3 resident instead of 4).
template <int kClass, int kRoots>
__global__ void __launch_bounds__(64) k_intersect_persistent(BatchCast B) {
    const int64_t total = B.items[kClass];
    for (;;) {
        unsigned int v = 0;
        if (threadIdx.x == 0) v = atomicAdd(B.queue + kClass, (unsigned)B.chunk);
        // lane 0's value, as a wave-uniform scalar
        const int64_t i0 = (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)v);
        if (i0 >= total) return;
        const int64_t i1 = min(total, i0 + B.chunk);
        for (int64_t i = i0; i < i1; ++i) {
            run_item<kClass, kRoots>(B, i);
            __syncthreads();  // the next item reuses the LDS queues
        }
    }
}

static void launch_class(hipStream_t s, const BatchCast &B, int cls, int roots, unsigned blocks) {
    if (cls == 0)
        hipLaunchKernelGGL((k_intersect_frames<0, kRootsClosed>), dim3(blocks), dim3(64), 0, s, B);
    else if (roots == kRootsMixed)
        hipLaunchKernelGGL((k_intersect_frames<1, kRootsMixed>), dim3(blocks), dim3(64), 0, s, B);
    else if (roots == kRootsSturm)
        hipLaunchKernelGGL((k_intersect_frames<1, kRootsSturm>), dim3(blocks), dim3(64), 0, s, B);
    else
        hipLaunchKernelGGL((k_intersect_frames<1, kRootsClosed>), dim3(blocks), dim3(64), 0, s, B);
}

int intersect_frames_occupancy(int cls, int roots) {
    int n = 0;
    const void *k = cls == 0                ? (const void *)k_intersect_persistent<0, kRootsClosed>
                    : roots == kRootsMixed   ? (const void *)k_intersect_persistent<1, kRootsMixed>
                    : roots == kRootsSturm   ? (const void *)k_intersect_persistent<1, kRootsSturm>
                                             : (const void *)k_intersect_persistent<1, kRootsClosed>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, 64, 0) != hipSuccess) return 0;
    return n;
}

void launch_intersect_frames(hipStream_t s, const BatchCast &B, int cls, int roots, int grid) {
    const int64_t n = B.items[cls];
    if (n <= 0) return;
    if (!B.persistent) {
        launch_class(s, B, cls, roots, (unsigned)n);
        return;
    }
    const dim3 blocks((unsigned)std::max<int64_t>(1, std::min<int64_t>(grid, n)));
    if (cls == 0)
        hipLaunchKernelGGL((k_intersect_persistent<0, kRootsClosed>), blocks, dim3(64), 0, s, B);
    else if (roots == kRootsMixed)
        hipLaunchKernelGGL((k_intersect_persistent<1, kRootsMixed>), blocks, dim3(64), 0, s, B);
    else if (roots == kRootsSturm)
        hipLaunchKernelGGL((k_intersect_persistent<1, kRootsSturm>), blocks, dim3(64), 0, s, B);
    else
        hipLaunchKernelGGL((k_intersect_persistent<1, kRootsClosed>), blocks, dim3(64), 0, s, B);
}

int64_t intersect_workgroups(const DevPatches &P, const DevRays &R) { return (int64_t)P.n * R.nslots; }
int64_t sort_blocks(const DevRays &R) { return ((int64_t)R.N * R.N * R.nslots + 255) / 256; }

// part: intersect_workgroups() rows, zeroed by the caller.  Two launches: the
// plain patches, then the tori (they need about twice the registers).
template <bool kCount, int kRoots>
static void launch_classes(hipStream_t s, const DevGrid &g, const DevPatches &P, const DevRays &R, double probe,
                           unsigned long long *part) {
    if (P.n == 0 || R.nslots == 0) return;
    if (P.n_plain > 0)
        hipLaunchKernelGGL((k_intersect_plain<kCount>), dim3((unsigned)P.n_plain), dim3(64), 0, s, g, P, R, part);
    if (P.n - P.n_skin > P.n_plain)
        hipLaunchKernelGGL((k_intersect_torus<kCount, kRoots>),
                           dim3((unsigned)(P.n - P.n_skin - P.n_plain) * (unsigned)R.nslots), dim3(64), 0, s, g,
                           P, R, probe, part);
    if (!kCount && P.n_skin > 0) launch_intersect_skin(s, g, P, R, part);
}
void launch_intersect(hipStream_t s, const DevGrid &g, const DevPatches &P, const DevRays &R, double probe,
                      unsigned long long *part, int roots) {
    if (roots == kRootsMixed)
        launch_classes<false, kRootsMixed>(s, g, P, R, probe, part);
    else if (roots == kRootsSturm)
        launch_classes<false, kRootsSturm>(s, g, P, R, probe, part);
    else
        launch_classes<false, kRootsClosed>(s, g, P, R, probe, part);
}
// Counting variant: same candidate walk, no hit output; part gets PC_N
// counters per workgroup (zeroed by the caller): how often each FP64 path of
// the reference's candidate set runs.
void launch_intersect_count(hipStream_t s, const DevGrid &g, const DevPatches &P, const DevRays &R, double probe,
                            unsigned long long *part) {
    launch_classes<true, kRootsClosed>(s, g, P, R, probe, part);
}

// part: sort_blocks() rows of 3
void launch_sort(hipStream_t s, const DevGrid &g, const DevRays &R, unsigned long long *part) {
    hipLaunchKernelGGL(k_sort, dim3((unsigned)sort_blocks(R)), dim3(256), 0, s, g, R, part);
}

void launch_reduce_cols(hipStream_t s, const unsigned long long *in, int64_t nrows, int ncols, unsigned int max_mask,
                        unsigned long long *out, bool zero_out) {
    if (zero_out) launch_fill32(s, out, 0u, 2 * (int64_t)ncols);
    const int64_t parts = std::min<int64_t>(std::max<int64_t>((nrows + 16383) / 16384, 1), 256);
    hipLaunchKernelGGL(k_reduce_cols, dim3((unsigned)ncols, (unsigned)parts), dim3(1024), 0, s, in, nrows, ncols,
                       max_mask, out);
}

}  // namespace nsg
