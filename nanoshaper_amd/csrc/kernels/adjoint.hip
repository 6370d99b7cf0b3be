// Adjoint of the enclosed volume and of the normal-based area estimate with
// respect to the atom centres and radii.
//
// Volume: V = A/3 * sum over the three grid-ray panels of the kept ray
// segments (lim2 - lim1), the getSurf volume with three panels
// (Surface.cpp:1677-1684).  Area estimate: S = A * sum over the same hits of
// 1/|n|_1 (n the unit surface normal): a patch of area dS is crossed by the
// rays of axis a at a density |n_a|/A, so the three panels together count it
// |n|_1 dS / A times.
//
// Every hit lies on the sphere or torus that its patch is cut from; the
// implicit function f(x; atoms) of that surface gives the depth derivative
// dt/dtheta = -(df/dtheta) / (grad f . ray) (implicit differentiation) and the
// normal derivative follows from grad f at the moved hit point:
// d(grad f)/dtheta + (dgrad f/dx_ray) L dt/dtheta.  Both are evaluated in
// forward mode with first-order dual numbers, one pass per parameter plus one
// pass seeded along the ray axis for the second term:
//   point cell: f = |x - c|^2 - r^2 of its atom;
//   probe cap (facet): f = |x - p|^2 - rp^2, the probe centre p tangent to its
//     three atoms (|p - c_m| = r_m + rp); dp follows from the three
//     linearised tangency conditions;
//   torus: f = (|y|^2 + R^2 - rp^2)^2 - 4 R^2 (|y|^2 - (y.a)^2), y = x - cc,
//     with axis a, circle centre cc and radius R in closed form from its two
//     atoms (the probe circle tangent to both inflated spheres).
#include "launch.hpp"

namespace nsg {

struct D1 {
    double v, d;
};
__device__ __forceinline__ D1 operator+(D1 a, D1 b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ D1 operator-(D1 a, D1 b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ D1 operator*(D1 a, D1 b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ D1 operator*(double s, D1 a) { return {s * a.v, s * a.d}; }
__device__ __forceinline__ D1 operator/(D1 a, D1 b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
__device__ __forceinline__ D1 dsqrt(D1 a) {
    const double s = sqrt(a.v);
    return {s, 0.5 * a.d / s};
}
__device__ __forceinline__ D1 dabs(D1 a) { return a.v < 0 ? D1{-a.v, -a.d} : a; }
__device__ __forceinline__ D1 dot3(const D1 a[3], const D1 b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

struct HitSurface {
    int kind, nat;
    int atom[3];
    double c[3][3], r[3];  // atom centres and radii
    double p[3];           // probe centre (facet)
    double minv[3][3];     // inverse of the facet tangency matrix, rows p - c_m
};

// f and grad_x f at x; parameter (atom ks, component kq: x, y, z, radius)
// carries the derivative seed (ks = 3: no parameter seed, only x's).  Inlined
// into loops over constant ks, kq, so every array index below is a constant.
__device__ __forceinline__ void surface_field(const HitSurface &h, double rp, int ks, int kq, const D1 x[3], D1 &f,
                                             D1 grad[3]) {
    D1 c[3][3], r[3];
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        for (int q = 0; q < 3; ++q) c[s][q] = {h.c[s][q], (s == ks && q == kq) ? 1. : 0.};
        r[s] = {h.r[s], (s == ks && kq == 3) ? 1. : 0.};
    }
    if (h.kind == NSGPU_PATCH_POINT) {
        D1 y[3] = {x[0] - c[0][0], x[1] - c[0][1], x[2] - c[0][2]};
        f = dot3(y, y) - r[0] * r[0];
        for (int q = 0; q < 3; ++q) grad[q] = 2. * y[q];
        return;
    }
    if (h.kind == NSGPU_PATCH_FACET_REG || h.kind == NSGPU_PATCH_FACET_SING) {
        // (p - c_m) . (dp - dc_m) = (r_m + rp) dr_m  =>  dp = Minv b, b only in row ks
        const double gk = ks >= 3 ? 0. : ((kq < 3) ? (h.p[kq] - h.c[ks][kq]) : (h.r[ks] + rp));
        D1 y[3];
        for (int q = 0; q < 3; ++q) y[q] = x[q] - D1{h.p[q], ks >= 3 ? 0. : h.minv[q][ks] * gk};
        f = dot3(y, y) - D1{rp * rp, 0.};
        for (int q = 0; q < 3; ++q) grad[q] = 2. * y[q];
        return;
    }
    D1 D[3] = {c[1][0] - c[0][0], c[1][1] - c[0][1], c[1][2] - c[0][2]};
    const D1 d = dsqrt(dot3(D, D));
    D1 a[3] = {D[0] / d, D[1] / d, D[2] / d};
    const D1 ri = r[0] + D1{rp, 0.}, rj = r[1] + D1{rp, 0.};
    const D1 hh = (ri * ri - rj * rj + d * d) / (2. * d);
    const D1 R2 = ri * ri - hh * hh;
    D1 y[3];
    for (int q = 0; q < 3; ++q) y[q] = x[q] - (c[0][q] + a[q] * hh);
    const D1 yy = dot3(y, y), ya = dot3(y, a);
    const D1 s2 = yy + R2 - D1{rp * rp, 0.};
    f = s2 * s2 - 4. * (R2 * (yy - ya * ya));
    for (int q = 0; q < 3; ++q) grad[q] = 4. * (s2 * y[q]) - 8. * (R2 * (y[q] - ya * a[q]));
}

__device__ __forceinline__ void inv3(const double m[3][3], double o[3][3]) {
    const double c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    const double c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    const double c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    const double det = m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02;
    const double id = 1.0 / det;
    o[0][0] = c00 * id;
    o[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
    o[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
    o[1][0] = c01 * id;
    o[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
    o[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
    o[2][0] = c02 * id;
    o[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
    o[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
}

// One hit: its volume and area derivatives for parameter (atom s, component
// q) in dv, da[4 s + q] (zero past the patch's atoms, whose ids go to atom[]),
// and its area term 1/|n|_1 as the result.  VC is the ray axis.
struct HitGrad {
    double dv[12], da[12];
    int atom[3];
    int nat;
};
template <int VC>
__device__ __forceinline__ double hit_adjoint(const DevPatches &P, const AdjointArgs &a, int pid, const double x0[3],
                                              double L, double sgn, HitGrad &out) {
    HitSurface h;
    h.kind = P.kind[pid];
    h.nat = h.kind == NSGPU_PATCH_POINT ? 1 : (h.kind <= NSGPU_PATCH_FACET_SING ? 3 : 2);
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const int id = s < h.nat ? P.atom[3 * pid + s] : 0;
        h.atom[s] = id;
        out.atom[s] = id;
        for (int q = 0; q < 3; ++q) h.c[s][q] = s < h.nat ? a.atoms[4 * id + q] : 0.;
        h.r[s] = s < h.nat ? a.atoms[4 * id + 3] : 0.;
    }
    if (h.kind == NSGPU_PATCH_FACET_REG || h.kind == NSGPU_PATCH_FACET_SING) {
        double m[3][3];
        for (int q = 0; q < 3; ++q) h.p[q] = P.bsph[4 * pid + q];
        for (int s = 0; s < 3; ++s)
            for (int q = 0; q < 3; ++q) m[s][q] = h.p[q] - h.c[s][q];
        inv3(m, h.minv);
    }
    out.nat = h.nat;
    const double vscale = sgn * a.area_px * L / a.flag;
    // d(grad f)/dx along the ray axis (no parameter seed)
    double hx[3];
    {
        D1 x[3] = {{x0[0], 0.}, {x0[1], 0.}, {x0[2], 0.}};
        x[VC].d = 1.;
        D1 f, gr[3];
        surface_field(h, a.probe, 3, 0, x, f, gr);
        for (int q = 0; q < 3; ++q) hx[q] = gr[q].d;
    }
    double inv_v = 0.;
#pragma unroll
    for (int ks = 0; ks < 3; ++ks) {
#pragma unroll
        for (int kq = 0; kq < 4; ++kq) {
            out.dv[4 * ks + kq] = 0.;
            out.da[4 * ks + kq] = 0.;
            if (ks >= h.nat) continue;
            // depth derivative at the fixed point: dt = -(df/dk) / (grad f . e L)
            const D1 x[3] = {{x0[0], 0.}, {x0[1], 0.}, {x0[2], 0.}};
            D1 f, gr[3];
            surface_field(h, a.probe, ks, kq, x, f, gr);
            const double ft = gr[VC].v * L;
            const double dt = (ft != 0.) ? -f.d / ft : 0.;
            // normal at the moved point: the parameter's own term plus the
            // ray-axis term at the depth change
            D1 gm[3];
            for (int q = 0; q < 3; ++q) gm[q] = D1{gr[q].v, gr[q].d + (L * dt) * hx[q]};
            const D1 gn = dsqrt(dot3(gm, gm));
            const D1 l1 = dabs(gm[0] / gn) + dabs(gm[1] / gn) + dabs(gm[2] / gn);
            const D1 inv = D1{1., 0.} / l1;
            inv_v = inv.v;
            out.dv[4 * ks + kq] = vscale * dt;
            out.da[4 * ks + kq] = a.area_px * inv.d;
        }
    }
    return inv_v;
}

// depth e of grid ray r (panel PANEL) when its segment is kept: the
// segment's length (from its entry depth) and the hit's area term
template <int PANEL>
__device__ __forceinline__ void hit_slot(const DevGrid &g, const DevPatches &P, const DevRays &R, const AdjointArgs &a,
                                         int64_t r, int64_t ray, int e, double &vsum, double &asum, int &key,
                                         HitGrad &hg) {
    constexpr int fd = PanelAxes<PANEL>::fd, ld = PanelAxes<PANEL>::ld, vc = PanelAxes<PANEL>::vc;
    const double *ga = vc == 0 ? g.x : (vc == 1 ? g.y : g.z);
    const double *gf = fd == 0 ? g.x : g.y;
    const double *gl = ld == 1 ? g.y : g.z;
    const int na = vc == 0 ? g.nx : (vc == 1 ? g.ny : g.nz);
    const int m = (int)(ray % R.N), n = (int)(ray / R.N);
    const double L = ga[na - 1] - ga[0];
    double pa[3];
    pa[vc] = ga[0];
    pa[fd] = gf[m];
    pa[ld] = gl[n];
    const double *T = R.t + r * R.K;
    const int i = e & ~1;
    if (fabs(T[i + 1] - T[i]) < kEpsInt) return;
    if (e == i) vsum += (pa[vc] + L * T[i + 1]) - (pa[vc] + L * T[i]);
    double x[3] = {pa[0], pa[1], pa[2]};
    x[vc] += L * T[e];
    key = R.pid[r * R.K + e];
    asum += hit_adjoint<vc>(P, a, key, x, L, (e & 1) ? 1. : -1., hg);
}

// One thread per depth slot, slot-major: thread t takes depth e = t / nrays of
// ray t % nrays, so a wave covers 64 neighbouring rays at the same depth index
// (mostly the same patches, so its lanes take the same branches), and the
// launch has K times the rays' parallelism.  Slots past a ray's count exit.
__global__ void __launch_bounds__(256) k_adjoint(DevGrid g, DevPatches P, DevRays R, AdjointArgs a) {
    __shared__ double sh[2][4];
    const int64_t nper = (int64_t)R.N * R.N, nrays = nper * R.panels0;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double vsum = 0., asum = 0.;
    int key = -1;  // patch of the lane's hit (-1: none)
    HitGrad hg;
    if (t < nrays * R.K) {
        const int e = (int)(t / nrays);
        const int64_t r = t - (int64_t)e * nrays;
        const int c = R.cnt[r];
        if (e < c && c <= R.K && (c & 1) == 0) {
            const int panel = (int)(r / nper);
            const int64_t ray = r - panel * nper;
            if (panel == 0) hit_slot<0>(g, P, R, a, r, ray, e, vsum, asum, key, hg);
            else if (panel == 1) hit_slot<1>(g, P, R, a, r, ray, e, vsum, asum, key, hg);
            else hit_slot<2>(g, P, R, a, r, ray, e, vsum, asum, key, hg);
        }
    }
    // Neighbouring lanes often hit the same patch, so the same atoms: the
    // derivatives are summed over each run of lanes with one patch (a
    // segmented scan across the wave) and the run's last lane adds them to
    // the gradients, one atomic per parameter instead of one per hit.
    {
        const int lane = threadIdx.x & 63;
        if (key < 0) {
#pragma unroll
            for (int q = 0; q < 12; ++q) hg.dv[q] = hg.da[q] = 0.;
        }
        const int prev = __shfl_up(key, 1, 64);
        const unsigned long long heads = __ballot(lane == 0 || prev != key);
        const int start = 63 - __builtin_clzll(heads & ((2ull << lane) - 1ull));
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const bool take = lane - off >= start;
#pragma unroll
            for (int q = 0; q < 12; ++q) {
                const double yv = __shfl_up(hg.dv[q], off, 64), ya = __shfl_up(hg.da[q], off, 64);
                if (take) {
                    hg.dv[q] += yv;
                    hg.da[q] += ya;
                }
            }
        }
        const int next = __shfl_down(key, 1, 64);
        if (key >= 0 && (lane == 63 || next != key)) {
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                if (s >= hg.nat) break;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    atomicAdd(a.grad_v + 4 * hg.atom[s] + q, hg.dv[4 * s + q]);
                    atomicAdd(a.grad_a + 4 * hg.atom[s] + q, hg.da[4 * s + q]);
                }
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        vsum += __shfl_down(vsum, off, 64);
        asum += __shfl_down(asum, off, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[0][w] = vsum;
        sh[1][w] = asum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a.sums[2 * blockIdx.x + 0] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
        a.sums[2 * blockIdx.x + 1] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
    }
}

int64_t adjoint_blocks(const DevRays &R) { return ((int64_t)R.N * R.N * R.panels0 * R.K + 255) / 256; }

void launch_adjoint(hipStream_t s, const DevGrid &g, const DevPatches &P, const DevRays &R, const AdjointArgs &a) {
    const int64_t nb = adjoint_blocks(R);
    if (nb > 0) hipLaunchKernelGGL(k_adjoint, dim3((unsigned)nb), dim3(256), 0, s, g, P, R, a);
}

}  // namespace nsg
