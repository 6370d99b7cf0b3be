// Triangle-mesh surfaces: the blobby build (scalar field, marching cubes on
// the field, one smoothing pass, the OFF text round trip) and the ray/triangle
// intersection that replaces the patch intersection for meshes.
//
// BlobbySurface::build (BlobbySurface.cpp:83-229) sums
//   phi(node) = sum over atoms of exp(B (|p - c|^2 / r^2 - 1)),  p = node - hside,
// over the nodes of each atom's box [ix - ng, ix + ng) (clamped to the grid),
// atom after atom.  Here every brick of 8^3 nodes gathers the atoms whose
// boxes overlap it, orders them by index in LDS and sums them in that order,
// so each node adds the same terms in the same order as the reference.
//
// MeshSurface ray casting (Möller-Trumbore, MeshSurface.cpp:1487-1553): one
// lane per triangle walks, panel slot after panel slot, the rays that test it:
//   ray-based (Patch_Based_Algorithm = false, the blobby path): the rays whose
//     auxiliary 2D cell (preProcessPanel :338-575, getRayIntersection
//     :1756-1855) lies in the triangle's cell range; direction pb - pa;
//   patch-based (getPatchIntersectionData :1264-1484): the rays of the
//     triangle's grid rectangle inside its bounding box; direction
//     x[N-1] - x[0].
#include <type_traits>

#include "launch.hpp"

namespace nsg {

static inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// ------------------------------------------------------------ blobby field
constexpr int kBrick = 8;         // nodes per brick side
constexpr int kBrickCap = 1024;   // atoms per brick list (LDS); a dense protein reaches ~400

// box [s, e) of nodes per axis of atom a (BlobbySurface.cpp:125-155)
__device__ __forceinline__ void atom_box(const DevGrid &g, const double *xyzr, int ng, int b[6]) {
    const int ix = (int)rintp((xyzr[0] - g.xmin) * g.scale);
    const int iy = (int)rintp((xyzr[1] - g.ymin) * g.scale);
    const int iz = (int)rintp((xyzr[2] - g.zmin) * g.scale);
    b[0] = max(0, ix - ng);
    b[1] = min(g.nx, ix + ng);
    b[2] = max(0, iy - ng);
    b[3] = min(g.ny, iy + ng);
    b[4] = max(0, iz - ng);
    b[5] = min(g.nz, iz + ng);
}

// per atom: its box (6 ints) and, per overlapped brick, a count or a list entry
__global__ void __launch_bounds__(256) k_blobby_pairs(DevGrid g, BlobbyArgs a, int scatter) {
    const int at = blockIdx.x * blockDim.x + threadIdx.x;
    if (at >= a.n) return;
    int b[6];
    atom_box(g, a.xyzr + 4 * (int64_t)at, a.ng, b);
    if (!scatter) {
        // the atom's record for the field pass: box and r^2 in 32 bytes
        for (int q = 0; q < 6; ++q) a.box[8 * at + q] = b[q];
        const double r = a.xyzr[4 * (int64_t)at + 3];
        *(double *)(a.box + 8 * at + 6) = r * r;
    }
    if (b[0] >= b[1] || b[2] >= b[3] || b[4] >= b[5]) return;
    const int bx0 = b[0] / kBrick, bx1 = (b[1] - 1) / kBrick;
    const int by0 = b[2] / kBrick, by1 = (b[3] - 1) / kBrick;
    const int bz0 = b[4] / kBrick, bz1 = (b[5] - 1) / kBrick;
    for (int bz = bz0; bz <= bz1; ++bz)
        for (int by = by0; by <= by1; ++by)
            for (int bx = bx0; bx <= bx1; ++bx) {
                const int br = (bz * a.nby + by) * a.nbx + bx;
                if (!scatter) {
                    atomicAdd(&a.bcount[br], 1);
                } else {
                    const int pos = a.boff[br] + atomicAdd(&a.bcur[br], 1);
                    a.blist[pos] = at;
                }
            }
}

// One workgroup of two waves per brick.  The brick's atom list is put in
// index order by ranking (ids are distinct: rank = number of smaller ids),
// then a lane sums the atoms for one (i, j) column and four z nodes; the
// partial dx^2 + dy^2 is shared by the four nodes (the reference's DIST2 adds
// the three squares left to right).
__global__ void __launch_bounds__(128) k_blobby_field(DevGrid g, BlobbyArgs a) {
    __shared__ int raw[kBrickCap];
    __shared__ int list[kBrickCap];
    const int br = blockIdx.x;
    const int c = a.bcount[br];
    if (c > kBrickCap) {
        if (threadIdx.x == 0) atomicMax(a.overflow, c);
        return;
    }
    for (int q = threadIdx.x; q < c; q += blockDim.x) raw[q] = a.blist[a.boff[br] + q];
    __syncthreads();
    for (int q = threadIdx.x; q < c; q += blockDim.x) {
        const int x = raw[q];
        int rank = 0;
        for (int u = 0; u < c; ++u) rank += raw[u] < x;
        list[rank] = x;
    }
    __syncthreads();
    const int bx = br % a.nbx, by = (br / a.nbx) % a.nby, bz = br / (a.nbx * a.nby);
    const int li = threadIdx.x & 7, lj = (threadIdx.x >> 3) & 7, lk = threadIdx.x >> 6;  // lk 0..1
    const int ii = bx * kBrick + li, jj = by * kBrick + lj, kb = bz * kBrick + 4 * lk;
    const bool inxy = ii < g.nx && jj < g.ny;
    double px = 0., py = 0., pz[4];
    if (inxy) {
        px = g.x[ii] - g.hside;
        py = g.y[jj] - g.hside;
    }
    for (int u = 0; u < 4; ++u) pz[u] = kb + u < g.nz ? g.z[kb + u] - g.hside : 0.;
    double f[4] = {0., 0., 0., 0.};
    // the next atom's record and centre are loaded while the current one is summed
    struct Rec {
        int4 b03;
        int2 b45;
        double r2, x, y, z;
    };
    auto load = [&](int q) {
        const int at = __builtin_amdgcn_readfirstlane(list[q]);
        const int *bp = a.box + 8 * at;
        const double *pos = a.xyzr + 4 * (int64_t)at;
        Rec r;
        r.b03 = *(const int4 *)bp;
        r.b45 = *(const int2 *)(bp + 4);
        r.r2 = *(const double *)(bp + 6);
        r.x = pos[0];
        r.y = pos[1];
        r.z = pos[2];
        return r;
    };
    Rec nx = c > 0 ? load(0) : Rec{};
    for (int q = 0; q < c; ++q) {
        const Rec cur = nx;
        if (q + 1 < c) nx = load(q + 1);
        const bool in = inxy & (ii >= cur.b03.x) & (ii < cur.b03.y) & (jj >= cur.b03.z) & (jj < cur.b03.w);
        if (!in) continue;
        const double dxy = (px - cur.x) * (px - cur.x) + (py - cur.y) * (py - cur.y);
        for (int u = 0; u < 4; ++u)
            if ((kb + u >= cur.b45.x) & (kb + u < cur.b45.y)) {
                const double d2 = dxy + (pz[u] - cur.z) * (pz[u] - cur.z);
                f[u] += exp(a.B * (d2 / cur.r2 - 1));
            }
    }
    if (inxy)
        for (int u = 0; u < 4; ++u)
            if (kb + u < g.nz) a.field[flat(g, ii, jj, kb + u)] = f[u];
}

// corner bits of the scalar marching cubes: bit 0 where phi < iso (the case
// bit classifyCube sets, Surface.cpp:14526-14547), 1 elsewhere and past the
// grid; one lane per node, a wave covers 64 consecutive nodes = two words
__global__ void __launch_bounds__(256) k_field_bits(const double *field, int64_t nvox, double iso, uint32_t *bits) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool out = f >= nvox || !(field[f] < iso);
    const uint64_t m = __ballot(out);
    const int64_t w = (f >> 6) << 1;
    const int64_t nw = (nvox + 31) >> 5;
    if ((threadIdx.x & 63) == 0) {
        if (w < nw) bits[w] = (uint32_t)m;
        if (w + 1 < nw) bits[w + 1] = (uint32_t)(m >> 32);
    }
}

void blobby_bricks(const DevGrid &g, int *nbx, int *nby, int *nbz) {
    *nbx = (g.nx + kBrick - 1) / kBrick;
    *nby = (g.ny + kBrick - 1) / kBrick;
    *nbz = (g.nz + kBrick - 1) / kBrick;
}
int blobby_brick_cap() { return kBrickCap; }

void launch_blobby_pairs(hipStream_t s, const DevGrid &g, const BlobbyArgs &a, int scatter) {
    if (a.n > 0) hipLaunchKernelGGL(k_blobby_pairs, dim3(nblk(a.n)), dim3(256), 0, s, g, a, scatter);
}
void launch_blobby_field(hipStream_t s, const DevGrid &g, const BlobbyArgs &a) {
    const int64_t nb = (int64_t)a.nbx * a.nby * a.nbz;
    if (nb > 0) hipLaunchKernelGGL(k_blobby_field, dim3((unsigned)nb), dim3(128), 0, s, g, a);
}
void launch_field_bits(hipStream_t s, const double *field, int64_t nvox, double iso, uint32_t *bits) {
    if (nvox > 0) hipLaunchKernelGGL(k_field_bits, dim3(nblk(nvox)), dim3(256), 0, s, field, nvox, iso, bits);
}

// ------------------------------------------------------------ smoothing
// Surface::smoothSurface (Surface.cpp:13089-13240): every triangle adds the
// sum of its other two vertices to each of its vertices, in triangle order;
// then v <- 0.5 (v + sum / (2 * triangles of v)).  A marching-cubes vertex
// lies on one grid edge and only the (at most four) cubes around that edge
// emit triangles using it; their triangles are contiguous runs of the
// triangle list in cube order, so visiting the cubes in increasing node order
// visits the vertex's triangles in triangle order, with no incidence lists.

// index of active cube `key` in the active-cube list, or -1: the map entry
// of the cube, confirmed by the list (entries of earlier calls, or never
// written, fail the check)
__device__ __forceinline__ int active_index(const int *anode, int nact, const int *cube_index, int key) {
    const int ci = cube_index[key];
    return (ci >= 0 && ci < nact && anode[ci] == key) ? ci : -1;
}

__global__ void __launch_bounds__(256) k_cube_index(const int *anode, int nact, int *cube_index) {
    const int ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci < nact) cube_index[anode[ci]] = ci;
}

__global__ void __launch_bounds__(256) k_smooth(DevGrid g, SmoothArgs a) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.nv) return;
    const int64_t slot = a.vslot[v];
    const int64_t f = slot / 3;
    const int dir = (int)(slot - 3 * f);
    const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
    // the cubes around the edge (node f, direction dir), increasing node index
    int64_t cubes[4];
    if (dir == 0) { cubes[0] = f - sy - sz; cubes[1] = f - sz; cubes[2] = f - sy; cubes[3] = f; }
    else if (dir == 1) { cubes[0] = f - 1 - sz; cubes[1] = f - sz; cubes[2] = f - 1; cubes[3] = f; }
    else { cubes[0] = f - 1 - sy; cubes[1] = f - sy; cubes[2] = f - 1; cubes[3] = f; }
    double s0 = 0., s1 = 0., s2 = 0., n0 = 0., n1 = 0., n2 = 0.;
    int c = 0;
    for (int q = 0; q < 4; ++q) {
        if (cubes[q] < 0) continue;
        const int ci = active_index(a.anode, a.nact, a.cube_index, (int)cubes[q]);
        if (ci < 0) continue;
        for (int t = a.atoff[ci]; t < a.atoff[ci] + a.ant[ci]; ++t) {
            const int *tv = a.tris + 3 * (int64_t)t;
            int pos = -1;
            if (tv[0] == v) pos = 0;
            else if (tv[1] == v) pos = 1;
            else if (tv[2] == v) pos = 2;
            if (pos < 0) continue;
            // the other two vertices in the reference's order: (v2, v3), (v1, v3), (v1, v2)
            const int oa = pos == 0 ? tv[1] : tv[0];
            const int ob = pos == 2 ? tv[1] : tv[2];
            const double *A = a.vin + 3 * (int64_t)oa, *B = a.vin + 3 * (int64_t)ob;
            s0 += A[0] + B[0];
            s1 += A[1] + B[1];
            s2 += A[2] + B[2];
            if (a.nin) {
                const double *NA = a.nin + 3 * (int64_t)oa, *NB = a.nin + 3 * (int64_t)ob;
                n0 += NA[0] + NB[0];
                n1 += NA[1] + NB[1];
                n2 += NA[2] + NB[2];
            }
            ++c;
        }
    }
    const double *V = a.vin + 3 * v;
    if (c == 0) {
        // no triangle names the vertex: it keeps its position and normal
        a.vout[3 * v + 0] = V[0];
        a.vout[3 * v + 1] = V[1];
        a.vout[3 * v + 2] = V[2];
        if (a.nin)
            for (int k = 0; k < 3; ++k) a.nout[3 * v + k] = a.nin[3 * v + k];
        return;
    }
    const double scale = 1.0 / (double)(2 * c);
    a.vout[3 * v + 0] = 0.5 * (V[0] + s0 * scale);
    a.vout[3 * v + 1] = 0.5 * (V[1] + s1 * scale);
    a.vout[3 * v + 2] = 0.5 * (V[2] + s2 * scale);
    if (a.nin) {
        // the normal plus the neighbours' average, normalised
        // (Surface.cpp:13156-13168)
        const double *Nv = a.nin + 3 * v;
        const double t0 = Nv[0] + n0 * scale, t1 = Nv[1] + n1 * scale, t2 = Nv[2] + n2 * scale;
        const double tt = 1.0 / sqrt(t0 * t0 + t1 * t1 + t2 * t2);
        a.nout[3 * v + 0] = t0 * tt;
        a.nout[3 * v + 1] = t1 * tt;
        a.nout[3 * v + 2] = t2 * tt;
    }
}

// "%.3f" (round half to even on the exact binary value, as printf) read back
// with sscanf("%f") into a float (MeshSurface::loadOFF, MeshSurface.cpp:
// 2436-2446): the nearest float to n / 1000, n the printed thousandths
__device__ double off_round_trip(double x) {
    if (isnan(x)) return x;
    double n = rint(x * 1000.0);
    // x*1000 against n +- 0.5 exactly: fma rounds once, so its sign is exact
    const double hi = fma(x, 1000.0, -(n + 0.5));
    const double lo = fma(x, 1000.0, -(n - 0.5));
    if (hi > 0. || (hi == 0. && fmod(n + 1., 2.) == 0.)) n += 1.;
    else if (lo < 0. || (lo == 0. && fmod(n - 1., 2.) == 0.)) n -= 1.;
    if (n == 0.) return signbit(x) ? -0.0 : 0.0;  // "-0.000" reads as -0.0f
    // nearest float to the decimal n/1000: the float of the nearest double,
    // corrected by comparing the exact distances of it and its neighbours
    // (f*1000 has at most 34 significant bits, so f*1000 - n is exact)
    const float f0 = (float)(n / 1000.0);
    float best = f0;
    double bd = fabs((double)f0 * 1000.0 - n);
    const float cand[2] = {nextafterf(f0, -INFINITY), nextafterf(f0, INFINITY)};
    for (int q = 0; q < 2; ++q) {
        const double d = fabs((double)cand[q] * 1000.0 - n);
        if (d < bd) {
            bd = d;
            best = cand[q];
        }
    }
    return (double)best;
}

__global__ void __launch_bounds__(256) k_quantize(const double *in, int64_t n, double *out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = off_round_trip(in[q]);
}

void launch_smooth(hipStream_t s, const DevGrid &g, const SmoothArgs &a) {
    if (a.nact > 0) hipLaunchKernelGGL(k_cube_index, dim3(nblk(a.nact)), dim3(256), 0, s, a.anode, a.nact, a.cube_index);
    if (a.nv > 0) hipLaunchKernelGGL(k_smooth, dim3(nblk(a.nv)), dim3(256), 0, s, g, a);
}
void launch_quantize(hipStream_t s, const double *in, int64_t n, double *out) {
    if (n > 0) hipLaunchKernelGGL(k_quantize, dim3(nblk(n)), dim3(256), 0, s, in, n, out);
}

// ------------------------------------------------------------ ray/triangle
// intersect_triangle, the non-culling branch (MeshSurface.cpp:1525-1550), EPS 1e-5 (MeshSurface.h:21)
__device__ __forceinline__ bool ray_triangle(const double o[3], const double d[3], const double *v0, const double *v1,
                                             const double *v2, double &t) {
    const double e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    const double e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    const double pv[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
    const double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
    if (det > -1e-5 && det < 1e-5) return false;
    const double inv = 1.0 / det;
    const double tv[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
    const double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * inv;
    if (u < 0.0 || u > 1.0) return false;
    const double qv[3] = {tv[1] * e1[2] - tv[2] * e1[1], tv[2] * e1[0] - tv[0] * e1[2], tv[0] * e1[1] - tv[1] * e1[0]};
    const double v = (d[0] * qv[0] + d[1] * qv[1] + d[2] * qv[2]) * inv;
    if (v < 0.0 || u + v > 1.0) return false;
    t = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * inv;
    return true;
}

__device__ __forceinline__ const double *axis_coords(const DevGrid &g, int d) { return d == 0 ? g.x : (d == 1 ? g.y : g.z); }
__device__ __forceinline__ int axis_n(const DevGrid &g, int d) { return d == 0 ? g.nx : (d == 1 ? g.ny : g.nz); }
__device__ __forceinline__ double axis_min(const DevGrid &g, int d) { return d == 0 ? g.xmin : (d == 1 ? g.ymin : g.zmin); }
__device__ __forceinline__ double axis_max(const DevGrid &g, int d) { return d == 0 ? g.xmax : (d == 1 ? g.ymax : g.zmax); }

// rays of axis d (coordinates c[m] + delta, c[m] = cmin + m side) whose
// coordinate may lie in [lo, hi]: a superset by one ray on each side (the
// caller compares exactly), clipped to [m0, m1]
__device__ __forceinline__ void ray_span(double lo, double hi, double cmin, double delta, double scale, int m0, int m1,
                                         int &a, int &b) {
    const double ea = fmin(fmax(floor((lo - delta - cmin) * scale), -2.), 2e9);
    const double eb = fmin(fmax(floor((hi - delta - cmin) * scale), -2.), 2e9);
    a = max(m0, (int)ea - 1);
    b = min(m1, (int)eb + 1);
}

// grid: ceil(nt / 64) workgroups of one wave; a lane takes one triangle
// through every panel slot; per-workgroup candidate counts to part
template <bool kRayBased>
__global__ void __launch_bounds__(64) k_intersect_mesh(DevGrid g, DevRays R, MeshArgs a, unsigned long long *part) {
    const int64_t it = (int64_t)blockIdx.x * 64 + threadIdx.x;
    unsigned long long cand = 0;
    if (it < a.nt) {
        const int *tv = a.tri + 3 * it;
        const double *p0 = a.v + 3 * (int64_t)tv[0], *p1 = a.v + 3 * (int64_t)tv[1], *p2 = a.v + 3 * (int64_t)tv[2];
        double down[3], up[3];
        for (int q = 0; q < 3; ++q) {
            down[q] = INFINITY;
            up[q] = -INFINITY;
        }
        const double *pp[3] = {p0, p1, p2};
        for (int pi = 0; pi < 3; ++pi)
            for (int q = 0; q < 3; ++q) {
                down[q] = down[q] < pp[pi][q] ? down[q] : pp[pi][q];
                up[q] = up[q] > pp[pi][q] ? up[q] : pp[pi][q];
            }
        bool culled = false;
        if (!kRayBased)  // CELL_CULLING (MeshSurface.cpp:1318-1331)
            for (int q = 0; q < 3; ++q) {
                culled |= down[q] > axis_max(g, q) + kEpsInt;
                culled |= up[q] < axis_min(g, q) - (g.hside - kDelta + kEpsInt);
            }
        for (int slot = 0; slot < R.nslots && !culled; ++slot) {
            const int phase = slot >= R.panels0;
            const int panel = phase ? slot - R.panels0 : slot;
            // one instance per panel: the axes are constants, so the ray origin
            // and the box stay in registers
            auto walk = [&](auto panel_c) {
            constexpr int first = PanelAxes<decltype(panel_c)::value>::fd;
            constexpr int last = PanelAxes<decltype(panel_c)::value>::ld;
            constexpr int vc = PanelAxes<decltype(panel_c)::value>::vc;
            const double delta = phase ? kDelta - g.hside : 0.;
            const double *cf = axis_coords(g, first), *cl = axis_coords(g, last), *cv = axis_coords(g, vc);
            const int nvc = axis_n(g, vc);
            double pa[3], dir[3] = {0., 0., 0.};
            pa[vc] = cv[0] + delta;
            dir[vc] = kRayBased ? (cv[nvc - 1] + delta) - pa[vc] : cv[nvc - 1] - cv[0];
            int m0, m1, n0, n1;
            if (kRayBased) {
                int64_t cs[2], ce[2];
                const int dd[2] = {first, last};
                for (int q = 0; q < 2; ++q) {
                    cs[q] = (int64_t)rintp(fmax(0., down[dd[q]] - a.min2d[dd[q]]) * a.scale2d);
                    ce[q] = (int64_t)rintp(fmax(0., up[dd[q]] - a.min2d[dd[q]]) * a.scale2d);
                    if (cs[q] >= a.n2d) cs[q] = a.n2d - 1;
                    if (ce[q] >= a.n2d) ce[q] = a.n2d - 1;
                }
                const int *rf = a.rr + (int64_t)((first * 2 + phase) * 2) * a.n2d;
                const int *rl = a.rr + (int64_t)((last * 2 + phase) * 2) * a.n2d;
                m0 = rf[cs[0]];
                m1 = rf[a.n2d + ce[0]];
                n0 = rl[cs[1]];
                n1 = rl[a.n2d + ce[1]];
                // every ray of the cell range is a candidate of the reference;
                // only those within 1e-9 A of the bounding box can pass the
                // barycentric tests (a ray d outside the box is d outside the
                // triangle's projection, and u, v carry rounding errors of
                // order 1e-16 |o - v0|, about 1e-14 A here)
                if (m1 >= m0 && n1 >= n0) cand += (unsigned long long)(m1 - m0 + 1) * (n1 - n0 + 1);
                const double mg = 1e-9;
                ray_span(down[first] - mg, up[first] + mg, axis_min(g, first), delta, g.scale, m0, m1, m0, m1);
                ray_span(down[last] - mg, up[last] + mg, axis_min(g, last), delta, g.scale, n0, n1, n0, n1);
                for (int n = n0; n <= n1; ++n) {
                    pa[last] = cl[n] + delta;
                    if (pa[last] < down[last] - mg || pa[last] > up[last] + mg) continue;
                    for (int m = m0; m <= m1; ++m) {
                        pa[first] = cf[m] + delta;
                        if (pa[first] < down[first] - mg || pa[first] > up[first] + mg) continue;
                        double t;
                        if (ray_triangle(pa, dir, p0, p1, p2, t))
                            push_hit(R, slot, (int64_t)n * R.N + m, t, (int)it);
                    }
                }
            } else {
                const int dd[2] = {first, last};
                int64_t s[2], e[2];
                for (int q = 0; q < 2; ++q) {
                    s[q] = (int64_t)rintp(fmax(0., down[dd[q]] - axis_min(g, dd[q])) * g.scale);
                    e[q] = (int64_t)rintp(fmax(0., up[dd[q]] - axis_min(g, dd[q])) * g.scale);
                    if (e[q] >= axis_n(g, dd[q])) e[q] = axis_n(g, dd[q]) - 1;
                }
                m0 = (int)s[0];
                m1 = (int)e[0];
                n0 = (int)s[1];
                n1 = (int)e[1];
                for (int n = n0; n <= n1; ++n) {
                    pa[last] = cl[n] + delta;
                    if (pa[last] < down[last] || pa[last] > up[last]) continue;
                    for (int m = m0; m <= m1; ++m) {
                        pa[first] = cf[m] + delta;
                        if (pa[first] < down[first] || pa[first] > up[first]) continue;
                        ++cand;
                        double t;
                        if (ray_triangle(pa, dir, p0, p1, p2, t))
                            push_hit(R, slot, (int64_t)n * R.N + m, t, (int)it);
                    }
                }
            }
            };
            if (panel == 0) walk(std::integral_constant<int, 0>());
            else if (panel == 1) walk(std::integral_constant<int, 1>());
            else walk(std::integral_constant<int, 2>());
        }
    }
    for (int off = 32; off > 0; off >>= 1) cand += __shfl_down(cand, off, 64);
    if (threadIdx.x == 0 && cand) part[blockIdx.x] = cand;
}

int64_t mesh_workgroups(int64_t nt, const DevRays &) { return (int64_t)nblk(nt, 64); }

void launch_intersect_mesh(hipStream_t s, const DevGrid &g, const DevRays &R, const MeshArgs &a,
                           unsigned long long *part) {
    if (a.nt <= 0 || R.nslots <= 0) return;
    if (a.ray_based)
        hipLaunchKernelGGL(k_intersect_mesh<true>, dim3(nblk(a.nt, 64)), dim3(64), 0, s, g, R, a, part);
    else
        hipLaunchKernelGGL(k_intersect_mesh<false>, dim3(nblk(a.nt, 64)), dim3(64), 0, s, g, R, a, part);
}

}  // namespace nsg
