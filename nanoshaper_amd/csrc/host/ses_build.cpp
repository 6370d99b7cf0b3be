// Host SES patch construction (the stage ConnollySurface::build performs on the
// CPU, ConnollySurface.cpp:385-2036) without CGAL.
//
// The reference builds the weighted alpha complex (alpha = 0) of the atoms with
// weights (r + probe)^2 and reads the SES cells off its classification.  The
// same classification follows from where the probe is free (outside every
// inflated atom): a facet (i,j,k) is SINGULAR when both probe placements
// tangent to the three atoms are free, REGULAR when exactly one is; an edge
// (i,j) is REGULAR when some of its circle of intersection is free and a free
// probe placement lies on it, SINGULAR when the whole circle is free; an atom
// is exposed when it has no neighbour or takes part in an exposed edge.  The
// cell geometry (probe centres, trimming planes, torus frames and clipping
// spheres, self-intersection planes) is computed with the formulas of
// BuildFacetCells (:787-1265), BuildEdgeCells (:1267-1926) and
// RemoveSelfIntersections (:1928-2036).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <vector>

#include "nsgpu.h"
#include "patch_set.hpp"

namespace {

inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// plane through three points, normalised (tools.cpp:258-274)
void plane3(const double p1[3], const double p2[3], const double p3[3], double w[4]) {
    w[0] = p1[1] * (p2[2] - p3[2]) + p2[1] * (p3[2] - p1[2]) + p3[1] * (p1[2] - p2[2]);
    w[1] = p1[2] * (p2[0] - p3[0]) + p2[2] * (p3[0] - p1[0]) + p3[2] * (p1[0] - p2[0]);
    w[2] = p1[0] * (p2[1] - p3[1]) + p2[0] * (p3[1] - p1[1]) + p3[0] * (p1[1] - p2[1]);
    w[3] = -dot3(w, p1);
    const double nrm = std::sqrt(dot3(w, w));
    w[0] /= nrm;
    w[1] /= nrm;
    w[2] /= nrm;
    w[3] /= nrm;
}

// unsigned point-plane distance (tools.cpp:277-287)
double point_plane_dist(const double p[3], const double w[4]) {
    const double den = dot3(w, w);
    const double val = dot3(w, p) + w[3];
    return std::fabs(val / std::sqrt(den));
}

// "safe" normalisation with the reference's 1e-20 guard (globals.h:243-253)
inline void normalize_s(double out[3], const double v[3]) {
    const double t = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    out[0] = v[0] / (t + 1e-20);
    out[1] = v[1] / (t + 1e-20);
    out[2] = v[2] / (t + 1e-20);
}

inline void cross3(double d[3], const double a[3], const double b[3]) {
    d[0] = a[1] * b[2] - a[2] * b[1];
    d[1] = a[2] * b[0] - a[0] * b[2];
    d[2] = a[0] * b[1] - a[1] * b[0];
}

// AND/OR decision between the two planes bounding a free torus arc
// (ConnollySurface::orientation, ConnollySurface.cpp:6108-6130)
bool orientation(const double *c1, const double *c2, const double *w1, const double *w2) {
    const double b = dot3(w1, w2), c = dot3(w2, w2);
    const double w0[3] = {c1[0] - c2[0], c1[1] - c2[1], c1[2] - c2[2]};
    const double d = dot3(w1, w0), e = dot3(w2, w0);
    return b * e < c * d;
}

struct Facet {
    double center[3];
    double planes[4][4];
    int id[3];
    int kind;    // NSGPU_PATCH_FACET_REG / _SING
    int mirror;  // index of the mirror facet or -1
    bool check;  // regular facet close enough to the atom plane to self-intersect
    std::vector<double> self;  // self-intersection planes, sign applied
};

struct Edge {
    int i, j;
    double center[3], invrot[9], R, self_r, clip_c[3], clip_r, clip_r2;
    double u[3], v[3], w[3], coi_pt[3];
    int kind;  // torus kind, or -1 when no free circle point
    std::vector<double> planes;  // pairs, sign applied
    uint32_t flags;
};

struct Builder {
    const double *xyzr;
    int n;
    double probe;
    std::vector<double> R;  // inflated radii
    std::vector<int> nb_off, nb;  // neighbour CSR, sorted
    std::vector<char> hidden;

    void neighbours() {
        R.resize(n);
        double rmax = 0;
        for (int a = 0; a < n; ++a) {
            R[a] = xyzr[4 * a + 3] + probe;
            rmax = std::max(rmax, R[a]);
        }
        const double cell = 2 * rmax;
        double lo[3] = {1e300, 1e300, 1e300};
        for (int a = 0; a < n; ++a)
            for (int d = 0; d < 3; ++d) lo[d] = std::min(lo[d], xyzr[4 * a + d]);
        auto key = [&](int a, int d) { return (int64_t)std::floor((xyzr[4 * a + d] - lo[d]) / cell); };
        std::unordered_map<int64_t, std::vector<int>> grid;
        grid.reserve(n * 2);
        auto hkey = [](int64_t x, int64_t y, int64_t z) { return (x * 73856093LL) ^ (y * 19349663LL) ^ (z * 83492791LL); };
        // exact cell coordinates are kept alongside to reject hash collisions
        std::vector<int64_t> cx(n), cy(n), cz(n);
        for (int a = 0; a < n; ++a) {
            cx[a] = key(a, 0);
            cy[a] = key(a, 1);
            cz[a] = key(a, 2);
            grid[hkey(cx[a], cy[a], cz[a])].push_back(a);
        }
        std::vector<std::vector<int>> lists(n);
        hidden.assign(n, 0);
        for (int a = 0; a < n; ++a) {
            const double *pa = xyzr + 4 * a;
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        auto it = grid.find(hkey(cx[a] + dx, cy[a] + dy, cz[a] + dz));
                        if (it == grid.end()) continue;
                        for (int b : it->second) {
                            if (b == a || cx[b] != cx[a] + dx || cy[b] != cy[a] + dy || cz[b] != cz[a] + dz)
                                continue;
                            const double *pb = xyzr + 4 * b;
                            const double d2 = (pa[0] - pb[0]) * (pa[0] - pb[0]) + (pa[1] - pb[1]) * (pa[1] - pb[1]) +
                                              (pa[2] - pb[2]) * (pa[2] - pb[2]);
                            const double s = R[a] + R[b];
                            if (d2 >= s * s) continue;
                            const double dr = R[a] - R[b];
                            if (d2 <= dr * dr) {
                                // one inflated sphere inside the other: the smaller is
                                // redundant in the power diagram and makes no patch
                                if (R[a] < R[b] || (R[a] == R[b] && a > b)) hidden[a] = 1;
                                continue;
                            }
                            lists[a].push_back(b);
                        }
                    }
        }
        nb_off.assign(n + 1, 0);
        for (int a = 0; a < n; ++a) {
            std::sort(lists[a].begin(), lists[a].end());
            nb_off[a + 1] = nb_off[a] + (int)lists[a].size();
        }
        nb.resize(nb_off[n]);
        for (int a = 0; a < n; ++a) std::copy(lists[a].begin(), lists[a].end(), nb.begin() + nb_off[a]);
    }

    // free = outside every other (non hidden) inflated atom, tested against the
    // neighbours of atom a (any sphere containing a point of sphere a intersects it)
    bool free_point(const double q[3], int a, int e1, int e2, int e3) const {
        for (int t = nb_off[a]; t < nb_off[a + 1]; ++t) {
            const int l = nb[t];
            if (l == e1 || l == e2 || l == e3 || hidden[l]) continue;
            const double *pl = xyzr + 4 * l;
            const double d2 = (q[0] - pl[0]) * (q[0] - pl[0]) + (q[1] - pl[1]) * (q[1] - pl[1]) +
                              (q[2] - pl[2]) * (q[2] - pl[2]);
            if (d2 < R[l] * R[l]) return false;
        }
        return true;
    }

    // probe placements tangent to three atoms: the closed form of
    // BuildFacetCells (ConnollySurface.cpp:869-965).  Returns false when the
    // three inflated spheres have no common point.
    bool probe_pair(int a, int b, int c, double plus[3], double minus[3]) const {
        const double *P1 = xyzr + 4 * a, *P2 = xyzr + 4 * b, *P3 = xyzr + 4 * c;
        const double P12[3] = {P2[0] - P1[0], P2[1] - P1[1], P2[2] - P1[2]};
        const double P13[3] = {P3[0] - P1[0], P3[1] - P1[1], P3[2] - P1[2]};
        const double P23[3] = {P3[0] - P2[0], P3[1] - P2[1], P3[2] - P2[2]};
        double nrm[3];
        cross3(nrm, P12, P13);
        const double A = dot3(P12, P12), B = dot3(P23, P23), C = dot3(P13, P13);
        const double D = R[a] * R[a], E = R[b] * R[b], F = R[c] * R[c];
        const double X = D + E - A, Y = E + F - B, Z = D + F - C;
        const double Varg = 4. * D * E * F - F * X * X - D * Y * Y - E * Z * Z + X * Y * Z;
        if (!(Varg >= 0.)) return false;
        const double V = 1. / 12. * std::sqrt(Varg);
        const double G = dot3(P12, P13), H = dot3(nrm, nrm);
        const double I = dot3(P2, P2) - dot3(P1, P1), J = dot3(P3, P3) - dot3(P1, P1);
        const double K = dot3(nrm, P1);
        const double L = I - E + D, M = J - F + D;
        const double S = C * L - G * M, T = A * M - G * L;
        double Q[3];
        for (int d = 0; d < 3; ++d) Q[d] = (K * nrm[d] + 0.5 * (S * P12[d] + T * P13[d])) / H;
        const double coeff = (6.0 * V) / H;
        for (int d = 0; d < 3; ++d) {
            plus[d] = Q[d] + coeff * nrm[d];
            minus[d] = Q[d] - coeff * nrm[d];
        }
        return true;
    }

    // trimming planes of a reentrant cap (ConnollySurface.cpp:1032-1121)
    void facet_planes(Facet &f, bool regular) const {
        const double *P[3] = {xyzr + 4 * f.id[0], xyzr + 4 * f.id[1], xyzr + 4 * f.id[2]};
        double PP[3][3];
        for (int q = 0; q < 3; ++q)
            for (int d = 0; d < 3; ++d)
                PP[q][d] = regular ? f.center[d] + 50000. * (P[q][d] - f.center[d]) : P[q][d];
        plane3(PP[0], PP[1], PP[2], f.planes[0]);
        plane3(PP[0], PP[1], f.center, f.planes[1]);
        plane3(PP[1], PP[2], f.center, f.planes[2]);
        plane3(PP[0], PP[2], f.center, f.planes[3]);
        double inner[3];
        for (int d = 0; d < 3; ++d) inner[d] = ((1. / 3.) * (PP[0][d] + PP[1][d] + PP[2][d]) + f.center[d]) * 0.5;
        for (int ff = 0; ff < 4; ++ff) {
            const double orient = dot3(inner, f.planes[ff]);
            if (!(orient < -f.planes[ff][3]))
                for (int d = 0; d < 4; ++d) f.planes[ff][d] = -f.planes[ff][d];
        }
    }

    std::vector<Facet> facets;
    std::vector<std::vector<int>> atom_facets;
    std::vector<Edge> edges;

    void build_facets(int nthreads) {
        // enumerate triples a<b<c with mutually intersecting inflated spheres
        std::vector<std::vector<Facet>> part(nthreads);
        auto work = [&](int tid) {
            std::vector<int> common;
            for (int a = tid; a < n; a += nthreads) {
                if (hidden[a]) continue;
                for (int s = nb_off[a]; s < nb_off[a + 1]; ++s) {
                    const int b = nb[s];
                    if (b <= a || hidden[b]) continue;
                    common.clear();
                    std::set_intersection(nb.begin() + s + 1, nb.begin() + nb_off[a + 1], nb.begin() + nb_off[b],
                                          nb.begin() + nb_off[b + 1], std::back_inserter(common));
                    for (int c : common) {
                        if (c <= b || hidden[c]) continue;
                        double pp[3], pm[3];
                        if (!probe_pair(a, b, c, pp, pm)) continue;
                        const bool fp = free_point(pp, a, a, b, c);
                        const bool fm = free_point(pm, a, a, b, c);
                        if (!fp && !fm) continue;
                        if (fp && fm) {
                            Facet f1, f2;
                            f1.id[0] = f2.id[0] = a;
                            f1.id[1] = f2.id[1] = b;
                            f1.id[2] = f2.id[2] = c;
                            f1.kind = f2.kind = NSGPU_PATCH_FACET_SING;
                            std::memcpy(f1.center, pp, sizeof pp);
                            std::memcpy(f2.center, pm, sizeof pm);
                            f1.check = f2.check = false;
                            facet_planes(f1, false);
                            facet_planes(f2, false);
                            f1.mirror = 1;  // relative: the next facet
                            f2.mirror = -1; // relative: the previous facet
                            part[tid].push_back(std::move(f1));
                            part[tid].push_back(std::move(f2));
                        } else {
                            Facet f;
                            f.id[0] = a;
                            f.id[1] = b;
                            f.id[2] = c;
                            f.kind = NSGPU_PATCH_FACET_REG;
                            std::memcpy(f.center, fp ? pp : pm, sizeof pp);
                            f.mirror = 0;
                            facet_planes(f, true);
                            double pl[4];
                            plane3(xyzr + 4 * a, xyzr + 4 * b, xyzr + 4 * c, pl);
                            f.check = point_plane_dist(f.center, pl) <= probe;
                            part[tid].push_back(std::move(f));
                        }
                    }
                }
            }
        };
        run(nthreads, work);
        // merge in atom order (deterministic for any thread count)
        std::vector<std::pair<int, int>> order;  // (first atom, thread) chunks
        size_t total = 0;
        for (auto &p : part) total += p.size();
        facets.clear();
        facets.reserve(total);
        std::vector<size_t> cursor(nthreads, 0);
        for (int a = 0; a < n; ++a) {
            auto &p = part[a % nthreads];
            size_t &cu = cursor[a % nthreads];
            while (cu < p.size() && p[cu].id[0] == a) {
                Facet f = std::move(p[cu]);
                const int here = (int)facets.size();
                f.mirror = (f.mirror == 0) ? -1 : here + f.mirror;
                facets.push_back(std::move(f));
                ++cu;
            }
        }
        atom_facets.assign(n, {});
        for (int f = 0; f < (int)facets.size(); ++f)
            for (int q = 0; q < 3; ++q) atom_facets[facets[f].id[q]].push_back(f);
    }

    // RemoveSelfIntersections (ConnollySurface.cpp:1928-2036): every facet is
    // clipped against each near-atom-plane regular facet whose probe sphere it
    // intersects, by the bisecting plane of the two probe centres.
    void self_intersections() {
        const double cell = 2 * probe;
        const double p2x4 = 4 * probe * probe;
        std::unordered_map<int64_t, std::vector<int>> grid;
        auto ck = [&](const double *c, int d) { return (int64_t)std::floor(c[d] / cell); };
        auto hkey = [](int64_t x, int64_t y, int64_t z) { return (x * 73856093LL) ^ (y * 19349663LL) ^ (z * 83492791LL); };
        for (int f = 0; f < (int)facets.size(); ++f)
            if (facets[f].check) grid[hkey(ck(facets[f].center, 0), ck(facets[f].center, 1), ck(facets[f].center, 2))].push_back(f);
        for (int f1 = 0; f1 < (int)facets.size(); ++f1) {
            const double *c1 = facets[f1].center;
            const int64_t x = ck(c1, 0), y = ck(c1, 1), z = ck(c1, 2);
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        auto it = grid.find(hkey(x + dx, y + dy, z + dz));
                        if (it == grid.end()) continue;
                        for (int f2 : it->second) {
                            if (f2 == f1) continue;
                            const double *c2 = facets[f2].center;
                            if (ck(c2, 0) != x + dx || ck(c2, 1) != y + dy || ck(c2, 2) != z + dz) continue;
                            const double d2 = (c1[0] - c2[0]) * (c1[0] - c2[0]) + (c1[1] - c2[1]) * (c1[1] - c2[1]) +
                                              (c1[2] - c2[2]) * (c1[2] - c2[2]);
                            if (!(d2 < p2x4)) continue;
                            double ref[3], dir[3];
                            for (int d = 0; d < 3; ++d) ref[d] = (c2[d] + c1[d]) * 0.5;
                            const double diff[3] = {c1[0] - c2[0], c1[1] - c2[1], c1[2] - c2[2]};
                            normalize_s(dir, diff);
                            const double bias = -dot3(dir, ref);
                            double pl[4] = {dir[0], dir[1], dir[2], bias};
                            if (!(dot3(c1, pl) < -pl[3])) {
                                pl[0] = -dir[0];
                                pl[1] = -dir[1];
                                pl[2] = -dir[2];
                                pl[3] = -bias;
                            }
                            facets[f1].self.insert(facets[f1].self.end(), pl, pl + 4);
                            const double neg[4] = {-pl[0], -pl[1], -pl[2], -pl[3]};
                            facets[f2].self.insert(facets[f2].self.end(), neg, neg + 4);
                        }
                    }
        }
    }

    // torus geometry of BuildEdgeCells (ConnollySurface.cpp:1292-1420, 1880-1893)
    bool torus_geometry(int i, int j, Edge &e) const {
        const double *c1 = xyzr + 4 * i, *c2 = xyzr + 4 * j;
        const double w1 = R[i] * R[i], w2 = R[j] * R[j];
        const double diff[3] = {c2[0] - c1[0], c2[1] - c1[1], c2[2] - c1[2]};
        const double dij2 = dot3(diff, diff);
        const double ratio = (w1 + dij2 - w2) * 0.5;
        const double tt2 = w1 - ratio * ratio / dij2;
        if (!(tt2 > 0.)) return false;
        e.i = i;
        e.j = j;
        e.R = std::sqrt(tt2);
        const double cof = ratio / dij2;
        const double rj2[3] = {diff[0] * cof, diff[1] * cof, diff[2] * cof};
        for (int d = 0; d < 3; ++d) e.center[d] = rj2[d] + c1[d];
        const double rcoi2 = w1 - dot3(rj2, rj2);
        if (!(rcoi2 > 0.)) return false;
        const double rcoi = std::sqrt(rcoi2);
        normalize_s(e.w, diff);
        const double va[3] = {0, e.w[2], -e.w[1]};
        const double vb[3] = {e.w[2], 0, -e.w[0]};
        normalize_s(e.u, va);
        double ct[3], vv[3];
        cross3(ct, va, vb);
        cross3(vv, ct, va);
        normalize_s(e.v, vv);
        double rp[3];
        for (int d = 0; d < 3; ++d) rp[d] = e.center[d] + rcoi * e.u[d];
        const double t1[3] = {rp[0] - c1[0], rp[1] - c1[1], rp[2] - c1[2]};
        const double t2[3] = {rp[0] - c2[0], rp[1] - c2[1], rp[2] - c2[2]};
        const double pai = std::sqrt(dot3(t1, t1)), paj = std::sqrt(dot3(t2, t2));
        const double coeff = pai / (pai + paj);
        const double ctemp[3] = {diff[0] * coeff, diff[1] * coeff, diff[2] * coeff};
        for (int d = 0; d < 3; ++d) e.clip_c[d] = ctemp[d] + c1[d];
        const double r1 = xyzr[4 * i + 3];
        const double xx[3] = {t1[0] * (r1 / pai), t1[1] * (r1 / pai), t1[2] * (r1 / pai)};
        const double tt[3] = {xx[0] - ctemp[0], xx[1] - ctemp[1], xx[2] - ctemp[2]};
        e.clip_r2 = dot3(tt, tt);
        e.clip_r = std::sqrt(e.clip_r2);
        e.self_r = (e.R >= probe) ? -1.0 : std::sqrt(probe * probe - e.R * e.R);
        // frame columns (v, u, w) and its inverse by adjugate/determinant (globals.h:340-380)
        const double m[3][3] = {{e.v[0], e.u[0], e.w[0]}, {e.v[1], e.u[1], e.w[1]}, {e.v[2], e.u[2], e.w[2]}};
        double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]);
        det -= m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]);
        det += m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
        const double s = 1.0 / det;
        double *b = e.invrot;
        b[0] = s * (m[1][1] * m[2][2] - m[1][2] * m[2][1]);
        b[3] = s * (m[1][2] * m[2][0] - m[1][0] * m[2][2]);
        b[6] = s * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
        b[1] = s * (m[0][2] * m[2][1] - m[0][1] * m[2][2]);
        b[4] = s * (m[0][0] * m[2][2] - m[0][2] * m[2][0]);
        b[7] = s * (m[0][1] * m[2][0] - m[0][0] * m[2][1]);
        b[2] = s * (m[0][1] * m[1][2] - m[0][2] * m[1][1]);
        b[5] = s * (m[0][2] * m[1][0] - m[0][0] * m[1][2]);
        b[8] = s * (m[0][0] * m[1][1] - m[0][1] * m[1][0]);
        // the circle point used for the clipping sphere doubles as the free-circle probe
        std::memcpy(e.coi_pt, rp, sizeof rp);
        return true;
    }

    // plane of facet f that contains the edge (a,b): planes[1..3] pair with
    // atom pairs (id0,id1), (id1,id2), (id0,id2)
    const double *edge_plane(const Facet &f, int a, int b) const {
        static const int pr[3][2] = {{0, 1}, {1, 2}, {0, 2}};
        for (int q = 0; q < 3; ++q) {
            const int i1 = f.id[pr[q][0]], i2 = f.id[pr[q][1]];
            if ((i1 == a && i2 == b) || (i1 == b && i2 == a)) return f.planes[q + 1];
        }
        return nullptr;
    }

    int errors = 0;

    // angular order of probe stations around the torus axis (ConnollySurface::sortProbes :6133-6170)
    void sort_probes(const Edge &e, const std::vector<int> &fcv, std::vector<int> &sorted) const {
        double ref[3], temp[3], vs[3], ws[3];
        const double *c0 = facets[fcv[0]].center, *c1 = facets[fcv[1]].center;
        const double d0[3] = {c0[0] - e.center[0], c0[1] - e.center[1], c0[2] - e.center[2]};
        const double d1[3] = {c1[0] - e.center[0], c1[1] - e.center[1], c1[2] - e.center[2]};
        normalize_s(ref, d0);
        normalize_s(temp, d1);
        cross3(vs, ref, temp);
        cross3(ws, vs, ref);
        std::vector<std::pair<double, int>> ang;
        for (int q = 1; q < (int)fcv.size(); ++q) {
            const double *c = facets[fcv[q]].center;
            const double dd[3] = {c[0] - e.center[0], c[1] - e.center[1], c[2] - e.center[2]};
            normalize_s(temp, dd);
            double a = std::atan2(dot3(ws, temp), dot3(ref, temp));
            if (a < 0) a = 2 * M_PI + a;
            ang.emplace_back(a, q);
        }
        std::sort(ang.begin(), ang.end(), [](auto &x, auto &y) { return x.first < y.first; });
        sorted.assign(1, 0);
        for (auto &p : ang) sorted.push_back(p.second);
    }

    void add_pair(Edge &e, const double *p1, const double *p2, bool acute) {
        // the reference keeps the facet planes and negates at test time; the
        // negation is applied here once
        for (int d = 0; d < 4; ++d) e.planes.push_back(-p1[d]);
        for (int d = 0; d < 4; ++d) e.planes.push_back(-p2[d]);
        if (acute) e.flags |= 1u << (uint32_t)(e.planes.size() / 8 - 1);
    }

    // clipping of a regular torus by its probe stations (ConnollySurface.cpp:1490-1830)
    void torus_planes(Edge &e, std::vector<int> &fcv) {
        const int a = e.i, b = e.j;
        const int index = (int)fcv.size();
        if (index == 2) {
            const double *p0 = edge_plane(facets[fcv[0]], a, b), *p1 = edge_plane(facets[fcv[1]], a, b);
            if (!p0 || !p1) { ++errors; return; }
            const bool acute = !orientation(facets[fcv[0]].center, facets[fcv[1]].center, p0, p1);
            add_pair(e, p0, p1, acute);
            return;
        }
        if (index % 2 != 0 || index > 50) { ++errors; return; }
        std::vector<int> sorted;
        sort_probes(e, fcv, sorted);
        int start = 0, direction = 0;
        for (int q = 0; q < index; ++q) {
            const Facet &f = facets[fcv[sorted[q]]];
            if (f.kind == NSGPU_PATCH_FACET_SING) {
                start = q;
                const int next = (q + 1) % index, prev = (index + q - 1) % index;
                if (f.mirror == fcv[sorted[next]]) direction = -2;
                else if (f.mirror == fcv[sorted[prev]]) direction = +2;
                else { ++errors; return; }
                break;
            }
        }
        if (direction == 0) {
            const double *c1 = facets[fcv[sorted[0]]].center, *c2 = facets[fcv[sorted[1]]].center;
            double mid[3];
            for (int d = 0; d < 3; ++d) mid[d] = (c1[d] + c2[d]) * 0.5 - e.center[d];
            const double *pl = edge_plane(facets[fcv[sorted[0]]], a, b);
            if (!pl) { ++errors; return; }
            const double plane[3] = {-pl[0], -pl[1], -pl[2]};
            const double test = -dot3(mid, plane);
            start = 0;
            direction = test > 0 ? +2 : -2;
        }
        int cp1 = start, cp2 = (index + start + (direction >> 1)) % index;
        for (int guard = 0; guard < index; ++guard) {
            const Facet &f1 = facets[fcv[sorted[cp1]]], &f2 = facets[fcv[sorted[cp2]]];
            const double *p1 = edge_plane(f1, a, b), *p2 = edge_plane(f2, a, b);
            if (!p1 || !p2) { ++errors; return; }
            const bool flag = !orientation(f1.center, f2.center, p1, p2);
            add_pair(e, p1, p2, flag);
            cp1 = (index + cp1 + direction) % index;
            cp2 = (index + cp1 + (direction >> 1)) % index;
            if (cp1 == start) break;
        }
    }

    void build_edges() {
        edges.clear();
        for (int a = 0; a < n; ++a) {
            if (hidden[a]) continue;
            for (int s = nb_off[a]; s < nb_off[a + 1]; ++s) {
                const int b = nb[s];
                if (b <= a || hidden[b]) continue;
                Edge e;
                if (!torus_geometry(a, b, e)) continue;
                e.flags = 0;
                std::vector<int> fcv;
                for (int f : atom_facets[a]) {
                    const Facet &F = facets[f];
                    if (F.id[0] == b || F.id[1] == b || F.id[2] == b) fcv.push_back(f);
                }
                if (!fcv.empty()) {
                    e.kind = NSGPU_PATCH_TORUS_REG;
                    torus_planes(e, fcv);
                } else {
                    // no probe station: the circle of intersection is either
                    // entirely free (singular torus) or entirely buried
                    e.kind = free_point(e.coi_pt, a, a, b, -1) ? NSGPU_PATCH_TORUS_SING : -1;
                }
                edges.push_back(std::move(e));
            }
        }
    }

    template <class F>
    static void run(int nthreads, F &&f) {
        if (nthreads <= 1) { f(0); return; }
        std::vector<std::thread> th;
        for (int t = 0; t < nthreads; ++t) th.emplace_back(f, t);
        for (auto &t : th) t.join();
    }
};

}  // namespace


extern "C" int nsgpu_ses_build(const double *xyzr, int32_t n, double probe_radius, nsgpu_patch_set **out) {
    if (!xyzr || n <= 0 || !out || !(probe_radius > 0)) return NSGPU_E_ARG;
    *out = nullptr;
    Builder B;
    B.xyzr = xyzr;
    B.n = n;
    B.probe = probe_radius;
    B.neighbours();
    int nth = (int)std::thread::hardware_concurrency();
    nth = std::max(1, std::min(nth, 16));
    if (n < 2000) nth = 1;
    B.build_facets(nth);
    B.self_intersections();
    B.build_edges();

    auto *ps = new nsgpu_patch_set();
    // exposed atoms: isolated, or on an exposed edge
    std::vector<char> exposed(n, 0);
    for (int a = 0; a < n; ++a) {
        if (B.hidden[a]) continue;
        bool lonely = true;
        for (int t = B.nb_off[a]; t < B.nb_off[a + 1] && lonely; ++t) lonely = B.hidden[B.nb[t]] != 0;
        if (lonely) exposed[a] = 1;
    }
    for (auto &e : B.edges)
        if (e.kind >= 0) exposed[e.i] = exposed[e.j] = 1;
    // clipping spheres of every edge with a circle of intersection
    std::vector<std::vector<int>> atom_edges(n);
    for (int q = 0; q < (int)B.edges.size(); ++q) {
        atom_edges[B.edges[q].i].push_back(q);
        atom_edges[B.edges[q].j].push_back(q);
    }
    auto push_patch = [&](int kind, const double *bs, double br, int cnt, uint32_t flags, int a0, int a1, int a2) {
        ps->kind.push_back(kind);
        ps->off.push_back((int32_t)ps->data.size());
        ps->cnt.push_back(cnt);
        ps->flags.push_back((int32_t)flags);
        ps->atom.push_back(a0);
        ps->atom.push_back(a1);
        ps->atom.push_back(a2);
        ps->bsph.insert(ps->bsph.end(), bs, bs + 3);
        ps->bsph.push_back(br);
        ps->counts[kind]++;
    };
    for (int a = 0; a < n; ++a) {
        if (!exposed[a]) continue;
        // Feasibility of a point-cell hit is the AND of all clipping spheres, so
        // their order only changes how soon a buried hit is rejected: the
        // spheres removing the largest caps go first.  The cap of neighbour b
        // on the inflated sphere of a has cos(half angle) =
        // (R_a^2 + d^2 - R_b^2) / (2 R_a d); smaller means larger.
        auto cap_cos = [&](int q) {
            const auto &e = B.edges[q];
            const int b = (e.i == a) ? e.j : e.i;
            const double *pa = xyzr + 4 * a, *pb = xyzr + 4 * b;
            const double d = std::sqrt((pa[0] - pb[0]) * (pa[0] - pb[0]) + (pa[1] - pb[1]) * (pa[1] - pb[1]) +
                                       (pa[2] - pb[2]) * (pa[2] - pb[2]));
            return (B.R[a] * B.R[a] + d * d - B.R[b] * B.R[b]) / (2 * B.R[a] * d);
        };
        std::stable_sort(atom_edges[a].begin(), atom_edges[a].end(),
                         [&](int p, int q) { return cap_cos(p) < cap_cos(q); });
        push_patch(NSGPU_PATCH_POINT, xyzr + 4 * a, xyzr[4 * a + 3], (int)atom_edges[a].size(), 0, a, -1, -1);
        for (int q : atom_edges[a]) {
            const auto &e = B.edges[q];
            ps->data.insert(ps->data.end(), e.clip_c, e.clip_c + 3);
            ps->data.push_back(e.clip_r2);
        }
    }
    for (auto &f : B.facets) {
        const int first = (f.kind == NSGPU_PATCH_FACET_SING) ? 0 : 1;
        const int np = 4 - first + (int)(f.self.size() / 4);
        push_patch(f.kind, f.center, probe_radius, np, 0, f.id[0], f.id[1], f.id[2]);
        for (int q = first; q < 4; ++q) ps->data.insert(ps->data.end(), f.planes[q], f.planes[q] + 4);
        ps->data.insert(ps->data.end(), f.self.begin(), f.self.end());
    }
    for (auto &e : B.edges) {
        if (e.kind < 0) continue;
        const int np = (int)(e.planes.size() / 8);
        push_patch(e.kind, e.clip_c, e.clip_r, np, e.flags, e.i, e.j, -1);
        ps->data.insert(ps->data.end(), e.center, e.center + 3);
        ps->data.insert(ps->data.end(), e.invrot, e.invrot + 9);
        ps->data.push_back(e.R);
        ps->data.push_back(e.self_r);
        ps->data.insert(ps->data.end(), e.planes.begin(), e.planes.end());
    }
    ps->errors = B.errors;
    *out = ps;
    return B.errors ? NSGPU_E_BUILD : NSGPU_OK;
}

extern "C" void nsgpu_patch_set_view(const nsgpu_patch_set *ps, nsgpu_patches *v) {
    if (!ps || !v) return;
    v->n = (int32_t)ps->kind.size();
    v->n_data = (int32_t)ps->data.size();
    v->kind = ps->kind.data();
    v->off = ps->off.data();
    v->cnt = ps->cnt.data();
    v->flags = ps->flags.data();
    v->atom = ps->atom.data();
    v->bsph = ps->bsph.data();
    v->data = ps->data.data();
}

extern "C" void nsgpu_patch_set_counts(const nsgpu_patch_set *ps, int64_t counts[5]) {
    if (!ps || !counts) return;
    for (int q = 0; q < 5; ++q) counts[q] = ps->counts[q];
}

extern "C" void nsgpu_patch_set_free(nsgpu_patch_set *ps) { delete ps; }
