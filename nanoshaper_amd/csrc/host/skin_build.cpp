// Host skin mixed-complex build (the stage SkinSurface::buildSkinCGAL performs
// on the CPU, SkinSurface.cpp:1271-1811) from a given regular triangulation.
//
// The reference inserts the atoms (weight r^2/s) and six far dummy points
// (weight -1) into CGAL's Regular_triangulation_3 and reads the mixed cells of
// the skin off it.  Here the caller supplies the triangulation of the same
// weighted points (nsgpu_skin_points): tetrahedra and their neighbours.  For
// points in general position the regular triangulation is unique, so any
// correct triangulator gives the cells the reference gets.  The cell geometry
// follows BuildDelaunayTetraCells (:348-480), BuildDelaunayPointCells
// (:482-566), BuildDelaunayEdgeCells (:568-892) and BuildDelaunayFacetCells
// (:894-1210):
//   Del3 x Vor0: sphere about the tetrahedron's orthocentre, inside the
//     tetrahedron of the shrunk vertices p + s (vor - p);
//   Del0 x Vor3: the atom's sphere of radius r, on the atom's side of the
//     upper/lower planes of its edge cells;
//   Del1 x Vor2, Del2 x Vor1: hyperboloids Q = C^-T diag(...) C^-1 in the
//     frame of the edge (facet) and its dual Voronoi face (edge), bounded by
//     upper, lower and lateral planes through the shrunk points.
// Cells touching a dummy point are dropped (QUADRIC_CULLING), as are
// tetrahedra with a negative orthoradius.  Every plane is stored so that a
// hit is rejected when a x + b y + c z + d > 0.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <vector>

#include "nsgpu.h"
#include "patch_set.hpp"

namespace {

inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// plane through three points (tools.cpp:258-274)
void plane3(const double p1[3], const double p2[3], const double p3[3], double w[4], bool normalize) {
    w[0] = p1[1] * (p2[2] - p3[2]) + p2[1] * (p3[2] - p1[2]) + p3[1] * (p1[2] - p2[2]);
    w[1] = p1[2] * (p2[0] - p3[0]) + p2[2] * (p3[0] - p1[0]) + p3[2] * (p1[0] - p2[0]);
    w[2] = p1[0] * (p2[1] - p3[1]) + p2[0] * (p3[1] - p1[1]) + p3[0] * (p1[1] - p2[1]);
    w[3] = -dot3(w, p1);
    if (normalize) {
        const double nrm = std::sqrt(dot3(w, w));
        for (int q = 0; q < 4; ++q) w[q] /= nrm;
    }
}

// orientation by an inner point: negative inside (SkinSurface.cpp:744-749)
void orient(double w[4], const double inner[3]) {
    if (dot3(w, inner) + w[3] > 0)
        for (int q = 0; q < 4; ++q) w[q] = -w[q];
}

void normalize3(double v[3]) {
    const double t = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] /= t;
    v[1] /= t;
    v[2] /= t;
}

void cross3(double d[3], const double a[3], const double b[3]) {
    d[0] = a[1] * b[2] - a[2] * b[1];
    d[1] = a[2] * b[0] - a[0] * b[2];
    d[2] = a[0] * b[1] - a[1] * b[0];
}

// first row of the inverse of a 3x3 matrix times v (INVERT_3X3, globals.h:371-377)
double inv_row0_dot(const double m[3][3], const double v[3]) {
    const double c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    const double c01 = m[0][2] * m[2][1] - m[0][1] * m[2][2];
    const double c02 = m[0][1] * m[1][2] - m[0][2] * m[1][1];
    const double det = m[0][0] * c00 + m[1][0] * c01 + m[2][0] * c02;
    const double id = 1.0 / det;
    return (c00 * id) * v[0] + (c01 * id) * v[1] + (c02 * id) * v[2];
}

// Q = Cinv^T M Cinv with C = [R | f] (R orthonormal, columns e0 e1 e2), M =
// diag(m): the quadric of a hyperboloid in the frame (SkinSurface.cpp:790-880)
void frame_quadric(const double e0[3], const double e1[3], const double e2[3], const double f[3], const double m[4],
                   double QQ[10]) {
    double Ci[4][4] = {};
    const double *E[3] = {e0, e1, e2};
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) Ci[r][c] = E[r][c];  // R^T
        Ci[r][3] = -dot3(E[r], f);
    }
    Ci[3][3] = 1.;
    double Q[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double acc = 0;
            for (int k = 0; k < 4; ++k) acc += Ci[k][i] * m[k] * Ci[k][j];
            Q[i][j] = acc;
        }
    // symmetric coding (SkinSurface.cpp:4508): q11 q22 q33 q44 q12 q23 q34 q13 q24 q14
    QQ[0] = Q[0][0];
    QQ[1] = Q[1][1];
    QQ[2] = Q[2][2];
    QQ[3] = Q[3][3];
    QQ[4] = Q[0][1];
    QQ[7] = Q[0][2];
    QQ[9] = Q[0][3];
    QQ[5] = Q[1][2];
    QQ[8] = Q[1][3];
    QQ[6] = Q[2][3];
}

struct Cell {
    int kind;
    int ids[3];
    double bsph[4];
    double box[6];
    double Q[10];
    std::vector<double> planes;  // 4 per plane, rejecting when > 0
};

void box_of(const std::vector<const double *> &pts, double box[6]) {
    for (int q = 0; q < 3; ++q) {
        box[q] = INFINITY;
        box[3 + q] = -INFINITY;
    }
    for (const double *p : pts)
        for (int q = 0; q < 3; ++q) {
            box[q] = box[q] < p[q] ? box[q] : p[q];
            box[3 + q] = box[3 + q] > p[q] ? box[3 + q] : p[q];
        }
}

}  // namespace

// Weighted points of the skin triangulation (SkinSurface.cpp:1290-1345): the
// atoms with weight r^2 / s, then six dummy points (weight -1) two extents
// outside the atoms' bounding box.  wpts receives (n + 6) * 4 doubles.
extern "C" int nsgpu_skin_points(const double *xyzr, int32_t n, double s, double *wpts) {
    if (!xyzr || n <= 0 || !wpts || !(s > 0 && s < 1)) return NSGPU_E_ARG;
    double max_x = -1e6, min_x = 1e6, max_y = -1e6, min_y = 1e6, max_z = -1e6, min_z = 1e6;
    for (int i = 0; i < n; ++i) {
        const double x = xyzr[4 * i], y = xyzr[4 * i + 1], z = xyzr[4 * i + 2], r = xyzr[4 * i + 3];
        wpts[4 * i] = x;
        wpts[4 * i + 1] = y;
        wpts[4 * i + 2] = z;
        wpts[4 * i + 3] = (r * r) / s;
        max_x = std::max(max_x, x + r);
        max_y = std::max(max_y, y + r);
        max_z = std::max(max_z, z + r);
        min_x = std::min(min_x, x - r);
        min_y = std::min(min_y, y - r);
        min_z = std::min(min_z, z - r);
    }
    const double mid_x = (max_x + min_x) * 0.5, mid_y = (max_y + min_y) * 0.5, mid_z = (max_z + min_z) * 0.5;
    min_x -= std::fabs(mid_x - min_x) * 2.;
    min_y -= std::fabs(mid_y - min_y) * 2.;
    min_z -= std::fabs(mid_z - min_z) * 2.;
    max_x += std::fabs(mid_x - max_x) * 2.;
    max_y += std::fabs(mid_y - max_y) * 2.;
    max_z += std::fabs(mid_z - max_z) * 2.;
    const double d[6][3] = {{min_x, mid_y, mid_z}, {max_x, mid_y, mid_z}, {mid_x, min_y, mid_z},
                            {mid_x, max_y, mid_z}, {mid_x, mid_y, min_z}, {mid_x, mid_y, max_z}};
    for (int q = 0; q < 6; ++q) {
        double *w = wpts + 4 * (int64_t)(n + q);
        w[0] = d[q][0];
        w[1] = d[q][1];
        w[2] = d[q][2];
        w[3] = -1.0;
    }
    return NSGPU_OK;
}

extern "C" int nsgpu_skin_build(const double *xyzr, int32_t n, double s, const int32_t *tets, const int32_t *nbrs,
                                int32_t ntets, nsgpu_patch_set **out) {
    if (!xyzr || n <= 0 || !tets || !nbrs || ntets <= 0 || !out || !(s > 0 && s < 1)) return NSGPU_E_ARG;
    *out = nullptr;
    const int np = n + 6;
    std::vector<double> wp((size_t)np * 4);
    nsgpu_skin_points(xyzr, n, s, wp.data());
    for (int64_t q = 0; q < 4 * (int64_t)ntets; ++q)
        if (tets[q] < 0 || tets[q] >= np || nbrs[q] < -1 || nbrs[q] >= ntets) return NSGPU_E_ARG;
    auto P = [&](int v) { return &wp[4 * (size_t)v]; };
    const auto is_dummy = [&](int v) { return v >= n; };

    // orthocentres and shrunk vertices of every tetrahedron
    std::vector<double> vor((size_t)ntets * 3), red((size_t)ntets * 12), rad2(ntets);
    for (int t = 0; t < ntets; ++t) {
        const int *v = tets + 4 * t;
        const double *p0 = P(v[0]);
        double A[3][3], b[3];
        for (int i = 1; i < 4; ++i) {
            const double *pi = P(v[i]);
            for (int q = 0; q < 3; ++q) A[i - 1][q] = 2. * (pi[q] - p0[q]);
            b[i - 1] = (dot3(pi, pi) - pi[3]) - (dot3(p0, p0) - p0[3]);
        }
        const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
                           A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                           A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
        double c[3];
        for (int q = 0; q < 3; ++q) {
            double M[3][3];
            for (int r = 0; r < 3; ++r)
                for (int k = 0; k < 3; ++k) M[r][k] = (k == q) ? b[r] : A[r][k];
            c[q] = (M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                    M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0])) / det;
        }
        for (int q = 0; q < 3; ++q) vor[3 * t + q] = c[q];
        for (int i = 0; i < 4; ++i) {
            const double *pi = P(v[i]);
            double *r = &red[12 * (size_t)t + 3 * i];
            for (int q = 0; q < 3; ++q) {
                const double a = c[q] - pi[q];
                r[q] = pi[q] + s * a;
            }
            if (i == 0) {
                const double a = c[0] - pi[0], bb = c[1] - pi[1], cc = c[2] - pi[2];
                rad2[t] = -(pi[3] - (a * a + bb * bb + cc * cc));
            }
        }
    }
    // the shrunk point of vertex `a` in tetrahedron t
    auto red_of = [&](int t, int a) -> const double * {
        for (int i = 0; i < 4; ++i)
            if (tets[4 * t + i] == a) return &red[12 * (size_t)t + 3 * i];
        return nullptr;
    };

    std::vector<Cell> tcells, pcells, ecells, fcells;
    // Del3 x Vor0 (BuildDelaunayTetraCells)
    for (int t = 0; t < ntets; ++t) {
        const int *v = tets + 4 * t;
        if (rad2[t] < 0) continue;
        if (is_dummy(v[0]) || is_dummy(v[1]) || is_dummy(v[2]) || is_dummy(v[3])) continue;
        Cell c{};
        c.kind = NSGPU_PATCH_SKIN_TETRA;
        c.ids[0] = v[0];
        c.ids[1] = v[1];
        c.ids[2] = v[2];
        const double *r[4] = {&red[12 * (size_t)t], &red[12 * (size_t)t + 3], &red[12 * (size_t)t + 6],
                              &red[12 * (size_t)t + 9]};
        double inner[3];
        for (int q = 0; q < 3; ++q) inner[q] = (r[0][q] + r[1][q] + r[2][q] + r[3][q]) / 4.;
        const int faces[4][3] = {{2, 1, 0}, {3, 2, 0}, {3, 1, 2}, {0, 1, 3}};
        for (auto &f : faces) {
            double w[4];
            plane3(r[f[0]], r[f[1]], r[f[2]], w, false);
            orient(w, inner);
            c.planes.insert(c.planes.end(), w, w + 4);
        }
        box_of({r[0], r[1], r[2], r[3]}, c.box);
        const double *vo = &vor[3 * (size_t)t];
        c.Q[0] = c.Q[1] = c.Q[2] = 1;
        c.Q[3] = -rad2[t] * (1 - s) + vo[0] * vo[0] + vo[1] * vo[1] + vo[2] * vo[2];
        c.Q[4] = c.Q[5] = c.Q[7] = 0;
        c.Q[9] = -vo[0];
        c.Q[8] = -vo[1];
        c.Q[6] = -vo[2];
        tcells.push_back(std::move(c));
    }
    // Del0 x Vor3 (BuildDelaunayPointCells): atoms that are vertices
    std::vector<int> pcell_of(n, -1);
    std::vector<char> present(n, 0);
    for (int64_t q = 0; q < 4 * (int64_t)ntets; ++q)
        if (tets[q] < n) present[tets[q]] = 1;
    for (int a = 0; a < n; ++a) {
        if (!present[a]) continue;
        Cell c{};
        c.kind = NSGPU_PATCH_SKIN_POINT;
        c.ids[0] = a;
        c.ids[1] = c.ids[2] = -1;
        const double *p = P(a);
        const double r = xyzr[4 * a + 3];
        c.bsph[0] = xyzr[4 * a];
        c.bsph[1] = xyzr[4 * a + 1];
        c.bsph[2] = xyzr[4 * a + 2];
        c.bsph[3] = r;
        for (int q = 0; q < 3; ++q) {
            c.box[q] = c.bsph[q] - r;
            c.box[3 + q] = c.bsph[q] + r;
        }
        c.Q[0] = c.Q[1] = c.Q[2] = 1;
        c.Q[3] = -s * (p[3]) + p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
        c.Q[4] = c.Q[5] = c.Q[7] = 0;
        c.Q[9] = -p[0];
        c.Q[8] = -p[1];
        c.Q[6] = -p[2];
        pcell_of[a] = (int)pcells.size();
        pcells.push_back(std::move(c));
    }
    // Del1 x Vor2 (BuildDelaunayEdgeCells): the tetrahedra around each edge in
    // cyclic order, walking across the faces that contain the edge
    std::map<std::pair<int, int>, int> seen;
    const int tedge[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    int errors = 0;
    for (int t0 = 0; t0 < ntets; ++t0)
        for (auto &te : tedge) {
            const int a = tets[4 * t0 + te[0]], b = tets[4 * t0 + te[1]];
            if (is_dummy(a) || is_dummy(b)) continue;
            const auto key = std::make_pair(std::min(a, b), std::max(a, b));
            if (seen.count(key)) continue;
            seen[key] = 1;
            std::vector<int> ring;
            int t = t0, prev = -1;
            bool closed = false;
            for (int guard = 0; guard < 4096; ++guard) {
                ring.push_back(t);
                // the two faces of t containing a-b are opposite its other two vertices
                int nxt = -1;
                for (int i = 0; i < 4; ++i) {
                    const int vi = tets[4 * t + i];
                    if (vi == a || vi == b) continue;
                    const int u = nbrs[4 * t + i];
                    if (u != prev) {
                        nxt = u;
                        break;
                    }
                }
                if (nxt < 0) break;  // the hull: dummies only, not reached for atom edges
                prev = t;
                t = nxt;
                if (t == t0) {
                    closed = true;
                    break;
                }
            }
            if (!closed || ring.size() < 3) {
                ++errors;
                continue;
            }
            const int num = (int)ring.size();
            std::vector<const double *> up(num), lo(num), pts;
            for (int i = 0; i < num; ++i) {
                up[i] = red_of(ring[i], a);
                lo[i] = red_of(ring[i], b);
                pts.push_back(up[i]);
                pts.push_back(lo[i]);
            }
            Cell c{};
            c.kind = NSGPU_PATCH_SKIN_EDGE;
            c.ids[0] = a;
            c.ids[1] = b;
            c.ids[2] = -1;
            box_of(pts, c.box);
            double upper[4], lower[4];
            plane3(up[0], up[1], up[2], upper, true);
            plane3(lo[0], lo[1], lo[2], lower, true);
            double inner[3];
            for (int q = 0; q < 3; ++q) inner[q] = (up[0][q] + up[1][q] + up[2][q] + lo[0][q] + lo[1][q] + lo[2][q]) / 6.;
            orient(upper, inner);
            orient(lower, inner);
            // the atoms' cells lie on the far sides of these planes
            double neg[4];
            for (int q = 0; q < 4; ++q) neg[q] = -upper[q];
            pcells[pcell_of[a]].planes.insert(pcells[pcell_of[a]].planes.end(), neg, neg + 4);
            for (int q = 0; q < 4; ++q) neg[q] = -lower[q];
            pcells[pcell_of[b]].planes.insert(pcells[pcell_of[b]].planes.end(), neg, neg + 4);
            for (int i = 0; i < num; ++i) {
                double w[4];
                plane3(up[i], lo[i], up[(i + 1) % num], w, true);
                orient(w, inner);
                c.planes.insert(c.planes.end(), w, w + 4);
            }
            c.planes.insert(c.planes.end(), upper, upper + 4);
            c.planes.insert(c.planes.end(), lower, lower + 4);
            // focus: the edge line meets the Voronoi face plane
            const double *v1p = P(a), *v2p = P(b);
            const double *f0 = &vor[3 * (size_t)ring[0]], *f1 = &vor[3 * (size_t)ring[1]], *f2 = &vor[3 * (size_t)ring[2]];
            double mat[3][3];
            for (int q = 0; q < 3; ++q) {
                mat[q][0] = v2p[q] - v1p[q];
                mat[q][1] = f1[q] - f0[q];
                mat[q][2] = f2[q] - f0[q];
            }
            const double v[3] = {v2p[0] - f0[0], v2p[1] - f0[1], v2p[2] - f0[2]};
            const double res = inv_row0_dot(mat, v);
            double focus[3];
            for (int q = 0; q < 3; ++q) focus[q] = v2p[q] + res * (v1p[q] - v2p[q]);
            const double aa = v1p[0] - focus[0], bb = v1p[1] - focus[1], cc = v1p[2] - focus[2];
            const double radius2 = v1p[3] - (aa * aa + bb * bb + cc * cc);
            double o1[3], o2[3], o3[3];
            for (int q = 0; q < 3; ++q) {
                o2[q] = f0[q] - focus[q];
                o1[q] = v2p[q] - v1p[q];
            }
            normalize3(o2);
            normalize3(o1);
            cross3(o3, o1, o2);
            normalize3(o3);
            const double m[4] = {-1. / (1. - s), 1. / s, 1. / s, -radius2};
            frame_quadric(o1, o3, o2, focus, m, c.Q);
            ecells.push_back(std::move(c));
        }
    // Del2 x Vor1 (BuildDelaunayFacetCells): facets between two tetrahedra
    for (int t = 0; t < ntets; ++t)
        for (int k = 0; k < 4; ++k) {
            const int u = nbrs[4 * t + k];
            if (u < 0 || u < t) continue;
            int d[3], ind = 0;
            for (int i = 0; i < 4; ++i)
                if (i != k) d[ind++] = tets[4 * t + i];
            if (is_dummy(d[0]) || is_dummy(d[1]) || is_dummy(d[2])) continue;
            const double *up[3], *lo[3];
            bool ok = true;
            for (int i = 0; i < 3; ++i) {
                up[i] = red_of(t, d[i]);
                lo[i] = red_of(u, d[i]);
                ok &= up[i] && lo[i];
            }
            if (!ok) {
                ++errors;
                continue;
            }
            Cell c{};
            c.kind = NSGPU_PATCH_SKIN_FACET;
            c.ids[0] = d[0];
            c.ids[1] = d[1];
            c.ids[2] = d[2];
            const double *mcp[6] = {up[0], lo[0], up[1], lo[1], up[2], lo[2]};
            box_of({mcp[0], mcp[1], mcp[2], mcp[3], mcp[4], mcp[5]}, c.box);
            double inner[3];
            for (int q = 0; q < 3; ++q) inner[q] = (up[0][q] + up[1][q] + up[2][q] + lo[0][q] + lo[1][q] + lo[2][q]) / 6.;
            double w[4];
            const int tri[5][3] = {{0, 2, 4}, {1, 3, 5}, {0, 1, 2}, {2, 3, 4}, {4, 5, 0}};
            for (auto &f : tri) {
                plane3(mcp[f[0]], mcp[f[1]], mcp[f[2]], w, true);
                orient(w, inner);
                c.planes.insert(c.planes.end(), w, w + 4);
            }
            const double *vor1 = &vor[3 * (size_t)t], *vor2 = &vor[3 * (size_t)u];
            const double *p0 = P(d[0]), *p1 = P(d[1]), *p2 = P(d[2]);
            double mat[3][3];
            for (int q = 0; q < 3; ++q) {
                mat[q][0] = vor2[q] - vor1[q];
                mat[q][1] = p1[q] - p0[q];
                mat[q][2] = p2[q] - p0[q];
            }
            const double v[3] = {vor2[0] - p0[0], vor2[1] - p0[1], vor2[2] - p0[2]};
            const double res = inv_row0_dot(mat, v);
            double focus[3];
            for (int q = 0; q < 3; ++q) focus[q] = vor2[q] + res * (vor1[q] - vor2[q]);
            const double aa = p0[0] - focus[0], bb = p0[1] - focus[1], cc = p0[2] - focus[2];
            const double radius2 = p0[3] - (aa * aa + bb * bb + cc * cc);
            double o1[3], o2[3], o3[3];
            for (int q = 0; q < 3; ++q) {
                o1[q] = vor1[q] - vor2[q];
                o2[q] = p0[q] - p1[q];
            }
            normalize3(o1);
            normalize3(o2);
            cross3(o3, o1, o2);
            normalize3(o3);
            const double m[4] = {-1. / (1. - s), -1. / (1. - s), 1. / s, -radius2};
            frame_quadric(o2, o3, o1, focus, m, c.Q);
            fcells.push_back(std::move(c));
        }

    auto *ps = new nsgpu_patch_set();
    for (auto *group : {&tcells, &pcells, &ecells, &fcells})
        for (auto &c : *group) {
            ps->kind.push_back(c.kind);
            ps->off.push_back((int32_t)ps->data.size());
            ps->cnt.push_back((int32_t)(c.planes.size() / 4));
            ps->flags.push_back(0);
            ps->atom.insert(ps->atom.end(), c.ids, c.ids + 3);
            ps->bsph.insert(ps->bsph.end(), c.bsph, c.bsph + 4);
            ps->data.insert(ps->data.end(), c.box, c.box + 6);
            ps->data.insert(ps->data.end(), c.Q, c.Q + 10);
            ps->data.insert(ps->data.end(), c.planes.begin(), c.planes.end());
        }
    ps->errors = errors;
    *out = ps;
    return errors ? NSGPU_E_BUILD : NSGPU_OK;
}
