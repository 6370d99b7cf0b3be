// Regular (weighted Delaunay) triangulation of the skin's weighted points,
// the host stage the reference takes from CGAL's Regular_triangulation_3
// (SkinSurface.cpp:1271-1345).
//
// The points are the atoms (x, y, z, r^2/s) and six dummy points of weight -1
// on the axes through the atoms' box centre, two extents outside it
// (nsgpu_skin_points).  The dummies span the hull: they are triangulated
// first, as an octahedron split along one of its three diagonals (the split
// whose tetrahedra are regular), and every atom lies strictly inside.  The
// atoms are then inserted one by one in Morton order (Bowyer-Watson): the
// cells whose power spheres the new point is in conflict with form a
// connected, star-shaped hole that is re-triangulated from the point;
// vertices inside the hole disappear (they became hidden), and a point in
// conflict with no cell is hidden itself.
//
// Predicates are exact: a floating-point evaluation with an error bound, and
// where that cannot decide the sign, exact expansion arithmetic (Shewchuk,
// "Adaptive precision floating-point arithmetic and fast robust geometric
// predicates", 1997: two-sum, two-product, grow- and scale-expansion).  The
// dummies are exactly coplanar in fours (each axis plane holds four of them);
// the octahedron start keeps those quadruples out of every predicate that
// decides a cell.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <tuple>
#include <vector>

#include "nsgpu.h"

namespace {

// ---- exact arithmetic: expansions (nonoverlapping components, increasing
// magnitude, zeros dropped; the value is the sum of the components)
using Exp = std::vector<double>;

inline void two_sum(double a, double b, double &x, double &y) {
    x = a + b;
    const double bv = x - a, av = x - bv;
    y = (a - av) + (b - bv);
}
inline void fast_two_sum(double a, double b, double &x, double &y) {  // |a| >= |b|
    x = a + b;
    y = b - (x - a);
}
inline void two_prod(double a, double b, double &x, double &y) {
    x = a * b;
    y = std::fma(a, b, -x);
}

// e + b
Exp grow(const Exp &e, double b) {
    Exp h;
    h.reserve(e.size() + 1);
    double q = b;
    for (double ei : e) {
        double Q, hh;
        two_sum(q, ei, Q, hh);
        if (hh != 0.0) h.push_back(hh);
        q = Q;
    }
    if (q != 0.0 || h.empty()) h.push_back(q);
    return h;
}

Exp add(const Exp &e, const Exp &f) {
    Exp h = e;
    for (double fi : f) h = grow(h, fi);
    return h;
}

Exp neg(Exp e) {
    for (double &x : e) x = -x;
    return e;
}

// e * b
Exp scale(const Exp &e, double b) {
    Exp h;
    if (e.empty() || b == 0.0) return Exp{0.0};
    h.reserve(2 * e.size());
    double Q, hh;
    two_prod(e[0], b, Q, hh);
    if (hh != 0.0) h.push_back(hh);
    for (size_t i = 1; i < e.size(); ++i) {
        double T, t, Q2;
        two_prod(e[i], b, T, t);
        two_sum(Q, t, Q2, hh);
        if (hh != 0.0) h.push_back(hh);
        fast_two_sum(T, Q2, Q, hh);
        if (hh != 0.0) h.push_back(hh);
    }
    if (Q != 0.0 || h.empty()) h.push_back(Q);
    return h;
}

Exp mul(const Exp &e, const Exp &f) {
    Exp h{0.0};
    for (double fi : f) h = add(h, scale(e, fi));
    return h;
}

int sign(const Exp &e) {
    for (size_t i = e.size(); i-- > 0;)
        if (e[i] != 0.0) return e[i] > 0 ? 1 : -1;
    return 0;
}

Exp diff(double a, double b) {  // a - b, exactly
    double x, y;
    two_sum(a, -b, x, y);
    return y != 0.0 ? Exp{y, x} : Exp{x};
}

// determinant of the 3x3 matrix of expansions (rows r[0..2])
Exp det3(const Exp r[3][3]) {
    const Exp m0 = add(mul(r[1][1], r[2][2]), neg(mul(r[1][2], r[2][1])));
    const Exp m1 = add(mul(r[1][0], r[2][2]), neg(mul(r[1][2], r[2][0])));
    const Exp m2 = add(mul(r[1][0], r[2][1]), neg(mul(r[1][1], r[2][0])));
    return add(add(mul(r[0][0], m0), neg(mul(r[0][1], m1))), mul(r[0][2], m2));
}

// ---- predicates
// orient(a, b, c, d) = sign of det[b - a; c - a; d - a]: positive when
// (b - a, c - a, d - a) is right-handed.
int orient(const double *a, const double *b, const double *c, const double *d) {
    const double bx = b[0] - a[0], by = b[1] - a[1], bz = b[2] - a[2];
    const double cx = c[0] - a[0], cy = c[1] - a[1], cz = c[2] - a[2];
    const double dx = d[0] - a[0], dy = d[1] - a[1], dz = d[2] - a[2];
    const double m0 = cy * dz - cz * dy, m1 = cx * dz - cz * dx, m2 = cx * dy - cy * dx;
    const double det = bx * m0 - by * m1 + bz * m2;
    const double perm = std::fabs(bx) * (std::fabs(cy * dz) + std::fabs(cz * dy)) +
                        std::fabs(by) * (std::fabs(cx * dz) + std::fabs(cz * dx)) +
                        std::fabs(bz) * (std::fabs(cx * dy) + std::fabs(cy * dx));
    // the rounding errors of the differences and of the products sum to
    // less than 8 ulps of the permanent
    const double bound = 1e-14 * perm;
    if (det > bound) return 1;
    if (det < -bound) return -1;
    if (perm == 0.0) return 0;
    const Exp r[3][3] = {{diff(b[0], a[0]), diff(b[1], a[1]), diff(b[2], a[2])},
                         {diff(c[0], a[0]), diff(c[1], a[1]), diff(c[2], a[2])},
                         {diff(d[0], a[0]), diff(d[1], a[1]), diff(d[2], a[2])}};
    return sign(det3(r));
}

// Power test of weighted points (x, y, z, w): the sign of the determinant of
// the rows (q - a, |q - a|^2 - (w_q - w_a)) for q = b, c, d, p (the lifted
// points relative to a).  For a positively oriented cell (a, b, c, d) it is
// negative exactly when p is in conflict with the cell: p's weighted point
// and the cell's orthogonal sphere have negative power (the lifted p lies
// below the lifted cell).
int power_test(const double *a, const double *b, const double *c, const double *d, const double *p) {
    const double *q[4] = {b, c, d, p};
    double m[4][4], mag[4];  // mag: the lifted entry's terms in magnitude (its rounding scale)
    for (int i = 0; i < 4; ++i) {
        const double x = q[i][0] - a[0], y = q[i][1] - a[1], z = q[i][2] - a[2];
        const double r2 = x * x + y * y + z * z, dw = q[i][3] - a[3];
        m[i][0] = x;
        m[i][1] = y;
        m[i][2] = z;
        m[i][3] = r2 - dw;
        mag[i] = r2 + std::fabs(dw);
    }
    // Laplace expansion along the last column, with its permanent
    double det = 0.0, perm = 0.0;
    for (int i = 0; i < 4; ++i) {
        int r[3], k = 0;
        for (int j = 0; j < 4; ++j)
            if (j != i) r[k++] = j;
        const double *u = m[r[0]], *v = m[r[1]], *w = m[r[2]];
        const double c0 = v[1] * w[2] - v[2] * w[1], c1 = v[0] * w[2] - v[2] * w[0], c2 = v[0] * w[1] - v[1] * w[0];
        const double d3 = u[0] * c0 - u[1] * c1 + u[2] * c2;
        const double p3 = std::fabs(u[0]) * (std::fabs(v[1] * w[2]) + std::fabs(v[2] * w[1])) +
                          std::fabs(u[1]) * (std::fabs(v[0] * w[2]) + std::fabs(v[2] * w[0])) +
                          std::fabs(u[2]) * (std::fabs(v[0] * w[1]) + std::fabs(v[1] * w[0]));
        const double sgn = ((i + 3) % 2 == 0) ? 1.0 : -1.0;
        det += sgn * m[i][3] * d3;
        perm += mag[i] * p3;
    }
    const double bound = 1e-13 * perm;
    if (det > bound) return 1;
    if (det < -bound) return -1;
    if (perm == 0.0) return 0;
    // exact: the same expansion over exact entries
    Exp e[4][4];
    for (int i = 0; i < 4; ++i) {
        for (int k = 0; k < 3; ++k) e[i][k] = diff(q[i][k], a[k]);
        Exp l = add(add(mul(e[i][0], e[i][0]), mul(e[i][1], e[i][1])), mul(e[i][2], e[i][2]));
        e[i][3] = add(l, neg(diff(q[i][3], a[3])));
    }
    Exp total{0.0};
    for (int i = 0; i < 4; ++i) {
        int r[3], k = 0;
        for (int j = 0; j < 4; ++j)
            if (j != i) r[k++] = j;
        const Exp minor[3][3] = {{e[r[0]][0], e[r[0]][1], e[r[0]][2]},
                                 {e[r[1]][0], e[r[1]][1], e[r[1]][2]},
                                 {e[r[2]][0], e[r[2]][1], e[r[2]][2]}};
        Exp t = mul(e[i][3], det3(minor));
        total = add(total, ((i + 3) % 2 == 0) ? t : neg(t));
    }
    return sign(total);
}

// ---- the triangulation
constexpr int kInf = -1;

struct Cell {
    int v[4];
    int n[4];
    unsigned stamp;
    bool alive;
};

struct Tri {
    const double *pts = nullptr;  // 4 per point
    int npts = 0;
    std::vector<Cell> cells;
    std::vector<int> freelist;
    unsigned stamp = 0;
    int last = -1;
    std::string err;

    const double *P(int v) const { return pts + 4 * (size_t)v; }
    static bool infinite(const Cell &c) { return c.v[0] == kInf || c.v[1] == kInf || c.v[2] == kInf || c.v[3] == kInf; }

    int new_cell(int a, int b, int c, int d) {
        int id;
        if (!freelist.empty()) {
            id = freelist.back();
            freelist.pop_back();
        } else {
            id = (int)cells.size();
            cells.push_back(Cell{});
        }
        Cell &x = cells[id];
        x.v[0] = a;
        x.v[1] = b;
        x.v[2] = c;
        x.v[3] = d;
        x.n[0] = x.n[1] = x.n[2] = x.n[3] = -1;
        x.stamp = 0;
        x.alive = true;
        return id;
    }

    // orientation of cell c with vertex slot i replaced by point q (an
    // infinite slot stands for q itself)
    int orient_with(const Cell &c, int i, const double *q) const {
        const double *p[4];
        for (int k = 0; k < 4; ++k) p[k] = (k == i) ? q : (c.v[k] == kInf ? q : P(c.v[k]));
        return orient(p[0], p[1], p[2], p[3]);
    }

    // conflict of point q with cell c
    bool conflict(const Cell &c, const double *q) const {
        for (int k = 0; k < 4; ++k)
            if (c.v[k] == kInf) return orient_with(c, k, q) > 0;  // beyond the hull facet
        return power_test(P(c.v[0]), P(c.v[1]), P(c.v[2]), P(c.v[3]), q) < 0;
    }

    // pairs up the facets of the given cells by their vertex triples
    void link(const std::vector<int> &ids) {
        struct F {
            int a, b, c, cell, slot;
        };
        std::vector<F> fs;
        for (int id : ids)
            for (int k = 0; k < 4; ++k) {
                int t[3], j = 0;
                for (int q = 0; q < 4; ++q)
                    if (q != k) t[j++] = cells[id].v[q];
                std::sort(t, t + 3);
                fs.push_back(F{t[0], t[1], t[2], id, k});
            }
        std::sort(fs.begin(), fs.end(), [](const F &x, const F &y) {
            return std::tie(x.a, x.b, x.c) < std::tie(y.a, y.b, y.c);
        });
        for (size_t i = 0; i + 1 < fs.size(); ++i)
            if (fs[i].a == fs[i + 1].a && fs[i].b == fs[i + 1].b && fs[i].c == fs[i + 1].c) {
                cells[fs[i].cell].n[fs[i].slot] = fs[i + 1].cell;
                cells[fs[i + 1].cell].n[fs[i + 1].slot] = fs[i].cell;
                ++i;
            }
    }

    // The six dummies (indices d[0..5]: -x, +x, -y, +y, -z, +z) as an
    // octahedron: four cells around one diagonal, the split whose cells are
    // regular (no dummy in conflict with a cell), and eight infinite cells
    // on its faces.
    bool octahedron(const int d[6]) {
        for (int axis = 0; axis < 3; ++axis) {
            const int A = d[2 * axis], B = d[2 * axis + 1];
            const int u = (axis + 1) % 3, w = (axis + 2) % 3;
            const int ring[4] = {d[2 * u], d[2 * w], d[2 * u + 1], d[2 * w + 1]};
            int cs[4][4];
            bool ok = true;
            for (int i = 0; i < 4 && ok; ++i) {
                int t[4] = {A, B, ring[i], ring[(i + 1) % 4]};
                const int o = orient(P(t[0]), P(t[1]), P(t[2]), P(t[3]));
                if (o == 0) ok = false;
                if (o < 0) std::swap(t[0], t[1]);
                std::memcpy(cs[i], t, sizeof t);
            }
            // regular: the two dummies outside each cell are not in conflict with it
            for (int i = 0; i < 4 && ok; ++i)
                for (int k = 0; k < 4 && ok; ++k) {
                    const int q = ring[k];
                    if (q == ring[i] || q == ring[(i + 1) % 4]) continue;
                    if (power_test(P(cs[i][0]), P(cs[i][1]), P(cs[i][2]), P(cs[i][3]), P(q)) < 0) ok = false;
                }
            if (!ok) continue;
            std::vector<int> ids;
            for (int i = 0; i < 4; ++i) ids.push_back(new_cell(cs[i][0], cs[i][1], cs[i][2], cs[i][3]));
            // hull faces: (apex, ring[i], ring[i+1]) for both apexes.  The
            // infinite vertex takes the place of the other apex, which lies
            // inside; two finite vertices are swapped, so that a point in
            // the infinite slot orients the cell positively exactly when it
            // lies beyond the face
            for (int i = 0; i < 4; ++i)
                for (int apex = 0; apex < 2; ++apex) {
                    Cell c = cells[ids[i]];
                    const int other = apex == 0 ? B : A;
                    int fin[3], j = 0;
                    for (int k = 0; k < 4; ++k) {
                        if (c.v[k] == other)
                            c.v[k] = kInf;
                        else
                            fin[j++] = k;
                    }
                    std::swap(c.v[fin[0]], c.v[fin[1]]);
                    ids.push_back(new_cell(c.v[0], c.v[1], c.v[2], c.v[3]));
                }
            link(ids);
            last = ids[0];
            return true;
        }
        err = "the dummy octahedron has no regular split";
        return false;
    }

    // finite cell containing q (or on whose boundary q lies), walking from
    // `start`; -1 when the walk leaves the hull
    int locate(const double *q, int start) const {
        int c = start;
        unsigned rot = 0;
        for (size_t guard = 0; guard < 4 * cells.size() + 64; ++guard) {
            const Cell &x = cells[c];
            if (infinite(x)) return -1;
            int next = -1;
            rot = rot * 1103515245u + 12345u;
            const int r0 = (int)((rot >> 16) & 3u);
            for (int t = 0; t < 4; ++t) {
                const int i = (r0 + t) & 3;
                if (orient_with(x, i, q) < 0) {
                    next = x.n[i];
                    break;
                }
            }
            if (next < 0) return c;
            c = next;
        }
        return -1;
    }

    // Inserts point id; false with err set on a failure.
    bool insert(int id) {
        const double *q = P(id);
        int c0 = locate(q, last);
        if (c0 < 0) {
            err = "a point lies outside the dummies' hull";
            return false;
        }
        if (!conflict(cells[c0], q)) return true;  // hidden
        ++stamp;
        std::vector<int> stack{c0}, hole{c0};
        struct B {
            int cell, slot;
        };
        std::vector<B> boundary;
        cells[c0].stamp = stamp;
        while (!stack.empty()) {
            const int c = stack.back();
            stack.pop_back();
            for (int k = 0; k < 4; ++k) {
                const int nb = cells[c].n[k];
                if (cells[nb].stamp == stamp) continue;
                if (conflict(cells[nb], q)) {
                    if (infinite(cells[nb])) {
                        err = "a point is in conflict with a hull cell";
                        return false;
                    }
                    cells[nb].stamp = stamp;
                    stack.push_back(nb);
                    hole.push_back(nb);
                } else {
                    boundary.push_back(B{c, k});
                }
            }
        }
        // star the hole from q: each boundary facet with q in place of the
        // hole cell's vertex opposite it
        std::vector<int> fresh;
        fresh.reserve(boundary.size());
        for (const B &b : boundary) {
            const Cell &h = cells[b.cell];
            int v[4] = {h.v[0], h.v[1], h.v[2], h.v[3]};
            v[b.slot] = id;
            if (orient(P(v[0]), P(v[1]), P(v[2]), P(v[3])) <= 0) {
                err = "a degenerate cell (a point on the plane of a hole facet)";
                return false;
            }
            const int outside = h.n[b.slot];
            const int nc = new_cell(v[0], v[1], v[2], v[3]);
            cells[nc].n[b.slot] = outside;
            Cell &o = cells[outside];
            for (int k = 0; k < 4; ++k)
                if (o.n[k] == b.cell) o.n[k] = nc;
            fresh.push_back(nc);
        }
        // the new cells' facets through q pair up along the hole's edges
        struct E {
            int a, b, cell, slot;
        };
        std::vector<E> es;
        es.reserve(3 * fresh.size());
        for (int nc : fresh) {
            const Cell &x = cells[nc];
            int qs = 0;
            for (int k = 0; k < 4; ++k)
                if (x.v[k] == id) qs = k;
            for (int k = 0; k < 4; ++k) {
                if (k == qs) continue;
                int t[2], j = 0;
                for (int m = 0; m < 4; ++m)
                    if (m != k && m != qs) t[j++] = x.v[m];
                es.push_back(E{std::min(t[0], t[1]), std::max(t[0], t[1]), nc, k});
            }
        }
        std::sort(es.begin(), es.end(), [](const E &x, const E &y) { return std::tie(x.a, x.b) < std::tie(y.a, y.b); });
        for (size_t i = 0; i < es.size(); i += 2) {
            if (i + 1 >= es.size() || es[i].a != es[i + 1].a || es[i].b != es[i + 1].b) {
                err = "the hole is not a closed star";
                return false;
            }
            cells[es[i].cell].n[es[i].slot] = es[i + 1].cell;
            cells[es[i + 1].cell].n[es[i + 1].slot] = es[i].cell;
        }
        for (int h : hole) {
            cells[h].alive = false;
            freelist.push_back(h);
        }
        last = fresh.empty() ? last : fresh[0];
        return true;
    }
};

// Morton order of the atoms (10 bits per axis over their box)
std::vector<int> morton_order(const double *pts, int n) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], pts[4 * (size_t)i + k]);
            hi[k] = std::max(hi[k], pts[4 * (size_t)i + k]);
        }
    auto spread = [](uint64_t x) {
        uint64_t r = 0;
        for (int b = 0; b < 10; ++b) r |= ((x >> b) & 1ull) << (3 * b);
        return r;
    };
    std::vector<std::pair<uint64_t, int>> key(n);
    for (int i = 0; i < n; ++i) {
        uint64_t m = 0;
        for (int k = 0; k < 3; ++k) {
            const double ext = hi[k] - lo[k];
            const double f = ext > 0 ? (pts[4 * (size_t)i + k] - lo[k]) / ext : 0.0;
            const uint64_t cell = (uint64_t)std::min(1023.0, std::max(0.0, std::floor(f * 1024.0)));
            m |= spread(cell) << k;
        }
        key[i] = {m, i};
    }
    std::sort(key.begin(), key.end());
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = key[i].second;
    return order;
}

}  // namespace

struct nsgpu_regular_tri {
    std::vector<int32_t> tets, nbrs;
};

extern "C" {

int nsgpu_orient3d(const double *a, const double *b, const double *c, const double *d) {
    return orient(a, b, c, d);
}

int nsgpu_power_test(const double *a, const double *b, const double *c, const double *d, const double *p) {
    return power_test(a, b, c, d, p);
}

}  // extern "C"

namespace {

// The regular triangulation of w[0 .. n + 6) (4 doubles per point) whose last
// six points are the dummies -x, +x, -y, +y, -z, +z around the others.
int triangulate(const std::vector<double> &w, int n, nsgpu_regular_tri **out) {
    for (double x : w)
        if (!std::isfinite(x)) return NSGPU_E_ARG;
    Tri T;
    T.pts = w.data();
    T.npts = n + 6;
    T.cells.reserve((size_t)8 * (n + 6));
    const int d[6] = {n, n + 1, n + 2, n + 3, n + 4, n + 5};
    if (!T.octahedron(d)) return NSGPU_E_BUILD;
    for (int id : morton_order(w.data(), n))
        if (!T.insert(id)) return NSGPU_E_BUILD;
    auto *r = new nsgpu_regular_tri();
    std::vector<int> out_id(T.cells.size(), -1);
    int m = 0;
    for (size_t c = 0; c < T.cells.size(); ++c)
        if (T.cells[c].alive && !Tri::infinite(T.cells[c])) out_id[c] = m++;
    r->tets.resize((size_t)m * 4);
    r->nbrs.resize((size_t)m * 4);
    for (size_t c = 0; c < T.cells.size(); ++c) {
        const int o = out_id[c];
        if (o < 0) continue;
        for (int k = 0; k < 4; ++k) {
            r->tets[4 * (size_t)o + k] = T.cells[c].v[k];
            r->nbrs[4 * (size_t)o + k] = out_id[T.cells[c].n[k]];
        }
    }
    *out = r;
    return NSGPU_OK;
}

}  // namespace

extern "C" {

int nsgpu_skin_triangulate(const double *xyzr, int32_t n, double s, nsgpu_regular_tri **out) {
    if (!xyzr || n <= 0 || !out || !(s > 0 && s < 1)) return NSGPU_E_ARG;
    *out = nullptr;
    std::vector<double> w((size_t)(n + 6) * 4);
    const int rc = nsgpu_skin_points(xyzr, n, s, w.data());
    if (rc) return rc;
    return triangulate(w, n, out);
}

int nsgpu_regular_triangulate(const double *wpts, int32_t n, nsgpu_regular_tri **out) {
    if (!wpts || n <= 0 || !out) return NSGPU_E_ARG;
    *out = nullptr;
    // six dummies of weight -1 two extents outside the box of the points'
    // spheres (radius sqrt(w), 0 for a negative weight), as for the skin
    std::vector<double> xyzr((size_t)n * 4);
    for (int i = 0; i < n; ++i) {
        for (int q = 0; q < 3; ++q) xyzr[4 * (size_t)i + q] = wpts[4 * (size_t)i + q];
        xyzr[4 * (size_t)i + 3] = std::sqrt(std::max(0.0, wpts[4 * (size_t)i + 3]));
    }
    std::vector<double> w((size_t)(n + 6) * 4);
    const int rc = nsgpu_skin_points(xyzr.data(), n, 0.5, w.data());
    if (rc) return rc;
    std::memcpy(w.data(), wpts, (size_t)n * 4 * sizeof(double));  // the caller's weights
    return triangulate(w, n, out);
}

int32_t nsgpu_regular_tri_count(const nsgpu_regular_tri *t) { return t ? (int32_t)(t->tets.size() / 4) : 0; }

void nsgpu_regular_tri_read(const nsgpu_regular_tri *t, int32_t *tets, int32_t *nbrs) {
    if (!t) return;
    if (tets) std::memcpy(tets, t->tets.data(), t->tets.size() * sizeof(int32_t));
    if (nbrs) std::memcpy(nbrs, t->nbrs.data(), t->nbrs.size() * sizeof(int32_t));
}

void nsgpu_regular_tri_free(nsgpu_regular_tri *t) { delete t; }

int nsgpu_skin_patches(const double *xyzr, int32_t n, double s, nsgpu_patch_set **out) {
    if (!out) return NSGPU_E_ARG;
    *out = nullptr;
    nsgpu_regular_tri *t = nullptr;
    int rc = nsgpu_skin_triangulate(xyzr, n, s, &t);
    if (rc) return rc;
    rc = nsgpu_skin_build(xyzr, n, s, t->tets.data(), t->nbrs.data(), nsgpu_regular_tri_count(t), out);
    nsgpu_regular_tri_free(t);
    return rc;
}

}  // extern "C"
