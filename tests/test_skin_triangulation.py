"""The library's regular triangulation of the skin's weighted points
(nsgpu_skin_triangulate, in place of CGAL's Regular_triangulation_3,
SkinSurface.cpp:1271-1345) and its exact predicates.

Pins: the tetrahedra equal, as a set, the lower hull of the lifted points that
Qhull computes (an independent algorithm) on chains, 1crn and 6xm4 (whose
4,559 hidden atoms both leave out); every tetrahedron's orthosphere is empty
(brute force over all points) and neighbours share faces; the predicates'
signs equal exact rational determinants (fractions.Fraction) on random,
near-degenerate and exactly degenerate inputs (the dummies are coplanar in
fours)."""
import os
from fractions import Fraction

import numpy as np
import pytest

import nanoshaper_amd as ns
from nanoshaper_amd import _lib
from nanoshaper_amd.surface import _ptr

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _sets(x, s=0.45):
    jit, _ = ns.jitter(x)
    tets, nbrs = ns.skin_triangulation(jit, s)
    qt, _ = ns.regular_triangulation(ns.skin_points(jit, s))
    return jit, tets, nbrs, {tuple(sorted(t)) for t in tets.tolist()}, {tuple(sorted(t)) for t in qt.tolist()}


@pytest.mark.parametrize("n,radius,seed", [(1, 1.0, 0), (30, 5.0, 1), (300, 11.0, 2), (2000, 21.0, 0)])
def test_tetrahedra_equal_the_lifted_lower_hull(n, radius, seed):
    x, _ = ns.generate_chain(n, radius, seed=seed)
    _, tets, _, mine, qhull = _sets(x)
    assert len(tets) == len(mine) and mine == qhull


def test_1crn_and_6xm4():
    for name, present in (("1crn.xyzr", 327), ("6xm4.xyzr", 43960)):
        x = ns.load_atoms(os.path.join(GOLDEN, name))
        _, tets, _, mine, qhull = _sets(x)
        assert mine == qhull, name
        # hidden atoms (6xm4: small weights inside larger ones) are left out by both
        assert len(set(tets.ravel().tolist()) & set(range(len(x)))) == present, name


def test_a_hidden_atom_is_left_out():
    # a small atom deep inside a large one: its weighted point is dominated
    x = np.array([[0.0, 0.0, 0.0, 2.0], [0.05, 0.0, 0.0, 0.5], [3.0, 0.0, 0.0, 1.7], [0.0, 3.1, 0.0, 1.6]])
    _, tets, _, mine, qhull = _sets(x)
    assert mine == qhull
    assert 1 not in set(tets.ravel().tolist())


def test_orthospheres_are_empty_and_neighbours_share_faces():
    x, _ = ns.generate_chain(120, 8.0, seed=5)
    jit, _ = ns.jitter(x)
    tets, nbrs = ns.skin_triangulation(jit, 0.45)
    w = ns.skin_points(jit, 0.45)
    P, W = w[:, :3], w[:, 3]
    for t, v in enumerate(tets):
        # orthocentre: |c - p_i|^2 - w_i equal for the four vertices
        A = 2 * (P[v[1:]] - P[v[0]])
        b = (P[v[1:]] ** 2).sum(1) - W[v[1:]] - ((P[v[0]] ** 2).sum() - W[v[0]])
        c = np.linalg.solve(A, b)
        pw = ((P - c) ** 2).sum(1) - W
        r2 = pw[v[0]]
        others = np.setdiff1d(np.arange(len(P)), v)
        assert (pw[others] >= r2 - 1e-7 * max(1.0, abs(r2))).all(), t
        for k in range(4):
            u = nbrs[t, k]
            if u >= 0:
                assert len(set(v) & set(tets[u])) == 3 and v[k] not in tets[u]
                assert t in nbrs[u]
    # every face of a finite tetrahedron is shared, or lies on the dummies' hull
    hull = sum(int((nbrs[t] < 0).sum()) for t in range(len(tets)))
    assert hull == 8


def _exact_orient(a, b, c, d):
    m = [[Fraction(q[i]) - Fraction(a[i]) for i in range(3)] for q in (b, c, d)]
    det = (m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
           m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]))
    return (det > 0) - (det < 0)


def _det(m):
    if len(m) == 1:
        return m[0][0]
    return sum((-1) ** j * m[0][j] * _det([row[:j] + row[j + 1:] for row in m[1:]]) for j in range(len(m)))


def _exact_power(a, b, c, d, p):
    rows = []
    for q in (b, c, d, p):
        dq = [Fraction(q[i]) - Fraction(a[i]) for i in range(3)]
        rows.append(dq + [sum(x * x for x in dq) - (Fraction(q[3]) - Fraction(a[3]))])
    det = _det(rows)
    return (det > 0) - (det < 0)


def _lib_orient(a, b, c, d):
    arr = [np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c, d)]
    return _lib.load().nsgpu_orient3d(*[_ptr(v) for v in arr])


def _lib_power(a, b, c, d, p):
    arr = [np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c, d, p)]
    return _lib.load().nsgpu_power_test(*[_ptr(v) for v in arr])


def test_orient_matches_exact_rationals():
    rng = np.random.default_rng(3)
    for trial in range(300):
        a, b, c = rng.normal(size=(3, 3)) * 10
        kind = trial % 3
        if kind == 0:
            d = rng.normal(size=3) * 10
        elif kind == 1:
            # nearly coplanar: an affine combination, rounded to doubles
            u, v = rng.uniform(-2, 2, 2)
            d = a + u * (b - a) + v * (c - a)
        else:
            # one ulp away from that
            u, v = rng.uniform(-2, 2, 2)
            d = np.nextafter(a + u * (b - a) + v * (c - a), np.inf)
        assert _lib_orient(a, b, c, d) == _exact_orient(a, b, c, d), trial
    # exactly coplanar (the dummies' axis planes): zero
    z = 1.234567
    pts = [np.array([rng.normal(), rng.normal(), z]) for _ in range(4)]
    assert _lib_orient(*pts) == 0 == _exact_orient(*pts)


def test_power_test_matches_exact_rationals_and_the_orthosphere():
    rng = np.random.default_rng(4)
    for trial in range(200):
        pts = [np.append(rng.normal(size=3) * 3, rng.uniform(0.5, 6.0)) for _ in range(4)]
        if _exact_orient(*pts) < 0:
            pts[0], pts[1] = pts[1], pts[0]
        P = np.array([q[:3] for q in pts])
        W = np.array([q[3] for q in pts])
        A = 2 * (P[1:] - P[0])
        bb = (P[1:] ** 2).sum(1) - W[1:] - ((P[0] ** 2).sum() - W[0])
        c = np.linalg.solve(A, bb)
        R2 = ((P[0] - c) ** 2).sum() - W[0]
        q = np.append(rng.normal(size=3) * 3, rng.uniform(0.5, 6.0))
        if trial % 2:
            # on the orthosphere up to rounding: power of q ~ 0
            q[3] = ((q[:3] - c) ** 2).sum() - R2
        e = _exact_power(*pts, q)
        assert _lib_power(*pts, q) == e, trial
        if trial % 2 == 0:
            power = ((q[:3] - c) ** 2).sum() - q[3] - R2
            if abs(power) > 1e-6:
                # negative power (conflict) <=> negative sign, for a positive cell
                assert (power < 0) == (e < 0), trial


def test_skin_cells_from_either_triangulation():
    x, _ = ns.generate_chain(200, 9.5, seed=7)
    jit, _ = ns.jitter(x)
    a = ns.skin_patches(jit, 0.45)
    b = ns.skin_patches(jit, 0.45, triangulation=ns.regular_triangulation(ns.skin_points(jit, 0.45)))
    # cells per kind (skin point, edge, facet, tetrahedron: kinds 5..8)
    assert list(np.bincount(a.kind, minlength=9)) == list(np.bincount(b.kind, minlength=9))
    assert np.bincount(a.kind, minlength=9)[5:].min() > 0

    # the point, edge and facet cells name the same atoms (a tetrahedron cell
    # keeps three of its four, which depend on the vertex order)
    def cells(ps):
        keep = ps.kind != 8  # NSGPU_PATCH_SKIN_TETRA
        return sorted((int(k), tuple(sorted(i for i in ids if i >= 0)))
                      for k, ids in zip(ps.kind[keep].tolist(), ps.atom[keep].tolist()))

    assert cells(a) == cells(b)


def test_weighted_points_of_any_weight():
    # nsgpu_regular_triangulate on points whose weights include negatives
    # and zero: its atom-only tetrahedra are tetrahedra of Qhull's lower hull
    # (the six dummies far out may only replace some near the hull), each
    # with an empty orthosphere, and every point not dominated by the others
    # is a vertex
    rng = np.random.default_rng(11)
    p = np.column_stack([rng.uniform(-6, 6, (300, 3)), rng.uniform(-1.0, 4.0, 300)])
    p[::17, 3] = 0.0
    n = len(p)
    tets, nbrs = ns.weighted_triangulation(p)
    inner = tets[(tets < n).all(axis=1)]
    mine = {tuple(sorted(t)) for t in inner.tolist()}
    qt, _ = ns.regular_triangulation(p)
    qhull = {tuple(sorted(t)) for t in qt.tolist()}
    assert mine <= qhull and len(mine) > 0.8 * len(qhull)
    assert set(tets.ravel().tolist()) >= set(qt.ravel().tolist())
    P, W = p[:, :3], p[:, 3]
    for v in inner:
        A = 2 * (P[v[1:]] - P[v[0]])
        b = (P[v[1:]] ** 2).sum(1) - W[v[1:]] - ((P[v[0]] ** 2).sum() - W[v[0]])
        c = np.linalg.solve(A, b)
        pw = ((P - c) ** 2).sum(1) - W
        assert (pw >= pw[v[0]] - 1e-7 * max(1.0, abs(pw[v[0]]))).all()
    for t in range(len(tets)):
        for k in range(4):
            u = nbrs[t, k]
            if u >= 0:
                assert len(set(tets[t]) & set(tets[u])) == 3
