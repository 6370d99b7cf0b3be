"""Pins the oracle's triangle-mesh and blobby restatements.

Independent checks: the ray/triangle test against a linear solve, the blobby
field against its closed form, the blobby mesh of one atom against the
sphere it must approximate, and the ray-cast maps of a closed mesh against
its exact polyhedral volume and a point-in-sphere classifier.
"""
from collections import Counter

import numpy as np
import pytest

import nanoshaper_amd as ns
from oracle import binding

OPTS = dict(build_epsmap=1, build_status=1, accurate_triangulation=1, fill_cavities=1, cav_vol=11.4, inside=5)


def _closed_manifold(t):
    ec = Counter()
    for tri in t:
        for a, b in ((tri[0], tri[1]), (tri[1], tri[2]), (tri[2], tri[0])):
            ec[(min(a, b), max(a, b))] += 1
    return all(c == 2 for c in ec.values()), len(ec)


def _signed_volume(v, t):
    p0, p1, p2 = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    return float(np.einsum("ij,ij->i", p0, np.cross(p1, p2)).sum() / 6)


def icosphere(radius, centre, levels):
    """Subdivided icosahedron projected on a sphere, outward winding."""
    p = (1 + 5 ** 0.5) / 2
    v = [(-1, p, 0), (1, p, 0), (-1, -p, 0), (1, -p, 0), (0, -1, p), (0, 1, p), (0, -1, -p), (0, 1, -p),
         (p, 0, -1), (p, 0, 1), (-p, 0, -1), (-p, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    v = [np.array(x, float) / np.linalg.norm(x) for x in v]
    for _ in range(levels):
        mid = {}

        def m(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                q = v[a] + v[b]
                v.append(q / np.linalg.norm(q))
                mid[k] = len(v) - 1
            return mid[k]

        f = [t for a, b, c in f for t in ((a, m(a, b), m(c, a)), (b, m(b, c), m(a, b)), (c, m(c, a), m(b, c)),
                                          (m(a, b), m(b, c), m(c, a)))]
    return np.array(v) * radius + np.asarray(centre), np.array(f, np.int32)


def test_ray_triangle_against_linear_solve():
    assert binding.ray_triangle([0, .25, .25], [4, 0, 0], [2, 0, 0], [2, 1, 0], [2, 0, 1]) == 0.5
    assert binding.ray_triangle([0, .8, .8], [4, 0, 0], [2, 0, 0], [2, 1, 0], [2, 0, 1]) is None
    # ray parallel to the plane: |det| < 1e-5
    assert binding.ray_triangle([0, .25, .25], [0, 4, 0], [2, 0, 0], [2, 1, 0], [2, 0, 1]) is None
    rng = np.random.default_rng(5)
    hits = 0
    for _ in range(2000):
        v0, v1, v2 = rng.normal(size=(3, 3)) * 2
        o = rng.normal(size=3) * 2 - np.array([6, 0, 0])
        d = np.array([12.0, 0, 0])
        # o + t d = v0 + u (v1 - v0) + w (v2 - v0)
        M = np.stack([-d, v1 - v0, v2 - v0], axis=1)
        if abs(np.linalg.det(M)) < 1e-3:
            continue
        t, u, w = np.linalg.solve(M, o - v0)
        margin = min(u, w, 1 - u - w)
        got = binding.ray_triangle(o, d, v0, v1, v2)
        if abs(margin) < 1e-9:
            continue
        if margin > 0:
            hits += 1
            assert got is not None and got == pytest.approx(t, abs=1e-10)
        else:
            assert got is None
    assert hits > 100


def test_blobby_field_closed_form():
    xyzr = np.array([[0.3, -0.2, 0.1, 1.7]])
    g = ns.build_grid(xyzr, 3.0, 50.0)
    B, cutoff = -2.5, 2.0
    b = binding.blobby_build(g, xyzr, B, cutoff)
    x = g.xmin + np.arange(g.nx) * g.side - g.hside
    y = g.ymin + np.arange(g.ny) * g.side - g.hside
    z = g.zmin + np.arange(g.nz) * g.side - g.hside
    ng = int(cutoff / g.side + 0.5)
    idx = [int(np.floor((xyzr[0, q] - m) * g.scale + 0.5)) for q, m in enumerate((g.xmin, g.ymin, g.zmin))]
    want = np.zeros((g.nz, g.ny, g.nx))
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    d2 = (X - xyzr[0, 0]) ** 2 + (Y - xyzr[0, 1]) ** 2 + (Z - xyzr[0, 2]) ** 2
    box = np.zeros_like(want, bool)
    box[max(0, idx[2] - ng):idx[2] + ng, max(0, idx[1] - ng):idx[1] + ng, max(0, idx[0] - ng):idx[0] + ng] = True
    want[box] = np.exp(B * (d2[box] / 1.7 ** 2 - 1))
    assert np.allclose(b["field"], want, rtol=1e-15, atol=0)
    assert (b["field"][~box] == 0).all()


def test_blobby_single_atom_is_a_sphere():
    r = 2.0
    xyzr = np.array([[0.11, 0.07, -0.05, r]])
    g = ns.build_grid(xyzr, 4.0, 50.0)
    b = binding.blobby_build(g, xyzr)
    v, t, raw = b["verts"], b["tris"], b["raw"]
    closed, ne = _closed_manifold(t)
    assert closed and len(v) - ne + len(t) == 2  # a sphere
    d = np.linalg.norm(v - xyzr[0, :3], axis=1)
    # the iso-surface phi = 1 of one atom is |p - c| = r; marching cubes and
    # one smoothing pass stay within a fraction of the 0.25 A spacing
    assert abs(d.mean() - r) < 0.05 and np.abs(d - r).max() < 0.1
    assert _signed_volume(v, t) == pytest.approx(4 / 3 * np.pi * r ** 3, rel=0.03)
    # the OFF round trip: 3 decimals, read as floats
    assert np.abs(v - raw).max() <= 5e-4 + 1e-6
    assert np.array_equal(v.astype(np.float32).astype(np.float64), v)


def test_mesh_ray_casting_volume_and_classification():
    R, c = 5.0, np.array([0.2, -0.1, 0.05])
    v, t = icosphere(R, c, 3)
    g = ns.build_grid(np.array([[*c, R]]), 3.0, 60.0)
    vol = _signed_volume(v, t)
    for ray_based in (True, False):
        o = binding.OracleSurface.from_mesh(g, v, t, ray_based=ray_based)
        r = o.get_surf(**OPTS)
        assert r["odd_rays"] == 0
        assert r["volume"] == pytest.approx(vol, rel=2e-3)
        st = o.status()
        x = g.xmin + np.arange(g.nx) * g.side
        y = g.ymin + np.arange(g.ny) * g.side
        z = g.zmin + np.arange(g.nz) * g.side
        Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
        dist = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2)
        # the polyhedron lies between radius R cos(edge angle) and R
        assert (st[dist < R - 0.2] == -1).all()
        assert (st[dist > R + 1e-6] != -1).all()
        m = o.triangulate()
        assert m["num_triangles"] > 0
    # both candidate rules find the same hits on this mesh
    a = binding.OracleSurface.from_mesh(g, v, t, ray_based=True)
    b = binding.OracleSurface.from_mesh(g, v, t, ray_based=False)
    a.get_surf(**OPTS)
    b.get_surf(**OPTS)
    assert np.array_equal(a.epsmap(), b.epsmap()) and np.array_equal(a.status(), b.status())


def test_blobby_chain_mesh_closed_and_consistent():
    xyzr, _ = ns.generate_chain(40, 5.6, seed=3)
    g = ns.build_grid(xyzr, 2.0, 80.0)
    b = binding.blobby_build(g, xyzr)
    closed, _ = _closed_manifold(b["tris"])
    assert closed
    o = binding.OracleSurface.from_mesh(g, b["verts"], b["tris"], ray_based=True)
    r = o.get_surf(**OPTS)
    assert r["odd_rays"] == 0
    assert r["volume"] == pytest.approx(_signed_volume(b["verts"], b["tris"]), rel=2e-3)
