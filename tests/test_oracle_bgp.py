"""Pins of the oracle's boundary-grid-point stage (detection and projection on
the SES), from the surface's definition rather than from the oracle's formulas.

* Detection: a node is a boundary grid point iff the six points half a grid
  side away from it along the axes (where its epsmap faces are sampled) are
  not all on one side of the surface; checked against the exact SES classifier
  at every node whose six points are further than MARGIN from the surface.
* Projection: one atom gives the closed form c + r (p - c) / |p - c|.  On
  chains every projected point lies on the SES (the classifier's signed
  distance is zero there) or on the tube of a torus patch outside that
  patch's clipping sphere: the reference's getProjection tests a torus point
  against the patch's planes and self-intersection sphere but not against
  the clipping sphere that bounds the saddle (ConnollySurface.cpp:5874-5890),
  so the outer side of a tube counts as well.  A projection on the SES is no
  closer to the node than the node's distance to the surface allows, and every
  projection is the nearest feasible point of all the patches' spheres and
  tori for nearly every node (the reference only looks at the patches of the
  node's auxiliary cell), found here by brute force with the textbook nearest
  point of a sphere and of a torus.
"""
import numpy as np
import pytest

import nanoshaper_amd as ns
from exact_surfaces import ExactSES
from oracle import binding

PROBE = 1.4
MARGIN = 1e-6


def _ses(n, radius, scale, perfil, seed):
    xyzr, _ = ns.generate_chain(n, radius, seed=seed)
    g = ns.build_grid(xyzr, scale, perfil)
    jit, _ = ns.jitter(xyzr)
    return g, jit, ns.ses_patches(jit, PROBE)


def _nodes(g, ib):
    xs, ys, zs = ns.grid_coords(g)
    return np.column_stack([xs[ib[:, 0]], ys[ib[:, 1]], zs[ib[:, 2]]])


@pytest.mark.parametrize("n,radius,scale,seed", [(5, 5.0, 2.0, 31), (20, 7.0, 2.5, 32)])
def test_detection_matches_exact_ses(n, radius, scale, seed):
    g, jit, ps = _ses(n, radius, scale, 70.0, seed)
    o = binding.OracleSurface(g, ps)
    o.get_surf(build_epsmap=True, build_status=True, fill_cavities=False)
    b = o.project_bgp()
    ex = ExactSES(jit[:, :3], jit[:, 3], PROBE)
    xs, ys, zs = ns.grid_coords(g)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    h = g.hside
    out = []
    sure = np.ones(X.shape, bool)
    for d in range(3):
        for sgn in (1.0, -1.0):
            P = [X, Y, Z]
            P = [a + (sgn * h if q == d else 0.0) for q, a in enumerate(P)]
            f = ex.signed(np.column_stack([a.ravel() for a in P])).reshape(X.shape)
            sure &= np.abs(f) > MARGIN
            out.append(f > 0)
    out = np.stack(out)
    want = ~(out.all(axis=0) | (~out).all(axis=0))
    # interior nodes only: [1, n-1] on every axis
    want[0, :, :] = want[:, 0, :] = want[:, :, 0] = False
    got = np.zeros_like(want)
    ib = b["ibgp"]
    got[ib[:, 2], ib[:, 1], ib[:, 0]] = True
    assert np.array_equal(got & sure, want & sure)
    assert sure.sum() > 0.95 * sure.size and want.sum() > 100
    # ascending flat order, as the reference's z, y, x loops emit them
    flat = (ib[:, 2].astype(np.int64) * g.ny + ib[:, 1]) * g.nx + ib[:, 0]
    assert np.all(np.diff(flat) > 0)
    assert b["counts"][1] == 0 and np.all(b["scsnor"] == 0)


def test_one_atom_closed_form():
    xyzr = np.array([[0.3, -0.2, 0.1, 1.7]])
    g = ns.build_grid(xyzr, 4.0, 70.0)
    jit, _ = ns.jitter(xyzr)
    ps = ns.ses_patches(jit, PROBE)
    o = binding.OracleSurface(g, ps)
    o.get_surf(build_epsmap=True, build_status=True)
    b = o.project_bgp()
    p = _nodes(g, b["ibgp"])
    c, r = jit[0, :3], jit[0, 3]
    v = p - c
    want = c + r * v / np.linalg.norm(v, axis=1, keepdims=True)
    assert len(p) > 400 and list(b["counts"][1:]) == [0, 0, 0]
    assert np.abs(b["scspos"] - want).max() < 1e-12


def _on_outer_tube(ps, q, probe):
    """q on the tube of some torus patch and outside its clipping sphere."""
    v = ps.view
    hit = np.zeros(len(q), bool)
    for it in range(v.n):
        if v.kind[it] < 3:
            continue
        off = v.off[it]
        td = np.array([v.data[off + k] for k in range(13)])
        loc = (q - td[:3]) @ td[3:12].reshape(3, 3).T
        tube = np.hypot(np.hypot(loc[:, 0], loc[:, 1]) - td[12], loc[:, 2]) - probe
        cs = np.array([v.bsph[4 * it + k] for k in range(4)])
        hit |= (np.abs(tube) < 1e-9) & (np.linalg.norm(q - cs[:3], axis=1) >= cs[3])
    return hit


def _near_self_intersecting_circle(ps, q, probe):
    """q within one probe radius of the circle of probe centres of a
    self-intersecting regular torus patch (where the reference samples probe
    positions when no patch projection is feasible)."""
    v = ps.view
    hit = np.zeros(len(q), bool)
    for it in range(v.n):
        if v.kind[it] != 3 or not v.data[v.off[it] + 13] >= 0:
            continue
        off = v.off[it]
        td = np.array([v.data[off + k] for k in range(13)])
        loc = (q - td[:3]) @ td[3:12].reshape(3, 3).T
        hit |= np.hypot(np.hypot(loc[:, 0], loc[:, 1]) - td[12], loc[:, 2]) <= probe + 1e-9
    return hit


def _nearest_feasible(ps, p, probe):
    """Brute force over every patch: the nearest point of its sphere or torus
    to p that the patch keeps (the feasibility test is the oracle's, pinned
    by the fill checks against the exact classifier)."""
    v = ps.view
    lib = binding.load()
    best = np.full(len(p), np.inf)
    for it in range(v.n):
        kind = v.kind[it]
        c = np.array([v.bsph[4 * it + q] for q in range(3)])
        if kind <= 2:
            r = v.bsph[4 * it + 3] if kind == 0 else probe
            u = p - c
            q = c + r * u / np.linalg.norm(u, axis=1, keepdims=True)
        else:
            off = v.off[it]
            td = np.array([v.data[off + q] for q in range(13)])
            M = td[3:12].reshape(3, 3)                      # world -> torus frame
            loc = (p - td[:3]) @ M.T
            rho = np.hypot(loc[:, 0], loc[:, 1])
            m = np.column_stack([td[12] * loc[:, 0] / rho, td[12] * loc[:, 1] / rho, np.zeros(len(p))])
            w = loc - m
            q = m + probe * w / np.linalg.norm(w, axis=1, keepdims=True)
            q = q @ M + td[:3]                              # back: M orthonormal
        d = np.linalg.norm(q - p, axis=1)
        for k in np.nonzero(d < best)[0]:
            pt = np.ascontiguousarray(q[k])
            if lib.orc_is_feasible(ps.view, it, pt.ctypes.data):
                best[k] = d[k]
    return best


@pytest.mark.parametrize("n,radius,scale,seed", [(6, 5.0, 3.0, 41), (25, 8.0, 2.5, 42), (60, 7.0, 3.0, 3)])
def test_projections_lie_on_the_ses_and_are_nearest(n, radius, scale, seed):
    g, jit, ps = _ses(n, radius, scale, 70.0, seed)
    o = binding.OracleSurface(g, ps)
    o.get_surf(build_epsmap=True, build_status=True, fill_cavities=True)
    b = o.project_bgp()
    p = _nodes(g, b["ibgp"])
    q = b["scspos"]
    moved = np.any(q != p, axis=1)
    assert (~moved).sum() == b["counts"][3] and b["counts"][3] <= 0.001 * len(p)
    ex = ExactSES(jit[:, :3], jit[:, 3], PROBE)
    on = np.abs(ex.signed(q[moved])) < 1e-7
    assert on.mean() > 0.5
    other = ~(on | _on_outer_tube(ps, q[moved], PROBE))
    assert other.sum() <= b["counts"][2]
    assert np.all(_near_self_intersecting_circle(ps, q[moved][other], PROBE))
    # no point of the surface is closer to p than |signed distance of p|
    dist = np.linalg.norm(q - p, axis=1)[moved]
    fp = np.abs(ex.signed(p[moved]))
    assert np.all(dist[on] >= fp[on] - 1e-9)
    # brute force over all patches: never farther, equal for nearly all nodes
    # (the sampled fallback points aside)
    bf = _nearest_feasible(ps, p[moved][~other], PROBE)
    assert np.all(bf <= dist[~other] + 1e-9)
    assert np.mean(np.abs(bf - dist[~other]) < 1e-9) > 0.97
