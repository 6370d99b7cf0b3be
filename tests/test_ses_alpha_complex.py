"""The SES builder's cell classification against the weighted alpha complex.

ConnollySurface::build (ConnollySurface.cpp:385-440) builds CGAL's
Fixed_alpha_shape_3 at alpha 0 over the atoms weighted (r + probe)^2 and reads
REGULAR and SINGULAR facets (BuildFacetCells, :787-1265) and edges
(BuildEdgeCells, :1267-1926) off it.  nsgpu_ses_build classifies the same cells
by where the probe is free instead.  Here the complex is computed on its own
terms: the library's exact regular triangulation of the weighted points
(nsgpu_regular_triangulate), each tetrahedron's orthosphere, each facet's and
edge's smallest orthosphere, and the classification rules of the alpha shape:

  cell      in K          iff its orthosphere has power radius^2 <= 0
  facet     INTERIOR      iff both incident cells are in K
            REGULAR       iff exactly one is
            SINGULAR      iff neither is, its smallest orthosphere has
                          radius^2 <= 0 and is not attached (no vertex of an
                          incident cell lies inside it in the power sense)
  edge      REGULAR       iff some incident facet is in K and not all of
                          them are INTERIOR
            SINGULAR      iff no incident facet is in K and the smallest
                          orthosphere has radius^2 <= 0 and is unattached
                          (tested against the edge's link vertices)

The builder's facet_regular / facet_singular atom triples and torus_regular /
torus_singular atom pairs must equal these sets exactly."""
import os

import numpy as np
import pytest

import nanoshaper_amd as ns

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PROBE = 1.4
# NSGPU_PATCH_* kinds
FACET_REG, FACET_SING, TORUS_REG, TORUS_SING = 1, 2, 3, 4


def _power(P, W, y):
    return ((P - y) ** 2).sum(-1) - W


def _alpha_complex(xyzr, probe):
    """Facet and edge classes of the alpha-0 complex of the atoms weighted
    (r + probe)^2, plus the largest |value| / scale margin seen at a decision
    (to show no decision was a rounding tie)."""
    n = len(xyzr)
    wp = np.column_stack([xyzr[:, :3], (xyzr[:, 3] + probe) ** 2])
    tets, nbrs = ns.weighted_triangulation(wp)
    P, W = wp[:, :3], wp[:, 3]
    finite = (tets < n).all(axis=1)
    margin = [np.inf]

    def decide(v, scale):
        margin[0] = min(margin[0], abs(v) / scale)
        return v <= 0

    # cells: orthosphere power radius^2 (cells with a dummy vertex are far out)
    cell_in = np.zeros(len(tets), bool)
    for c in np.flatnonzero(finite):
        v = tets[c]
        A = 2 * (P[v[1:]] - P[v[0]])
        b = (P[v[1:]] ** 2).sum(1) - W[v[1:]] - ((P[v[0]] ** 2).sum() - W[v[0]])
        y = np.linalg.solve(A, b)
        cell_in[c] = decide(_power(P[v[0]], W[v[0]], y), W[v].max())

    # facets: each atom triple with its incident cells and opposite vertices
    facets = {}
    for c, v in enumerate(tets):
        for k in range(4):
            f = tuple(sorted(int(q) for q in np.delete(v, k)))
            if max(f) >= n:
                continue
            facets.setdefault(f, []).append((c, int(v[k])))
    fclass = {}
    for f, inc in facets.items():
        assert len(inc) == 2, f  # atom facets are interior to the dummies' hull
        ins = sum(cell_in[c] for c, _ in inc)
        if ins == 2:
            fclass[f] = "interior"
        elif ins == 1:
            fclass[f] = "regular"
        else:
            p = P[list(f)]
            w = W[list(f)]
            nrm = np.cross(p[1] - p[0], p[2] - p[0])
            A = np.vstack([2 * (p[1:] - p[0]), nrm])
            b = np.append((p[1:] ** 2).sum(1) - w[1:] - ((p[0] ** 2).sum() - w[0]), nrm @ p[0])
            y = np.linalg.solve(A, b)
            rho2 = _power(p[0], w[0], y)
            scale = w.max()
            small = decide(rho2, scale)
            # a dummy (two extents out, weight -1) attaches nothing
            free = all(not decide(_power(P[o], W[o], y) - rho2, scale) for _, o in inc if o < n)
            fclass[f] = "singular" if small and free else "exterior"

    # edges: incident facets and link vertices
    inc_f, link = {}, {}
    for f in facets:
        for e in ((f[0], f[1]), (f[0], f[2]), (f[1], f[2])):
            inc_f.setdefault(e, []).append(f)
    for v in tets:
        vs = [int(q) for q in v]
        for a in range(4):
            for b in range(a + 1, 4):
                e = tuple(sorted((vs[a], vs[b])))
                if e[1] < n:
                    link.setdefault(e, set()).update(q for q in vs if q not in e)
    eclass = {}
    for e, lk in link.items():
        cls = [fclass[f] for f in inc_f.get(e, [])]
        if any(c != "exterior" for c in cls):
            eclass[e] = "interior" if all(c == "interior" for c in cls) else "regular"
            continue
        i, j = e
        d = P[j] - P[i]
        t = (d @ d + W[i] - W[j]) / (2 * (d @ d))
        y = P[i] + t * d
        rho2 = _power(P[i], W[i], y)
        scale = max(W[i], W[j])
        small = decide(rho2, scale)
        free = all(not decide(_power(P[q], W[q], y) - rho2, scale) for q in lk if q < n)
        eclass[e] = "singular" if small and free else "exterior"
    return fclass, eclass, margin[0]


def _builder_cells(xyzr, probe):
    ps = ns.ses_patches(xyzr, probe)
    kind, atom = ps.kind.copy(), ps.atom.copy()
    ps.close()

    def pick(k, m):
        return [tuple(sorted(int(q) for q in a[:m])) for a in atom[kind == k]]

    freg, fsing = pick(FACET_REG, 3), pick(FACET_SING, 3)
    treg, tsing = pick(TORUS_REG, 2), pick(TORUS_SING, 2)
    # a singular facet is two cells (its two free probe placements)
    assert sorted(fsing) == sorted(fsing[0::2] * 2) if fsing else True
    assert len(set(freg)) == len(freg) and len(set(treg)) == len(treg) and len(set(tsing)) == len(tsing)
    return set(freg), set(fsing), set(treg), set(tsing)


def _compare(xyzr, probe=PROBE):
    fclass, eclass, margin = _alpha_complex(xyzr, probe)
    # every decision was clear of rounding (the inputs are jittered)
    assert margin > 1e-9, margin
    freg, fsing, treg, tsing = _builder_cells(xyzr, probe)
    want = {c: {f for f, k in fclass.items() if k == c} for c in ("regular", "singular")}
    wante = {c: {e for e, k in eclass.items() if k == c} for c in ("regular", "singular")}
    assert freg == want["regular"], (len(freg ^ want["regular"]), sorted(freg ^ want["regular"])[:5])
    assert fsing == want["singular"], sorted(fsing ^ want["singular"])[:5]
    assert treg == wante["regular"], sorted(treg ^ wante["regular"])[:5]
    assert tsing == wante["singular"], sorted(tsing ^ wante["singular"])[:5]
    return {k: len(v) for k, v in (("freg", freg), ("fsing", fsing), ("treg", treg), ("tsing", tsing))}


@pytest.mark.parametrize("n,radius,seed", [(40, 5.0, 1), (150, 8.0, 2), (400, 12.0, 3)])
def test_chain_cells_equal_the_alpha_complex(n, radius, seed):
    x, _ = ns.generate_chain(n, radius, seed=seed)
    jit, _ = ns.jitter(x)
    counts = _compare(jit)
    assert counts["freg"] > 0 and counts["treg"] > 0


def test_1crn_cells_equal_the_alpha_complex():
    x = ns.load_atoms(os.path.join(GOLDEN, "1crn.xyzr"))
    jit, _ = ns.jitter(x)
    counts = _compare(jit)
    assert counts["freg"] > 100


@pytest.mark.parametrize("probe", [0.6, 3.0])
def test_other_probes(probe):
    x, _ = ns.generate_chain(150, 8.0, seed=4)
    jit, _ = ns.jitter(x)
    _compare(jit, probe)


def test_singular_cells_occur_and_match():
    # three spheres in a row-triangle, probe small enough to pass through the
    # middle (a singular facet), plus a pair far apart from them whose
    # inflated spheres just overlap (a singular torus, no facet)
    x = np.array([[0.0, 0.0, 0.0, 1.5], [3.4, 0.0, 0.0, 1.5], [1.7, 2.9, 0.0, 1.5],
                  [20.0, 0.0, 0.0, 1.5], [20.0, 4.6, 0.0, 1.5]])
    jit, _ = ns.jitter(x)
    counts = _compare(jit, 1.0)
    assert counts["fsing"] == 1 and counts["tsing"] == 1, counts
