"""Pins of the oracle's pocket detection (the grid part of pocket mode)
against an independent pipeline written with scipy.ndimage:

* the difference map: nodes not outside the fat-probe surface and outside
  the small-probe one;
* the reference's shape filter (a probe ball must fit around a support point
  with only the cavity's nodes inside) is a morphological opening: supports =
  erosion by the digital ball {d : |d|^2 side^2 <= r^2} (nodes beyond the grid
  count as inside the set), kept nodes = dilation of the supports;
* components by 6-connectivity, numbered from 4 in the reference's
  alternating scan order (forward on even passes, k descending on odd ones);
* volumes = node count x side^3, filled when <= 11.4 x Num_Wat_Pocket;
* the filter again on the union of the kept components (a ball that fits in
  the union is connected, so it lies in one component);
* a kept component is a pocket when one of its nodes is in the small
  surface's outside component.
"""
import numpy as np
import pytest
from scipy import ndimage

import nanoshaper_amd as ns
from oracle import binding

SIX = ndimage.generate_binary_structure(3, 1)


def _ball(side, r):
    m = int(np.floor(r / side))
    d = np.arange(-m, m + 1)
    Z, Y, X = np.meshgrid(d, d, d, indexing="ij")
    return (X * X + Y * Y + Z * Z) * side * side <= r * r


def _replay_ids(mask):
    lab, n = ndimage.label(mask, structure=SIX)
    if n == 0:
        return lab, np.zeros(0, int)
    NZ, NY, NX = mask.shape
    flat = np.arange(mask.size).reshape(mask.shape)
    brank = (NZ - 1 - np.arange(NZ)[:, None, None]) * NY * NX + flat % (NY * NX)
    fkey = ndimage.minimum(flat, lab, np.arange(1, n + 1))
    bkey = ndimage.minimum(brank, lab, np.arange(1, n + 1))
    remaining = list(range(n))
    ids = np.zeros(n, int)
    for t in range(n):
        key = fkey if t % 2 == 0 else bkey
        pick = min(remaining, key=lambda q: key[q])
        remaining.remove(pick)
        ids[pick] = 4 + t
    return lab, ids


def _opening(mask, ball):
    sup = ndimage.binary_erosion(mask, ball, border_value=1)
    return sup, ndimage.binary_dilation(sup, ball) & mask


def _surfaces(n, radius, scale, seed, cav_and_pockets):
    xyzr, _ = ns.generate_chain(n, radius, seed=seed)
    g = ns.build_grid(xyzr, scale, 80.0)
    seed_state = 1234
    jit1, seed_state = ns.jitter(xyzr, seed=seed_state)
    jit2, _ = ns.jitter(xyzr, seed=seed_state)
    fat = binding.OracleSurface(g, ns.ses_patches(jit1, 3.0))
    small = binding.OracleSurface(g, ns.ses_patches(jit2, 1.4))
    fat.get_surf(build_epsmap=False, build_status=True, fill_cavities=True, cav_vol=1e30)
    small.get_surf(build_epsmap=False, build_status=True, fill_cavities=True, cav_vol=0.0)
    return g, fat, small


@pytest.mark.parametrize("n,radius,scale,seed,cap", [(300, 11.0, 2.0, 3, True), (500, 13.0, 2.5, 4, True),
                                                    (500, 13.0, 2.5, 4, False)])
def test_pockets_match_morphology(n, radius, scale, seed, cap):
    g, fat, small = _surfaces(n, radius, scale, seed, cap)
    s1, s2 = fat.status(), small.status()
    out2 = (s2 == 2) | (s2 == 3) | ((s2 >= 4) if cap else False)
    D = (s1 != 2) & (s1 != 3) & out2
    ball = _ball(g.side, 1.4)
    _, K = _opening(D, ball)
    lab, ids = _replay_ids(K)
    r = fat.pockets(small, cav_and_pockets=cap)
    st = fat.status()
    lab0 = fat.pocket_labels()
    assert len(ids) == len(r["volumes"]) > 0
    counts = np.bincount(lab.ravel(), minlength=len(ids) + 1)[1:]
    order = np.argsort(ids)
    assert np.allclose(r["volumes"], counts[order] * g.side ** 3, rtol=1e-12)
    filled = r["volumes"] <= 11.4 * 2
    # second filter on the kept components' union
    id_of = np.zeros(len(ids) + 1, int)
    id_of[1:] = ids
    L = id_of[lab]                      # component id per node (0 outside)
    assert np.array_equal(lab0, np.where(L > 0, L, -1))
    U = (L > 0) & ~filled[np.maximum(L - 4, 0)]
    sup, K2 = _opening(U, ball)
    exp = np.full(st.shape, -1, np.int32)
    keep_filled = (L > 0) & filled[np.maximum(L - 4, 0)]
    exp[keep_filled] = L[keep_filled]
    exp[K2] = L[K2]
    exp[sup] = -L[sup]
    assert np.array_equal(exp, st)
    now_filled = filled.copy()
    for c in range(len(ids)):
        if not filled[c] and not (K2 & (L == 4 + c)).any():
            now_filled[c] = True
    assert np.array_equal(now_filled, r["filled"])
    kept = np.nonzero(~now_filled)[0]
    assert len(r["support"]) == len(kept)
    for q, c in enumerate(kept):
        want = np.nonzero((sup & (L == 4 + c)).ravel())[0]
        assert np.array_equal(np.sort(r["support"][q]), want)
        assert r["is_pocket"][c] == bool((s2.ravel()[np.nonzero((L == 4 + c).ravel())[0]] == 3).any())
    assert r["is_pocket"][kept].any()
