"""GPU: the cell grid and its ray tiles.

* get_surf with ray tiles over the cell grid (NSGPU_CAST_TILES) gives the
  patch-centric cast's results exactly: hit lists, maps, cavities, mesh and
  normals, for SES (all torus root modes) and skin, and in batches.
* nsgpu_cast_rays (castAxisOrientedRay for a batch): along grid lines it
  returns the oracle's depths of those rays as coordinates; off the grid it
  crosses the exact SES where the surface's definition says, with unit normals
  pointing out of it.
"""
import numpy as np
import pytest

import nanoshaper_amd as ns
from tests.exact_surfaces import ExactSES

pytestmark = pytest.mark.gpu

OPTS = dict(build_epsmap=True, build_status=True, accurate_triangulation=True, fill_cavities=True)


def _state(dev, opts):
    r = dev.get_surf(**opts)
    nslots = (3 if opts.get("build_epsmap", True) else 1) + (3 if opts.get("accurate_triangulation", True) else 0)
    out = {"r": r, "eps": dev.epsmap(), "st": dev.status(), "rays": [dev.ray_hits(s) for s in range(nslots)]}
    if opts.get("accurate_triangulation", True):
        out["bits"] = dev.corner_bits()
        out["m"] = dev.triangulate()
        out["mesh"] = dev.mesh()
    return out


def _same(a, b):
    assert a["r"] == b["r"]
    assert np.array_equal(a["eps"], b["eps"]) and np.array_equal(a["st"], b["st"])
    for (ca, ta), (cb, tb) in zip(a["rays"], b["rays"]):
        assert np.array_equal(ca, cb) and np.array_equal(ta, tb)
    if "m" in a:
        assert a["m"] == b["m"] and np.array_equal(a["bits"], b["bits"])
        assert all(np.array_equal(x, y) for x, y in zip(a["mesh"], b["mesh"]))


@pytest.mark.parametrize("n,radius,scale,precision", [(330, 12.0, 2.0, 0), (2000, 21.0, 3.0, 0), (600, 14.0, 3.0, 1),
                                                      (300, 11.0, 2.0, 2)])
def test_tiles_match_patch_cast_ses(device_ok, n, radius, scale, precision):
    xyzr, _ = ns.generate_chain(n, radius, seed=1)
    g = ns.build_grid(xyzr, scale, 80.0)
    jit, _ = ns.jitter(xyzr)
    ps = ns.ses_patches(jit, 1.4)
    print("tiles ses n=%d grid=%d precision=%d" % (n, g.nx, precision), flush=True)
    dev = ns.DeviceSurface(0)
    dev.set_grid(g)
    dev.set_patches(ps)
    dev.set_vertex_normals(True)
    opts = dict(OPTS, precision=precision)
    a = _state(dev, opts)
    na = dev.normals()
    dev.set_cast_mode(dev.CAST_TILES)
    b = _state(dev, opts)
    _same(a, b)
    # depths of equal value may come from two patches in either order
    assert np.abs(dev.normals() - na).max() < 1e-9
    # the second call reuses the built cell grid
    _same(a, _state(dev, opts))
    dev.close()


def test_tiles_match_patch_cast_skin(device_ok):
    xyzr, _ = ns.generate_chain(1000, 1.65 * 1000 ** (1 / 3), seed=2)
    g = ns.build_grid(xyzr, 2.0, 80.0)
    jit, _ = ns.jitter(xyzr)
    ps = ns.skin_patches(jit, 0.45)
    dev = ns.DeviceSurface(0)
    dev.set_grid(g)
    dev.set_patches(ps)
    a = _state(dev, OPTS)
    dev.set_cast_mode(dev.CAST_TILES)
    _same(a, _state(dev, OPTS))
    dev.close()


def test_tiles_in_a_batch(device_ok):
    frames = ns.perturbed_frames(4, n_min=200, n_max=900, seed=9)
    opts = dict(build_epsmap=True, build_status=True, accurate_triangulation=False, fill_cavities=True)
    ref = []
    for xyzr in frames:
        g = ns.build_grid(xyzr, 2.0, 80.0)
        jit, _ = ns.jitter(xyzr)
        ps = ns.ses_patches(jit, 1.4)
        d = ns.DeviceSurface(0)
        d.set_grid(g)
        d.set_patches(ps)
        ref.append((g, ps, d.get_surf(**opts), d.status()))
        d.close()
    b = ns.Batch(len(frames))
    for f, (g, ps, _, _) in enumerate(ref):
        b.set_frame(f, g, ps)
        b.frame(f).set_cast_mode(ns.DeviceSurface.CAST_TILES)
    b.prepare(chains=2, **opts)
    for rep in range(2):
        b.run()
        for f, (_, _, r, st) in enumerate(ref):
            assert b.result(f) == r
            assert np.array_equal(b.frame(f).status(), st)
    b.close()


def _dedupe(a):
    out, last = [], np.inf
    for x in a:
        if abs(x - last) >= 1e-8:
            out.append(x)
            last = x
    return np.array(out)


def test_cast_rays_along_grid_lines_match_oracle(device_ok):
    from oracle.binding import OracleSurface
    xyzr, _ = ns.generate_chain(400, 12.0, seed=3)
    g = ns.build_grid(xyzr, 2.0, 80.0)
    jit, _ = ns.jitter(xyzr)
    ps = ns.ses_patches(jit, 1.4)
    orc = OracleSurface(g, ps)
    orc.get_surf(build_epsmap=True, build_status=True, accurate_triangulation=False, fill_cavities=False)
    dev = ns.DeviceSurface(0)
    dev.set_grid(g)
    dev.set_patches(ps)
    dev.get_surf(build_epsmap=True, build_status=True, accurate_triangulation=False, fill_cavities=False)
    xs, ys, zs = ns.grid_coords(g)
    N = g.nx
    rng = np.random.default_rng(0)
    checked = 0
    # slot 0: x rays over (y_m, z_n); slot 1: z rays over (x_m, y_n); slot 2: y rays over (x_m, z_n)
    for slot, direction in ((0, 0), (1, 2), (2, 1)):
        cnt, T = orc.ray_hits(slot)
        dcnt, dT = dev.ray_hits(slot)
        assert np.array_equal(cnt, dcnt)
        rays = rng.choice(np.nonzero(cnt > 0)[0], size=60, replace=False)
        axis = (xs, zs, ys)[slot]
        o = np.zeros((len(rays), 3))
        for q, ray in enumerate(rays):
            n_, m_ = divmod(int(ray), N)
            if slot == 0:
                o[q] = (xs[0], ys[m_], zs[n_])
            elif slot == 1:
                o[q] = (xs[m_], ys[n_], zs[0])
            else:
                o[q] = (xs[m_], ys[0], zs[n_])
        counts, hits = dev.cast_rays(direction, o, axis[-1], max_hits=32)
        for q, ray in enumerate(rays):
            # the device's own grid ray: the same arithmetic, bit for bit
            want = _dedupe(axis[0] + (axis[-1] - axis[0]) * dT[ray, :dcnt[ray]])
            assert counts[q] == len(want), (slot, ray, counts[q], want)
            assert np.array_equal(hits[q, :counts[q]], want)
            # the oracle's: device and host cbrt/acos differ in the last bits
            ref = _dedupe(axis[0] + (axis[-1] - axis[0]) * T[ray, :cnt[ray]])
            assert len(ref) == len(want) and np.abs(ref - want).max() < 1e-9
            checked += 1
    # refusals: non-progressive ray, ray outside the grid
    counts, _ = dev.cast_rays(0, [[xs[5], ys[5], zs[5]], [xs[0], ys[0] - 10 * g.side, zs[5]]], [xs[2], xs[-1]])
    assert list(counts) == [-1, -2]
    assert checked == 180
    dev.close()


def test_cast_rays_off_grid_cross_the_exact_ses(device_ok):
    xyzr, _ = ns.generate_chain(12, 5.0, seed=4)
    g = ns.build_grid(xyzr, 2.0, 60.0)
    jit, _ = ns.jitter(xyzr)
    dev = ns.DeviceSurface(0)
    dev.set_grid(g)
    dev.set_patches(ns.ses_patches(jit, 1.4))
    ex = ExactSES(jit[:, :3], jit[:, 3], 1.4)
    rng = np.random.default_rng(5)
    xs, ys, zs = ns.grid_coords(g)
    lo = np.array([xs[0], ys[0], zs[0]])
    hi = np.array([xs[-1], ys[-1], zs[-1]])
    found = 0
    for direction in (0, 1, 2):
        o = lo + rng.random((40, 3)) * (hi - lo)
        o[:, direction] = lo[direction]
        counts, hits, nrm = dev.cast_rays(direction, o, hi[direction], max_hits=16, normals=True)
        s = np.linspace(lo[direction], hi[direction], 4001)
        for q in range(len(o)):
            pts = np.repeat(o[q][None], len(s), 0)
            pts[:, direction] = s
            f = ex.signed(pts)
            ch = np.nonzero((f[1:] > 0) != (f[:-1] > 0))[0]
            assert counts[q] == len(ch), (direction, q, counts[q], len(ch))
            for h in range(counts[q]):
                a, b = s[ch[h]], s[ch[h] + 1]
                assert a - 1e-6 <= hits[q, h] <= b + 1e-6
                p = o[q].copy()
                p[direction] = hits[q, h]
                # unit normal pointing to the outside (the signed function grows along it)
                assert abs(np.linalg.norm(nrm[q, h]) - 1) < 1e-9
                assert ex.signed((p + 1e-4 * nrm[q, h])[None])[0] > 0 > ex.signed((p - 1e-4 * nrm[q, h])[None])[0]
                found += 1
    assert found > 20
    dev.close()


def test_python_facade_cast_and_grid_queries(device_ok):
    xyzr, _ = ns.generate_chain(200, 10.0, seed=6)
    shaper = ns.NanoShaper(xyzr, "ses", scale=2.0, perfil=80.0)
    shaper.build_analytical_surface()
    shaper.colour_grid()
    g = shaper.grid
    pa = list(shaper.get_grid_point_coordinates(0, g.ny // 2, g.nz // 2))
    assert shaper.cast_axis_oriented_ray(pa, pa[0] + 40.0, 0)[0] == -1  # no direction set
    assert shaper.set_direction(0)
    assert shaper.cast_axis_oriented_ray(pa, pa[0] - 1.0, 0)[0] == -1  # not progressive
    status, hits = shaper.cast_axis_oriented_ray(pa, pa[0] + 40.0, 0, compute_normals=True)
    assert status == 1 and len(hits) % 2 == 0
    assert all(abs(np.linalg.norm(n) - 1) < 1e-9 for _, n in hits)
    idx, sat = shaper.get_nearest_grid_point([pa[0] + 0.1, pa[1] - 0.2, pa[2]])
    assert idx == (0, g.ny // 2, g.nz // 2) and not sat
    assert shaper.get_nearest_grid_point([-1e4, 0, 0])[1]
