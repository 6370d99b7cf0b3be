"""GPU parity of the skin surface: the device's ray casting of the mixed
cells against the oracle's on the same cells.

Bars: hit counts and depths, maps, cavity labels, corner bitmap and the
marching-cubes mesh bit-exact (same arithmetic on the same cell data);
volume and area relative 1e-12 (summation order).
"""
import numpy as np
import pytest

import nanoshaper_amd as ns

pytestmark = pytest.mark.gpu

OPTS = dict(build_epsmap=True, build_status=True, accurate_triangulation=True, fill_cavities=True, cav_vol=11.4,
            inside=5)


def _skin(n, scale, seed, s=0.45):
    from oracle.binding import OracleSurface
    if n == 1:
        xyzr = np.array([[0.1, 0.2, -0.05, 1.7]])
    else:
        xyzr, _ = ns.generate_chain(n, 1.65 * n ** (1 / 3), seed=seed)
    jit, _ = ns.jitter(xyzr)
    ps = ns.skin_patches(jit, s)
    grid = ns.build_grid(xyzr, scale, 80.0)
    print("skin case n=%d grid=%d s=%.2f cells=%d" % (n, grid.nx, s, len(ps)), flush=True)
    dev = ns.DeviceSurface(0)
    dev.set_grid(grid)
    dev.set_patches(ps)
    return grid, ps, dev, OracleSurface(grid, ps)


def _compare(dev, orc, opts):
    r = dev.get_surf(**opts)
    ro = orc.get_surf(**opts)
    for k in ("total_rays", "odd_rays", "total_hits", "candidates", "num_cavities", "max_hits_per_ray"):
        assert r[k] == ro[k], (k, r[k], ro[k])
    assert r["volume"] == pytest.approx(ro["volume"], rel=1e-12)
    nslots = (3 if opts["build_epsmap"] else (1 if opts["build_status"] else 0)) + \
             (3 if opts["accurate_triangulation"] else 0)
    for s in range(nslots):
        ca, ta = dev.ray_hits(s)
        cb, tb = orc.ray_hits(s)
        assert np.array_equal(ca, cb) and np.array_equal(ta, tb), "slot %d" % s
    if opts["build_epsmap"]:
        assert np.array_equal(dev.epsmap(), orc.epsmap())
        assert np.array_equal(dev.idebmap(), orc.idebmap())
    if opts["build_status"]:
        assert np.array_equal(dev.status(), orc.status())
        if opts["fill_cavities"]:
            sa, fa = dev.cavities()
            sb, fb = orc.cavities()
            assert np.array_equal(sa, sb) and np.array_equal(fa, fb)
    if opts["accurate_triangulation"]:
        assert np.array_equal(dev.corner_bits(), orc.corner_bits())
        m, mo = dev.triangulate(), orc.triangulate()
        assert (m["num_vertices"], m["num_triangles"]) == (mo["num_vertices"], mo["num_triangles"])
        assert m["area"] == pytest.approx(mo["area"], rel=1e-12)
        va, ta = dev.mesh()
        vb, tb = orc.mesh()
        assert np.array_equal(ta, tb) and np.array_equal(va, vb)
    return r


@pytest.mark.parametrize("n,scale,seed,s", [(1, 4.0, 0, 0.45), (40, 2.0, 1, 0.45), (330, 2.0, 0, 0.45),
                                            (330, 3.0, 2, 0.8), (2000, 2.0, 0, 0.45)])
def test_skin_matches_oracle(device_ok, n, scale, seed, s):
    grid, ps, dev, orc = _skin(n, scale, seed, s)
    r = _compare(dev, orc, OPTS)
    assert r["odd_rays"] == 0


@pytest.mark.parametrize("opts", [
    dict(build_epsmap=True, build_status=False, accurate_triangulation=False, fill_cavities=False, inside=5),
    dict(build_epsmap=False, build_status=True, accurate_triangulation=True, fill_cavities=True, inside=5),
])
def test_skin_flags(device_ok, opts):
    grid, ps, dev, orc = _skin(330, 2.0, 3)
    _compare(dev, orc, opts)


@pytest.mark.parametrize("n,scale,seed,s", [(1, 4.0, 0, 0.45), (330, 2.0, 0, 0.45), (330, 3.0, 2, 0.8)])
def test_skin_vertex_normals_match_oracle(device_ok, n, scale, seed, s):
    grid, ps, dev, orc = _skin(n, scale, seed, s)
    dev.set_vertex_normals(True)
    _compare(dev, orc, OPTS)
    na, no = dev.normals(), orc.normals()
    assert na.shape == no.shape
    # analytic normals: the same quadric expression at points equal to 1e-11 A
    assert np.abs(na - no).max() < 1e-9


def test_skin_refuses_gradients(device_ok):
    grid, ps, dev, orc = _skin(40, 2.0, 1)
    with pytest.raises(ns.NsgpuError):
        dev.surface_gradient()


@pytest.mark.parametrize("surface_type", ["ses", "skin", "blobby"])
def test_facade_surface_types(device_ok, surface_type):
    xyzr, _ = ns.generate_chain(200, 10.0, seed=5)
    shaper = ns.NanoShaper(xyzr, surface_type, scale=2.0, perfil=80.0)
    shaper.build_analytical_surface()
    vol = shaper.colour_grid()
    area = shaper.triangulate()
    print("facade", surface_type, vol, area, flush=True)
    assert vol > 0 and area > 0
    assert shaper.get_grid_point_status(0, 0, 0) == 1
