"""GPU parity: the HIP path against the oracle on the same seeded inputs.

Bars (per the hot-path contract):
* integer maps (epsmap, idebmap, status), the corner bitmap, cavity labels,
  per-ray hit counts and the triangle index lists: bit-exact;
* hit depths: |dt| * ray length <= 1e-6 A;
* volume and area: relative 1e-12 (the device sums in a different order);
* vertex coordinates: <= 1e-9 A (they are hit depths mapped to space).
"""
import numpy as np
import pytest

import nanoshaper_amd as ns

pytestmark = pytest.mark.gpu

HIT_TOL_A = 1e-6


def _surfaces(n, radius, scale, perfil, seed=0, probe=1.4):
    from oracle.binding import OracleSurface
    xyzr, _ = ns.generate_chain(n, radius, seed=seed)
    grid = ns.build_grid(xyzr, scale, perfil)
    jit, _ = ns.jitter(xyzr)
    ps = ns.ses_patches(jit, probe)
    dev = ns.DeviceSurface(0)
    dev.set_grid(grid)
    dev.set_patches(ps)
    return grid, ps, dev, OracleSurface(grid, ps)


def _ray_length(grid):
    return (grid.nx - 1) * grid.side


def _compare(grid, dev, orc, opts, mesh=True):
    print("case grid=%d opts=%s" % (grid.nx, opts), flush=True)
    r = dev.get_surf(**opts)
    ro = orc.get_surf(**opts)
    for k in ("total_rays", "odd_rays", "total_hits", "candidates", "num_cavities", "max_hits_per_ray"):
        assert r[k] == ro[k], (k, r[k], ro[k])
    assert r["volume"] == pytest.approx(ro["volume"], rel=1e-12)
    nslots = (3 if (opts.get("build_epsmap", True)) else (1 if opts.get("build_status", True) else 0)) + \
             (3 if opts.get("accurate_triangulation", True) else 0)
    L = _ray_length(grid)
    for s in range(nslots):
        ca, ta = dev.ray_hits(s)
        cb, tb = orc.ray_hits(s)
        assert np.array_equal(ca, cb), "slot %d hit counts differ at %d rays" % (s, int((ca != cb).sum()))
        assert np.abs(ta - tb).max() * L <= HIT_TOL_A
    if opts.get("build_epsmap", True):
        assert np.array_equal(dev.epsmap(), orc.epsmap())
        assert np.array_equal(dev.idebmap(), orc.idebmap())
    if opts.get("build_status", True):
        assert np.array_equal(dev.status(), orc.status())
    if opts.get("fill_cavities", True) and opts.get("build_status", True):
        sa, fa = dev.cavities()
        sb, fb = orc.cavities()
        assert np.array_equal(sa, sb) and np.array_equal(fa, fb)
    if opts.get("accurate_triangulation", True):
        assert np.array_equal(dev.corner_bits(), orc.corner_bits())
        if mesh:
            m = dev.triangulate()
            mo = orc.triangulate()
            assert m["num_vertices"] == mo["num_vertices"]
            assert m["num_triangles"] == mo["num_triangles"]
            assert m["num_analytic_vertices"] == mo["num_analytic_vertices"]
            assert m["area"] == pytest.approx(mo["area"], rel=1e-12)
            va, ta = dev.mesh()
            vb, tb = orc.mesh()
            assert np.array_equal(ta, tb)
            assert np.abs(va - vb).max() <= 1e-9
    return r


@pytest.mark.parametrize("n,radius,scale,perfil,seed", [
    (1, 2.0, 2.0, 80.0, 0),        # single atom: one point cell
    (2, 2.0, 3.0, 80.0, 1),        # bonded pair: torus seam
    (40, 6.0, 2.0, 80.0, 0),
    (330, 12.0, 2.0, 90.0, 0),     # the CPU-runnable configuration
    (330, 12.0, 2.0, 90.0, 3),
    (800, 16.0, 3.0, 80.0, 5),
])
def test_full_pipeline_matches_oracle(device_ok, n, radius, scale, perfil, seed):
    grid, ps, dev, orc = _surfaces(n, radius, scale, perfil, seed)
    _compare(grid, dev, orc, {})


@pytest.mark.parametrize("opts", [
    dict(build_epsmap=True, build_status=False, accurate_triangulation=False, fill_cavities=False),
    dict(build_epsmap=False, build_status=True, accurate_triangulation=False, fill_cavities=True),
    dict(build_epsmap=True, build_status=True, accurate_triangulation=False, fill_cavities=True),
    dict(build_epsmap=False, build_status=True, accurate_triangulation=True, fill_cavities=True),
    dict(build_epsmap=True, build_status=True, accurate_triangulation=True, fill_cavities=False),
    dict(build_epsmap=True, build_status=True, accurate_triangulation=True, fill_cavities=True, cav_vol=1000.0),
])
def test_flag_combinations(device_ok, opts):
    grid, ps, dev, orc = _surfaces(330, 12.0, 2.0, 90.0, seed=2)
    _compare(grid, dev, orc, opts)


def test_cavities_labels_and_fill(device_ok):
    # 2,000 atoms at 3 points/A has buried cavities; labels must follow the
    # reference's alternating scan order exactly
    grid, ps, dev, orc = _surfaces(2000, 21.0, 3.0, 80.0, seed=0)
    r = _compare(grid, dev, orc, dict(cav_vol=30.0))
    assert r["num_cavities"] > 0
    st = dev.status()
    labels = np.unique(st[st >= ns.STATUS_FIRST_CAV])
    assert np.array_equal(labels, np.arange(4, 4 + r["num_cavities"]))


def test_repeat_calls_are_identical(device_ok):
    grid, ps, dev, orc = _surfaces(330, 12.0, 2.0, 90.0, seed=4)
    r1 = dev.get_surf()
    e1, s1 = dev.epsmap(), dev.status()
    m1 = dev.triangulate()
    r2 = dev.get_surf()
    m2 = dev.triangulate()
    assert r1 == r2 and m1 == m2
    assert np.array_equal(e1, dev.epsmap()) and np.array_equal(s1, dev.status())


def test_c2_scale_parity(device_ok):
    # the benchmark workload, 5,000 atoms at 4 points/A
    grid, ps, dev, orc = _surfaces(5000, 28.0, 4.0, 80.0, seed=0)
    _compare(grid, dev, orc, {})


def test_no_patches_in_grid_is_background(device_ok):
    # an empty patch set leaves every map at background
    grid, ps, dev, orc = _surfaces(1, 2.0, 2.0, 80.0)
    empty = ns.ses_patches(np.array([[1e4, 1e4, 1e4, 1.5]]), 1.4)
    dev.set_patches(empty)
    r = dev.get_surf()
    assert r["total_hits"] == 0
    assert (dev.status() == ns.STATUS_OUT).all()
    assert (dev.epsmap() == 0).all() and (dev.idebmap() == 1).all()
