"""GPU parity of the triangle-mesh surfaces: the blobby build and mesh ray casting.

Bars:
* blobby field: within 4 ulp of the oracle's (the device exp and glibc's exp
  may differ in the last bit; the sums are taken in the same atom order);
* smoothed vertices before the text round trip: 1e-12 A;
* vertices after the round trip, triangle lists, hit counts, maps, cavity
  labels and the marching-cubes mesh of the mesh surface: bit-exact;
* hit depths: bit-exact (same arithmetic on the same floats);
* areas and volumes: relative 1e-12 (summation order).
"""
import numpy as np
import pytest

import nanoshaper_amd as ns

pytestmark = pytest.mark.gpu

OPTS = dict(build_epsmap=True, build_status=True, accurate_triangulation=True, fill_cavities=True, cav_vol=11.4,
            inside=5)


def _blobby(n, scale, seed, perfil=80.0, cutoff=6.0, B=-2.5):
    from oracle import binding
    xyzr, _ = ns.generate_chain(n, 1.65 * n ** (1 / 3), seed=seed)
    grid = ns.build_grid(xyzr, scale, perfil)
    print("blobby case n=%d grid=%d seed=%d" % (n, grid.nx, seed), flush=True)
    dev = ns.DeviceSurface(0)
    dev.set_grid(grid)
    info = dev.blobby_build(xyzr, B, cutoff)
    ref = binding.blobby_build(grid, xyzr, B, cutoff)
    return grid, dev, info, ref


def _compare_surface(grid, dev, orc, opts, area_rel=1e-12):
    r = dev.get_surf(**opts)
    ro = orc.get_surf(**opts)
    for k in ("total_rays", "odd_rays", "total_hits", "candidates", "num_cavities", "max_hits_per_ray"):
        assert r[k] == ro[k], (k, r[k], ro[k])
    assert r["volume"] == pytest.approx(ro["volume"], rel=1e-12)
    for s in range(6):
        ca, ta = dev.ray_hits(s)
        cb, tb = orc.ray_hits(s)
        assert np.array_equal(ca, cb), "slot %d hit counts differ at %d rays" % (s, int((ca != cb).sum()))
        assert np.array_equal(ta, tb), "slot %d depths differ" % s
    assert np.array_equal(dev.epsmap(), orc.epsmap())
    assert np.array_equal(dev.idebmap(), orc.idebmap())
    assert np.array_equal(dev.status(), orc.status())
    sa, fa = dev.cavities()
    sb, fb = orc.cavities()
    assert np.array_equal(sa, sb) and np.array_equal(fa, fb)
    assert np.array_equal(dev.corner_bits(), orc.corner_bits())
    m, mo = dev.triangulate(), orc.triangulate()
    assert (m["num_vertices"], m["num_triangles"]) == (mo["num_vertices"], mo["num_triangles"])
    assert m["area"] == pytest.approx(mo["area"], rel=area_rel)
    va, ta = dev.mesh()
    vb, tb = orc.mesh()
    assert np.array_equal(ta, tb) and np.array_equal(va, vb)
    return r


@pytest.mark.parametrize("n,scale,seed", [(1, 4.0, 0), (40, 2.0, 3), (200, 2.0, 1), (300, 3.0, 2)])
def test_blobby_build_matches_oracle(device_ok, n, scale, seed):
    grid, dev, info, ref = _blobby(n, scale, seed)
    f = dev.blobby_field()
    fo = ref["field"]
    assert np.all(np.abs(f - fo) <= 4 * np.spacing(np.maximum(np.abs(fo), 1e-300)))
    v_raw, t = dev.surface_mesh(raw=True)
    v, t2 = dev.surface_mesh()
    assert info["num_vertices"] == len(ref["verts"]) and info["num_triangles"] == len(ref["tris"])
    assert np.array_equal(t, ref["tris"]) and np.array_equal(t2, ref["tris"])
    assert np.abs(v_raw - ref["raw"]).max() <= 1e-12
    # "%.3f" then a float: the device's exact emulation equals printf/sscanf
    assert np.array_equal(v, ref["verts"]), "%d coordinates differ" % int((v != ref["verts"]).sum())
    assert info["area"] == pytest.approx(ref["area"], rel=1e-12)


@pytest.mark.parametrize("ray_based", [True, False])
def test_blobby_surface_matches_oracle(device_ok, ray_based):
    from oracle.binding import OracleSurface
    grid, dev, info, ref = _blobby(200, 2.0, 1)
    if not ray_based:
        dev.set_mesh(ref["verts"], ref["tris"], ray_based=False)
    orc = OracleSurface.from_mesh(grid, ref["verts"], ref["tris"], ray_based=ray_based)
    _compare_surface(grid, dev, orc, OPTS)


@pytest.mark.parametrize("opts", [
    dict(build_epsmap=True, build_status=False, accurate_triangulation=False, fill_cavities=False, inside=5),
    dict(build_epsmap=False, build_status=True, accurate_triangulation=True, fill_cavities=True, inside=5),
])
def test_mesh_surface_flags(device_ok, opts):
    from oracle.binding import OracleSurface
    grid, dev, info, ref = _blobby(120, 2.0, 4)
    orc = OracleSurface.from_mesh(grid, ref["verts"], ref["tris"], ray_based=True)
    r = dev.get_surf(**opts)
    ro = orc.get_surf(**opts)
    for k in ("total_hits", "candidates", "odd_rays", "num_cavities"):
        assert r[k] == ro[k], k
    if opts["build_epsmap"]:
        assert np.array_equal(dev.epsmap(), orc.epsmap())
    if opts["build_status"]:
        assert np.array_equal(dev.status(), orc.status())
    if opts["accurate_triangulation"]:
        assert np.array_equal(dev.corner_bits(), orc.corner_bits())


@pytest.mark.parametrize("ray_based,aux", [(True, 100), (True, 20), (False, 100)])
def test_loaded_mesh_matches_oracle(device_ok, ray_based, aux):
    from oracle.binding import OracleSurface
    from test_oracle_mesh import icosphere
    v, t = icosphere(6.0, np.array([0.31, -0.17, 0.08]), 3)
    v = v.astype(np.float32).astype(np.float64)  # loadOFF reads floats
    grid = ns.build_grid(np.array([[0.31, -0.17, 0.08, 6.0]]), 3.0, 60.0)
    print("icosphere grid=%d ray_based=%d aux=%d" % (grid.nx, ray_based, aux), flush=True)
    dev = ns.DeviceSurface(0)
    dev.set_grid(grid)
    dev.set_mesh(v, t, ray_based=ray_based, aux_grid_dim=aux)
    orc = OracleSurface.from_mesh(grid, v, t, ray_based=ray_based, aux_grid_dim=aux)
    _compare_surface(grid, dev, orc, OPTS)


def test_blobby_larger_molecule(device_ok):
    from oracle.binding import OracleSurface
    grid, dev, info, ref = _blobby(2000, 2.0, 0)
    v, t = dev.surface_mesh()
    assert np.array_equal(v, ref["verts"]) and np.array_equal(t, ref["tris"])
    orc = OracleSurface.from_mesh(grid, ref["verts"], ref["tris"], ray_based=True)
    _compare_surface(grid, dev, orc, OPTS)


def test_mesh_errors(device_ok):
    grid, dev, info, ref = _blobby(40, 2.0, 3)
    with pytest.raises(ns.NsgpuError):
        dev.set_mesh(ref["verts"], ref["tris"] + len(ref["verts"]))  # vertex index out of range
    dev.set_vertex_normals(True)
    with pytest.raises(ns.NsgpuError):
        dev.get_surf(**OPTS)  # analytic normals need SES patches
    with pytest.raises(ns.NsgpuError):
        dev.surface_gradient()
