"""GPU parity at the sizes of the benchmark configurations, against the
oracle (run on all host cores with the reference's thread decomposition; its
results do not depend on the thread count).

* skin (s = 0.45) and blobby (B = -2.5) surfaces of 20,000 atoms at 2 points/A:
  hit lists, maps, cavities, corner bits and the mesh;
* a batch of 32 frames of 500-8,000 atoms (log-uniform, sigma 0.3 A
  perturbations), SES at 2 points/A, maps and cavities from one captured graph
  launch, every frame against the oracle.
"""
import os

import numpy as np
import pytest

import nanoshaper_amd as ns

pytestmark = pytest.mark.gpu

OPTS = dict(build_epsmap=True, build_status=True, accurate_triangulation=True, fill_cavities=True, cav_vol=11.4,
            inside=5)


@pytest.fixture()
def oracle_threads():
    from oracle import binding
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    binding.set_threads(max(1, min(16, n)))
    yield binding
    binding.set_threads(1)


def _compare(dev, orc, opts, depth_tol=0.0):
    r = dev.get_surf(**opts)
    ro = orc.get_surf(**opts)
    for k in ("total_rays", "odd_rays", "total_hits", "candidates", "num_cavities", "max_hits_per_ray"):
        assert r[k] == ro[k], (k, r[k], ro[k])
    assert r["volume"] == pytest.approx(ro["volume"], rel=1e-12)
    for s in range(6):
        ca, ta = dev.ray_hits(s)
        cb, tb = orc.ray_hits(s)
        assert np.array_equal(ca, cb), "slot %d counts" % s
        assert np.abs(ta - tb).max() <= depth_tol, "slot %d depths" % s
    assert np.array_equal(dev.epsmap(), orc.epsmap())
    assert np.array_equal(dev.idebmap(), orc.idebmap())
    assert np.array_equal(dev.status(), orc.status())
    sa, fa = dev.cavities()
    sb, fb = orc.cavities()
    assert np.array_equal(sa, sb) and np.array_equal(fa, fb)
    assert np.array_equal(dev.corner_bits(), orc.corner_bits())
    m, mo = dev.triangulate(), orc.triangulate()
    assert (m["num_vertices"], m["num_triangles"]) == (mo["num_vertices"], mo["num_triangles"])
    assert m["area"] == pytest.approx(mo["area"], rel=1e-12)
    va, ta = dev.mesh()
    vb, tb = orc.mesh()
    assert np.array_equal(ta, tb) and np.array_equal(va, vb)
    return r, m


def test_skin_20000_atoms(device_ok, oracle_threads):
    xyzr, _ = ns.generate_chain(20000, 48.0, seed=0)
    grid = ns.build_grid(xyzr, 2.0, 80.0)
    jit, _ = ns.jitter(xyzr)
    ps = ns.skin_patches(jit, 0.45)
    print("skin 20000: grid %d, cells %d" % (grid.nx, len(ps)), flush=True)
    dev = ns.DeviceSurface(0)
    dev.set_grid(grid)
    dev.set_patches(ps)
    r, m = _compare(dev, oracle_threads.OracleSurface(grid, ps), OPTS)
    assert r["odd_rays"] == 0 and m["num_triangles"] > 1000000
    dev.close()


def test_blobby_20000_atoms(device_ok, oracle_threads):
    xyzr, _ = ns.generate_chain(20000, 48.0, seed=0)
    grid = ns.build_grid(xyzr, 2.0, 80.0)
    print("blobby 20000: grid %d" % grid.nx, flush=True)
    dev = ns.DeviceSurface(0)
    dev.set_grid(grid)
    dev.blobby_build(xyzr, -2.5)
    ref = oracle_threads.blobby_build(grid, xyzr, -2.5)
    v, t = dev.surface_mesh()
    assert np.array_equal(v, ref["verts"]) and np.array_equal(t, ref["tris"])
    orc = oracle_threads.OracleSurface.from_mesh(grid, ref["verts"], ref["tris"], ray_based=True)
    r, m = _compare(dev, orc, OPTS)
    assert r["odd_rays"] == 0
    dev.close()


def test_batch_of_32_uneven_frames(device_ok, oracle_threads):
    opts = dict(build_epsmap=True, build_status=True, accurate_triangulation=False, fill_cavities=True)
    frames = ns.perturbed_frames(32, n_min=500, n_max=8000, sigma=0.3, seed=11)
    sizes = [len(f) for f in frames]
    print("batch: %d frames, %d-%d atoms" % (len(frames), min(sizes), max(sizes)), flush=True)
    inputs = []
    for xyzr in frames:
        g = ns.build_grid(xyzr, 2.0, 80.0)
        jit, _ = ns.jitter(xyzr)
        inputs.append((g, ns.ses_patches(jit, 1.4)))
    # the autotuned graph, and the persistent multi-frame cast (workgroups
    # taking 16-item chunks of every frame's patches from one queue)
    batches = []
    for shape in (dict(chains=0), dict(chains=16, cast="persistent", chunk=16)):
        b = ns.Batch(len(inputs))
        for f, (g, ps) in enumerate(inputs):
            b.set_frame(f, g, ps)
        b.prepare(**shape, **opts)
        print("shape", b.shape(), flush=True)
        lat, redone = b.run()
        assert (lat > 0).all()
        batches.append(b)
    for f, (g, ps) in enumerate(inputs):
        orc = oracle_threads.OracleSurface(g, ps)
        ro = orc.get_surf(**opts)
        for b in batches:
            r = b.result(f)
            for k in ("total_rays", "odd_rays", "total_hits", "candidates", "num_cavities", "max_hits_per_ray"):
                assert r[k] == ro[k], (f, k, r[k], ro[k])
            assert r["volume"] == pytest.approx(ro["volume"], rel=1e-12)
            v = b.frame(f)
            assert np.array_equal(v.epsmap(), orc.epsmap()), f
            assert np.array_equal(v.idebmap(), orc.idebmap()), f
            assert np.array_equal(v.status(), orc.status()), f
            sa, fa = v.cavities()
            sb, fb = orc.cavities()
            assert np.array_equal(sa, sb) and np.array_equal(fa, fb), f
        orc.close()
    for b in batches:
        b.close()
