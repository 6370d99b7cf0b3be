"""GPU: the allocation-free getSurf (nsgpu_get_surf_noalloc, called through
the C ABI) writes the same maps and results as nsgpu_get_surf, on two calls in
a row (the second smaller), into the caller's device arrays."""
import ctypes

import numpy as np
import pytest

import nanoshaper_amd as ns
from nanoshaper_amd import _lib
from nanoshaper_amd.surface import SurfResult, surf_opts

pytestmark = pytest.mark.gpu

FIELDS = ("volume", "total_rays", "odd_rays", "total_hits", "candidates", "num_cavities", "max_hits_per_ray")


def _surface(n, radius, scale, seed):
    xyzr, _ = ns.generate_chain(n, radius, seed=seed)
    g = ns.build_grid(xyzr, scale, 80.0)
    jit, _ = ns.jitter(xyzr)
    return g, ns.ses_patches(jit, 1.4)


def _noalloc(dev, g, **kw):
    import torch
    lib = _lib.load()
    nvox = g.nx * g.ny * g.nz
    o = surf_opts(**kw)
    eps = torch.full((nvox * 3,), -7, dtype=torch.int32, device="cuda")
    ideb = torch.full((nvox,), 9, dtype=torch.uint8, device="cuda")
    st = torch.full((nvox,), -7, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    r = SurfResult()
    rc = lib.nsgpu_get_surf_noalloc(dev._h, ctypes.byref(o), ctypes.byref(r),
                                    ctypes.c_void_p(eps.data_ptr() if o.build_epsmap else 0),
                                    ctypes.c_void_p(ideb.data_ptr() if o.build_epsmap else 0),
                                    ctypes.c_void_p(st.data_ptr() if o.build_status else 0))
    msg = lib.nsgpu_last_error(dev._h).decode() if rc else ""
    res = {f: getattr(r, f) for f in FIELDS}
    shape = (g.nz, g.ny, g.nx)
    return rc, msg, res, eps.cpu().numpy().reshape(shape + (3,)), ideb.cpu().numpy().reshape(shape), \
        st.cpu().numpy().reshape(shape)


def _check(dev, g, **kw):
    rc, msg, res, eps, ideb, st = _noalloc(dev, g, **kw)
    assert rc == 0, msg
    ref = dev.get_surf(**kw)
    for f in FIELDS:
        assert res[f] == ref[f], (f, res[f], ref[f])
    if kw.get("build_epsmap", True):
        assert np.array_equal(eps, dev.epsmap())
        assert np.array_equal(ideb, dev.idebmap())
    if kw.get("build_status", True):
        assert np.array_equal(st, dev.status())
    return res


@pytest.mark.parametrize("opts", [dict(accurate_triangulation=False),
                                  dict(accurate_triangulation=False, fill_cavities=False),
                                  dict(build_epsmap=False, accurate_triangulation=False, fill_cavities=False),
                                  dict(build_status=False, accurate_triangulation=False, fill_cavities=False),
                                  dict(accurate_triangulation=False, precision=1)])
def test_noalloc_matches_get_surf_on_two_calls(device_ok, opts):
    dev = ns.DeviceSurface(0)
    g1, ps1 = _surface(900, 12.0, 3.0, 2)  # 135^3, with cavities
    g2, ps2 = _surface(200, 8.0, 2.0, 3)    # smaller: every buffer of the second call is shorter
    print("noalloc %s: %d^3 then %d^3" % (opts, g1.nx, g2.nx), flush=True)
    dev.set_grid(g1)
    dev.set_patches(ps1)
    r1 = _check(dev, g1, **opts)
    assert r1["total_hits"] > 0
    dev.set_grid(g2)
    dev.set_patches(ps2)
    _check(dev, g2, **opts)
    dev.close()


def test_noalloc_refusals(device_ok):
    dev = ns.DeviceSurface(0)
    g, ps = _surface(200, 8.0, 2.0, 3)
    dev.set_grid(g)
    dev.set_patches(ps)
    rc, msg, *_ = _noalloc(dev, g, accurate_triangulation=True)
    assert rc == -1 and "edge-ray" in msg, (rc, msg)
    # a grid whose ray lists and runs exceed the scratch array
    xyzr, _ = ns.generate_chain(5000, 28.0, seed=0)
    big = ns.build_grid(xyzr, 4.0, 80.0)
    jit, _ = ns.jitter(xyzr)
    dev.set_grid(big)
    dev.set_patches(ns.ses_patches(jit, 1.4))
    rc, msg, *_ = _noalloc(dev, big, accurate_triangulation=False)
    assert rc == -4 and "bytes of scratch" in msg, (rc, msg)
    # the context still works normally afterwards, and reads refuse nothing stale
    dev.set_grid(g)
    dev.set_patches(ps)
    _check(dev, g, accurate_triangulation=False)
    dev.close()
