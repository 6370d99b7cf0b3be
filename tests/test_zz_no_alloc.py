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


class _DeviceArray:
    """A device array of the caller, made with the HIP runtime the library
    itself runs on (already loaded by it), filled from a host array."""
    H2D, D2H = 1, 2

    def __init__(self, host):
        path = [p for p in _lib.hip_runtime() if "/opt/rocm" in p][0]
        self.hip = hip = ctypes.CDLL(path)
        hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        hip.hipFree.argtypes = [ctypes.c_void_p]
        self.host = np.ascontiguousarray(host)
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), self.host.nbytes) == 0
        assert hip.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, self.H2D) == 0

    def read(self):
        out = np.empty_like(self.host)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, self.D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _noalloc(dev, g, **kw):
    lib = _lib.load()
    nvox = g.nx * g.ny * g.nz
    o = surf_opts(**kw)
    eps = _DeviceArray(np.full(nvox * 3, -7, np.int32))
    ideb = _DeviceArray(np.full(nvox, 9, np.uint8))
    st = _DeviceArray(np.full(nvox, -7, np.int32))
    r = SurfResult()
    rc = lib.nsgpu_get_surf_noalloc(dev._h, ctypes.byref(o), ctypes.byref(r),
                                    eps.ptr if o.build_epsmap else ctypes.c_void_p(0),
                                    ideb.ptr if o.build_epsmap else ctypes.c_void_p(0),
                                    st.ptr if o.build_status else ctypes.c_void_p(0))
    msg = lib.nsgpu_last_error(dev._h).decode() if rc else ""
    res = {f: getattr(r, f) for f in FIELDS}
    shape = (g.nz, g.ny, g.nx)
    out = (eps.read().reshape(shape + (3,)), ideb.read().reshape(shape), st.read().reshape(shape))
    for a in (eps, ideb, st):
        a.free()
    return (rc, msg, res) + out


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
