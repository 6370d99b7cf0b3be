"""GPU: the reference's own test inputs and configurations (tests/golden,
copied from the reference's tests/): 1crn (327 atoms) as SES and skin, and
6xm4 (48,519 atoms) as skin, each run as the reference CLI's normal mode
would run its .prm, against the oracle on the same patches.

Bars: hit counts and depths per panel, status/epsmap/idebmap, corner bits,
the mesh (vertices and triangles) bit-exact; vertex normals within 1e-9;
volume and area relative 1e-12."""
import os

import numpy as np
import pytest

from nanoshaper_amd import config

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _compare(run):
    from oracle.binding import OracleSurface
    conf, dev, r = run["config"], run["device"], run["result"]
    opts = conf.surf_opts()
    orc = OracleSurface(run["grid"], run["patches"])
    ro = orc.get_surf(**opts)
    for k in ("total_rays", "odd_rays", "total_hits", "candidates", "num_cavities", "max_hits_per_ray"):
        assert r[k] == ro[k], (k, r[k], ro[k])
    assert r["volume"] == pytest.approx(ro["volume"], rel=1e-12)
    nslots = (3 if opts["build_epsmap"] else 1) + (3 if opts["accurate_triangulation"] else 0)
    for s in range(nslots):
        ca, ta = dev.ray_hits(s)
        cb, tb = orc.ray_hits(s)
        assert np.array_equal(ca, cb), "slot %d counts" % s
        assert np.abs(ta - tb).max() <= 3e-13, "slot %d depths" % s
    assert np.array_equal(dev.status(), orc.status())
    if opts["build_epsmap"]:
        assert np.array_equal(dev.epsmap(), orc.epsmap())
        assert np.array_equal(dev.idebmap(), orc.idebmap())
    assert np.array_equal(dev.corner_bits(), orc.corner_bits())
    mo = orc.triangulate()
    m = run["mesh"]
    assert (m["num_vertices"], m["num_triangles"]) == (mo["num_vertices"], mo["num_triangles"])
    assert m["area"] == pytest.approx(mo["area"], rel=1e-12)
    va, ta = dev.mesh()
    vb, tb = orc.mesh()
    assert np.array_equal(ta, tb)
    assert np.abs(va - vb).max() <= 1e-9
    if conf.normals:
        assert np.abs(dev.normals() - orc.normals()).max() <= 1e-9
    orc.close()
    return r, m


@pytest.mark.parametrize("prm", ["1crn_ses_conf.prm", "1crn_skin_conf.prm"])
def test_1crn_configurations(device_ok, prm):
    print("config", prm, flush=True)
    run = config.normal_mode(os.path.join(GOLDEN, prm))
    assert len(run["atoms"]) == 327 and run["grid"].nx == 67
    r, m = _compare(run)
    assert r["odd_rays"] == 0 and m["num_triangles"] > 10000
    run["device"].close()


def test_6xm4_skin_configuration(device_ok):
    print("config 6xm4_skin_conf.prm", flush=True)
    run = config.normal_mode(os.path.join(GOLDEN, "6xm4_skin_conf.prm"))
    assert len(run["atoms"]) == 48519
    print("6xm4 grid %d, cells %d" % (run["grid"].nx, len(run["patches"])), flush=True)
    r, m = _compare(run)
    assert m["num_triangles"] > 1000000
    run["device"].close()
