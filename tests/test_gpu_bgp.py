"""GPU parity of the boundary grid points (Project_boundary_grid_points):
the HIP detection and SES projection against the oracle on the same maps.

Bars: the point list (node indices, in order) and the counts bit-exact; the
projections within 1e-12 A (the same FP64 expressions; only the sampled
fallback calls cos/sin, whose last bit may differ between the device and
glibc); normals zero, as the reference leaves them for the SES.
"""
import os

import numpy as np
import pytest

import nanoshaper_amd as ns

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
POS_TOL_A = 1e-12


def _check(grid, ps, opts, aux=50):
    from oracle.binding import OracleSurface
    print("bgp case grid=%d patches=%d aux=%d opts=%s" % (grid.nx, len(ps), aux, opts), flush=True)
    dev = ns.DeviceSurface(0)
    dev.set_grid(grid)
    dev.set_patches(ps)
    orc = OracleSurface(grid, ps)
    dev.get_surf(**opts)
    orc.get_surf(**opts)
    assert np.array_equal(dev.epsmap(), orc.epsmap())
    a = dev.project_bgp(aux)
    b = orc.project_bgp(aux)
    assert list(a["counts"]) == list(b["counts"]), (a["counts"], b["counts"])
    assert np.array_equal(a["ibgp"], b["ibgp"])
    assert np.abs(a["scspos"] - b["scspos"]).max(initial=0.0) <= POS_TOL_A
    assert np.all(a["scsnor"] == 0)
    dev.close()
    return a


@pytest.mark.parametrize("n,radius,scale,perfil,seed,aux", [(120, 9.0, 2.0, 80.0, 7, 50),
                                                           (330, 12.0, 2.0, 90.0, 0, 50),
                                                           (60, 7.0, 3.0, 70.0, 3, 50),
                                                           (60, 7.0, 3.0, 70.0, 3, 100),
                                                           (2000, 22.0, 3.0, 80.0, 1, 50)])
def test_bgp_matches_oracle(n, radius, scale, perfil, seed, aux):
    xyzr, _ = ns.generate_chain(n, radius, seed=seed)
    grid = ns.build_grid(xyzr, scale, perfil)
    jit, _ = ns.jitter(xyzr)
    ps = ns.ses_patches(jit, 1.4)
    a = _check(grid, ps, dict(build_epsmap=True, build_status=True, fill_cavities=True), aux)
    assert a["counts"][0] > 0
    if (n, seed) == (60, 3):
        assert a["counts"][2] > 0  # the sampled projection near self-intersecting tori runs


def test_bgp_1crn():
    atoms = ns.load_atoms(os.path.join(GOLDEN, "1crn.xyzr"))
    grid = ns.build_grid(atoms, 2.0, 80.0)
    jit, _ = ns.jitter(atoms)
    _check(grid, ns.ses_patches(jit, 1.4), dict(build_epsmap=True, build_status=True, fill_cavities=True))


def test_bgp_benchmark_size():
    # the 5,000-atom chain at 4 points/A (325^3 nodes), cavities filled
    xyzr, _ = ns.generate_chain(5000, 28.0, seed=0)
    grid = ns.build_grid(xyzr, 4.0, 80.0)
    jit, _ = ns.jitter(xyzr)
    a = _check(grid, ns.ses_patches(jit, 1.4), dict(build_epsmap=True, build_status=True, fill_cavities=True))
    assert a["counts"][0] > 500000


def test_bgp_tables_follow_the_inputs():
    # one context: the cached auxiliary grid follows the aux size and a new
    # patch set on the same grid
    from oracle.binding import OracleSurface
    xyzr, _ = ns.generate_chain(150, 9.0, seed=11)
    grid = ns.build_grid(xyzr, 3.0, 80.0)
    dev = ns.DeviceSurface(0)
    dev.set_grid(grid)
    opts = dict(build_epsmap=True, build_status=True, fill_cavities=True)
    for seed, aux in ((11, 50), (11, 100), (12, 100)):
        print("bgp tables seed=%d aux=%d" % (seed, aux), flush=True)
        x, _ = ns.generate_chain(150, 9.0, seed=seed)
        jit, _ = ns.jitter(x)
        ps = ns.ses_patches(jit, 1.4)
        dev.set_patches(ps)
        dev.get_surf(**opts)
        orc = OracleSurface(grid, ps)
        orc.get_surf(**opts)
        a, b = dev.project_bgp(aux), orc.project_bgp(aux)
        assert np.array_equal(a["ibgp"], b["ibgp"]) and list(a["counts"]) == list(b["counts"])
        assert np.abs(a["scspos"] - b["scspos"]).max() <= POS_TOL_A
        orc.close()
    dev.close()


def test_bgp_refusals():
    xyzr, _ = ns.generate_chain(40, 6.0, seed=2)
    grid = ns.build_grid(xyzr, 2.0, 80.0)
    jit, _ = ns.jitter(xyzr)
    dev = ns.DeviceSurface(0)
    dev.set_grid(grid)
    dev.set_patches(ns.ses_patches(jit, 1.4))
    with pytest.raises(ns.NsgpuError):
        dev.project_bgp()  # no surface yet
    dev.get_surf(build_epsmap=False, build_status=True, fill_cavities=False)
    with pytest.raises(ns.NsgpuError):
        dev.project_bgp()  # no epsilon map
    dev.close()
