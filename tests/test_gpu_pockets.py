"""GPU parity of pocket detection (the grid part of pocket mode): the bit-plane
morphology, union-find components and labels on the device against the
oracle's restatement of the reference's loops, on the same two status maps.
Bars: the final status map, component volumes, filled and pocket flags and
the support point lists all bit-exact."""
import numpy as np
import pytest

import nanoshaper_amd as ns

pytestmark = pytest.mark.gpu


def _pair(xyzr, scale, perfil):
    g = ns.build_grid(xyzr, scale, perfil)
    jit1, st = ns.jitter(xyzr)
    jit2, _ = ns.jitter(xyzr, seed=st)
    return g, ns.ses_patches(jit1, 3.0), ns.ses_patches(jit2, 1.4)


def _run(g, ps_fat, ps_small, cap, num_mol=2):
    from oracle.binding import OracleSurface
    print("pockets grid=%d cav_and_pockets=%s" % (g.nx, cap), flush=True)
    fat_opts = dict(build_epsmap=False, build_status=True, fill_cavities=True, cav_vol=1e30)
    small_opts = dict(build_epsmap=False, build_status=True, fill_cavities=True, cav_vol=0.0)
    df, ds = ns.DeviceSurface(0), ns.DeviceSurface(0)
    for d, ps, o in ((df, ps_fat, fat_opts), (ds, ps_small, small_opts)):
        d.set_grid(g)
        d.set_patches(ps)
        d.get_surf(**o)
    of, os_ = OracleSurface(g, ps_fat), OracleSurface(g, ps_small)
    of.get_surf(**fat_opts)
    os_.get_surf(**small_opts)
    assert np.array_equal(df.status(), of.status()) and np.array_equal(ds.status(), os_.status())
    a = df.pockets(ds, num_mol=num_mol, cav_and_pockets=cap)
    b = of.pockets(os_, num_mol=num_mol, cav_and_pockets=cap)
    assert np.array_equal(df.status(), of.status())
    assert np.array_equal(df.pocket_labels(), of.pocket_labels())
    assert np.array_equal(a["volumes"], b["volumes"])
    assert np.array_equal(a["filled"], b["filled"])
    kept = ~b["filled"]
    assert np.array_equal(a["is_pocket"][kept], b["is_pocket"][kept])
    assert len(a["support"]) == len(b["support"])
    for x, y in zip(a["support"], b["support"]):
        assert np.array_equal(x, y)
    df.close()
    ds.close()
    return a


@pytest.mark.parametrize("n,radius,scale,seed,cap", [(300, 11.0, 2.0, 3, True), (500, 13.0, 2.5, 4, False),
                                                    (2000, 22.0, 2.0, 1, True), (2000, 22.0, 3.0, 1, True)])
def test_pockets_match_oracle(n, radius, scale, seed, cap):
    xyzr, _ = ns.generate_chain(n, radius, seed=seed)
    g, pf, psm = _pair(xyzr, scale, 80.0)
    a = _run(g, pf, psm, cap)
    assert len(a["volumes"]) > 0 and len(a["support"]) > 0


def test_pockets_1crn():
    import os
    atoms = ns.load_atoms(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "1crn.xyzr"))
    g, pf, psm = _pair(atoms, 2.0, 80.0)
    _run(g, pf, psm, True)


def test_pockets_refusals():
    xyzr, _ = ns.generate_chain(60, 7.0, seed=2)
    g, pf, psm = _pair(xyzr, 2.0, 80.0)
    df, ds = ns.DeviceSurface(0), ns.DeviceSurface(0)
    df.set_grid(g)
    df.set_patches(pf)
    ds.set_grid(g)
    ds.set_patches(psm)
    with pytest.raises(ns.NsgpuError):
        df.pockets(ds)  # no status maps yet
    df.close()
    ds.close()


def test_pocket_mode_envelopes(tmp_path):
    """Pocket mode end to end (config.run_pockets): the device pockets, the
    files, and each envelope's volume and area against the oracle's skin
    surface of the same jittered support balls."""
    from nanoshaper_amd import config
    from oracle.binding import OracleSurface
    conf = config.Configuration(config.ConfigFile(text="Operative_Mode = pocket\nBuild_status_map = true\n"
                                                       "Triangulation = true\nAccurate_Triangulation = true\n"
                                                       "Grid_scale = 2.0\nGrid_perfil = 80\n"))
    xyzr, _ = ns.generate_chain(300, 11.0, seed=3)
    print("pocket mode 300 atoms", flush=True)
    run = config.run_pockets(conf, xyzr)
    run["config"] = conf
    pk, kept = run["pockets"], run["kept"]
    assert len(kept) >= 3 and pk["is_pocket"][kept].any() and not pk["is_pocket"][kept].all()
    paths, lines = config.write_pocket_outputs(run, str(tmp_path))
    assert open(tmp_path / "numcav.txt").read() == str(len(kept))
    assert len(lines) == len(kept) + 1 and sum(l.startswith("Pocket") for l in lines[1:]) == int(pk["is_pocket"][kept].sum())
    rows = open(tmp_path / "cavitiesSize.txt").read().splitlines()
    assert len(rows) == len(pk["volumes"]) + 1
    # the envelopes, rebuilt on the host with the oracle: same random sequence
    _, seed = ns.jitter(xyzr)
    _, seed = ns.jitter(xyzr, seed=seed)
    g = run["grid"]
    x, y, z = ns.grid_coords(g)
    for q, c in enumerate(kept):
        f = pk["support"][q]
        pts = np.stack([x[f % g.nx], y[(f // g.nx) % g.ny], z[f // (g.nx * g.ny)]], 1)
        a = np.array([[float("%f" % v) for v in p] + [1.4] for p in pts])
        if len(a) < 4:
            a = np.vstack([a, np.repeat(a[-1:], 4 - len(a), axis=0)])
        ge = ns.build_grid(a, conf.scale, 50.0)
        jit, seed = ns.jitter(a, amplitude=0.1, seed=seed)
        orc = OracleSurface(ge, ns.skin_patches(jit, conf.skin_s))
        r = orc.get_surf(build_epsmap=False, build_status=False, accurate_triangulation=True, fill_cavities=False)
        m = orc.triangulate()
        assert run["envelope_volumes"][q] == pytest.approx(r["volume"], rel=1e-12)
        assert run["envelope_areas"][q] == pytest.approx(m["area"], rel=1e-12)
        orc.close()
    run["fat"].close()
    run["small"].close()
