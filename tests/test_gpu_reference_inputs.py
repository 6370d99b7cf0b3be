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


@pytest.mark.parametrize("extra,files", [
    ({"Vertex_Atom_Info": "true", "Save_Status_map": "true", "Cavity_Detection_Filling": "true"},
     ["triangulatedSurf.off", "triangulatedSurftriangleAreas.txt", "1crn.status.txt", "cavities.txt",
      "cavitiesSize.txt", "cavAtomsSerials.txt"]),
    ({"Save_Mesh_MSMS_Format": "true", "Vertex_Atom_Info": "true", "Build_epsilon_maps": "true",
      "Save_eps_maps": "true", "Save_ideb_map": "true"},
     ["triangulatedSurf.face", "triangulatedSurf.vert", "1crn.epsmapx.txt", "1crn.epsmapy.txt",
      "1crn.epsmapz.txt", "1crn.idebmap.txt"]),
    ({"Save_Mesh_PLY_Format": "true", "Smooth_Mesh": "true"}, ["triangulatedSurf.ply"]),
    ({"Build_epsilon_maps": "true", "Project_boundary_grid_points": "true", "Save_bgps": "true"},
     ["1crn.projections.txt", "surf.txt"]),
])
def test_1crn_normal_mode_files(device_ok, tmp_path, extra, files):
    # the 1crn SES configuration with output keys switched on: the files the
    # reference's normal mode writes, holding the device's maps and mesh
    from nanoshaper_amd import io
    text = open(os.path.join(GOLDEN, "1crn_ses_conf.prm")).read()
    lines = [l for l in text.splitlines() if l.split("=")[0].strip() not in set(extra) | {"XYZR_FileName",
                                                                                       "Sys_Name"}]
    lines += ["%s = %s" % kv for kv in extra.items()]
    lines += ["XYZR_FileName = %s" % os.path.join(GOLDEN, "1crn.xyzr"), "Sys_Name = 1crn"]
    prm = tmp_path / "run.prm"
    prm.write_text("\n".join(lines) + "\n")
    print("config", extra, flush=True)
    run = config.normal_mode(str(prm))
    out_dir = tmp_path / "out"
    out_dir.mkdir()
    paths = config.write_outputs(run, str(out_dir))
    for f in files:
        assert os.path.exists(out_dir / f), f
    dev = run["device"]
    v, t = dev.mesh()
    if "Save_Mesh_PLY_Format" in extra:
        v2, t2, n2 = io.load_ply(str(out_dir / "triangulatedSurf.ply"))
        # Smooth_Mesh: the saved mesh is the smoothed one, with its normals
        assert np.array_equal(v2, v) and np.array_equal(t2, t) and np.array_equal(n2, dev.normals())
    elif "Save_bgps" in extra:
        b = run["bgp"]
        rows = open(out_dir / "1crn.projections.txt").read().splitlines()
        assert int(rows[0]) == len(b["ibgp"]) == len(rows) - 1 > 1000
        first = rows[1].split()
        assert [int(x) - 1 for x in first[:3]] == list(b["ibgp"][0])
        assert np.allclose([float(x) for x in first[3:6]], b["scspos"][0], atol=5e-7)
        assert len(open(out_dir / "surf.txt").read().splitlines()) == len(b["ibgp"])
    elif "Save_Mesh_MSMS_Format" in extra:
        vert = open(out_dir / "triangulatedSurf.vert").read().splitlines()
        va = np.array([int(l.split()[7]) for l in vert[3:]]) - 1
        assert np.array_equal(va, io.vertex_atoms(run["grid"], run["jittered"], v))
        e = np.loadtxt(out_dir / "1crn.epsmapx.txt").ravel()
        assert e.size == run["grid"].nx * run["grid"].ny * run["grid"].nz
    else:
        v3, t3, _ = io.load_off(str(out_dir / "triangulatedSurf.off"))
        assert np.array_equal(t3, t) and np.abs(v3 - v).max() <= 5e-4
        head = open(out_dir / "triangulatedSurf.off").readline().strip()
        assert head == "OFF+N+A"
        st = np.loadtxt(out_dir / "1crn.status.txt").ravel()
        assert st.size == run["grid"].nx * run["grid"].ny * run["grid"].nz
        rows = open(out_dir / "cavitiesSize.txt").read().splitlines()
        assert float(rows[0]) == run["result"]["num_cavities"] == len(rows) - 1
    run["device"].close()
