"""The C++ plugin layer (include/nsgpu_surface.hpp) compiles against the C ABI
and drives the path the way reference callers do."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "nanoshaper_amd")


@pytest.fixture(scope="module")
def client(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("cpp") / "surface_api_client")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "surface_api_client.cpp"), "-L" + PKG, "-lnsgpu",
                    "-Wl,-rpath," + PKG, "-o", exe], check=True)
    return exe


def test_cpp_client_host_stages(client):
    out = subprocess.run([client, "host"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "atoms 330 grid 61" in out.stdout


def test_cpp_skin_surface_builds_its_own_triangulation(client):
    # "skin" through the factory with no triangulator plugged in: the
    # library's regular triangulation, the same cells as the Python path
    import nanoshaper_amd as ns
    out = subprocess.run([client, "skinhost"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    xyzr, _ = ns.generate_chain(330, 12.0, seed=0)
    jit, _ = ns.jitter(xyzr)
    assert "skin patches %d" % len(ns.skin_patches(jit, 0.45)) in out.stdout


@pytest.mark.gpu
def test_cpp_client_full_path_matches_python(client, device_ok):
    import numpy as np

    import nanoshaper_amd as ns
    out = subprocess.run([client, "gpu"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    line = [l for l in out.stdout.splitlines() if l.startswith("volume")][0].split()
    vol, area = float(line[1]), float(line[3])
    xyzr, _ = ns.generate_chain(330, 12.0, seed=0)
    g = ns.build_grid(xyzr, 2.0, 90.0)
    jit, _ = ns.jitter(xyzr)
    dev = ns.DeviceSurface(0)
    dev.set_grid(g)
    dev.set_patches(ns.ses_patches(jit, 1.4))
    r = dev.get_surf()
    m = dev.triangulate()
    assert vol == pytest.approx(r["volume"], rel=1e-12)
    assert area == pytest.approx(m["area"], rel=1e-12)
    assert int(line[7]) == int((dev.status() == -1).sum())


@pytest.mark.gpu
def test_cpp_client_blobby_and_mesh(client, device_ok, tmp_path):
    import numpy as np

    import nanoshaper_amd as ns
    from nanoshaper_amd.io import save_off
    xyzr, _ = ns.generate_chain(330, 12.0, seed=0)
    g = ns.build_grid(xyzr, 2.0, 80.0)
    dev = ns.DeviceSurface(0)
    dev.set_grid(g)
    dev.blobby_build(xyzr)
    r = dev.get_surf(inside=5)
    m = dev.triangulate()
    out = subprocess.run([client, "blobby"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    f = out.stdout.split()
    assert float(f[1]) == pytest.approx(r["volume"], rel=1e-12)
    assert float(f[3]) == pytest.approx(m["area"], rel=1e-12)
    assert int(f[7]) == int((dev.status() == -1).sum())
    # the blobby mesh written as OFF and loaded as a "mesh" surface: the same
    # numbers (the file holds the same 3-decimal coordinates)
    v, t = dev.surface_mesh()
    off = str(tmp_path / "blobby.off")
    save_off(off, v, t)
    out = subprocess.run([client, "mesh", off], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert float(out.stdout.split()[1]) == pytest.approx(r["volume"], rel=1e-12)


@pytest.mark.gpu
def test_cpp_facade(client, device_ok):
    """NS::NanoShaper over the device: volume and area equal the device path's,
    castAxisOrientedRay returns the grid line's depths with unit normals, and
    the collected grid-ray crossings are the kept segment ends of all panels."""
    import numpy as np

    import nanoshaper_amd as ns
    out = subprocess.run([client, "facade"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    f = out.stdout.split()
    vol, area, nhits, first, rays, crossings = float(f[2]), float(f[4]), int(f[6]), float(f[8]), int(f[12]), int(f[14])
    xyzr, _ = ns.generate_chain(330, 12.0, seed=0)
    g = ns.build_grid(xyzr, 2.0, 90.0)
    jit, _ = ns.jitter(xyzr)
    dev = ns.DeviceSurface(0)
    dev.set_grid(g)
    dev.set_patches(ns.ses_patches(jit, 1.4))
    r = dev.get_surf(build_epsmap=True, build_status=True, accurate_triangulation=True, fill_cavities=False)
    m = dev.triangulate()
    assert vol == pytest.approx(r["volume"], rel=1e-12) and area == pytest.approx(m["area"], rel=1e-12)
    xs, ys, zs = ns.grid_coords(g)
    counts, hits = dev.cast_rays(0, [[xs[0], ys[g.ny // 2], zs[g.nz // 2]]], xs[-1])
    assert counts[0] == nhits and float("%.12f" % hits[0, 0]) == pytest.approx(first, abs=1e-12)
    kept = 0
    for slot in range(3):
        cnt, t = dev.ray_hits(slot)
        for ray in np.nonzero((cnt > 0) & (cnt % 2 == 0))[0]:
            T = t[ray, :cnt[ray]]
            kept += 2 * int((np.abs(T[1::2] - T[0::2]) >= 1e-8).sum())
    assert crossings == kept and rays == 3 * g.nx * g.ny


@pytest.mark.gpu
def test_cpp_configuration_run(client, device_ok):
    from nanoshaper_amd import config
    prm = os.path.join(ROOT, "tests", "golden", "1crn_ses_conf.prm")
    out = subprocess.run([client, "prm", prm], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    f = out.stdout.split()
    run = config.normal_mode(prm)
    assert float(f[2]) == pytest.approx(run["result"]["volume"], rel=1e-12)
    assert float(f[4]) == pytest.approx(run["mesh"]["area"], rel=1e-12)
    assert int(f[6]) == run["mesh"]["num_triangles"] and int(f[10]) == 327


@pytest.mark.gpu
def test_cpp_surface_follows_a_rebuilt_grid(client, device_ok):
    out = subprocess.run([client, "rescale"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr


@pytest.mark.gpu
def test_cpp_bgp_files_match_python(client, device_ok, tmp_path):
    """getSurf with Project_boundary_grid_points through the C++ layer, saved
    by DelPhiShared::saveBGP: the same files as the Python path writes."""
    import nanoshaper_amd as ns
    from nanoshaper_amd import io as IO
    cdir = tmp_path / "cpp"
    cdir.mkdir()
    out = subprocess.run([client, "bgp", str(cdir)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    nb = int(out.stdout.split()[1])
    xyzr, _ = ns.generate_chain(330, 12.0, seed=0)
    g = ns.build_grid(xyzr, 2.0, 90.0)
    jit, _ = ns.jitter(xyzr)
    dev = ns.DeviceSurface(0)
    dev.set_grid(g)
    dev.set_patches(ns.ses_patches(jit, 1.4))
    dev.get_surf(build_epsmap=True, build_status=True, accurate_triangulation=False, fill_cavities=True)
    b = dev.project_bgp()
    assert b["counts"][0] == nb
    pdir = tmp_path / "py"
    pdir.mkdir()
    IO.save_bgp(str(pdir) + "/", "mol", b["ibgp"], b["scspos"], b["scsnor"])
    for f in ("mol.projections.txt", "surf.txt"):
        assert (cdir / f).read_bytes() == (pdir / f).read_bytes()
