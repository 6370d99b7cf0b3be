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
