"""Host side without a GPU: the C ABI library loads and exports every symbol
the header declares, argument checking, the SES patch builder's geometry and
the synthetic input generator."""
import ctypes
import os
import re

import numpy as np
import pytest

import nanoshaper_amd as ns
from nanoshaper_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "nsgpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(nsgpu_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)
    assert lib.nsgpu_abi_version() == 1


def test_library_runs_on_the_rocm_runtime_it_is_built_for():
    # loaded first (as the test session and bench.py load it), the library
    # maps ROCm's libamdhip64, not the copy torch's wheel bundles under the
    # same soname
    import subprocess
    import sys
    code = "import nanoshaper_amd as ns; ns._lib.load(); print(ns._lib.hip_runtime())"
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120,
                         cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    assert out.returncode == 0, out.stderr
    assert "/opt/rocm" in out.stdout and "torch" not in out.stdout, out.stdout


def test_argument_errors_without_device():
    lib = _lib.load()
    assert lib.nsgpu_build_grid(None, 0, 2.0, 80.0, None) == -1
    a = np.zeros((2, 4))
    a[:, 3] = 1.5
    g = _lib.Grid()
    assert lib.nsgpu_build_grid(a.ctypes.data_as(ctypes.c_void_p), 2, 2.0, 120.0, ctypes.byref(g)) == -1
    assert lib.nsgpu_create(0, None) == -1
    with pytest.raises(ValueError):
        ns.build_grid(np.zeros((3, 3)))
    with pytest.raises(ns.NsgpuError):
        ns.NanoShaper(a, surface_type="mesh")  # the facade takes ses, skin and blobby
    sk = ns.NanoShaper(a, surface_type="skin")
    assert sk.parameter == 0.45 and ns.NanoShaper(a, "ses", probe=1.2).probe == 1.2


def test_generator_is_deterministic_and_self_avoiding():
    a, r1 = ns.generate_chain(400, 13.0, seed=3)
    b, r2 = ns.generate_chain(400, 13.0, seed=3)
    c, _ = ns.generate_chain(400, 13.0, seed=4)
    assert np.array_equal(a, b) and r1 == r2
    assert not np.array_equal(a, c)
    step = np.linalg.norm(np.diff(a[:, :3], axis=0), axis=1)
    assert np.allclose(step, 1.5)
    d = np.linalg.norm(a[:, None, :3] - a[None, :, :3], axis=2)
    far = np.abs(np.arange(400)[:, None] - np.arange(400)[None, :]) >= 3
    assert d[far].min() >= 3.0 - 1e-9
    assert np.linalg.norm(a[:, :3], axis=1).max() <= r1 + 1e-9
    assert set(np.unique(a[:, 3])) <= {1.70, 1.55, 1.52, 1.80}


@pytest.fixture(scope="module")
def patches():
    xyzr, _ = ns.generate_chain(330, 12.0, seed=0)
    jit, _ = ns.jitter(xyzr)
    return jit, ns.ses_patches(jit, 1.4)


def test_patch_counts_and_layout(patches):
    jit, ps = patches
    c = ps.counts()
    assert c["point"] > 0 and c["facet_regular"] > 0 and c["torus_regular"] > 0
    assert sum(c.values()) == len(ps)
    for p in range(len(ps)):
        k = ps.kind[p]
        need = 14 + 8 * ps.cnt[p] if k >= 3 else 4 * ps.cnt[p]
        assert ps.off[p] + need <= len(ps.data)
    # singular facets come in mirror pairs
    assert c["facet_singular"] % 2 == 0


def test_facet_geometry(patches):
    jit, ps = patches
    R = jit[:, 3] + 1.4
    for p in np.nonzero(ps.kind == 1)[0][:200]:
        probe = ps.bsph[p, :3]
        # the probe touches its three atoms and is free of all others
        ids = ps.atom[p]
        d = np.linalg.norm(jit[ids, :3] - probe, axis=1)
        assert np.allclose(d, R[ids], atol=1e-9)
        dall = np.linalg.norm(jit[:, :3] - probe, axis=1)
        assert (dall >= R - 1e-9).all()
        planes = ps.data[ps.off[p]:ps.off[p] + 4 * ps.cnt[p]].reshape(-1, 4)
        assert np.allclose(np.linalg.norm(planes[:, :3], axis=1), 1.0)
        # the probe centre lies on its three trimming planes (regular facets)
        assert np.allclose(planes[:3, :3] @ probe + planes[:3, 3], 0.0, atol=1e-6)


def test_torus_geometry(patches):
    jit, ps = patches
    for p in np.nonzero(ps.kind == 3)[0][:200]:
        d = ps.data[ps.off[p]:ps.off[p] + 14]
        center, inv, Rmaj = d[:3], d[3:12].reshape(3, 3), d[12]
        assert np.allclose(inv @ inv.T, np.eye(3), atol=1e-9)  # orthonormal frame
        i, j = ps.atom[p, :2]
        axis = jit[j, :3] - jit[i, :3]
        axis /= np.linalg.norm(axis)
        assert np.allclose(np.abs(inv[2] @ axis), 1.0)  # frame z is the atom axis
        # the circle of probe centres is at distance r_i + probe from atom i
        ri = jit[i, 3] + 1.4
        assert np.hypot(Rmaj, np.linalg.norm(center - jit[i, :3])) == pytest.approx(ri, abs=1e-9)


def test_point_cells_are_exposed_atoms(patches):
    jit, ps = patches
    pts = ps.atom[ps.kind == 0, 0]
    assert len(np.unique(pts)) == len(pts)
    # every atom on a torus or facet has a point cell
    on_surface = set(ps.atom[ps.kind >= 1, :].ravel()) - {-1}
    assert on_surface <= set(pts.tolist())


def test_writers_follow_the_reference_formats(tmp_path):
    from nanoshaper_amd import io
    ideb = np.zeros((2, 3, 4), np.uint8)  # (nz, ny, nx)
    ideb[1, 2, 3] = 1
    p = tmp_path / "m.idebmap.txt"
    io.save_idebmap(str(p), ideb)
    text = p.read_text()
    # i outermost (4 blocks), j lines, k values: node (i=3, j=2, k=1) is the last value
    blocks = text.split("\n\n")
    assert len([b for b in blocks if b.strip()]) == 4
    assert blocks[3].splitlines()[2] == "0 1 "
    eps = np.zeros((2, 3, 4, 3), np.int32)
    eps[0, 1, 2, 1] = 5
    paths = io.save_epsmaps(str(tmp_path / "m"), eps)
    assert [pth.rsplit(".", 2)[-2] for pth in paths] == ["epsmapx", "epsmapy", "epsmapz"]
    assert open(paths[1]).read().split("\n\n")[2].splitlines()[1] == "5 0 "
    v = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.25, 0.0]])
    t = np.array([[0, 1, 2]])
    n = np.array([[0.0, 0.0, 1.0]] * 3)
    io.save_off(str(tmp_path / "m.off"), v, t, normals=n, revert=True)
    lines = (tmp_path / "m.off").read_text().splitlines()
    assert lines[0] == "OFF+N" and lines[2] == "3 1 0" and lines[5] == "0.000 1.250 0.000 0.000 0.000 1.000"
    assert lines[-1] == "3 2 1 0"
    vv, tt, nn = io.load_off(str(tmp_path / "m.off"))
    assert np.allclose(vv, v) and np.array_equal(tt, t[:, ::-1]) and np.allclose(nn, n)


def test_bench_intersect_kernels_name_the_profiled_kernels():
    # the roofline's traffic comes from the committed counter summary: the
    # bench must name the intersection kernels as rocprof lists them
    import importlib.util
    import json
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(root, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    kern = json.load(open(os.path.join(root, "profiles", "ses_c2_5000atoms_traffic.json")))["kernels"]
    for k in bench.INTERSECT_KERNELS:
        assert k in kern, k
