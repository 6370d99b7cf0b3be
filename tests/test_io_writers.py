"""Writers of the reference's output files and the nearest-atom search behind
the vertex-atom map and the cavity atoms (CPU: the status maps and meshes come
from the oracle; the nearest atoms from the library's host stage)."""
import os

import numpy as np
import pytest

import nanoshaper_amd as ns
from nanoshaper_amd import io


def _power_nearest(xyzr, pts):
    # brute force over every atom: smallest |p - c|^2 - r^2
    pw = ((pts[:, None, :] - xyzr[None, :, :3]) ** 2).sum(-1) - xyzr[None, :, 3] ** 2
    return pw.argmin(1)


def test_nearest_atoms_is_the_power_nearest_near_the_atoms():
    xyzr, _ = ns.generate_chain(80, 9.0, seed=3)
    g = ns.build_grid(xyzr, 2.0, 80.0)
    rng = np.random.default_rng(0)
    k = rng.integers(0, len(xyzr), 2000)
    d = rng.normal(size=(2000, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    # points up to 1.5 A off an atom's sphere (where mesh vertices lie)
    pts = xyzr[k, :3] + d * (xyzr[k, 3:4] + rng.uniform(0, 1.5, (2000, 1)))
    near = io.nearest_atoms(g, xyzr, pts)
    assert (near == _power_nearest(xyzr, pts)).all()
    # far outside the acceleration grid's reach: no atom
    far = io.nearest_atoms(g, xyzr, np.array([[1e4, 1e4, 1e4]]))
    assert far[0] == -1


def test_save_status_layout(tmp_path):
    st = np.arange(2 * 3 * 4).reshape(2, 3, 4) - 1  # (nz, ny, nx)
    p = str(tmp_path / "m.status.txt")
    io.save_status(p, st)
    blocks = open(p).read().split("\n\n")
    # i outermost, then one line per j holding k = 0..nz-1
    for i in range(4):
        lines = blocks[i].strip("\n").split("\n")
        assert len(lines) == 3
        for j in range(3):
            assert lines[j] == " ".join(str(st[k, j, i]) for k in range(2)) + " "


def _oracle_surface_with_cavities():
    from oracle.binding import OracleSurface
    xyzr, _ = ns.generate_chain(300, 10.0, seed=1)
    grid = ns.build_grid(xyzr, 2.0, 80.0)
    jit, _ = ns.jitter(xyzr)
    orc = OracleSurface(grid, ns.ses_patches(jit, 1.4))
    r = orc.get_surf(fill_cavities=True, cav_vol=11.4)
    return grid, jit, orc, r


def test_save_cavities_files(tmp_path):
    grid, jit, orc, r = _oracle_surface_with_cavities()
    assert r["num_cavities"] >= 2
    st = orc.status()
    sizes, filled = orc.cavities()
    root = str(tmp_path / "x")
    catoms = io.cavity_atoms(grid, jit, st, r["num_cavities"])
    io.save_cavities(root, grid, st, sizes, filled, cav_atoms=catoms)
    x, y, z = ns.grid_coords(grid)
    lines = open(root + "cavities.txt").read().split("\n")
    assert lines[0] == ""
    kk, jj, ii = np.nonzero(st > ns.STATUS_OUT)
    assert len(lines) - 1 == len(kk)
    got = np.array([[float(v) for v in l.split()] for l in lines[1:]])
    assert np.abs(got - np.stack([x[ii], y[jj], z[kk]], 1)).max() <= 5e-7
    rows = open(root + "cavitiesSize.txt").read().splitlines()
    assert float(rows[0]) == len(sizes)
    for c, row in enumerate(rows[1:]):
        bx, by, bz, vol, flag = row.split()
        sel = st == ns.STATUS_FIRST_CAV + c
        k_, j_, i_ = np.nonzero(sel)
        assert float(bx) == pytest.approx(x[i_].mean(), abs=1e-6)
        assert float(by) == pytest.approx(y[j_].mean(), abs=1e-6)
        assert float(bz) == pytest.approx(z[k_].mean(), abs=1e-6)
        assert float(vol) == pytest.approx(sizes[c], abs=1e-6)
        assert int(flag) == int(filled[c])
    ser = open(root + "cavAtomsSerials.txt").read().splitlines()
    assert len(ser) == r["num_cavities"]
    for c, line in enumerate(ser):
        assert [int(v) - 1 for v in line.split()] == catoms[c]
    # each listed atom is the power-nearest atom of a boundary node of its cavity
    for c in range(r["num_cavities"]):
        sel = st == ns.STATUS_FIRST_CAV + c
        pad = np.pad(st, 1, mode="edge")
        core = pad[1:-1, 1:-1, 1:-1]
        edge = np.zeros(st.shape, bool)
        for ax in range(3):
            for sh in (1, -1):
                edge |= np.roll(pad, sh, axis=ax)[1:-1, 1:-1, 1:-1] != core
        k_, j_, i_ = np.nonzero(sel & edge)
        want = sorted(set(_power_nearest(jit, np.stack([x[i_], y[j_], z[k_]], 1)).tolist()))
        assert catoms[c] == want
    # normal mode computes no cavity atoms: the serials file is empty
    io.save_cavities(root, grid, st, sizes, filled)
    assert open(root + "cavAtomsSerials.txt").read() == ""


def _mesh():
    from oracle.binding import OracleSurface
    xyzr, _ = ns.generate_chain(40, 6.0, seed=2)
    grid = ns.build_grid(xyzr, 2.0, 80.0)
    jit, _ = ns.jitter(xyzr)
    orc = OracleSurface(grid, ns.ses_patches(jit, 1.4))
    orc.get_surf()
    orc.triangulate()
    v, t = orc.mesh()
    return grid, jit, v, t, orc.normals()


@pytest.mark.parametrize("fmt", io.MESH_FORMATS)
def test_save_mesh_formats_round_trip(tmp_path, fmt):
    grid, jit, v, t, n = _mesh()
    va = io.vertex_atoms(grid, jit, v)
    assert (va >= 0).all() and (va == _power_nearest(jit, v)).all()
    root = str(tmp_path / "r_")
    paths = io.save_mesh(root, v, t, normals=n, vertex_atoms=va, fmt=fmt)
    assert all(os.path.exists(p) for p in paths)
    if fmt == "PLY":
        v2, t2, n2 = io.load_ply(paths[0])
        assert np.array_equal(v2, v) and np.array_equal(t2, t) and np.array_equal(n2, n)
    elif fmt.startswith("OFF"):
        lines = open(paths[0]).read().split("\n")
        assert lines[0] == fmt and lines[1].startswith("# File created by") and lines[2] == ""
        assert lines[3] == "%d %d 0" % (len(v), len(t))
        rows = [l.split() for l in lines[4:4 + len(v)]]
        cols = np.array([[float(x) for x in r] for r in rows])
        assert np.abs(cols[:, :3] - v).max() <= 5e-4
        if "N" in fmt:
            assert np.abs(cols[:, 3:6] - n).max() <= 5e-4
        if fmt.endswith("A"):
            assert np.array_equal(cols[:, -1].astype(int), va)
        tri = np.array([[int(x) for x in l.split()] for l in lines[4 + len(v):4 + len(v) + len(t)]])
        assert (tri[:, 0] == 3).all() and np.array_equal(tri[:, 1:], t)
        v3, t3, _ = io.load_off(paths[0])
        assert np.array_equal(t3, t)
    else:
        face = open(paths[0]).read().splitlines()
        vert = open(paths[1]).read().splitlines()
        assert face[1] == "#faces" and int(face[2]) == len(t)
        ft = np.array([[int(x) for x in l.split()] for l in face[3:]])
        assert np.array_equal(ft[:, :3] - 1, t) and (ft[:, 3:] == 1).all()
        assert vert[1] == "#vertex" and int(vert[2]) == len(v)
        vv = np.array([[float(x) for x in l.split()] for l in vert[3:]])
        assert np.abs(vv[:, :3] - v).max() <= 5e-4 and np.abs(vv[:, 3:6] - n).max() <= 5e-4
        want = va + 1 if fmt == "MSMS" else np.ones(len(v))
        assert np.array_equal(vv[:, 7], want) and (vv[:, 6] == 0).all() and (vv[:, 8] == 0).all()
    areas = np.array([float(x) for x in open(paths[-1]).read().split()])
    e1, e2 = v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]]
    assert np.abs(areas - 0.5 * np.linalg.norm(np.cross(e1, e2), axis=1)).max() <= 1e-5


def test_deduce_format():
    assert io.deduce_format() == "OFF"
    assert io.deduce_format(normals=True) == "OFF+N"
    assert io.deduce_format(vertex_atoms=True) == "OFF+A"
    assert io.deduce_format(vertex_atoms=True, normals=True) == "OFF+N+A"
    assert io.deduce_format(save_msms=True) == "MSMS_NO_A"
    assert io.deduce_format(save_msms=True, vertex_atoms=True) == "MSMS"
    assert io.deduce_format(save_ply=True, save_msms=True) == "PLY"


def test_save_bgp_format(tmp_path):
    # DelPhiShared::saveBGP: count line, 1-based indices and "%lf" fields, no
    # newline after the last line; surf.txt without the indices
    ib = np.array([[1, 2, 3], [4, 5, 6]], np.int32)
    pos = np.array([[0.5, -1.25, 2.0], [1e-7, 3.0, -0.0]])
    nor = np.zeros((2, 3))
    p1, p2 = io.save_bgp(str(tmp_path) + "/", "mol", ib, pos, nor)
    assert open(p1).read() == ("2\n2 3 4 0.500000 -1.250000 2.000000 0.000000 0.000000 0.000000\n"
                               "5 6 7 0.000000 3.000000 -0.000000 0.000000 0.000000 0.000000")
    assert open(p2).read() == ("0.500000 -1.250000 2.000000 0.000000 0.000000 0.000000\n"
                               "0.000000 3.000000 -0.000000 0.000000 0.000000 0.000000")
