"""Text writers in the reference's formats, for inspecting results.

* DelPhiShared::saveIdebMap / saveEpsMaps (DelphiShared.cpp:1255-1345):
  one value per node followed by a space, k fastest, a newline after every
  (i, j) line and an empty line after every i.
* Surface::saveMesh, OFF and OFF+N (Surface.cpp:10860-10970): "OFF" (or
  "OFF+N"), a comment line, "nv nt 0", vertices with three decimals (and
  normals), then "3 a b c" per triangle, reversed when `revert`.
* save_mesh: every format of Surface::saveMesh (Surface.cpp:10857-11069) and
  savePLYMesh (:10739-10850): OFF, OFF+A, OFF+N, OFF+N+A, MSMS (.face /
  .vert, with or without atoms), binary little-endian PLY, and the
  triangleAreas.txt it writes beside them; deduce_format is
  Surface::deduceFormat (:13355-13385).
* save_status (DelPhiShared::saveStatus, DelphiShared.cpp:1458-1503) and
  save_cavities (DelPhiShared::saveCavities, DelphiShared.cpp:1889-1990):
  cavities.txt, cavitiesSize.txt and cavAtomsSerials.txt.
* save_bgp (DelPhiShared::saveBGP, DelphiShared.cpp:1421-1455): the
  boundary grid points as root + name + ".projections.txt" and root +
  "surf.txt".
* vertex_atoms / cavity_atoms: the nearest atom of every mesh vertex
  (vertexAtomsMap) and the atoms of every cavity (Surface::getCavitiesAtoms,
  Surface.cpp:2129-2245), through nsgpu_nearest_atoms.
"""
import ctypes
import time

import numpy as np

from . import _lib

PROGNAME = "nanoshaper_amd"
VERSION = "1.0"
STATUS_OUT, STATUS_FIRST_CAV = 3, 4


def _write_grid(path, a):
    """a: (nz, ny, nx) integer array in the DelPhi layout."""
    nz, ny, nx = a.shape
    t = np.ascontiguousarray(a.transpose(2, 1, 0))  # (i, j, k)
    with open(path, "w") as f:
        for i in range(nx):
            lines = [" ".join(str(int(v)) for v in t[i, j]) + " \n" for j in range(ny)]
            f.write("".join(lines))
            f.write("\n")


def save_idebmap(path, idebmap):
    """idebmap: (nz, ny, nx), 1 = out (DelPhiShared::saveIdebMap)."""
    _write_grid(path, np.asarray(idebmap))


def save_epsmaps(prefix, epsmap):
    """epsmap: (nz, ny, nx, 3); writes prefix + .epsmapx/y/z.txt (DelPhiShared::saveEpsMaps)."""
    e = np.asarray(epsmap)
    paths = []
    for l, axis in enumerate("xyz"):
        p = "%s.epsmap%s.txt" % (prefix, axis)
        _write_grid(p, e[..., l])
        paths.append(p)
    return paths


def save_off(path, verts, tris, normals=None, revert=False, comment="# File created by nanoshaper_amd"):
    """OFF (or OFF+N with normals) mesh as Surface::saveMesh writes it."""
    v = np.asarray(verts, dtype=np.float64).reshape(-1, 3)
    t = np.asarray(tris, dtype=np.int64).reshape(-1, 3)
    if revert:
        t = t[:, ::-1]
    with open(path, "w") as f:
        f.write("OFF+N\n" if normals is not None else "OFF\n")
        f.write(comment + "\n")
        f.write("%d %d 0\n" % (len(v), len(t)))
        if normals is not None:
            n = np.asarray(normals, dtype=np.float64).reshape(-1, 3)
            rows = np.hstack([v, n])
            f.write("".join("%.3f %.3f %.3f %.3f %.3f %.3f\n" % tuple(r) for r in rows))
        else:
            f.write("".join("%.3f %.3f %.3f\n" % tuple(r) for r in v))
        f.write("".join("3 %d %d %d\n" % tuple(r) for r in t))


def load_off(path, as_float32=False):
    """(verts, tris, normals or None) of an OFF / OFF+N file written by save_off.
    as_float32: coordinates rounded to float as MeshSurface::loadOFF reads them
    (sscanf "%f", MeshSurface.cpp:2436-2446), the input of a "mesh" surface."""
    with open(path) as f:
        head = f.readline().strip()
        line = f.readline()
        while line.startswith("#") or not line.strip():
            line = f.readline()
        nv, nt, _ = (int(x) for x in line.split())
        rows = np.array([[float(x) for x in f.readline().split()] for _ in range(nv)]).reshape(nv, -1)
        tris = np.array([[int(x) for x in f.readline().split()[1:4]] for _ in range(nt)], dtype=np.int64).reshape(nt, 3)
    normals = rows[:, 3:6] if head == "OFF+N" else None
    verts = rows[:, :3]
    if as_float32:
        # float(str) is the nearest double; the nearest float to the decimal
        # text is the float nearest that double except at exact float
        # midpoints, which a 3-decimal text cannot be
        verts = np.array([[float(np.float32(x)) for x in r] for r in verts]).reshape(-1, 3)
    return verts, tris, normals


def _stamp():
    # ctime(): "Thu Dec  4 11:23:30 2026\n", newline included
    return time.asctime() + "\n"


def save_status(path, status):
    """status: (nz, ny, nx) in the DelPhi layout (DelPhiShared::saveStatus:
    i outermost, a value and a space per node, k fastest)."""
    _write_grid(path, np.asarray(status))


def save_bgp(root, name, ibgp, scspos, scsnor):
    """root + name + '.projections.txt': the count, then per point its 1-based
    node indices, projection and normal; root + 'surf.txt': projection and
    normal per point.  "%lf" fields, no newline after the last line.
    Returns the two paths."""
    n = len(ibgp)
    ib = np.asarray(ibgp, dtype=np.int64) + 1
    pos, nor = np.asarray(scspos, float), np.asarray(scsnor, float)
    p1, p2 = root + name + ".projections.txt", root + "surf.txt"
    six = ["%f %f %f %f %f %f" % (tuple(pos[i]) + tuple(nor[i])) for i in range(n)]
    with open(p1, "w") as f:
        f.write("%d\n" % n)
        f.write("\n".join("%d %d %d " % tuple(ib[i]) + six[i] for i in range(n)))
    with open(p2, "w") as f:
        f.write("\n".join(six))
    return [p1, p2]


def _grid_axes(grid):
    x = grid.xmin + np.arange(grid.nx) * grid.side
    y = grid.ymin + np.arange(grid.ny) * grid.side
    z = grid.zmin + np.arange(grid.nz) * grid.side
    return x, y, z


def save_cavities(root, grid, status, sizes, filled, cav_atoms=None, only_non_filled=False, axes=None):
    """DelPhiShared::saveCavities: root + cavities.txt (the coordinates of
    every node whose status is a cavity label, k outermost), cavitiesSize.txt
    (the count, then per cavity its nodes' barycenter, its volume and its
    filled flag) and cavAtomsSerials.txt (the serials, index + 1, of the
    atoms of each cavity when cav_atoms is given, one line per cavity; empty
    otherwise, as after a normal-mode run).  axes: the node coordinates
    (x, y, z) when the caller has them exactly (they are xmin + i side
    otherwise)."""
    st = np.asarray(status)
    x, y, z = axes if axes is not None else _grid_axes(grid)
    kk, jj, ii = np.nonzero(st > STATUS_OUT)  # C order: k, then j, then i
    with open(root + "cavities.txt", "w") as f:
        f.write("".join("\n %f %f %f" % (x[i], y[j], z[k]) for k, j, i in zip(kk, jj, ii)))
    sizes = np.asarray(sizes, dtype=np.float64)
    lab = st[kk, jj, ii] - STATUS_FIRST_CAV
    with open(root + "cavitiesSize.txt", "w") as f:
        f.write("%f\n" % float(np.float32(len(sizes))))
        for c in range(len(sizes)):
            sel = lab == c
            n = int(sel.sum())
            bx = float(np.cumsum(x[ii[sel]])[-1]) / n if n else float("nan")
            by = float(np.cumsum(y[jj[sel]])[-1]) / n if n else float("nan")
            bz = float(np.cumsum(z[kk[sel]])[-1]) / n if n else float("nan")
            f.write("%f %f %f %f %d\n" % (bx, by, bz, sizes[c], 1 if filled[c] else 0))
    with open(root + "cavAtomsSerials.txt", "w") as f:
        for c, atoms in enumerate(cav_atoms or []):
            if only_non_filled and filled[c]:
                continue
            f.write("".join("%d " % (a + 1) for a in atoms) + "\n")


def nearest_atoms(grid, xyzr, points):
    """nsgpu_nearest_atoms: the nearest atom (|p - c|^2 - r^2) of each point
    over the reference's atom acceleration grid, -1 where none is in reach."""
    lib = _lib.load()
    a = np.ascontiguousarray(xyzr, dtype=np.float64).reshape(-1, 4)
    p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    out = np.empty(len(p), np.int32)
    rc = lib.nsgpu_nearest_atoms(ctypes.byref(grid), a.ctypes.data_as(ctypes.c_void_p), a.shape[0],
                                 p.ctypes.data_as(ctypes.c_void_p), p.shape[0], out.ctypes.data_as(ctypes.c_void_p))
    if rc != 0:
        raise ValueError("nsgpu_nearest_atoms failed (%d)" % rc)
    return out


def vertex_atoms(grid, xyzr, verts):
    """vertexAtomsMap: the nearest atom of every mesh vertex (Surface.cpp:10348-10390)."""
    return nearest_atoms(grid, xyzr, verts)


def cavity_atoms(grid, xyzr, status, num_cavities, axes=None):
    """Surface::getCavitiesAtoms: for cavity c (status label 4 + c), the
    sorted atoms nearest to its nodes that have a 6-neighbour of another
    status (-1 where no atom is in reach)."""
    st = np.asarray(status)
    nz, ny, nx = st.shape
    x, y, z = axes if axes is not None else _grid_axes(grid)
    pad = np.pad(st, 1, mode="edge")
    core = pad[1:-1, 1:-1, 1:-1]
    edge = np.zeros(st.shape, bool)
    for d, sl in enumerate(((slice(2, None), slice(1, -1), slice(1, -1)), (slice(None, -2), slice(1, -1), slice(1, -1)),
                            (slice(1, -1), slice(2, None), slice(1, -1)), (slice(1, -1), slice(None, -2), slice(1, -1)),
                            (slice(1, -1), slice(1, -1), slice(2, None)), (slice(1, -1), slice(1, -1), slice(None, -2)))):
        edge |= pad[sl] != core
    kk, jj, ii = np.nonzero(edge & (st >= STATUS_FIRST_CAV) & (st < STATUS_FIRST_CAV + num_cavities))
    near = nearest_atoms(grid, xyzr, np.stack([x[ii], y[jj], z[kk]], 1)) if len(ii) else np.zeros(0, np.int32)
    lab = st[kk, jj, ii] - STATUS_FIRST_CAV
    return [sorted(set(near[lab == c].tolist())) for c in range(num_cavities)]


MESH_FORMATS = ("OFF", "OFF+A", "OFF+N", "OFF+N+A", "MSMS", "MSMS_NO_A", "PLY")


def deduce_format(save_ply=False, save_msms=False, vertex_atoms=False, normals=False):
    """Surface::deduceFormat (Surface.cpp:13355-13385)."""
    if save_ply:
        return "PLY"
    if not save_msms:
        if vertex_atoms and normals:
            return "OFF+N+A"
        return "OFF+A" if vertex_atoms else ("OFF+N" if normals else "OFF")
    return "MSMS" if vertex_atoms else "MSMS_NO_A"


def triangle_areas(verts, tris):
    """Heron's area per triangle as Surface::saveMesh computes it (0 for degenerate ones)."""
    v = np.asarray(verts, dtype=np.float64).reshape(-1, 3)
    t = np.asarray(tris, dtype=np.int64).reshape(-1, 3)

    def dist(p, q):
        d = v[p] - v[q]
        return np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])

    a, b, c = dist(t[:, 0], t[:, 1]), dist(t[:, 0], t[:, 2]), dist(t[:, 1], t[:, 2])
    ttt = (a + b + c) * (b + c - a) * (c + a - b) * (a + b - c)
    return np.where(ttt > 0, 0.25 * np.sqrt(np.maximum(ttt, 0.0)), 0.0), ttt > 0


def save_mesh(root, verts, tris, normals=None, vertex_atoms=None, fmt="OFF", revert=False,
              file_name="triangulatedSurf"):
    """Surface::saveMesh: root + file_name + .off / .face and .vert / .ply in
    format fmt (MESH_FORMATS), then root + file_name + triangleAreas.txt.
    Vertex atoms are written as stored in the map (OFF+A) or + 1 (MSMS).
    Returns the paths written."""
    if fmt not in MESH_FORMATS:
        raise ValueError("unknown mesh format %r" % fmt)
    v = np.asarray(verts, dtype=np.float64).reshape(-1, 3)
    t = np.asarray(tris, dtype=np.int64).reshape(-1, 3)
    tw = t[:, ::-1] if revert else t
    n = None if normals is None else np.asarray(normals, dtype=np.float64).reshape(-1, 3)
    va = None if vertex_atoms is None else np.asarray(vertex_atoms, dtype=np.int64).reshape(-1)
    if fmt in ("OFF+N", "OFF+N+A", "MSMS", "MSMS_NO_A") and n is None:
        raise ValueError("%s needs normals" % fmt)
    if fmt in ("OFF+A", "OFF+N+A", "MSMS") and va is None:
        raise ValueError("%s needs the vertex atoms" % fmt)
    base = root + file_name
    paths = []
    if fmt == "PLY":
        path = base + ".ply"
        head = ["ply", "format binary_little_endian 1.0", "element vertex %d" % len(v),
                "property double x", "property double y", "property double z"]
        if n is not None:
            head += ["property double nx", "property double ny", "property double nz"]
        head += ["element face %d" % len(t), "property list uchar int vertex_indices", "end_header"]
        rows = np.hstack([v, n]) if n is not None else v
        faces = np.zeros(len(t), dtype=[("k", "u1"), ("i", "<i4", 3)])
        faces["k"] = 3
        faces["i"] = tw
        with open(path, "wb") as f:
            f.write(("\n".join(head) + "\n").encode())
            f.write(np.ascontiguousarray(rows, dtype="<f8").tobytes())
            f.write(faces.tobytes())
        paths.append(path)
    elif fmt.startswith("OFF"):
        path = base + ".off"
        with open(path, "w") as f:
            f.write(fmt + "\n")
            f.write("# File created by %s version %s date %s\n" % (PROGNAME, VERSION, _stamp()))
            f.write("%d %d 0\n" % (len(v), len(t)))
            if fmt == "OFF+A":
                f.write("".join("%.3f %.3f %.3f %d\n" % (r[0], r[1], r[2], a) for r, a in zip(v, va)))
            elif fmt == "OFF+N":
                f.write("".join("%.3f %.3f %.3f %.3f %.3f %.3f\n" % (tuple(r) + tuple(q)) for r, q in zip(v, n)))
            elif fmt == "OFF+N+A":
                f.write("".join("%.3f %.3f %.3f %.3f %.3f %.3f %d\n" % (tuple(r) + tuple(q) + (a,))
                                for r, q, a in zip(v, n, va)))
            else:
                f.write("".join("%.3f %.3f %.3f\n" % tuple(r) for r in v))
            f.write("".join("3 %d %d %d\n" % tuple(r) for r in tw))
        paths.append(path)
    else:
        stamp = _stamp()
        fpath, vpath = base + ".face", base + ".vert"
        with open(fpath, "w") as f:
            f.write("# File created by %s version %s date %s" % (PROGNAME, VERSION, stamp))
            f.write("#faces\n%d\n" % len(t))
            f.write("".join("%6d %6d %6d %2d %6d\n" % (r[0] + 1, r[1] + 1, r[2] + 1, 1, 1) for r in tw))
        with open(vpath, "w") as f:
            f.write("# File created by %s version %s date %s" % (PROGNAME, VERSION, stamp))
            f.write("#vertex\n%d\n" % len(v))
            atoms = va + 1 if fmt == "MSMS" else np.ones(len(v), np.int64)
            f.write("".join("%9.3f %9.3f %9.3f %9.3f %9.3f %9.3f %7d %7d %2d \n" % (tuple(r) + tuple(q) + (0, a, 0))
                            for r, q, a in zip(v, n, atoms)))
        paths += [fpath, vpath]
    areas, ok = triangle_areas(v, t)
    apath = base + "triangleAreas.txt"
    with open(apath, "w") as f:
        f.write("".join(("%f\n" % a) if k else "0.\n" for a, k in zip(areas, ok)))
    paths.append(apath)
    return paths


def load_ply(path):
    """(verts, tris, normals or None) of a binary little-endian PLY written by save_mesh."""
    with open(path, "rb") as f:
        data = f.read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    head = data[:end].decode().splitlines()
    nv = int([l for l in head if l.startswith("element vertex")][0].split()[2])
    nt = int([l for l in head if l.startswith("element face")][0].split()[2])
    nprop = sum(1 for l in head if l.startswith("property double"))
    rows = np.frombuffer(data, dtype="<f8", count=nv * nprop, offset=end).reshape(nv, nprop)
    faces = np.frombuffer(data, dtype=[("k", "u1"), ("i", "<i4", 3)], count=nt, offset=end + 8 * nv * nprop)
    assert (faces["k"] == 3).all()
    return rows[:, :3].copy(), faces["i"].astype(np.int64), (rows[:, 3:6].copy() if nprop >= 6 else None)
