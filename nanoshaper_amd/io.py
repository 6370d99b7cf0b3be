"""Text writers in the reference's formats, for inspecting results.

* DelPhiShared::saveIdebMap / saveEpsMaps (DelphiShared.cpp:1255-1345):
  one value per node followed by a space, k fastest, a newline after every
  (i, j) line and an empty line after every i.
* Surface::saveMesh, OFF and OFF+N (Surface.cpp:10860-10970): "OFF" (or
  "OFF+N"), a comment line, "nv nt 0", vertices with three decimals (and
  normals), then "3 a b c" per triangle, reversed when `revert`.
"""
import numpy as np


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
        while line.startswith("#"):
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
