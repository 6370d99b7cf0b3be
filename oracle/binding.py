"""ctypes binding of the oracle library (oracle/libns_oracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the product package.  The interface
mirrors nanoshaper_amd.DeviceSurface so results can be compared field by field.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libns_oracle.so")
VP = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    from nanoshaper_amd._lib import Grid, MeshInfo, Patches, SurfOpts, SurfResult
    lib = ctypes.CDLL(LIB_PATH)
    sig = {
        "orc_create": (VP, [ctypes.POINTER(Grid), ctypes.POINTER(Patches), ctypes.c_double]),
        "orc_free": (None, [VP]),
        "orc_get_surf": (ctypes.c_int, [VP, ctypes.POINTER(SurfOpts), ctypes.POINTER(SurfResult)]),
        "orc_triangulate": (ctypes.c_int, [VP, ctypes.POINTER(MeshInfo)]),
        "orc_read_epsmap": (None, [VP, VP]),
        "orc_read_idebmap": (None, [VP, VP]),
        "orc_read_status": (None, [VP, VP]),
        "orc_read_corner_bits": (None, [VP, VP]),
        "orc_read_cavities": (None, [VP, VP, VP]),
        "orc_read_mesh": (None, [VP, VP, VP]),
        "orc_read_normals": (None, [VP, VP]),
        "orc_read_ray_hits": (ctypes.c_int, [VP, ctypes.c_int, ctypes.c_int, VP, VP]),
        "orc_ray_sphere": (ctypes.c_int, [VP, VP, VP, ctypes.c_double, VP, VP]),
        "orc_cubic": (ctypes.c_double, [ctypes.c_double] * 3),
        "orc_quartic": (None, [VP, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, VP]),
        "orc_ray_torus": (ctypes.c_int, [VP, VP, VP, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                         VP, VP, VP, VP]),
        "orc_is_feasible": (ctypes.c_int, [ctypes.POINTER(Patches), ctypes.c_int, VP]),
        "orc_set_root_refinement": (None, [ctypes.c_int]),
        "orc_set_threads": (None, [ctypes.c_int]),
        "orc_smooth_surface": (ctypes.c_int, [VP, ctypes.c_int]),
        "orc_sturm_quartic": (ctypes.c_int, [VP, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
        "orc_create_mesh": (VP, [ctypes.POINTER(Grid), VP, ctypes.c_int64, VP, ctypes.c_int64, ctypes.c_int,
                                 ctypes.c_int]),
        "orc_ray_triangle": (ctypes.c_int, [VP, VP, VP, VP, VP, VP]),
        "orc_blobby_build": (ctypes.c_int, [ctypes.POINTER(Grid), VP, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                            ctypes.POINTER(VP), ctypes.POINTER(VP), ctypes.POINTER(VP),
                                            ctypes.POINTER(VP), ctypes.POINTER(ctypes.c_int64),
                                            ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(VP)]),
        "orc_free_buf": (None, [VP]),
        "orc_project_bgp": (ctypes.c_int, [VP, ctypes.c_int, VP]),
        "orc_read_bgp": (None, [VP, VP, VP, VP]),
        "orc_pockets": (ctypes.c_int, [VP, VP, ctypes.c_double, ctypes.c_double, ctypes.c_int, VP]),
        "orc_read_pockets": (None, [VP, VP, VP, VP, VP, VP]),
        "orc_read_pocket_labels": (None, [VP, VP]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_threads(n):
    """Threads of the oracle's patch loop, ray rows and marching-cubes slices
    (the reference's decomposition); results do not depend on the count."""
    load().orc_set_threads(int(n))


def blobby_build(grid, xyzr, blobbyness=-2.5, cutoff=6.0):
    """BlobbySurface::build restated: {'field' (nz, ny, nx), 'mc' (nv, 3): the
    marching-cubes vertices, 'raw' and 'verts' (nv, 3): smoothed vertices
    before and after the OFF text round trip,
    'tris' (nt, 3), 'area' of the marching cubes before smoothing}."""
    lib = load()
    a = np.ascontiguousarray(xyzr, dtype=np.float64).reshape(-1, 4)
    f, raw, v, t, mc = VP(), VP(), VP(), VP(), VP()
    nv, nt, area = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double()
    lib.orc_blobby_build(ctypes.byref(grid), _ptr(a), a.shape[0], float(blobbyness), float(cutoff), ctypes.byref(f),
                         ctypes.byref(raw), ctypes.byref(v), ctypes.byref(t), ctypes.byref(nv), ctypes.byref(nt),
                         ctypes.byref(area), ctypes.byref(mc))

    def take(p, n, dt):
        out = np.empty(n, dt)
        if n:
            ctypes.memmove(out.ctypes.data, p.value, out.nbytes)
        lib.orc_free_buf(p)
        return out

    field = take(f, grid.nx * grid.ny * grid.nz, np.float64).reshape(grid.nz, grid.ny, grid.nx)
    return {"field": field, "mc": take(mc, 3 * nv.value, np.float64).reshape(-1, 3),
            "raw": take(raw, 3 * nv.value, np.float64).reshape(-1, 3),
            "verts": take(v, 3 * nv.value, np.float64).reshape(-1, 3),
            "tris": take(t, 3 * nt.value, np.int32).reshape(-1, 3), "area": area.value}


def ray_triangle(orig, d, v0, v1, v2):
    """MeshSurface::intersect_triangle: t or None."""
    lib = load()
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (orig, d, v0, v1, v2)]
    t = np.zeros(1)
    ok = lib.orc_ray_triangle(*[_ptr(x) for x in arrs], _ptr(t))
    return float(t[0]) if ok else None


class OracleSurface:
    """CPU restatement of Surface::getSurf / triangulateSurface on one patch set."""

    @classmethod
    def from_mesh(cls, grid, verts, tris, ray_based=True, aux_grid_dim=100):
        """The same on a triangle mesh (MeshSurface)."""
        o = cls.__new__(cls)
        o._lib = load()
        o.grid = grid
        o.refine_roots = False
        v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
        o._keep = (v, t)
        o._h = o._lib.orc_create_mesh(ctypes.byref(grid), _ptr(v), v.shape[0], _ptr(t), t.shape[0],
                                      int(bool(ray_based)), int(aux_grid_dim))
        o.result = None
        o.mesh_info = None
        return o

    def __init__(self, grid, patchset, refine_roots=False):
        """refine_roots: torus roots refined to the true root of the computed
        quartic (extended-precision Newton), the yardstick of the device's
        mixed-precision mode; False restates the reference exactly."""
        from nanoshaper_amd._lib import Grid  # noqa: F401
        self._lib = load()
        self.grid = grid
        self.refine_roots = bool(refine_roots)
        self._h = self._lib.orc_create(ctypes.byref(grid), ctypes.byref(patchset.view), patchset.probe)
        self.result = None
        self.mesh_info = None

    def get_surf(self, **kw):
        from nanoshaper_amd._lib import SurfResult
        from nanoshaper_amd.surface import surf_opts
        o = surf_opts(**kw)
        r = SurfResult()
        self._lib.orc_set_root_refinement(int(self.refine_roots))
        try:
            rc = self._lib.orc_get_surf(self._h, ctypes.byref(o), ctypes.byref(r))
        finally:
            self._lib.orc_set_root_refinement(0)
        if rc != 0:
            raise RuntimeError("oracle get_surf failed")
        self.result = {f: getattr(r, f) for f, _ in SurfResult._fields_}
        return self.result

    def triangulate(self):
        from nanoshaper_amd._lib import MeshInfo
        m = MeshInfo()
        if self._lib.orc_triangulate(self._h, ctypes.byref(m)) != 0:
            raise RuntimeError("oracle triangulate needs accurate_triangulation")
        self.mesh_info = {f: getattr(m, f) for f, _ in MeshInfo._fields_}
        return self.mesh_info

    def _shape(self):
        g = self.grid
        return g.nz, g.ny, g.nx

    def epsmap(self):
        out = np.empty(self._shape() + (3,), np.int32)
        self._lib.orc_read_epsmap(self._h, _ptr(out))
        return out

    def idebmap(self):
        out = np.empty(self._shape(), np.uint8)
        self._lib.orc_read_idebmap(self._h, _ptr(out))
        return out

    def status(self):
        out = np.empty(self._shape(), np.int32)
        self._lib.orc_read_status(self._h, _ptr(out))
        return out

    def corner_bits(self):
        nz, ny, nx = self._shape()
        out = np.empty((nx * ny * nz + 31) // 32, np.uint32)
        self._lib.orc_read_corner_bits(self._h, _ptr(out))
        return out

    def cavities(self):
        n = self.result["num_cavities"]
        sizes = np.zeros(max(n, 1))
        filled = np.zeros(max(n, 1), np.uint8)
        self._lib.orc_read_cavities(self._h, _ptr(sizes), _ptr(filled))
        return sizes[:n], filled[:n].astype(bool)

    def mesh(self):
        m = self.mesh_info
        v = np.empty((m["num_vertices"], 3))
        t = np.empty((m["num_triangles"], 3), np.int32)
        self._lib.orc_read_mesh(self._h, _ptr(v), _ptr(t))
        return v, t

    def smooth_surface(self, normals=False):
        self._lib.orc_smooth_surface(self._h, int(bool(normals)))

    def normals(self):
        out = np.empty((self.mesh_info["num_vertices"], 3))
        self._lib.orc_read_normals(self._h, _ptr(out))
        return out

    def project_bgp(self, aux_grid_dim=50):
        """Boundary grid points and their projections on the SES:
        {'ibgp' (n, 3) node indices, 'scspos' (n, 3), 'scsnor' (n, 3),
        'counts' [points, internal, sampled fallback, left at the node]}."""
        counts = np.zeros(4, np.int64)
        if self._lib.orc_project_bgp(self._h, int(aux_grid_dim), _ptr(counts)) != 0:
            raise RuntimeError("oracle BGP projection needs an SES get_surf with the epsilon map")
        n = int(counts[0])
        ib = np.empty((n, 3), np.int32)
        pos = np.empty((n, 3))
        nor = np.empty((n, 3))
        self._lib.orc_read_bgp(self._h, _ptr(ib), _ptr(pos), _ptr(nor))
        return {"ibgp": ib, "scspos": pos, "scsnor": nor, "counts": counts}

    def pockets(self, small, filter_probe=1.4, num_mol=2, cav_and_pockets=True):
        """The grid part of pocket mode on this (fat-probe) surface and the
        small-probe surface `small` (both after get_surf on the same grid):
        rewrites this surface's status to the pocket labels and returns
        {'volumes', 'filled', 'is_pocket'} per component (ids 4, 5, ...) and
        'support': for each kept component, its support nodes (flat indices)."""
        counts = np.zeros(3, np.int64)
        if self._lib.orc_pockets(self._h, small._h, float(filter_probe), float(num_mol), int(bool(cav_and_pockets)),
                                 _ptr(counts)) != 0:
            raise RuntimeError("oracle pockets need two status maps on one grid")
        nc, nk, ns_ = (int(c) for c in counts)
        vol = np.zeros(max(nc, 1))
        filled = np.zeros(max(nc, 1), np.uint8)
        pocket = np.zeros(max(nc, 1), np.uint8)
        off = np.zeros(nk + 1, np.int64)
        sup = np.zeros(max(ns_, 1), np.int64)
        self._lib.orc_read_pockets(self._h, _ptr(vol), _ptr(filled), _ptr(pocket), _ptr(off), _ptr(sup))
        return {"volumes": vol[:nc], "filled": filled[:nc].astype(bool), "is_pocket": pocket[:nc].astype(bool),
                "support": [sup[off[q]:off[q + 1]] for q in range(nk)]}

    def pocket_labels(self):
        out = np.empty(self._shape(), np.int32)
        self._lib.orc_read_pocket_labels(self._h, _ptr(out))
        return out

    def ray_hits(self, slot):
        g = self.grid
        N = max(g.nx, g.ny, g.nz)
        maxh = max(1, self.result["max_hits_per_ray"])
        counts = np.empty(N * N, np.int32)
        t = np.zeros((N * N, maxh))
        self._lib.orc_read_ray_hits(self._h, int(slot), maxh, _ptr(counts), _ptr(t))
        return counts, t

    def close(self):
        if getattr(self, "_h", None):
            self._lib.orc_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
