"""Host-side mirror of NanoShaper's surface API over the C ABI.

* ``build_grid``        DelPhiShared::buildGrid (DelphiShared.cpp:1038-1201)
* ``jitter``            ConnollySurface.cpp:2484-2489 with randnum() (tools.cpp:64-84)
* ``ses_patches``       ConnollySurface::build (host stage, ConnollySurface.cpp:2326-3316)
* ``DeviceSurface``     Surface::getSurf / triangulateSurface on the GPU
* ``NanoShaper``        the NS::NanoShaper facade (nanoshaper.h:294-424)

Arrays cross the boundary as NumPy buffers in the DelPhi layouts.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import Grid, MeshInfo, Patches, SurfOpts, SurfResult

PATCH_KINDS = ("point", "facet_regular", "facet_singular", "torus_regular", "torus_singular")
STATUS_INSIDE, STATUS_TEMP_OUT, STATUS_OUT, STATUS_FIRST_CAV = -1, 2, 3, 4
STAGES = ("intersect", "sort", "fill", "corners", "cavities", "vertices", "marching_cubes", "adjoint")


class NsgpuError(RuntimeError):
    pass


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc, ctx=None, what=""):
    if rc != 0:
        lib = _lib.load()
        msg = lib.nsgpu_last_error(ctx).decode() if ctx else ""
        raise NsgpuError("%s failed (%d): %s" % (what, rc, msg))


def as_xyzr(atoms):
    a = np.ascontiguousarray(np.asarray(atoms, dtype=np.float64))
    if a.ndim != 2 or a.shape[1] != 4 or a.shape[0] == 0:
        raise ValueError("atoms must be an (n, 4) array of x, y, z, radius")
    return a


def load_atoms(path, max_atoms=-1, domain_shrinkage=0.0, charges=False):
    """DelPhiShared::loadAtoms: (n, 4) x, y, z, radius of an .xyzr or .pqr
    file (and the (n,) charges when charges=True)."""
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.nsgpu_load_atoms(str(path).encode(), int(max_atoms), float(domain_shrinkage), ctypes.byref(h))
    try:
        if rc != 0:
            msg = lib.nsgpu_atom_set_error(h).decode() if h else ""
            raise NsgpuError("load_atoms %s failed (%d): %s" % (path, rc, msg))
        px, pq = ctypes.POINTER(ctypes.c_double)(), ctypes.POINTER(ctypes.c_double)()
        n = lib.nsgpu_atom_set_view(h, ctypes.byref(px), ctypes.byref(pq))
        xyzr = np.ctypeslib.as_array(px, (n, 4)).copy()
        q = np.ctypeslib.as_array(pq, (n,)).copy()
    finally:
        if h:
            lib.nsgpu_atom_set_free(h)
    return (xyzr, q) if charges else xyzr


def build_grid(xyzr, scale=2.0, perfil=80.0):
    """DelPhiShared::buildGrid: cubic odd grid centred on the atom bounding box."""
    lib = _lib.load()
    a = as_xyzr(xyzr)
    g = Grid()
    _check(lib.nsgpu_build_grid(_ptr(a), a.shape[0], float(scale), float(perfil), ctypes.byref(g)), None,
           "build_grid")
    return g


def grid_coords(g):
    """Node coordinates x[i] = xmin + i*side (DelphiShared.cpp:1135-1142)."""
    return (g.xmin + np.arange(g.nx) * g.side, g.ymin + np.arange(g.ny) * g.side,
            g.zmin + np.arange(g.nz) * g.side)


def jitter(xyzr, amplitude=1e-4, seed=1234):
    """Return (jittered copy, new generator state)."""
    lib = _lib.load()
    a = as_xyzr(xyzr).copy()
    s = ctypes.c_int32(seed)
    lib.nsgpu_jitter(_ptr(a), a.shape[0], float(amplitude), ctypes.byref(s))
    return a, s.value


def generate_chain(n, radius, seed=0):
    """Seeded self-avoiding chain; returns (xyzr, confinement radius used)."""
    lib = _lib.load()
    out = np.zeros((n, 4))
    ru = ctypes.c_double()
    got = lib.nsgpu_generate_chain(int(n), float(radius), int(seed), _ptr(out), ctypes.byref(ru))
    return out[:got].copy(), ru.value


class PatchSet:
    """SES patches built on the host (library-owned), with NumPy views."""

    def __init__(self, handle, probe):
        self._lib = _lib.load()
        self._h = handle
        self.probe = probe
        self.view = Patches()
        self._lib.nsgpu_patch_set_view(self._h, ctypes.byref(self.view))
        n, nd = self.view.n, self.view.n_data
        self.kind = np.ctypeslib.as_array(self.view.kind, (n,)) if n else np.zeros(0, np.int32)
        self.off = np.ctypeslib.as_array(self.view.off, (n,)) if n else np.zeros(0, np.int32)
        self.cnt = np.ctypeslib.as_array(self.view.cnt, (n,)) if n else np.zeros(0, np.int32)
        self.flags = np.ctypeslib.as_array(self.view.flags, (n,)) if n else np.zeros(0, np.int32)
        self.atom = np.ctypeslib.as_array(self.view.atom, (n, 3)) if n else np.zeros((0, 3), np.int32)
        self.bsph = np.ctypeslib.as_array(self.view.bsph, (n, 4)) if n else np.zeros((0, 4))
        self.data = np.ctypeslib.as_array(self.view.data, (nd,)) if nd else np.zeros(0)

    def counts(self):
        c = np.zeros(5, np.int64)
        self._lib.nsgpu_patch_set_counts(self._h, _ptr(c))
        return dict(zip(PATCH_KINDS, c.tolist()))

    def __len__(self):
        return self.view.n

    def close(self):
        if self._h:
            self._lib.nsgpu_patch_set_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ses_patches(xyzr, probe=1.4):
    """Build the SES patch set of (already jittered) atoms."""
    lib = _lib.load()
    a = as_xyzr(xyzr)
    h = ctypes.c_void_p()
    rc = lib.nsgpu_ses_build(_ptr(a), a.shape[0], float(probe), ctypes.byref(h))
    if rc != 0:
        if h:
            lib.nsgpu_patch_set_free(h)
        raise NsgpuError("ses_build failed (%d)" % rc)
    return PatchSet(h, float(probe))


SKIN_KINDS = ("skin_point", "skin_edge", "skin_facet", "skin_tetra")


def regular_triangulation(wpts):
    """Finite tetrahedra of the regular (weighted Delaunay) triangulation of
    weighted points (x, y, z, w) and their neighbours (-1 on the hull): the
    lower convex hull of the points lifted to (x, y, z, |p|^2 - w), by Qhull.
    Coordinates are centred first (the triangulation does not move)."""
    from scipy.spatial import ConvexHull
    p = np.asarray(wpts, dtype=np.float64).reshape(-1, 4)
    xyz = p[:, :3] - p[:, :3].mean(axis=0)
    lifted = np.column_stack([xyz, (xyz * xyz).sum(axis=1) - p[:, 3]])
    hull = ConvexHull(lifted)
    lower = hull.equations[:, 3] < 0
    idx = np.full(len(lower), -1, np.int64)
    idx[lower] = np.arange(int(lower.sum()))
    tets = np.ascontiguousarray(hull.simplices[lower], dtype=np.int32)
    nb = hull.neighbors[lower]
    nbrs = np.ascontiguousarray(np.where(nb >= 0, idx[np.maximum(nb, 0)], -1), dtype=np.int32)
    return tets, nbrs


def _read_tri(lib, h):
    m = lib.nsgpu_regular_tri_count(h)
    tets = np.zeros((m, 4), np.int32)
    nbrs = np.zeros((m, 4), np.int32)
    lib.nsgpu_regular_tri_read(h, _ptr(tets), _ptr(nbrs))
    lib.nsgpu_regular_tri_free(h)
    return tets, nbrs


def skin_triangulation(xyzr, s=0.45):
    """The library's regular triangulation of the skin's weighted points
    (nsgpu_skin_triangulate): (tets [m, 4], nbrs [m, 4]) as nsgpu_skin_build
    takes them."""
    lib = _lib.load()
    a = as_xyzr(xyzr)
    h = ctypes.c_void_p()
    rc = lib.nsgpu_skin_triangulate(_ptr(a), a.shape[0], float(s), ctypes.byref(h))
    if rc != 0:
        raise NsgpuError("skin_triangulate failed (%d)" % rc)
    return _read_tri(lib, h)


def weighted_triangulation(wpts):
    """The library's regular triangulation of weighted points (x, y, z, w)
    (nsgpu_regular_triangulate); indices n .. n+5 in the tetrahedra are the
    six dummies that close the hull."""
    lib = _lib.load()
    p = np.ascontiguousarray(wpts, dtype=np.float64).reshape(-1, 4)
    h = ctypes.c_void_p()
    rc = lib.nsgpu_regular_triangulate(_ptr(p), p.shape[0], ctypes.byref(h))
    if rc != 0:
        raise NsgpuError("regular_triangulate failed (%d)" % rc)
    return _read_tri(lib, h)


def skin_patches(xyzr, s=0.45, triangulation=None):
    """Skin mixed cells of (already jittered) atoms with shrink factor s
    (Skin_Surface_Parameter): the weighted points and dummies of the
    reference, their regular triangulation, then the cells.  triangulation:
    None for the library's own (nsgpu_skin_patches), or (tets, nbrs) from
    elsewhere (e.g. regular_triangulation's Qhull lifting)."""
    lib = _lib.load()
    a = as_xyzr(xyzr)
    n = a.shape[0]
    h = ctypes.c_void_p()
    if triangulation is None:
        rc = lib.nsgpu_skin_patches(_ptr(a), n, float(s), ctypes.byref(h))
    else:
        tets = np.ascontiguousarray(triangulation[0], dtype=np.int32)
        nbrs = np.ascontiguousarray(triangulation[1], dtype=np.int32)
        rc = lib.nsgpu_skin_build(_ptr(a), n, float(s), _ptr(tets), _ptr(nbrs), tets.shape[0], ctypes.byref(h))
    if rc != 0:
        if h:
            lib.nsgpu_patch_set_free(h)
        raise NsgpuError("skin patch construction failed (%d)" % rc)
    return PatchSet(h, 1.4)


def skin_points(xyzr, s=0.45):
    """The (n + 6) weighted points of the skin triangulation (nsgpu_skin_points)."""
    lib = _lib.load()
    a = as_xyzr(xyzr)
    w = np.zeros(((a.shape[0] + 6), 4))
    if lib.nsgpu_skin_points(_ptr(a), a.shape[0], float(s), _ptr(w)) != 0:
        raise NsgpuError("skin_points failed")
    return w


def surf_opts(build_epsmap=True, build_status=True, accurate_triangulation=True, fill_cavities=True,
              cav_vol=11.4, inside=5, precision=0):
    return SurfOpts(int(build_epsmap), int(build_status), int(accurate_triangulation), int(fill_cavities),
                    float(cav_vol), int(inside), int(precision))


def _result_dict(r):
    return {f: getattr(r, f) for f, _ in SurfResult._fields_}


class DeviceSurface:
    """One device context: grid + patches -> maps, cavities and mesh on the GPU."""

    def __init__(self, device=0):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        rc = self._lib.nsgpu_create(int(device), ctypes.byref(h))
        if rc != 0:
            raise NsgpuError("nsgpu_create failed (%d): no usable HIP device %d" % (rc, device))
        self._h = h
        self.grid = None
        self.result = None
        self.mesh_info = None

    @classmethod
    def _view(cls, handle, grid):
        """A non-owning view of a context owned elsewhere (a batch frame)."""
        d = cls.__new__(cls)
        d._lib = _lib.load()
        d._h = ctypes.c_void_p(handle)
        d._owned = False
        d.grid = grid
        d.result = None
        d.mesh_info = None
        return d

    def _chk(self, rc, what):
        _check(rc, self._h, what)

    def set_grid(self, g):
        self._chk(self._lib.nsgpu_set_grid(self._h, ctypes.byref(g)), "set_grid")
        self.grid = g

    def set_patches(self, ps):
        self._chk(self._lib.nsgpu_set_patches(self._h, ctypes.byref(ps.view), ps.probe), "set_patches")

    def set_mesh(self, verts, tris, ray_based=True, aux_grid_dim=100):
        """A triangle mesh as the surface (MeshSurface::load): the rays are cast
        against its triangles.  ray_based: Patch_Based_Algorithm = false (the
        auxiliary 2D grid of at most aux_grid_dim cells per side,
        Max_mesh_auxiliary_grid_2d_size)."""
        v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
        self._chk(self._lib.nsgpu_set_mesh(self._h, _ptr(v), v.shape[0], _ptr(t), t.shape[0], int(bool(ray_based)),
                                           int(aux_grid_dim)), "set_mesh")

    def blobby_build(self, xyzr, blobbyness=-2.5, cutoff=6.0, aux_grid_dim=100):
        """BlobbySurface::build on the device: field, marching cubes at iso 1,
        one smoothing pass and the OFF text round trip; the mesh becomes the
        surface (ray-based).  Returns the blobby mesh info (area before
        smoothing)."""
        a = np.ascontiguousarray(as_xyzr(xyzr), dtype=np.float64)
        m = MeshInfo()
        self._chk(self._lib.nsgpu_blobby_build(self._h, _ptr(a), a.shape[0], float(blobbyness), float(cutoff),
                                               int(aux_grid_dim), ctypes.byref(m)), "blobby_build")
        return {f: getattr(m, f) for f, _ in MeshInfo._fields_}

    def surface_mesh(self, raw=False):
        """(verts, tris) of the mesh surface; raw: the blobby build's smoothed
        vertices before the text round trip."""
        nv, nt = ctypes.c_int64(), ctypes.c_int64()
        self._chk(self._lib.nsgpu_surface_mesh_size(self._h, ctypes.byref(nv), ctypes.byref(nt)), "mesh_size")
        v = np.empty((nv.value, 3))
        t = np.empty((nt.value, 3), np.int32)
        self._chk(self._lib.nsgpu_read_surface_mesh(self._h, int(bool(raw)), _ptr(v), _ptr(t)), "read_surface_mesh")
        return v, t

    def blobby_field(self):
        g = self.grid
        out = np.empty((g.nz, g.ny, g.nx))
        self._chk(self._lib.nsgpu_read_blobby_field(self._h, _ptr(out)), "read_blobby_field")
        return out

    def set_atoms(self, xyzr):
        """Atoms (after jitter) the current patch set was built from (adjoint)."""
        a = np.ascontiguousarray(as_xyzr(xyzr), dtype=np.float64)
        self._chk(self._lib.nsgpu_set_atoms(self._h, _ptr(a), a.shape[0]), "set_atoms")
        self._natoms = a.shape[0]

    def surface_gradient(self):
        """Three-panel volume and area estimate with their gradients:
        {'volume', 'area', 'dvolume' (n, 4), 'darea' (n, 4)}, d/d(x, y, z, r)."""
        n = getattr(self, "_natoms", 0)
        v, s = ctypes.c_double(), ctypes.c_double()
        dv = np.zeros((max(n, 1), 4))
        da = np.zeros((max(n, 1), 4))
        r = SurfResult()
        self._chk(self._lib.nsgpu_surface_gradient(self._h, ctypes.byref(v), ctypes.byref(s), _ptr(dv), _ptr(da),
                                                   ctypes.byref(r)), "surface_gradient")
        self.result = _result_dict(r)
        return {"volume": v.value, "area": s.value, "dvolume": dv[:n], "darea": da[:n]}

    CAST_PATCHES, CAST_TILES = 0, 1

    def set_cast_mode(self, mode):
        """Intersection of get_surf: CAST_PATCHES (one workgroup per patch and
        panel) or CAST_TILES (ray tiles over a cell grid of the patches)."""
        self._chk(self._lib.nsgpu_set_cast_mode(self._h, int(mode)), "set_cast_mode")

    def cast_rays(self, direction, origins, coords, max_hits=32, normals=False):
        """castAxisOrientedRay for a batch: rays from origins (n, 3) along
        direction (0 x, 1 y, 2 z) to coordinate coords (n,).  Returns
        (counts (n,), hits (n, max_hits) axis coordinates ascending[, normals
        (n, max_hits, 3)]); counts -1: coord <= origin, -2: outside the grid."""
        o = np.ascontiguousarray(origins, dtype=np.float64).reshape(-1, 3)
        n = o.shape[0]
        c = np.ascontiguousarray(np.broadcast_to(np.asarray(coords, dtype=np.float64), (n,)))
        counts = np.zeros(n, np.int32)
        hits = np.zeros((n, max_hits))
        nrm = np.zeros((n, max_hits, 3)) if normals else None
        self._chk(self._lib.nsgpu_cast_rays(self._h, int(direction), n, _ptr(o), _ptr(c), int(max_hits), _ptr(counts),
                                            _ptr(hits), _ptr(nrm) if normals else None), "cast_rays")
        return (counts, hits, nrm) if normals else (counts, hits)

    def set_profiling(self, on=True):
        self._chk(self._lib.nsgpu_set_profiling(self._h, int(on)), "set_profiling")

    def stage_times(self):
        ms = np.zeros(len(STAGES))
        self._lib.nsgpu_stage_times(self._h, _ptr(ms), len(STAGES))
        return dict(zip(STAGES, ms.tolist()))

    def intersect_path_counts(self, **kw):
        """Counts of each FP64 path of the ray/patch candidate set (no hits written)."""
        o = surf_opts(**kw)
        c = np.zeros(8, np.int64)
        self._chk(self._lib.nsgpu_intersect_path_counts(self._h, ctypes.byref(o), _ptr(c), 8), "path_counts")
        return dict(zip(("candidates", "disc_pass", "sphere_hits", "torus_quartics", "roots_tested",
                         "feasible_hits", "feas_items", "feas_calls"), c.tolist()))

    def get_surf(self, **kw):
        o = surf_opts(**kw)
        r = SurfResult()
        self._chk(self._lib.nsgpu_get_surf(self._h, ctypes.byref(o), ctypes.byref(r)), "get_surf")
        self.result = _result_dict(r)
        return self.result

    def triangulate(self):
        m = MeshInfo()
        self._chk(self._lib.nsgpu_triangulate(self._h, ctypes.byref(m)), "triangulate")
        self.mesh_info = {f: getattr(m, f) for f, _ in MeshInfo._fields_}
        return self.mesh_info

    def synchronize(self):
        self._chk(self._lib.nsgpu_synchronize(self._h), "synchronize")

    def _shape(self):
        g = self.grid
        return g.nz, g.ny, g.nx

    def epsmap(self):
        out = np.empty(self._shape() + (3,), np.int32)
        self._chk(self._lib.nsgpu_read_epsmap(self._h, _ptr(out)), "read_epsmap")
        return out

    def idebmap(self):
        out = np.empty(self._shape(), np.uint8)
        self._chk(self._lib.nsgpu_read_idebmap(self._h, _ptr(out)), "read_idebmap")
        return out

    def status(self):
        out = np.empty(self._shape(), np.int32)
        self._chk(self._lib.nsgpu_read_status(self._h, _ptr(out)), "read_status")
        return out

    def corner_bits(self):
        nz, ny, nx = self._shape()
        out = np.empty((nx * ny * nz + 31) // 32, np.uint32)
        self._chk(self._lib.nsgpu_read_corner_bits(self._h, _ptr(out)), "read_corner_bits")
        return out

    def cavities(self):
        n = self.result["num_cavities"] if self.result else 0
        sizes = np.zeros(max(n, 1))
        filled = np.zeros(max(n, 1), np.uint8)
        self._lib.nsgpu_read_cavities(self._h, _ptr(sizes), _ptr(filled))
        return sizes[:n], filled[:n].astype(bool)

    def mesh(self):
        m = self.mesh_info
        v = np.empty((m["num_vertices"], 3))
        t = np.empty((m["num_triangles"], 3), np.int32)
        self._chk(self._lib.nsgpu_read_mesh(self._h, _ptr(v), _ptr(t)), "read_mesh")
        return v, t

    def smooth_surface(self):
        """Smooth_Mesh: one smoothing pass over the last mesh (and its normals
        when vertex normals are on)."""
        self._chk(self._lib.nsgpu_smooth_surface(self._h), "smooth_surface")

    def project_bgp_device(self, aux_grid_dim=50):
        """The device stage of project_bgp alone (results stay on the device
        for read_bgp); returns the counts."""
        counts = np.zeros(4, np.int64)
        self._chk(self._lib.nsgpu_project_bgp(self._h, int(aux_grid_dim), _ptr(counts)), "project_bgp")
        return counts

    def project_bgp(self, aux_grid_dim=50):
        """Project_boundary_grid_points on the maps of the last get_surf (with
        the epsilon map; SES patches): {'ibgp' (n, 3) 0-based node indices in
        z, y, x order, 'scspos' (n, 3) projections on the surface, 'scsnor'
        (n, 3) normals (zero, as the reference's), 'counts' [points, points
        with no face outside, points placed by the singular-edge sampling,
        points left at their node]}."""
        counts = self.project_bgp_device(aux_grid_dim)
        n = int(counts[0])
        ib = np.empty((n, 3), np.int32)
        pos = np.empty((n, 3))
        nor = np.empty((n, 3))
        self._chk(self._lib.nsgpu_read_bgp(self._h, _ptr(ib), _ptr(pos), _ptr(nor)), "read_bgp")
        return {"ibgp": ib, "scspos": pos, "scsnor": nor, "counts": counts}

    def pockets(self, small, filter_probe=1.4, num_mol=2, cav_and_pockets=True):
        """The grid part of pocket mode with this surface as the fat-probe one
        and `small` the small-probe one (both after get_surf with the status
        map on the same grid; see nsgpu_pockets): this surface's status map
        becomes the pocket labels.  Returns {'volumes', 'filled', 'is_pocket'}
        per component (ids 4, 5, ...) and 'support': the flat node indices of
        each kept component's support points."""
        o = _lib.PocketOpts(float(filter_probe), float(num_mol), int(bool(cav_and_pockets)), 0)
        r = _lib.PocketResult()
        self._chk(self._lib.nsgpu_pockets(self._h, small._h, ctypes.byref(o), ctypes.byref(r)), "pockets")
        nc, nk, ns_ = r.num_components, r.num_kept, r.support_points
        vol = np.zeros(max(nc, 1))
        filled = np.zeros(max(nc, 1), np.uint8)
        pocket = np.zeros(max(nc, 1), np.uint8)
        off = np.zeros(nk + 1, np.int64)
        sup = np.zeros(max(ns_, 1), np.int64)
        self._chk(self._lib.nsgpu_read_pockets(self._h, _ptr(vol), _ptr(filled), _ptr(pocket), _ptr(off), _ptr(sup)),
                  "read_pockets")
        return {"volumes": vol[:nc], "filled": filled[:nc].astype(bool), "is_pocket": pocket[:nc].astype(bool),
                "support": [sup[off[q]:off[q + 1]] for q in range(nk)]}

    def pocket_labels(self):
        """(nz, ny, nx) component ids of the last pockets() before the second
        filter (-1 off the components)."""
        out = np.empty(self._shape(), np.int32)
        self._chk(self._lib.nsgpu_read_pocket_labels(self._h, _ptr(out)), "read_pocket_labels")
        return out

    def set_vertex_normals(self, on=True):
        """Compute_Vertex_Normals: keep what normals() needs from the next get_surf on."""
        self._chk(self._lib.nsgpu_set_vertex_normals(self._h, int(on)), "set_vertex_normals")

    def normals(self):
        """(num_vertices, 3) vertex normals of the last mesh (set_vertex_normals first)."""
        m = self.mesh_info
        out = np.empty((m["num_vertices"], 3))
        self._chk(self._lib.nsgpu_read_normals(self._h, _ptr(out)), "read_normals")
        return out

    def ray_hits(self, slot):
        g = self.grid
        N = max(g.nx, g.ny, g.nz)
        maxh = max(1, self.result["max_hits_per_ray"])
        counts = np.empty(N * N, np.int32)
        t = np.zeros((N * N, maxh))
        self._chk(self._lib.nsgpu_read_ray_hits(self._h, int(slot), maxh, _ptr(counts), _ptr(t)), "read_ray_hits")
        return counts, t

    def close(self):
        if getattr(self, "_h", None):
            if getattr(self, "_owned", True):
                self._lib.nsgpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class NanoShaper:
    """Mirror of NS::NanoShaper (nanoshaper.h:294-424).

    ``surface_type`` "ses" (parameter: Probe_Radius) or "skin" (parameter:
    Skin_Surface_Parameter), the two types the reference's constructor takes
    (nanoshaper.cpp:74-90); "blobby" (parameter: Blobbyness) is added here.
    ``build_analytical_surface`` runs the host stages (grid, jitter, patches or
    the blobby build), ``colour_grid`` runs Surface::getSurf on the device and
    returns the volume, ``triangulate`` returns the mesh area.
    """

    DEFAULT_PARAMETER = {"ses": 1.4, "skin": 0.45, "blobby": -2.5}

    def __init__(self, atoms, surface_type="ses", parameter=None, scale=2.0, perfil=80.0, device=0,
                 rand_displacement=1e-4, fill_cavities=True, cav_vol=11.4, accurate_triangulation=True, probe=None):
        if surface_type not in self.DEFAULT_PARAMETER:
            raise NsgpuError("Unknown surface type %r" % surface_type)
        if parameter is None:
            parameter = probe if (probe is not None and surface_type == "ses") else self.DEFAULT_PARAMETER[surface_type]
        self.surface_type = surface_type
        self.parameter = float(parameter)
        self.atoms = as_xyzr(atoms)
        probe = self.parameter if surface_type == "ses" else 1.4
        self.probe, self.scale, self.perfil = probe, scale, perfil
        self.rand_displacement = rand_displacement
        self.opts = dict(fill_cavities=fill_cavities, cav_vol=cav_vol,
                         accurate_triangulation=accurate_triangulation)
        self.device = device
        self.surface = None

    def build_analytical_surface(self):
        self.grid = build_grid(self.atoms, self.scale, self.perfil)
        self.surface = DeviceSurface(self.device)
        self.surface.set_grid(self.grid)
        if self.surface_type == "blobby":
            # BlobbySurface::build: no jitter, the atoms as given
            self.surface.blobby_build(self.atoms, self.parameter)
            return True
        self.jittered, _ = jitter(self.atoms, self.rand_displacement, 1234)
        if self.surface_type == "skin":
            self.patches = skin_patches(self.jittered, self.parameter)
        else:
            self.patches = ses_patches(self.jittered, self.probe)
        self.surface.set_patches(self.patches)
        return True

    def colour_grid(self):
        r = self.surface.get_surf(**self.opts)
        return r["volume"]

    def triangulate(self):
        return self.surface.triangulate()["area"]

    def get_grid_point_status(self, ix, iy, iz):
        """+1 outside, -1 inside (NS::NanoShaper::getGridPointStatus)."""
        st = self.surface.status()
        return -1 if st[iz, iy, ix] == STATUS_INSIDE else 1

    def get_grid_size(self):
        g = self.grid
        return g.nx, g.ny, g.nz

    def get_grid_point_coordinates(self, ix, iy, iz):
        """(x, y, z) of node (ix, iy, iz) (NS::NanoShaper::getGridPointCoordinates)."""
        g = self.grid
        return g.xmin + ix * g.side, g.ymin + iy * g.side, g.zmin + iz * g.side

    def get_nearest_grid_point(self, p):
        """(indices, saturated): the nearest node, clamped to the grid; saturated
        when p lies outside it (NS::NanoShaper::getNearestGridPoint)."""
        g = self.grid
        out = False
        idx = []
        for d, (lo, n) in enumerate(((g.xmin, g.nx), (g.ymin, g.ny), (g.zmin, g.nz))):
            i = int(np.floor((p[d] - lo) / g.side + 0.5))
            if i < 0 or i >= n:
                out = True
                i = min(max(i, 0), n - 1)
            idx.append(i)
        return tuple(idx), out

    def set_direction(self, direction):
        """NS::NanoShaper::setDirection: 0 x, 1 y, 2 z."""
        if direction not in (0, 1, 2) or self.surface is None:
            return False
        self.direction = direction
        return True

    def cast_axis_oriented_ray(self, pa, coord, direction, compute_normals=False):
        """NS::NanoShaper::castAxisOrientedRay: (status, [(coordinate, normal or
        None), ...]); status -1 for a wrong direction, a non-progressive ray or
        no intersection, +1 otherwise."""
        if getattr(self, "direction", None) != direction or pa[direction] >= coord or self.surface is None:
            return -1, []
        maxh = 64
        while True:
            res = self.surface.cast_rays(direction, [pa], coord, max_hits=maxh, normals=compute_normals)
            cnt = int(res[0][0])
            if cnt > maxh:
                maxh = cnt
                continue
            if cnt <= 0:
                return -1, []
            nrm = res[2][0] if compute_normals else [None] * cnt
            return 1, [(float(res[1][0, q]), nrm[q]) for q in range(cnt)]


class Batch:
    """Frames of uneven size (one grid and patch set each) computed by one
    captured graph launch; maps and cavities per frame (no triangulation)."""

    def __init__(self, n_frames, device=0):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        rc = self._lib.nsgpu_batch_create(int(device), int(n_frames), ctypes.byref(h))
        if rc != 0:
            raise NsgpuError("nsgpu_batch_create failed (%d)" % rc)
        self._h = h
        self.n = int(n_frames)
        self.grids = [None] * self.n
        self._keep = [None] * self.n

    def _chk(self, rc, what):
        if rc != 0:
            raise NsgpuError("%s failed (%d): %s" % (what, rc, self._lib.nsgpu_batch_last_error(self._h).decode()))

    def set_frame(self, f, grid, patchset):
        self._chk(self._lib.nsgpu_batch_set_frame(self._h, int(f), ctypes.byref(grid), ctypes.byref(patchset.view),
                                                  patchset.probe), "batch_set_frame")
        self.grids[f] = grid
        self._keep[f] = patchset

    CASTS = {"frames": 0, "joint": 1, "persistent": 2}

    def prepare(self, chains=0, cast=None, workgroups=0, chunk=0, **kw):
        """Size every frame and capture the graph.  cast None: per-frame
        casts on `chains` chains, or (chains <= 0) the autotuned shape;
        cast "frames", "joint" or "persistent" with chains > 0: that shape
        (workgroups: persistent workgroups per class, 0 the resident limit;
        chunk: items per queue take, 0 for 16).  Returns the chain count used."""
        kw.setdefault("accurate_triangulation", False)
        o = surf_opts(**kw)
        if cast is None:
            self._chk(self._lib.nsgpu_batch_prepare(self._h, ctypes.byref(o), int(chains)), "batch_prepare")
        else:
            sh = _lib.BatchShape(self.CASTS[cast], int(chains), int(workgroups), int(chunk))
            self._chk(self._lib.nsgpu_batch_prepare_shape(self._h, ctypes.byref(o), ctypes.byref(sh)),
                      "batch_prepare_shape")
        return int(self._lib.nsgpu_batch_chains(self._h))

    def shape(self):
        """Launch shape of the captured graph: {"cast", "chains", "workgroups", "chunk"}."""
        sh = _lib.BatchShape()
        self._chk(self._lib.nsgpu_batch_get_shape(self._h, ctypes.byref(sh)), "batch_get_shape")
        names = {v: k for k, v in self.CASTS.items()}
        return {"cast": names[sh.cast], "chains": sh.chains, "workgroups": sh.workgroups, "chunk": sh.chunk}

    def run(self):
        """One graph launch: (per-frame completion ms after launch, frames redone on the host path)."""
        lat = np.zeros(self.n)
        redone = ctypes.c_int32()
        self._chk(self._lib.nsgpu_batch_run(self._h, _ptr(lat), ctypes.byref(redone)), "batch_run")
        return lat, redone.value

    def run_loop(self):
        """The per-frame launch loop over the same frames: per-frame wall ms."""
        lat = np.zeros(self.n)
        self._chk(self._lib.nsgpu_batch_run_loop(self._h, _ptr(lat)), "batch_run_loop")
        return lat

    def result(self, f):
        r = SurfResult()
        self._chk(self._lib.nsgpu_batch_result(self._h, int(f), ctypes.byref(r)), "batch_result")
        return _result_dict(r)

    def frame(self, f):
        """DeviceSurface view of frame f (maps, status, cavities)."""
        h = self._lib.nsgpu_batch_frame(self._h, int(f))
        d = DeviceSurface._view(h, self.grids[f])
        r = SurfResult()
        if self._lib.nsgpu_batch_result(self._h, int(f), ctypes.byref(r)) == 0:
            d.result = _result_dict(r)  # None until the frame has been computed
        return d

    def close(self):
        if getattr(self, "_h", None):
            self._lib.nsgpu_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def perturbed_frames(n_frames, n_min=500, n_max=8000, sigma=0.3, seed=0):
    """Frames of uneven size: atom counts log-uniform in [n_min, n_max], each
    a seeded chain (confinement radius 1.65 n^(1/3) A) perturbed by per-atom
    Gaussian displacements of standard deviation sigma (A)."""
    rng = np.random.default_rng(seed)
    frames = []
    for f in range(n_frames):
        n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        xyzr, _ = generate_chain(n, 1.65 * n ** (1.0 / 3.0), seed=seed * 100003 + f)
        xyzr[:, :3] += rng.normal(0.0, sigma, size=(len(xyzr), 3))
        frames.append(xyzr)
    return frames
