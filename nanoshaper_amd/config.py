"""Configuration-file runs: the reference CLI's normal mode over the device path.

* ``ConfigFile``   "key = value" lines, '#' comments (ConfigFile/ConfigFile.h:59-269)
* ``Configuration`` the keys normal mode reads, with the reference's defaults
  and its refusals of inconsistent flag combinations (init, main.cpp:1032-1204)
* ``normal_mode``  DelPhiShared from the atom file and grid keys, the surface
  of key "Surface", build, getSurf, triangulateSurface (normalMode,
  main.cpp:1253-1320)
* ``write_outputs`` the files normal mode saves (main.cpp:1359-1394 and the
  mesh of triangulateSurface / smoothSurface, Surface.cpp:10478-10515)
* ``pocket_mode`` / ``run_pockets`` / ``write_pocket_outputs`` the CLI's
  pocket mode (pocketMode, main.cpp:1411-1873): the fat- and small-probe
  surfaces, the device's pocket detection (nsgpu_pockets) and the skin
  envelope of each pocket or cavity
"""
import math
import os

import numpy as np

from . import io as IO
from . import surface as S


class ConfigError(ValueError):
    pass


class ConfigFile:
    def __init__(self, path=None, text=None):
        self.kv = {}
        self.path = path
        if path is not None:
            with open(path) as f:
                text = f.read()
        for line in (text or "").splitlines():
            self.parse_line(line)

    def parse_line(self, line):
        line = line.split("#", 1)[0]
        if "=" not in line:
            return
        k, v = line.split("=", 1)
        k = k.strip()
        if k:
            self.kv[k] = v.strip()

    def read(self, key, default):
        if key not in self.kv:
            return default
        v = self.kv[key]
        if isinstance(default, bool):
            return v.lower() in ("true", "1", "yes")
        if isinstance(default, int):
            return int(v.split()[0])
        if isinstance(default, float):
            return float(v.split()[0])
        return v.split()[0] if v else v

    def set(self, key, value):
        self.kv[key] = str(value).lower() if isinstance(value, bool) else str(value)


class Configuration:
    """The normal-mode keys (main.cpp:1066-1193); raises ConfigError on the
    flag combinations the reference refuses (main.cpp:1090-1145)."""

    def __init__(self, cf):
        r = cf.read
        self.build_epsmaps = r("Build_epsilon_maps", False)
        self.save_epsmaps = r("Save_eps_maps", False)
        self.build_status = r("Build_status_map", False)
        self.save_status = r("Save_Status_map", False)
        self.save_idebmap = r("Save_ideb_map", False)
        self.fill_cavities = r("Cavity_Detection_Filling", False)
        self.proj_bgp = r("Project_boundary_grid_points", False)
        self.acc_tri = r("Accurate_Triangulation", False)
        self.tri = r("Triangulation", False)
        self.operative_mode = r("Operative_Mode", "normal")
        self.patch_based = r("Patch_Based_Algorithm", True)
        self.analytical_torus = r("Analytical_Ray_Vs_Torus_Intersection", True)
        self.max_num_atoms = r("Max_Num_Atoms", -1)
        self.domain_shrinkage = r("Domain_Shrinkage", 0.0)
        self.optimize_grids = r("Optimize_Grids", True)
        self.surface = r("Surface", "ses")
        self.cav_vol = r("Conditional_Volume_Filling_Value", 11.4)
        self.scale = r("Grid_scale", 2.0)
        self.perfil = r("Grid_perfil", 80.0)
        self.mol_file = r("XYZR_FileName", "temp.xyzr")
        self.smoothing = r("Smooth_Mesh", False)
        self.probe = r("Probe_Radius", 1.4)
        self.skin_s = r("Skin_Surface_Parameter", 0.45)
        self.blobbyness = r("Blobbyness", -2.5)
        self.normals = r("Compute_Vertex_Normals", False)
        self.aux_grid_2d = r("Max_mesh_auxiliary_grid_2d_size", 100)
        self.sys_name = r("Sys_Name", "mol")
        self.root_file = r("Root_FileName", "")
        self.save_cavities = r("Save_Cavities", False)
        self.save_msms = r("Save_Mesh_MSMS_Format", False)
        self.save_ply = r("Save_Mesh_PLY_Format", False)
        self.vertex_atom_info = r("Vertex_Atom_Info", False)
        self.save_bgps = r("Save_bgps", False)
        # pocket mode (main.cpp:1166, :1187-1194)
        self.num_mol = r("Num_Wat_Pocket", 2)
        self.cav_and_pockets = r("Pockets_And_Cavities", True)
        self.link_pockets = r("Link_Pockets", False)
        self.pocket_big = r("Pocket_Radius_Big", 3.0)
        self.pocket_small = r("Pocket_Radius_Small", 1.4)
        self.aux_grid_3d = r("Max_ses_patches_auxiliary_grid_size", 50)
        if not self.build_epsmaps and self.save_epsmaps:
            raise ConfigError("Asked to save epsmap without builiding it: set Build_epsilon_maps = true")
        if not self.build_epsmaps and self.proj_bgp:
            raise ConfigError("Cannot project boundary grid points without an epsilon map")
        if not self.acc_tri and not self.build_status and self.tri:
            raise ConfigError("If non analytical triangulation is enabled status map is needed")
        if self.fill_cavities and not self.build_status:
            raise ConfigError("If cavity detection is enabled status map is needed")
        if self.save_idebmap and not self.build_epsmaps:
            raise ConfigError("Idebmap is computed only if epsilon map is enabled")
        if self.patch_based and self.surface not in ("ses", "skin", "mesh"):
            raise ConfigError("This surface type does not support patch based ray-tracing: "
                              "set Patch_Based_Algorithm = false")
        if self.operative_mode not in ("normal", "pocket"):
            raise ConfigError("operative mode %r is not built (normal and pocket only)" % self.operative_mode)
        if self.operative_mode == "pocket" and self.link_pockets:
            raise ConfigError("Link_Pockets is not built")
        if self.surface not in ("ses", "skin", "blobby"):
            raise ConfigError("Unknown surface type %r" % self.surface)

    def surf_opts(self):
        prec = 0 if self.analytical_torus else 2  # NSGPU_PREC_FP64 / NSGPU_PREC_STURM
        return dict(build_epsmap=self.build_epsmaps, build_status=self.build_status,
                    accurate_triangulation=self.acc_tri, fill_cavities=self.fill_cavities, cav_vol=self.cav_vol,
                    inside=5, precision=prec)


def normal_mode(prm_path, device=0):
    """Run a configuration file's normal mode on the device.  The atom file is
    taken relative to the file's directory.  Returns a dict with the
    configuration, atoms (as loaded), jittered atoms, grid, patch set, the
    DeviceSurface, the getSurf result and (with Triangulation) the mesh info."""
    cf = ConfigFile(prm_path)
    conf = Configuration(cf)
    mol = conf.mol_file
    if not os.path.isabs(mol):
        mol = os.path.join(os.path.dirname(os.path.abspath(prm_path)), mol)
    out = run(conf, S.load_atoms(mol, conf.max_num_atoms, conf.domain_shrinkage), device)
    out["config"] = conf
    return out


def build_surface(conf, atoms):
    """The host stages: grid (DelPhiShared::buildGrid), jitter and patches
    (ses, skin) as Surface::build does; returns (grid, jittered atoms,
    patch set or None for blobby)."""
    grid = S.build_grid(atoms, conf.scale, conf.perfil)
    if conf.surface == "blobby":
        return grid, atoms, None
    jit, _ = S.jitter(atoms)
    ps = S.skin_patches(jit, conf.skin_s) if conf.surface == "skin" else S.ses_patches(jit, conf.probe)
    return grid, jit, ps


def run(conf, atoms, device=0):
    grid, jit, ps = build_surface(conf, atoms)
    dev = S.DeviceSurface(device)
    dev.set_grid(grid)
    if ps is None:
        dev.blobby_build(atoms, conf.blobbyness, aux_grid_dim=conf.aux_grid_2d)
    else:
        dev.set_patches(ps)
    dev.set_vertex_normals(conf.normals and conf.acc_tri and conf.tri)
    res = dev.get_surf(**conf.surf_opts())
    # phase 2 of getSurf: boundary grid points (Surface.cpp:1476-1588)
    bgp = dev.project_bgp(conf.aux_grid_3d) if conf.proj_bgp else None
    mesh = dev.triangulate() if (conf.tri and conf.acc_tri) else None
    return {"atoms": atoms, "jittered": jit, "grid": grid, "patches": ps, "device": dev, "result": res,
            "mesh": mesh, "bgp": bgp}


def write_outputs(out, directory="."):
    """The files of a normal-mode run (out from normal_mode or run): the mesh
    (triangulatedSurf in the format of Surface::deduceFormat, smoothed when
    Smooth_Mesh, with triangleAreas.txt), then the maps and cavity files the
    Save_* keys ask for, each named root + Sys_Name as the reference names
    them.  Returns the paths written."""
    conf = out["config"]
    dev, grid = out["device"], out["grid"]
    root = os.path.join(directory, conf.root_file)
    name = conf.sys_name
    paths = []
    if out.get("mesh") is not None:
        if conf.smoothing:
            dev.smooth_surface()
        v, t = dev.mesh()
        normals = dev.normals() if conf.normals else None
        va = IO.vertex_atoms(grid, out["jittered"], v) if conf.vertex_atom_info else None
        fmt = IO.deduce_format(conf.save_ply, conf.save_msms, va is not None, normals is not None)
        paths += IO.save_mesh(root, v, t, normals=normals, vertex_atoms=va, fmt=fmt)
    if conf.save_epsmaps:
        paths += IO.save_epsmaps(root + name, dev.epsmap())
    if conf.save_status or conf.save_cavities:
        st = dev.status()
        sizes, filled = dev.cavities()
        if conf.save_status:
            paths.append(root + name + ".status.txt")
            IO.save_status(paths[-1], st)
        IO.save_cavities(root, grid, st, sizes, filled)
        paths += [root + "cavities.txt", root + "cavitiesSize.txt", root + "cavAtomsSerials.txt"]
    if conf.save_bgps and out.get("bgp") is not None:
        b = out["bgp"]
        paths += IO.save_bgp(root, name, b["ibgp"], b["scspos"], b["scsnor"])
    if conf.save_idebmap:
        paths.append(root + name + ".idebmap.txt")
        IO.save_idebmap(paths[-1], dev.idebmap())
    return paths


# ------------------------------------------------------------------ pocket mode
def pocket_mode(prm_path, device=0):
    """Run a configuration file's pocket mode on the device (the atom file
    relative to the file's directory); see run_pockets."""
    cf = ConfigFile(prm_path)
    conf = Configuration(cf)
    mol = conf.mol_file
    if not os.path.isabs(mol):
        mol = os.path.join(os.path.dirname(os.path.abspath(prm_path)), mol)
    out = run_pockets(conf, S.load_atoms(mol, conf.max_num_atoms, conf.domain_shrinkage), device)
    out["config"] = conf
    return out


def _envelope(conf, pts, seed, device):
    """The skin envelope of one pocket or cavity (pocketWrapper::
    buildPocketCavityEnvelope, main.cpp:398-430): its support points as
    balls of 1.4 A read back from their "%f" text (cavitiesToAtoms), at least
    four (the last repeated), grid at perfil 50, jitter 0.1 A continuing the
    run's random sequence, getSurf with the edge-ray panels only (no maps),
    the volume and, with Triangulation, the mesh and its area."""
    a = np.array([[float("%f" % v) for v in p] + [1.4] for p in pts])
    if len(a) < 4:
        a = np.vstack([a, np.repeat(a[-1:], 4 - len(a), axis=0)])
    g = S.build_grid(a, conf.scale, 50.0)
    jit, seed = S.jitter(a, amplitude=0.1, seed=seed)
    ps = S.skin_patches(jit, conf.skin_s)
    dev = S.DeviceSurface(device)
    dev.set_grid(g)
    dev.set_patches(ps)
    r = dev.get_surf(build_epsmap=False, build_status=False, accurate_triangulation=True, fill_cavities=False)
    mesh = dev.mesh() if conf.tri and dev.triangulate() else None
    area = dev.mesh_info["area"] if conf.tri else 0.0
    dev.close()
    return r["volume"], area, mesh, seed


def run_pockets(conf, atoms, device=0, envelopes=True):
    """pocketMode (main.cpp:1411-1660): the fat-probe surface (probe
    Pocket_Radius_Big, cavities filled) and the small-probe one (probe
    Pocket_Radius_Small; its volume, and its area with Triangulation), both
    jittered from one random sequence as the reference's builds draw from
    one generator, the device's pocket detection, then the skin envelope of
    every kept pocket or cavity in turn (the reference runs them on parallel
    threads that share that generator).  Returns the grid, the two device
    surfaces, the pockets (nsgpu_pockets), the components' labels before the
    second filter, and per kept component its envelope volume and area."""
    grid = S.build_grid(atoms, conf.scale, conf.perfil)
    jit1, seed = S.jitter(atoms)
    ps1 = S.ses_patches(jit1, conf.pocket_big)
    jit2, seed = S.jitter(atoms, seed=seed)
    ps2 = S.ses_patches(jit2, conf.pocket_small)
    fat = S.DeviceSurface(device)
    fat.set_grid(grid)
    fat.set_patches(ps1)
    fat.get_surf(build_epsmap=False, build_status=True, accurate_triangulation=False, fill_cavities=True,
                 cav_vol=math.inf)
    small = S.DeviceSurface(device)
    small.set_grid(grid)
    small.set_patches(ps2)
    # the small surface as the reference computes it (cavities detected and
    # filled only when pockets are wanted alone, main.cpp:1494), triangulated
    # and smoothed with Triangulation
    r2 = small.get_surf(build_epsmap=False, build_status=True, accurate_triangulation=conf.tri,
                        fill_cavities=not conf.cav_and_pockets, cav_vol=math.inf)
    area2 = 0.0
    if conf.tri:
        area2 = small.triangulate()["area"]
        small.smooth_surface()
    if conf.cav_and_pockets:
        # its outside component, for markPockets (getCavities, main.cpp:1629-1646)
        small.get_surf(build_epsmap=False, build_status=True, accurate_triangulation=False, fill_cavities=True,
                       cav_vol=0.0)
    pk = fat.pockets(small, filter_probe=1.4, num_mol=conf.num_mol, cav_and_pockets=conf.cav_and_pockets)
    labels = fat.pocket_labels()
    x, y, z = IO._grid_axes(grid)
    nz, ny, nx = labels.shape
    kept = np.nonzero(~pk["filled"])[0]
    vols, areas, meshes = [], [], []
    if envelopes:
        for q, c in enumerate(kept):
            f = pk["support"][q]
            pts = np.stack([x[f % nx], y[(f // nx) % ny], z[f // (nx * ny)]], 1)
            v, a, m, seed = _envelope(conf, pts, seed, device)
            vols.append(v)
            areas.append(a)
            meshes.append(m)
    return {"atoms": atoms, "jittered": jit1, "grid": grid, "fat": fat, "small": small, "pockets": pk,
            "labels": labels, "kept": kept, "small_volume": r2["volume"], "small_area": area2,
            "envelope_volumes": vols, "envelope_areas": areas, "envelope_meshes": meshes}


def write_pocket_outputs(out, directory="."):
    """The files of a pocket-mode run without atom information
    (main.cpp:1590-1660, 1752-1800): cavities.txt, cavitiesSize.txt and
    cavAtomsSerials.txt of the components after the volume threshold (only
    the kept ones' atoms, DelPhiShared::saveCavities(true)), numcav.txt, and
    with Triangulation each envelope's mesh as cav_tri<i>.off.  The
    cav<i>.txt / all_cav<i>.txt inputs of the envelopes are deleted after use
    by the reference and are not written.  Returns (paths, summary lines)."""
    conf = out["config"]
    grid, pk, labels = out["grid"], out["pockets"], out["labels"]
    root = os.path.join(directory, conf.root_file)
    n = len(pk["volumes"])
    by_volume = pk["volumes"] <= 11.4 * conf.num_mol
    cav_atoms = IO.cavity_atoms(grid, out["jittered"], labels, n)
    IO.save_cavities(root, grid, labels, pk["volumes"], by_volume, cav_atoms=cav_atoms, only_non_filled=True)
    paths = [root + "cavities.txt", root + "cavitiesSize.txt", root + "cavAtomsSerials.txt"]
    nc = len(out["kept"])
    with open(root + "numcav.txt", "w") as f:
        f.write("%d" % nc)
    paths.append(root + "numcav.txt")
    lines = ["Detected a total of %d pockets/cavities having at least the volume of %d water molecules" %
             (nc, conf.num_mol)]
    for q, c in enumerate(out["kept"]):
        v = out["envelope_volumes"][q] if out["envelope_volumes"] else 0.0
        kind = "Pocket" if (pk["is_pocket"][c] or not conf.cav_and_pockets) else "Cavity"
        if conf.tri and out["envelope_meshes"] and out["envelope_meshes"][q] is not None:
            a = out["envelope_areas"][q]
            lines.append("%s %d vol %s area %s" % (kind, q, repr(v), repr(a)))
            mv, mt = out["envelope_meshes"][q]
            p = root + "cav_tri%d.off" % q
            IO.save_off(p, mv, mt)
            paths.append(p)
        else:
            lines.append("%s %d vol %s" % (kind, q, repr(v)))
    return paths, lines
