"""Configuration-file runs: the reference CLI's normal mode over the device path.

* ``ConfigFile``   "key = value" lines, '#' comments (ConfigFile/ConfigFile.h:59-269)
* ``Configuration`` the keys normal mode reads, with the reference's defaults
  and its refusals of inconsistent flag combinations (init, main.cpp:1032-1204)
* ``normal_mode``  DelPhiShared from the atom file and grid keys, the surface
  of key "Surface", build, getSurf, triangulateSurface (normalMode,
  main.cpp:1253-1320)
* ``write_outputs`` the files normal mode saves (main.cpp:1359-1394 and the
  mesh of triangulateSurface / smoothSurface, Surface.cpp:10478-10515)
"""
import os

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
        if self.operative_mode != "normal":
            raise ConfigError("operative mode %r is not built (normal only)" % self.operative_mode)
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
