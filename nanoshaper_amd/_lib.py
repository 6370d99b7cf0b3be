"""ctypes binding of libnsgpu.so (the C ABI declared in include/nsgpu.h).

The library is loaded from the package directory only; a missing library is
an error, never a fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# NSGPU_LIBRARY selects another build of the same library (e.g. an
# instrumented one); there is no fallback when the file is missing.
LIB_PATH = os.environ.get("NSGPU_LIBRARY") or os.path.join(HERE, "libnsgpu.so")

c_int32_p = ctypes.POINTER(ctypes.c_int32)
c_double_p = ctypes.POINTER(ctypes.c_double)


class Grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32), ("pad_", ctypes.c_int32),
                ("xmin", ctypes.c_double), ("ymin", ctypes.c_double), ("zmin", ctypes.c_double),
                ("xmax", ctypes.c_double), ("ymax", ctypes.c_double), ("zmax", ctypes.c_double),
                ("scale", ctypes.c_double), ("side", ctypes.c_double), ("hside", ctypes.c_double),
                ("A", ctypes.c_double), ("rmaxdim", ctypes.c_double), ("baricenter", ctypes.c_double * 3)]


class Patches(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("n_data", ctypes.c_int32),
                ("kind", c_int32_p), ("off", c_int32_p), ("cnt", c_int32_p), ("flags", c_int32_p),
                ("atom", c_int32_p), ("bsph", c_double_p), ("data", c_double_p)]


class SurfOpts(ctypes.Structure):
    _fields_ = [("build_epsmap", ctypes.c_int32), ("build_status", ctypes.c_int32),
                ("accurate_triangulation", ctypes.c_int32), ("fill_cavities", ctypes.c_int32),
                ("cav_vol", ctypes.c_double), ("inside", ctypes.c_int32), ("precision", ctypes.c_int32)]


class SurfResult(ctypes.Structure):
    _fields_ = [("volume", ctypes.c_double), ("total_rays", ctypes.c_int64), ("odd_rays", ctypes.c_int64),
                ("total_hits", ctypes.c_int64), ("candidates", ctypes.c_int64), ("num_cavities", ctypes.c_int32),
                ("max_hits_per_ray", ctypes.c_int32)]


class MeshInfo(ctypes.Structure):
    _fields_ = [("num_vertices", ctypes.c_int64), ("num_triangles", ctypes.c_int64),
                ("num_analytic_vertices", ctypes.c_int64), ("area", ctypes.c_double)]


VP = ctypes.c_void_p

# name -> (restype, argtypes); every symbol declared in include/nsgpu.h
class PocketOpts(ctypes.Structure):
    _fields_ = [("filter_probe", ctypes.c_double), ("num_mol", ctypes.c_double),
                ("cav_and_pockets", ctypes.c_int32), ("pad_", ctypes.c_int32)]


class PocketResult(ctypes.Structure):
    _fields_ = [("num_components", ctypes.c_int32), ("num_kept", ctypes.c_int32),
                ("support_points", ctypes.c_int64)]


class BatchShape(ctypes.Structure):
    _fields_ = [("cast", ctypes.c_int32), ("chains", ctypes.c_int32), ("workgroups", ctypes.c_int32),
                ("chunk", ctypes.c_int32)]


SIGNATURES = {
    "nsgpu_abi_version": (ctypes.c_int, []),
    "nsgpu_device_count": (ctypes.c_int, []),
    "nsgpu_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(VP)]),
    "nsgpu_destroy": (None, [VP]),
    "nsgpu_last_error": (ctypes.c_char_p, [VP]),
    "nsgpu_load_atoms": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_double, ctypes.POINTER(VP)]),
    "nsgpu_atom_set_view": (ctypes.c_int32, [VP, ctypes.POINTER(c_double_p), ctypes.POINTER(c_double_p)]),
    "nsgpu_atom_set_error": (ctypes.c_char_p, [VP]),
    "nsgpu_atom_set_free": (None, [VP]),
    "nsgpu_build_grid": (ctypes.c_int, [VP, ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.POINTER(Grid)]),
    "nsgpu_nearest_atoms": (ctypes.c_int, [ctypes.POINTER(Grid), VP, ctypes.c_int64, VP, ctypes.c_int64, VP]),
    "nsgpu_jitter": (None, [VP, ctypes.c_int32, ctypes.c_double, c_int32_p]),
    "nsgpu_ses_build": (ctypes.c_int, [VP, ctypes.c_int32, ctypes.c_double, ctypes.POINTER(VP)]),
    "nsgpu_skin_points": (ctypes.c_int, [VP, ctypes.c_int32, ctypes.c_double, VP]),
    "nsgpu_skin_build": (ctypes.c_int, [VP, ctypes.c_int32, ctypes.c_double, VP, VP, ctypes.c_int32,
                                        ctypes.POINTER(VP)]),
    "nsgpu_patch_set_view": (None, [VP, ctypes.POINTER(Patches)]),
    "nsgpu_patch_set_free": (None, [VP]),
    "nsgpu_patch_set_counts": (None, [VP, VP]),
    "nsgpu_set_grid": (ctypes.c_int, [VP, ctypes.POINTER(Grid)]),
    "nsgpu_set_patches": (ctypes.c_int, [VP, ctypes.POINTER(Patches), ctypes.c_double]),
    "nsgpu_get_surf": (ctypes.c_int, [VP, ctypes.POINTER(SurfOpts), ctypes.POINTER(SurfResult)]),
    "nsgpu_get_surf_noalloc": (ctypes.c_int, [VP, ctypes.POINTER(SurfOpts), ctypes.POINTER(SurfResult), VP, VP, VP]),
    "nsgpu_triangulate": (ctypes.c_int, [VP, ctypes.POINTER(MeshInfo)]),
    "nsgpu_synchronize": (ctypes.c_int, [VP]),
    "nsgpu_set_cast_mode": (ctypes.c_int, [VP, ctypes.c_int]),
    "nsgpu_cast_rays": (ctypes.c_int, [VP, ctypes.c_int32, ctypes.c_int32, VP, VP, ctypes.c_int32, VP, VP, VP]),
    "nsgpu_set_profiling": (ctypes.c_int, [VP, ctypes.c_int]),
    "nsgpu_stage_times": (ctypes.c_int, [VP, VP, ctypes.c_int]),
    "nsgpu_intersect_path_counts": (ctypes.c_int, [VP, ctypes.POINTER(SurfOpts), VP, ctypes.c_int]),
    "nsgpu_read_epsmap": (ctypes.c_int, [VP, VP]),
    "nsgpu_read_idebmap": (ctypes.c_int, [VP, VP]),
    "nsgpu_read_status": (ctypes.c_int, [VP, VP]),
    "nsgpu_read_corner_bits": (ctypes.c_int, [VP, VP]),
    "nsgpu_read_cavities": (ctypes.c_int, [VP, VP, VP]),
    "nsgpu_read_mesh": (ctypes.c_int, [VP, VP, VP]),
    "nsgpu_read_normals": (ctypes.c_int, [VP, VP]),
    "nsgpu_smooth_surface": (ctypes.c_int, [VP]),
    "nsgpu_project_bgp": (ctypes.c_int, [VP, ctypes.c_int32, VP]),
    "nsgpu_read_bgp": (ctypes.c_int, [VP, VP, VP, VP]),
    "nsgpu_pockets": (ctypes.c_int, [VP, VP, ctypes.POINTER(PocketOpts), ctypes.POINTER(PocketResult)]),
    "nsgpu_read_pockets": (ctypes.c_int, [VP, VP, VP, VP, VP, VP]),
    "nsgpu_read_pocket_labels": (ctypes.c_int, [VP, VP]),
    "nsgpu_set_vertex_normals": (ctypes.c_int, [VP, ctypes.c_int]),
    "nsgpu_read_ray_hits": (ctypes.c_int, [VP, ctypes.c_int32, ctypes.c_int32, VP, VP]),
    "nsgpu_set_mesh": (ctypes.c_int, [VP, VP, ctypes.c_int32, VP, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "nsgpu_blobby_build": (ctypes.c_int, [VP, VP, ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                          ctypes.POINTER(MeshInfo)]),
    "nsgpu_surface_mesh_size": (ctypes.c_int, [VP, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "nsgpu_read_surface_mesh": (ctypes.c_int, [VP, ctypes.c_int32, VP, VP]),
    "nsgpu_read_blobby_field": (ctypes.c_int, [VP, VP]),
    "nsgpu_set_atoms": (ctypes.c_int, [VP, VP, ctypes.c_int32]),
    "nsgpu_surface_gradient": (ctypes.c_int, [VP, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                              VP, VP, ctypes.POINTER(SurfResult)]),
    "nsgpu_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int32, ctypes.POINTER(VP)]),
    "nsgpu_batch_destroy": (None, [VP]),
    "nsgpu_batch_last_error": (ctypes.c_char_p, [VP]),
    "nsgpu_batch_frame": (VP, [VP, ctypes.c_int32]),
    "nsgpu_batch_set_frame": (ctypes.c_int, [VP, ctypes.c_int32, ctypes.POINTER(Grid), ctypes.POINTER(Patches),
                                             ctypes.c_double]),
    "nsgpu_batch_prepare": (ctypes.c_int, [VP, ctypes.POINTER(SurfOpts), ctypes.c_int32]),
    "nsgpu_batch_chains": (ctypes.c_int32, [VP]),
    "nsgpu_skin_triangulate": (ctypes.c_int, [VP, ctypes.c_int32, ctypes.c_double, ctypes.POINTER(VP)]),
    "nsgpu_regular_triangulate": (ctypes.c_int, [VP, ctypes.c_int32, ctypes.POINTER(VP)]),
    "nsgpu_regular_tri_count": (ctypes.c_int32, [VP]),
    "nsgpu_regular_tri_read": (None, [VP, VP, VP]),
    "nsgpu_regular_tri_free": (None, [VP]),
    "nsgpu_skin_patches": (ctypes.c_int, [VP, ctypes.c_int32, ctypes.c_double, ctypes.POINTER(VP)]),
    "nsgpu_orient3d": (ctypes.c_int, [VP, VP, VP, VP]),
    "nsgpu_power_test": (ctypes.c_int, [VP, VP, VP, VP, VP]),
    "nsgpu_batch_prepare_shape": (ctypes.c_int, [VP, ctypes.POINTER(SurfOpts), ctypes.POINTER(BatchShape)]),
    "nsgpu_batch_get_shape": (ctypes.c_int, [VP, ctypes.POINTER(BatchShape)]),
    "nsgpu_batch_run": (ctypes.c_int, [VP, VP, VP]),
    "nsgpu_batch_run_loop": (ctypes.c_int, [VP, VP]),
    "nsgpu_batch_result": (ctypes.c_int, [VP, ctypes.c_int32, ctypes.POINTER(SurfResult)]),
    "nsgpu_generate_chain": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_double, ctypes.c_uint64, VP,
                                              ctypes.POINTER(ctypes.c_double)]),
}

_lib = None


def hip_runtime():
    """Paths of the HIP runtime libraries mapped into this process."""
    try:
        with open("/proc/self/maps") as f:
            return sorted({l.split()[-1] for l in f if "libamdhip64" in l})
    except OSError:
        return []


def load():
    """Load libnsgpu.so from the package directory (raises if it is absent).

    The library is built against ROCm's libamdhip64.so.7.  Torch's wheel
    bundles another libamdhip64 with the same soname; a process that loads
    torch's first serves this library with it (the loader reuses the soname).
    Load this package before torch: a warning says when that did not happen."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libnsgpu.so is not built: run __graft_entry__.build() or "
                           "python nanoshaper_amd/build.py (looked in %s)" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.nsgpu_abi_version() != 1:
        raise RuntimeError("libnsgpu ABI version mismatch")
    rt = hip_runtime()
    if rt and not any(p.startswith("/opt/rocm") for p in rt):
        import warnings
        warnings.warn("libnsgpu.so runs on the HIP runtime loaded before it (%s), not the ROCm one it is "
                      "built for: import nanoshaper_amd before torch" % ", ".join(rt), RuntimeWarning)
    _lib = lib
    return lib
