"""MI355X (gfx950) grid ray-casting surface builder with NanoShaper's semantics.

The device path lives in libnsgpu.so (C ABI: include/nsgpu.h); this package
is its host-side mirror.
"""
from .surface import (PATCH_KINDS, STAGES, STATUS_FIRST_CAV, STATUS_INSIDE, STATUS_OUT, STATUS_TEMP_OUT, Batch,
                      DeviceSurface, NanoShaper, NsgpuError, PatchSet, as_xyzr, build_grid, generate_chain,
                      grid_coords, jitter, load_atoms, perturbed_frames, regular_triangulation, ses_patches, skin_patches,
                      skin_points, skin_triangulation, surf_opts, weighted_triangulation)

__all__ = ["PATCH_KINDS", "STAGES", "STATUS_FIRST_CAV", "STATUS_INSIDE", "STATUS_OUT", "STATUS_TEMP_OUT", "Batch",
           "DeviceSurface", "perturbed_frames", "NanoShaper", "NsgpuError", "PatchSet", "as_xyzr", "build_grid", "generate_chain",
           "grid_coords", "jitter", "load_atoms", "regular_triangulation", "ses_patches", "skin_patches",
           "skin_points", "skin_triangulation", "surf_opts", "weighted_triangulation"]
