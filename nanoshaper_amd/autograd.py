"""Differentiable SES volume and area estimate (torch.autograd.Function).

The forward pass builds the SES patches of the atoms on the host, traces the
three grid-ray panels on the GPU and runs the hand-written adjoint kernel
alongside (one trace serves both): the volume and area gradients with
respect to every atom centre and radius come back with the values.  The
backward pass applies the chain rule to those cached gradients.
"""
import numpy as np
import torch

from .surface import DeviceSurface, as_xyzr, ses_patches


class SurfaceVolumeArea(torch.autograd.Function):
    """(volume, area) of the SES of xyzr (n, 4) float64 on a fixed grid."""

    @staticmethod
    def forward(ctx, xyzr, grid, probe, dev):
        a = np.ascontiguousarray(as_xyzr(xyzr.detach().cpu().numpy()), dtype=np.float64)
        ps = ses_patches(a, probe)
        dev.set_grid(grid)
        dev.set_patches(ps)
        dev.set_atoms(a)
        out = dev.surface_gradient()
        ctx.save_for_backward(torch.from_numpy(out["dvolume"]).to(xyzr.device),
                              torch.from_numpy(out["darea"]).to(xyzr.device))
        return (torch.tensor(out["volume"], dtype=torch.float64, device=xyzr.device),
                torch.tensor(out["area"], dtype=torch.float64, device=xyzr.device))

    @staticmethod
    def backward(ctx, g_volume, g_area):
        dv, da = ctx.saved_tensors
        return g_volume * dv + g_area * da, None, None, None


def surface_volume_area(xyzr, grid, probe=1.4, dev=None):
    """Differentiable (volume, area) of the SES of xyzr ((n, 4) float64 tensor:
    x, y, z, radius, already jittered) on grid; dev: a DeviceSurface to reuse."""
    if dev is None:
        dev = DeviceSurface(0)
    return SurfaceVolumeArea.apply(xyzr, grid, float(probe), dev)
