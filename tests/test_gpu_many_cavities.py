"""More cavities than the device replay holds (4,096): the host numbers them.

4,335 thick hollow shells (an outer and an inner icosahedron each) enclose one
cavity apiece; every cavity is below the fill threshold, so fillCavities
rewrites the maps and the corner bitmap after the host numbering, and the
vertices and triangles must follow that final bitmap.  Everything is compared
with the oracle on the same mesh, to the bars of test_gpu_blobby.py.
"""
import numpy as np
import pytest

import nanoshaper_amd as ns

pytestmark = pytest.mark.gpu


def _shells(nx, ny, nz, spacing=6.2, ro=2.9, ri=1.0):
    from test_oracle_mesh import icosphere
    vs, ts, centres = [], [], []
    off = 0
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = np.array([i, j, k], float) * spacing + np.array([0.137, 0.291, 0.063])
                for r in (ro, ri):
                    v, t = icosphere(r, c, 0)
                    vs.append(v)
                    ts.append(t + off)
                    off += len(v)
                centres.append([*c, ro])
    # loadOFF reads floats
    return np.vstack(vs).astype(np.float32).astype(np.float64), np.vstack(ts).astype(np.int32), np.array(centres)


def test_more_cavities_than_the_device_replay(device_ok):
    from oracle.binding import OracleSurface
    from test_gpu_blobby import OPTS, _compare_surface
    v, t, centres = _shells(17, 17, 15)
    grid = ns.build_grid(centres, 2.0, 80.0)
    print("shells: %d triangles, grid %d^3" % (len(t), grid.nx), flush=True)
    dev = ns.DeviceSurface(0)
    dev.set_grid(grid)
    dev.set_mesh(v, t, ray_based=True)
    orc = OracleSurface.from_mesh(grid, v, t, ray_based=True)
    # 4 M triangle areas summed in another order: worst case n eps ~ 4e-10
    r = _compare_surface(grid, dev, orc, OPTS, area_rel=1e-10)
    assert r["num_cavities"] == 4335 > 4096
    sizes, filled = dev.cavities()
    assert filled.all()
