"""The configuration reader follows the reference CLI (init, main.cpp:1032-1204):
the keys and defaults of normal mode and the flag combinations it refuses;
the reference's own test configurations read as the reference reads them."""
import os

import pytest

from nanoshaper_amd import config

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_reference_configurations():
    c = config.Configuration(config.ConfigFile(os.path.join(GOLDEN, "1crn_ses_conf.prm")))
    assert (c.surface, c.scale, c.perfil, c.probe) == ("ses", 2.0, 90.0, 1.4)
    assert (c.build_epsmaps, c.build_status, c.acc_tri, c.tri, c.fill_cavities) == (False, True, True, True, False)
    assert c.normals and c.mol_file == "1crn.xyzr" and c.cav_vol == 11.4
    s = config.Configuration(config.ConfigFile(os.path.join(GOLDEN, "1crn_skin_conf.prm")))
    assert s.surface == "skin" and s.skin_s == 0.45
    x = config.Configuration(config.ConfigFile(os.path.join(GOLDEN, "6xm4_skin_conf.prm")))
    assert x.surface == "skin" and x.build_epsmaps and x.mol_file == "6xm4.xyzr"
    assert x.surf_opts()["build_epsmap"] and x.surf_opts()["accurate_triangulation"]


def test_defaults_and_comments():
    cf = config.ConfigFile(text="Grid_scale = 3.5 # fine\n# Surface = skin\nSurface=ses\n")
    c = config.Configuration(cf)
    assert c.scale == 3.5 and c.surface == "ses" and c.perfil == 80.0
    assert not c.build_status and not c.tri and c.patch_based and c.analytical_torus


@pytest.mark.parametrize("lines", [
    "Save_eps_maps = true",
    "Project_boundary_grid_points = true",
    "Triangulation = true\nAccurate_Triangulation = false",
    "Cavity_Detection_Filling = true",
    "Save_ideb_map = true",
    "Surface = blobby",
])
def test_refused_combinations(lines):
    with pytest.raises(config.ConfigError):
        config.Configuration(config.ConfigFile(text=lines))


def test_blobby_needs_ray_based_casting():
    c = config.Configuration(config.ConfigFile(text="Surface = blobby\nPatch_Based_Algorithm = false"))
    assert c.surface == "blobby" and not c.patch_based


def test_sturm_flag():
    c = config.Configuration(config.ConfigFile(text="Analytical_Ray_Vs_Torus_Intersection = false"))
    assert c.surf_opts()["precision"] == 2
