"""Generated terrains (ictop 1-7 and -7 of the terrain file, Topo_main
"2 SPH_FEM_topo_2019.f90":262-444, get_cone 2:1113-1151) through the
library's reader, against the terrain the reference built from the same
files (tests/golden/g_terrain_gen.npz): the grid and Z bit for bit (finite-
difference slopes: host code only), and on the GPU the nodal slopes and
curvature terms of Get_Topol as well."""
import os

import numpy as np
import pytest

import goldens
from gfsph import cases as C
from gfsph.dem import plane_runout_dem
from oracle import refio


def _terrains():
    """ictop -> (extent, nx, ny, parameters) as the fixture was made."""
    f = goldens.load("g_terrain_gen")
    out = {}
    for ictop in (1, 2, 3, 4, 5, 6, 7, -7):
        k = f"t{ictop}".replace("-", "m")
        nx, ny = (int(v) for v in f[k + "_nxny"])
        out[ictop] = (tuple(f[k + "_extent"]), nx, ny, tuple(f[k + "_params"]))
    return out


def _read(tmp_path, ictop, ext, nx, ny, prm):
    from gfsph.inputs import Inputs
    release = C.lattice_block(10.0, 10.0, 20.0, 20.0, 10, 10, 2.0, 1.5)
    case = C.Case("gen", C.Params(integrator=4, dt=0.05, h_inf=1e-3, const=C.voellmy(0.2, 500.0)),
                  plane_runout_dem(nx=81, ny=41, d=5.0, x_break=250.0, blend=50.0), release, 0)
    refio.write_inputs(str(tmp_path), case)
    with open(os.path.join(tmp_path, "topo.top"), "w") as f:
        f.write(refio.generated_top(ictop, ext, nx, ny, prm))
    return Inputs(str(tmp_path), "case", "topo")


@pytest.mark.parametrize("ictop", [1, 2, 3, 4, 5, 6, 7])
def test_generated_terrain_matches_reference(ictop, tmp_path):
    f = goldens.load("g_terrain_gen")
    ext, nx, ny, prm = _terrains()[ictop]
    key = f"t{ictop + 100}"
    with _read(tmp_path, ictop + 100, ext, nx, ny, prm) as inp:
        g = inp.grid
        assert (g.nx, g.ny) == (nx, ny)
        np.testing.assert_array_equal([g.xmin, g.xmax, g.ymin, g.ymax, g.dx, g.dy], f[key + "_grid"])
        d = inp.dem()
        np.testing.assert_array_equal(d["Z"].ravel(), f[key + "_topol"][:, 0])
        np.testing.assert_array_equal(d["in_model"].ravel(), f[key + "_in_model"])
    assert np.ptp(f[key + "_topol"][:, 0]) > 0 or ictop == 1


def test_generated_terrain_outside_its_segments_is_refused(tmp_path):
    from gfsph import GfsphError
    ext, nx, ny, prm = _terrains()[3]
    with pytest.raises(GfsphError):
        _read(tmp_path, 103, (-10.0, 400.0, 0.0, 200.0), nx, ny, prm)


@pytest.mark.gpu
@pytest.mark.parametrize("ictop", [1, 2, 3, 4, 5, 6, 7, -7])
def test_gpu_generated_terrain_slopes_match_reference(ictop, tmp_path):
    """Get_Topol's nodal terms (the finite-element projection the reader runs
    on the device) for every generated form."""
    f = goldens.load("g_terrain_gen")
    ext, nx, ny, prm = _terrains()[ictop]
    key = f"t{ictop}".replace("-", "m")
    with _read(tmp_path, ictop, ext, nx, ny, prm) as inp:
        d = inp.dem()
    ref = f[key + "_topol"]
    np.testing.assert_array_equal(d["Z"].ravel(), ref[:, 0])
    for k in range(1, 9):
        err = np.max(np.abs(d["topol"][:, k] - ref[:, k])) / max(np.max(np.abs(ref[:, k])), 1e-300)
        assert err <= 1e-12, (ictop, k, err)
