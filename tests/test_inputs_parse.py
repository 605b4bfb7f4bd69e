"""The host half of the input path (gfsph_inputs_parse): the reference's
positional ASCII files read with Fortran list-directed rules, the terrain
grid's geometry and the run control, checked against what the reference's
own Init_sph built from the same files (tests/golden/picacho.npz) and against
synthetic inputs written in the reference's formats (oracle/refio.py).  No
device work: these run without a GPU."""
import gzip
import os
import shutil

import numpy as np
import pytest

import goldens
from gfsph import GfsphError, Inputs
from oracle import refio

HERE = os.path.dirname(os.path.abspath(__file__))
PICACHO = os.path.join(HERE, "golden", "picacho")


def unpack_picacho(d):
    for f in os.listdir(PICACHO):
        with gzip.open(os.path.join(PICACHO, f), "rb") as a, open(os.path.join(d, f[:-3]), "wb") as b:
            shutil.copyfileobj(a, b)
    return str(d)


def test_picacho_grid_and_run_control(tmp_path):
    d = unpack_picacho(tmp_path)
    f = np.load(os.path.join(HERE, "golden", "picacho.npz"))
    with Inputs(d, "Picacho", build=False) as inp:
        g = inp.grid
        # Get_DTM's grid, bit for bit as the reference computed it
        assert (g.nx, g.ny) == tuple(int(v) for v in f["grid_n"])
        got = [g.xmin, g.xmax, g.ymin, g.ymax, g.dx, g.dy, g.xshift, g.yshift]
        assert got == [float(v) for v in f["grid"]]
        assert g.fdiff == 0  # ictop 10: Get_Topol projection
        r = inp.run
        assert (r.present, r.t_integ_alg, r.adapt_dt, r.maxtimestep) == (1, 4, 1, 1000000)
        assert (r.dt, r.time_end, r.dt_sph) == (0.5, 1000.0, 0.1)
        assert (r.print_step, r.save_step, r.plot_step) == (50, 50, 50)
        p = inp.params
        assert (p.pa_sph, p.nnps, p.sle, p.skf, p.summation_density, p.average_velocity, p.nor_density) == \
            (2, 4, 1, 1, 1, 1, 0)
        assert (p.icpwp, p.pwp0rel, p.h_inf, p.dt, p.integrator, p.init_input) == (3, 1.0, 0.01, 0.1, 4, 1)
        # const(1:15) are REAL(4): each decimal string rounded once to single precision
        text = "9.8   2000.  300.          0.0       7      0.      0.0     0.    0.6     1.e-2      0.    0.  " \
               "   0.001     1.0     1.e-2"
        want = [float(np.float32(t)) for t in text.split()]
        assert [p.cst[k] for k in range(1, 16)] == want
        assert inp.n == 0  # the gridded release is built on the device


@pytest.mark.parametrize("name", ["g_frictional_pwp", "g_icpwp1_sle1", "g_bingham", "g_hb25", "g_pwp_column_step",
                                  "g_erosion_hungr", "g_k0", "g_costheta_k0"])
def test_restart_inputs_round_trip(name, tmp_path):
    """Inputs written in the reference's formats (explicit-grid .top with
    finite-difference slopes, restart.dat) read back exactly."""
    case, _ = goldens.case_of(name)
    refio.write_inputs(str(tmp_path), case, name="case", topo="topo")
    with Inputs(str(tmp_path), "case", "topo", build=False) as inp:
        p, cp = inp.params, case.params
        assert (p.pa_sph, p.nnps, p.sle, p.skf, p.icpwp) == (cp.pa_sph, cp.nnps, cp.sle, cp.skf, cp.icpwp)
        assert (p.summation_density, p.average_velocity, p.nor_density) == \
            (cp.summation_density, cp.average_velocity, cp.nor_density)
        assert p.h_inf == cp.h_inf
        if cp.icpwp in (1, 2, 3):
            assert p.pwp0rel == cp.pwp0rel
        if cp.ic_k0:  # K0 coefficients (6:1383-1389)
            assert (p.ic_k0, p.k0_active, p.k0_pasive) == (1, cp.k0_active, cp.k0_pasive)
        if cp.icpwp == 2:  # the column block of Initialize_Pwp
            assert (p.nuaux, tuple(p.pwp_bc_type), tuple(p.pwp_bc_val), p.pwp0_type, p.xirel) == \
                (cp.nuaux, tuple(cp.pwp_bc_type), tuple(cp.pwp_bc_val), cp.pwp0_type, cp.xirel)
        c = cp.cst()
        assert all(p.cst[k] == c[k] for k in range(1, 16))
        g, dem = inp.grid, case.dem
        assert (g.nx, g.ny, g.fdiff) == (dem.nx, dem.ny, 1)
        assert (g.xmin, g.ymin, g.dx, g.dy) == (dem.xmin, dem.ymin, dem.dx, dem.dy)
        assert inp.run.present == 0 and inp.run.time_restart == 0.0
        q = inp.particles()
        P = case.particles
        for k in ("x", "v", "h", "mass", "hsml", "u"):
            np.testing.assert_array_equal(q[k], getattr(P, k), err_msg=k)
        np.testing.assert_array_equal(q["itype"], P.itype)


def test_list_directed_reading(tmp_path):
    """Values continued over records, r*v repeats, '/' ending a read, D
    exponents, commas and .TRUE. spellings, as Fortran list-directed input."""
    case, _ = goldens.case_of("g_frictional_pwp")
    refio.write_inputs(str(tmp_path), case, name="case", topo="topo")
    path = os.path.join(str(tmp_path), "case.dat")
    lines = open(path).read().split("\n")
    k = lines.index("const 1..15") + 1
    vals = lines[k].split()
    # const(1:15) split over three records, two zeros as a repeat, a D exponent
    zero_pos = [i for i, v in enumerate(vals) if float(v) == 0.0]
    assert len(zero_pos) >= 2
    vals[zero_pos[0]] = "2*0.0"
    del vals[zero_pos[1]]
    vals[0] = vals[0].replace("e", "D") if "e" in vals[0] else vals[0] + "D0"
    lines[k] = " ".join(vals[:5]) + ","
    lines.insert(k + 1, " ".join(vals[5:9]))
    lines.insert(k + 2, " ".join(vals[9:]) + "   trailing words are ignored")
    j = lines.index("sum_den av_vel virt_part nor_dens") + 1
    lines[j] = ".TRUE., .true. ,.FALSE. F"
    i = lines.index("pa_sph nnps sle skf") + 1
    lines[i] = lines[i] + " / 99 99"
    open(path, "w").write("\n".join(lines))
    with Inputs(str(tmp_path), "case", "topo", build=False) as inp:
        c = case.params.cst()
        assert all(inp.params.cst[k] == c[k] for k in range(1, 16))
        assert (inp.params.summation_density, inp.params.average_velocity, inp.params.nor_density) == (1, 1, 0)


@pytest.mark.parametrize("edit,what", [
    (("100", "20"), "ictop"),  # Get_DTM_FEM (ictop 20, 21) is not read
    (("200 ", "4 "), "icunk"),  # a 3-D point cloud (Read_3DCloud) is not read
])
def test_unsupported_options_are_rejected(edit, what, tmp_path):
    case, _ = goldens.case_of("g_voellmy_plane")
    refio.write_inputs(str(tmp_path), case, name="case", topo="topo")
    f = "topo.top" if what == "ictop" else "case.dat"
    path = os.path.join(str(tmp_path), f)
    s = open(path).read().replace(edit[0], edit[1], 1)
    open(path, "w").write(s)
    with pytest.raises(GfsphError) as e:
        Inputs(str(tmp_path), "case", "topo", build=False)
    assert e.value.code == -2
    assert what in str(e.value)


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(GfsphError) as e:
        Inputs(str(tmp_path), "nothing", build=False)
    assert e.value.code == -1 and "cannot open" in str(e.value)
