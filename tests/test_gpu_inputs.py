"""The input path on the GPU (gfsph_inputs_read / gfsph_inputs_create)
against the reference's own Init_sph and Driver_Flow on the same files.

* Picacho (the reference's example: scattered 10 m DTM, ictop 10; 2402
  release points, icunk 7; MASTER.dat with dt_sph 0.1 and ic_adapt_dt_sph 1):
  the terrain (Get_DTM gridding, Get_Topol projection) bit for bit, the
  gridded release and Input_SW's initialisation, then 100 adaptive steps.
* Synthetic cases written in the reference's formats (explicit grid with
  finite-difference slopes, restart.dat): the context the file path creates
  advances exactly as the one built from the arrays."""
import hashlib
import os

import numpy as np
import pytest

import goldens
from gfsph import Inputs, SPHModel
from oracle import refio
from test_inputs_parse import unpack_picacho

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def picacho(tmp_path_factory):
    d = unpack_picacho(tmp_path_factory.mktemp("picacho"))
    return d, np.load(os.path.join(HERE, "golden", "picacho.npz"))


def test_picacho_terrain_matches_reference(picacho):
    d, f = picacho
    with Inputs(d, "Picacho") as inp:
        dem = inp.dem()
        top = dem["topol"]
        s = f["sample"]
        # sampled values first (a readable failure), then the whole arrays
        np.testing.assert_array_equal(dem["in_model"].ravel()[s], f["sample_in_model"])
        np.testing.assert_array_equal(top[s, 0], f["sample_topol"][:, 0])
        err = np.max(np.abs(top[s, 1:] - f["sample_topol"][:, 1:]) / (np.abs(f["sample_topol"][:, 1:]) + 1e-300))
        assert err == 0.0, f"topol differs at sampled nodes by {err:.3e} (relative)"
        assert sha(top[:, 0]) == str(f["sha_Z"])
        assert sha(top[:, 1:9]) == str(f["sha_topol"])
        assert sha(dem["in_model"].ravel().astype(np.int32)) == str(f["sha_in_model"])


def test_picacho_release_and_run_match_reference(picacho):
    d, f = picacho
    with Inputs(d, "Picacho") as inp:
        q = inp.particles()
        n = inp.n
        assert n == f["s0_x"].shape[0]
        # the gridded release: node positions, masses and smoothing lengths exactly
        np.testing.assert_array_equal(q["x"], f["s0_x"])
        np.testing.assert_array_equal(q["mass"], f["s0_mass"])
        np.testing.assert_array_equal(q["hsml"], f["s0_hsml"])
        with inp.model() as m:
            st = m.state()
            # Input_SW's summed depth and pore pressure (sums in another order)
            assert goldens.rel_err(st["h"], f["s0_h"]) <= 1e-13
            assert goldens.rel_err(st["u"], f["s0_u"]) <= 1e-13
            done = 0
            for s in [int(v) for v in f["dumps"] if v > 0]:
                m.time_integration_sph(s - done)
                done = s
                st = m.state()
                tol = 1e-10 if s >= 100 else 1e-8
                for k in ("x", "v", "h", "u", "hsml"):
                    err = goldens.rel_err(st[k], f[f"s{s}_{k}"])
                    assert err <= tol, f"Picacho step {s} field {k}: {err:.3e}"
                np.testing.assert_array_equal(st["flag"], f[f"s{s}_flag"])
                if f"s{s}_uaux" in f:
                    assert goldens.rel_err(st["uaux"], f[f"s{s}_uaux"]) <= tol
                t, _ = m.time()
                assert t == pytest.approx(float(f[f"s{s}_time"]), rel=1e-9)


@pytest.mark.parametrize("name", ["g_frictional_pwp", "g_curvature7", "g_icpwp1_sle1", "g_adaptive_dt10",
                                  "g_pwp_column_step", "g_erosion_hungr", "g_k0"])
def test_file_path_equals_array_path(name, tmp_path):
    case, _ = goldens.case_of(name)
    refio.write_inputs(str(tmp_path), case, name="case", topo="topo")
    refio.write_master(str(tmp_path), "case", case.params.dt, case.params.adapt_dt)
    with Inputs(str(tmp_path), "case", "topo") as inp, inp.model() as a, SPHModel(case) as b:
        for _ in range(2):
            a.time_integration_sph(25)
            b.time_integration_sph(25)
            sa, sb = a.state(), b.state()
            for k in ("x", "v", "h", "u", "uaux", "hsml", "flag"):
                np.testing.assert_array_equal(sa[k], sb[k], err_msg=f"{name} {k}")
            assert a.time() == b.time()
            if case.params.icpwp == 2:
                np.testing.assert_array_equal(a.pwp_column(), b.pwp_column())
            if name.startswith("g_erosion"):
                np.testing.assert_array_equal(a.mass(), b.mass())
