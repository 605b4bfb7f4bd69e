"""Releases the reference generates from the .dat file (icunk_s 11, a
rectangle: Read_rect_SW "6 SPH_SW_2019.f90":4426-4497; icunk_s 12, an
ellipse: Read_ellipsoidal_SW 6:3818-3899; icunk_s 13, a rectangle with a
paraboloid depth: Read_rect_ellipsoid 6:3903-3979; icunk_s 1, a cylinder
with a constant, hemispherical or two-level depth: Read_Cylinder_SW
6:3983-4089) through the library's reader,
against the reference's own run from the same files
(tests/golden/g_release_gen.npz): the particles as generated, bit for bit,
and on the GPU the state after Init_sph and after 1 and 10 RK4 steps."""
import os

import numpy as np
import pytest

import goldens
from gfsph import cases as C
from gfsph.dem import plane_runout_dem
from oracle import refio

RELEASES = {"rect": ("rect", (12, 8), (60.0, 80.0, 100.0, 84.0, 12.0, 3.0, 1.3)),
            "ellipse": ("ellipse", (80.0, 100.0, 25.0, 15.0, 0.3, 4.0), (2.5, 1.3)),
            "rect_ellipsoid": ("rect_ellipsoid", (14, 9), (60.0, 84.0, 100.0, 80.0, 12.0, 3.0, 1.3)),
            "cylinder1": ("cylinder", (80.0, 100.0, 15.0, 4.0, 1), None, (6, 1.2)),
            "cylinder2": ("cylinder", (80.0, 100.0, 15.0, 4.0, 2), (7.0, 6.0), (6, 1.2))}


def _files(tmp_path, tag):
    case = C.Case("rel", C.Params(integrator=4, dt=0.05, h_inf=1e-3, const=C.voellmy(0.2, 500.0)),
                  plane_runout_dem(nx=81, ny=41, d=5.0, x_break=250.0, blend=50.0),
                  C.lattice_block(60.0, 80.0, 30.0, 20.0, 15, 10, 5.0, 1.5), 10)
    refio.write_inputs(str(tmp_path), case, release=RELEASES[tag])
    refio.write_master(str(tmp_path), "case", case.params.dt, 0)


@pytest.mark.parametrize("tag", list(RELEASES))
def test_generated_release_matches_reference(tag, tmp_path):
    from gfsph.inputs import Inputs
    f = goldens.load("g_release_gen")
    _files(tmp_path, tag)
    with Inputs(str(tmp_path), "case", "topo", build=False) as inp:
        q = inp.particles()
    for k in ("x", "mass", "hsml"):
        np.testing.assert_array_equal(q[k], f[f"{tag}_s0_{k}"], err_msg=k)
    np.testing.assert_array_equal(q["v"], 0.0)


@pytest.mark.gpu
@pytest.mark.parametrize("tag", list(RELEASES))
def test_gpu_generated_release_run_matches_reference(tag, tmp_path):
    from gfsph.inputs import Inputs
    f = goldens.load("g_release_gen")
    _files(tmp_path, tag)
    with Inputs(str(tmp_path), "case", "topo") as inp, inp.model() as m:
        done = 0
        for s in (0, 1, 10):
            if s > done:
                m.time_integration_sph(s - done)
            done = s
            st = m.state()
            for k in ("x", "v", "h", "u", "hsml"):
                err = goldens.rel_err(st[k], f[f"{tag}_s{s}_{k}"])
                assert err <= (1e-13 if s == 0 else 1e-8), (tag, s, k, err)
