"""MASTER.dat case tables (ic_cases_mat 1 and 2, Driver_Input
"8 SPH_FEM_Driver_2019.f90":475-639) expanded by the host reader into OC_mat,
checked against the closed forms of the reference's construction (linear and
logarithmic value grids, the statistic lists, the mixed-radix order with the
first variable fastest).  No device work."""
import math

import numpy as np
import pytest

import goldens
from gfsph import GfsphError, Inputs
from oracle import refio


def _inputs(tmp_path, **kw):
    case, _ = goldens.case_of("g_voellmy_plane")
    refio.write_inputs(str(tmp_path), case, name="case", topo="topo")
    refio.write_master(str(tmp_path), "case", case.params.dt, **kw)
    return Inputs(str(tmp_path), "case", "topo", build=False)


def test_full_factorial_table(tmp_path):
    rows = [[9, 0.2, 3, 0.1, 0.3, 1],            # mu: linear, 3 values
            [3, 500.0, 2, 200.0, 800.0, 2],      # xi: logarithmic, 2 values
            [501, 1.0, 2, 0.0, 0.0, 100]]        # depth factor: statistic list
    lists = [([501, 0.8, 1.25], [501, 0.5, 0.5])]
    with _inputs(tmp_path, cases=2, case_rows=rows, lists=lists, stop=(0.05, 60.0, 1e-4)) as inp:
        t = inp.cases()
        assert (t["mode"], t["ncases"]) == (2, 3)
        assert (t["vstop"], t["timestop"], t["dtimestop"]) == (0.05, 60.0, 1e-4)
        T = t["table"]
        assert T.shape == (12, 3 + 3 + 2)
        mu = [0.1 + (0.3 - 0.1) / 2 * k for k in range(3)]
        lx = [math.log(200.0), math.log(800.0)]
        xi = [math.exp(lx[0] + (lx[1] - lx[0]) / 1 * k) for k in range(2)]
        fac = [0.8, 1.25]
        p = 0
        for c in range(2):
            for b in range(2):
                for a in range(3):  # variable 1 fastest
                    assert list(T[p, 3:6]) == [mu[a], xi[b], fac[c]], p
                    p += 1


def test_one_at_a_time_table(tmp_path):
    rows = [[9, 0.2, 3, 0.1, 0.3, 1, 0, 0], [521, 0.0, 2, -5.0, 5.0, 1, 2.0, 4.0]]
    with _inputs(tmp_path, cases=1, case_rows=rows) as inp:
        T = inp.cases()["table"]
        assert T.shape == (5, 3 + 2 + 2)
        np.testing.assert_array_equal(T[:, 0], [1, 1, 1, 2, 2])
        np.testing.assert_array_equal(T[:, 1], [1, 2, 3, 1, 2])
        np.testing.assert_array_equal(T[:, 2], [9, 9, 9, 521, 521])
        np.testing.assert_array_equal(T[:3, 3], [0.1 + (0.3 - 0.1) / 2 * k for k in range(3)])
        np.testing.assert_array_equal(T[3:, 3], [0.2, 0.2])   # the other variable at its default
        np.testing.assert_array_equal(T[3:, 4], [-5.0, 5.0])
        np.testing.assert_array_equal(T[:3, 4], [0.0, 0.0, 0.0])


def test_sampled_sets_are_rejected(tmp_path):
    with pytest.raises(GfsphError) as e:
        _inputs(tmp_path, cases=4, case_rows=[[9, 0.2, 3, 0.1, 0.3, 1]])
    assert e.value.code == -2 and "ic_cases_mat 4" in str(e.value)
