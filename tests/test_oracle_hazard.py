"""The oracle's Monte Carlo hazard statistics (oracle.mc_hazard, the
restatement of Norm_MC_pixel_Vars_SW, "6 SPH_SW_2019.f90":20838-20908)
against the reference's own: the reference build ran the problems of
tests/golden/make_golden.py:hazard_cases one after another with
Get_MC_pixel_Vars_SW every 10 steps and Norm_MC_pixel_Vars_SW(1) after
each (tests/golden/g_mc_hazard_ref.npz); the oracle runs the same problems
with its own step and pixel accumulation and reduces them with
oracle.mc_hazard.  Also a closed form: one problem gives mean = its value and
standard deviation 0; two problems with values a, b give |a - b| / 2."""
import os
import sys

import numpy as np

import goldens
from oracle import oracle

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))


def test_oracle_hazard_matches_reference():
    from make_golden import hazard_cases

    f = goldens.load("g_mc_hazard_ref")
    cs = hazard_cases()
    nsteps, every = int(f["nsteps"]), int(f["plot_every"])
    npg = cs[0].dem.nx * cs[0].dem.ny
    A = np.zeros((len(cs), npg, 4))
    N = np.zeros((len(cs), npg), np.int32)
    for r, c in enumerate(cs):
        st = oracle.init(c)
        for _ in range(nsteps // every):
            oracle.run(c, st, every)
            oracle.mc_pixel_accum(c.dem, st, A[r], N[r])
    out = oracle.mc_hazard(A, N, 0.5)
    node = f["node"]
    np.testing.assert_array_equal(np.flatnonzero(out["nprob"] > 0), node)
    np.testing.assert_array_equal(out["nprob"][node], f["nprob"])
    pr = f["mc_pr"]
    for row, tol in ((0, 1e-10), (1, 1e-10), (2, 1e-10), (3, 1e-8)):
        scale = np.max(np.abs(pr[:, row, :]), axis=0)
        err = np.max(np.abs(out["pr"][node, row, :] - pr[:, row, :]), axis=0) / np.maximum(scale, 1e-300)
        assert np.all(err <= tol), (row, err)
    # not a tautology: some nodes were reached by only some of the problems,
    # and the spread between problems is not zero there
    assert len(set(f["nprob"].tolist())) > 1 and np.max(pr[:, 3, :]) > 0


def test_oracle_hazard_closed_forms():
    npg = 3
    A = np.zeros((2, npg, 4))
    N = np.zeros((2, npg), np.int32)
    # node 0: one problem, 4 visits summing h = 2.0, |v| = 6.0, maxima 0.9, 2.5
    A[0, 0] = [2.0, 0.9, 6.0, 2.5]
    N[0, 0] = 4
    # node 1: two problems, averages 1 and 3 (h), maxima 2 and 5 (h)
    A[0, 1] = [2.0, 2.0, 2.0, 1.0]
    N[0, 1] = 2
    A[1, 1] = [9.0, 5.0, 3.0, 4.0]
    N[1, 1] = 3
    out = oracle.mc_hazard(A, N, 1.0)
    pr = out["pr"]
    np.testing.assert_array_equal(out["nprob"], [1, 2, 0])
    np.testing.assert_allclose(pr[0, 2], [0.5, 0.9, 1.5, 2.5], rtol=1e-15)
    np.testing.assert_array_equal(pr[0, 3], [0.0, 0.0, 0.0, 0.0])
    np.testing.assert_allclose(pr[1, 2], [2.0, 3.5, 1.0, 2.5], rtol=1e-15)
    np.testing.assert_allclose(pr[1, 3], [1.0, 1.5, 0.0, 1.5], rtol=1e-12)
    np.testing.assert_array_equal(out["maxs"][1], [5.0, 4.0])
    np.testing.assert_allclose(out["pexc"], [0.0, 1.0, 0.0])
    assert not pr[2].any()
