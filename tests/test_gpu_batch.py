"""Batched Monte Carlo realisations against the same realisations run one
by one through the plain drop-in (the batch must not change any
realisation's result), and the hazard statistics against the oracle's
Norm_MC_pixel_Vars_SW ("6 SPH_SW_2019.f90":20838-20908) and against the
reference's own statistics over the same problems."""
import numpy as np
import pytest

import goldens
from oracle import oracle

from gfsph import BatchModel, SPHModel
from gfsph import cases as C
from gfsph.dem import fractal_dem

pytestmark = pytest.mark.gpu


def realisations(R=5, seed=11, integrator=2):
    dem = fractal_dem(129, 2.0, 31, relief=60.0)
    rng = np.random.default_rng(seed)
    out = []
    sizes = [300, 900, 150, 1200, 500, 60, 700][:R]
    for r in range(R):
        mu = float(rng.uniform(0.08, 0.35))
        xi = float(np.exp(rng.normal(np.log(400.0), 0.5)))
        cx = 60.0 + rng.uniform(-20, 20)
        cy = 128.0 + rng.uniform(-30, 30)
        parts = C.elliptical_scar(cx, cy, sizes[r], 1.0, 2.0, seed + r, h_range=(1.0, 4.0))
        prm = C.Params(integrator=integrator, dt=0.02, h_inf=1e-3, const=C.voellmy(mu, xi))
        out.append(C.Case(f"r{r}", prm, dem, parts, 0))
    return out


@pytest.mark.parametrize("integrator", [2, 4])
def test_batch_equals_single_runs(integrator):
    cases = realisations(integrator=integrator)
    nsteps = 80
    with BatchModel(cases) as b:
        b.time_integration_sph(nsteps)
        per = b.split(b.state())
    for r, case in enumerate(cases):
        with SPHModel(case) as m:
            m.time_integration_sph(nsteps)
            st = m.state()
        for k in ("x", "v", "h", "u", "hsml"):
            np.testing.assert_array_equal(per[r][k], st[k], err_msg=f"realisation {r} field {k}")
        np.testing.assert_array_equal(per[r]["flag"], st["flag"])


@pytest.mark.parametrize("variant", ["sle3", "nfric23", "nfric3"])
def test_batch_laws_and_sle_equal_single_runs(variant):
    """Smoothing lengths evolved by h_upgrade (sle = 3, each realisation's grid
    from its own new lengths), the exact Herschel-Bulkley law and the
    Bingham-erosion law (mass0 per realisation) in a batch: every realisation
    equals its run alone bit for bit."""
    cases = realisations(R=3, seed=5, integrator=2)
    for r, c in enumerate(cases):
        if variant == "sle3":
            c.params.sle = 3
        elif variant == "nfric23":
            c.params.const = {1: 9.81, 2: 1800.0, 5: 23, 6: 300.0 + 100 * r, 7: 20.0, 8: (1.0, 2.0, 0.5)[r],
                              10: 1e-2, 14: 1.0, 15: 1e-2}
        else:
            c.params.const = {1: 9.81, 2: 2000.0, 3: 5.0, 5: 3, 6: 400.0, 7: 30.0, 8: 0.04, 10: 1e-2, 14: 1.0,
                              15: 1e-2, 11: -1.0, 21: 1, 22: 0.004 * (r + 1)}
    nsteps = 60
    with BatchModel(cases) as b:
        b.time_integration_sph(nsteps)
        per = b.split(b.state())
    for r, case in enumerate(cases):
        with SPHModel(case) as m:
            m.time_integration_sph(nsteps)
            st = m.state()
        for k in ("x", "v", "h", "u", "hsml"):
            np.testing.assert_array_equal(per[r][k], st[k], err_msg=f"{variant} realisation {r} field {k}")
        np.testing.assert_array_equal(per[r]["flag"], st["flag"])


@pytest.mark.parametrize("mode", [2, 7])
def test_batch_adaptive_dt_equals_single_runs(mode):
    """Each realisation adapts its own dt_sph (ic_adapt_dt_sph = 2, or 7 over
    consecutive particles of the realisation) exactly as it would alone; the
    minimum over particles does not depend on order."""
    cases = realisations(R=4, seed=3, integrator=4)
    for c in cases:
        c.params.adapt_dt = mode
    nsteps = 60
    with BatchModel(cases) as b:
        b.time_integration_sph(nsteps)
        per = b.split(b.state())
        tb, dtb = b.time()
    assert len(set(dtb.tolist())) > 1  # the realisations' time steps differ
    for r, case in enumerate(cases):
        with SPHModel(case) as m:
            m.time_integration_sph(nsteps)
            st = m.state()
            t, dt = m.time()
        assert (tb[r], dtb[r]) == (t, dt)
        for k in ("x", "v", "h", "u", "hsml"):
            np.testing.assert_array_equal(per[r][k], st[k], err_msg=f"realisation {r} field {k}")


def test_hazard_netcdf(tmp_path):
    """The hazard rasters written as NetCDF read back (scipy's reader) equal
    the arrays gfsph_batch_hazard returns."""
    from scipy.io import netcdf_file
    cases = realisations(R=3, seed=9)
    with BatchModel(cases) as b:
        for _ in range(4):
            b.time_integration_sph(10)
            b.pixel_accumulate()
        hz = b.write_hazard(str(tmp_path / "hazard.nc"), 0.5)
    dem = cases[0].dem
    with netcdf_file(str(tmp_path / "hazard.nc"), "r", mmap=False) as f:
        g = lambda k: f.variables[k][:].reshape(-1)  # noqa: E731
        np.testing.assert_array_equal(g("max_h"), hz["max_h"])
        np.testing.assert_array_equal(g("p_exceed"), hz["p_exceed"])
        np.testing.assert_array_equal(g("mean_h"), hz["mean"][:, 0])
        np.testing.assert_array_equal(g("std_v"), hz["std"][:, 2])
        np.testing.assert_array_equal(g("n_reached"), hz["nprob"])
        assert f.dimensions == {"x": dem.nx, "y": dem.ny}


def test_batch_pixels_and_hazard():
    cases = realisations(R=4, seed=5)
    plots, every = 6, 15
    with BatchModel(cases) as b:
        for _ in range(plots):
            b.time_integration_sph(every)
            b.pixel_accumulate()
        accs = [b.pixel_read(r) for r in range(b.R)]
        hz = b.hazard(0.5)
    for r, case in enumerate(cases):
        with SPHModel(case) as m:
            for _ in range(plots):
                m.time_integration_sph(every)
                m.mc_pixel_accumulate()
            acc, cnt = m.mc_pixel_read()
        np.testing.assert_array_equal(accs[r][1], cnt)
        np.testing.assert_allclose(accs[r][0], acc, rtol=1e-12, atol=0)
    # hazard statistics against the oracle's Norm_MC_pixel_Vars_SW (pinned to
    # the reference's own statistics by tests/test_oracle_hazard.py)
    A = np.stack([a for a, _ in accs])            # [R, np, 4]
    N = np.stack([c for _, c in accs])            # [R, np]
    ref = oracle.mc_hazard(A, N, 0.5)
    np.testing.assert_array_equal(hz["nprob"], ref["nprob"])
    np.testing.assert_allclose(hz["mean"], ref["pr"][:, 2, :], rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(hz["std"], ref["pr"][:, 3, :], rtol=1e-9, atol=1e-9)
    np.testing.assert_array_equal(hz["max_h"], ref["maxs"][:, 0])
    np.testing.assert_array_equal(hz["max_v"], ref["maxs"][:, 1])
    np.testing.assert_allclose(hz["p_exceed"], ref["pexc"], rtol=0, atol=1e-15)
    assert hz["p_exceed"].max() > 0


def bingham_realisations(R=4, seed=5, integrator=2):
    """Bingham realisations with tau_y ~ U(500, 5000) Pa and mu_B ~ U(10, 200)
    Pa s (the calibration workload's ranges) and releases of uneven size."""
    dem = fractal_dem(129, 2.0, 31, relief=60.0)
    rng = np.random.default_rng(seed)
    sizes = [400, 1100, 250, 800, 600][:R]
    out = []
    for r in range(R):
        tau, mub = float(rng.uniform(500.0, 5000.0)), float(rng.uniform(10.0, 200.0))
        parts = C.elliptical_scar(60.0 + rng.uniform(-20, 20), 128.0 + rng.uniform(-30, 30), sizes[r], 1.0, 2.0,
                                  seed + r, h_range=(1.0, 4.0))
        prm = C.Params(integrator=integrator, dt=0.02, h_inf=1e-3, const=C.bingham(tau, mub))
        out.append(C.Case(f"b{r}", prm, dem, parts, 0))
    return out


@pytest.mark.parametrize("integrator,adapt", [(2, 0), (4, 0), (2, 2)])
def test_bingham_batch_equals_single_runs(integrator, adapt):
    """Bingham realisations as one batch: each realisation's Tau0 is carried
    from particle to particle and call to call within that realisation only
    (a segmented last-valid scan), so every realisation ends exactly where
    it ends alone."""
    cases = bingham_realisations(integrator=integrator)
    for c in cases:
        c.params.adapt_dt = adapt
    with BatchModel(cases) as b:
        b.time_integration_sph(60)
        per = b.split(b.state())
        tb, dtb = b.time()
    for r, case in enumerate(cases):
        with SPHModel(case) as m:
            m.time_integration_sph(60)
            st = m.state()
            t, dt = m.time()
        assert (t, dt) == (tb[r], dtb[r]), r
        for k in ("x", "v", "h", "u", "hsml", "flag"):
            np.testing.assert_array_equal(per[r][k], st[k], err_msg=f"realisation {r} {k}")


def test_batch_rejects_bingham_and_mismatched_gravity():
    cases = realisations(R=2)
    bing = C.Case("b", C.Params(integrator=2, dt=0.02, h_inf=1e-3, const=C.bingham(1000.0, 50.0)), cases[0].dem,
                  cases[0].particles, 0)
    with pytest.raises(Exception, match="unsupported|Bingham"):  # Bingham mixed with Voellmy
        BatchModel([cases[0], bing])
    g2 = dict(C.voellmy(0.2, 400.0))
    g2[1] = 9.7
    other = C.Case("g", C.Params(integrator=2, dt=0.02, h_inf=1e-3, const=g2), cases[0].dem, cases[1].particles, 0)
    with pytest.raises(Exception, match="const\\(1\\)"):
        BatchModel([cases[0], other])


def test_autotune_keeps_results():
    cases = realisations(R=3, seed=21)
    with BatchModel(cases) as a, BatchModel(cases) as b:
        a.time_integration_sph(9)
        b.time_integration_sph(9)
        chosen = b.autotune(reps=2)
        assert all(c in (64, 128, 256, 512) for c in chosen)
        a.time_integration_sph(17)
        b.time_integration_sph(17)
        sa, sb = a.state(), b.state()
    for k in ("x", "v", "h", "u"):
        np.testing.assert_array_equal(sa[k], sb[k])


def test_batch_hazard_matches_reference_statistics():
    """The reference's own Monte Carlo pixel statistics (the reference build
    ran Get_MC_pixel_Vars_SW every 10 steps and Norm_MC_pixel_Vars_SW(1)
    after each of 4 problems, tests/golden/g_mc_hazard_ref.npz) against the
    batch of the same 4 problems: nodes reached, problems per node, mean and
    standard deviation of average h, max h, average |v|, max |v|."""
    import sys
    import os
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    from make_golden import hazard_cases

    f = goldens.load("g_mc_hazard_ref")
    cs = hazard_cases()
    nsteps, every = int(f["nsteps"]), int(f["plot_every"])
    with BatchModel(cs) as b:
        for _ in range(nsteps // every):
            b.time_integration_sph(every)
            b.pixel_accumulate()
        hz = b.hazard(0.5)
    node = f["node"]
    reached = np.flatnonzero(hz["nprob"] > 0)
    np.testing.assert_array_equal(reached, node)
    np.testing.assert_array_equal(hz["nprob"][node], f["nprob"])
    pr = f["mc_pr"]
    scale = np.max(np.abs(pr[:, 2, :]), axis=0)
    for q in range(4):
        err_m = np.max(np.abs(hz["mean"][node, q] - pr[:, 2, q])) / scale[q]
        err_s = np.max(np.abs(hz["std"][node, q] - pr[:, 3, q])) / scale[q]
        assert err_m <= 1e-10 and err_s <= 1e-8, (q, err_m, err_s)
