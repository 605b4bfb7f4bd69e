"""Case control on the device (gfsph_batch_cases): the step loop the
reference runs for every case of a case table or Monte Carlo set
(Driver_Flow_23, "8 SPH_FEM_Driver_2019.f90":1444-1548) with Stop_case_SW
("6 SPH_SW_2019.f90":20114-20177), per realisation.

* One case against the oracle's restatement of that loop (oracle.run_case):
  stop iteration, time_sph, the state, the plot-time pixel accumulators.
* A batch of realisations that stop at different iterations against each
  run alone under the same control: bit for bit."""
import numpy as np
import pytest

import goldens
from gfsph import BatchModel, SPHModel
from oracle import oracle
from test_gpu_batch import realisations

pytestmark = pytest.mark.gpu

TOL = 1e-8


@pytest.mark.parametrize("name,adapt,integrator", [("g_voellmy_plane", 2, 4), ("g_frictional_pwp", 0, 2),
                                                   ("g_curvature7", 1, 4)])
def test_case_loop_matches_oracle(name, adapt, integrator):
    case, _ = goldens.case_of(name)
    case.params.adapt_dt = adapt
    case.params.integrator = integrator
    ctl = dict(vstop=0.05, timestop=3.0, dtimestop=0.0, dt=0.5, time_end=1e9, time_plot=0.7)
    st = oracle.init(case)
    it, stopped, acc, cnt = oracle.run_case(case, st, ctl["vstop"], ctl["timestop"], ctl["dtimestop"], ctl["dt"],
                                            ctl["time_end"], ctl["time_plot"], max_steps=500)
    assert stopped
    with SPHModel(case) as m:
        m.case_control(**ctl, pixels=True)
        m.run_cases(500, chunk=16)
        stp, its, k = m.case_status()
        assert k == 1 and int(its[0]) == it
        got = m.state()
        t, dt = m.time()
        ga, gc = m.pixel_read(0)
    assert t == pytest.approx(st.dt_time[1], rel=1e-12)
    assert dt == pytest.approx(st.dt_time[0], rel=1e-9)
    for k in ("x", "v", "h", "u"):
        assert goldens.rel_err(got[k], getattr(st, k)) <= TOL, k
    np.testing.assert_array_equal(gc, cnt)
    np.testing.assert_allclose(ga, acc, rtol=1e-8, atol=1e-12)


def test_batch_cases_equal_single_runs():
    cases = realisations(R=4, seed=21, integrator=4)
    for c in cases:
        c.params.adapt_dt = 2
    ctl = dict(vstop=0.6, timestop=1e9, dtimestop=0.0, dt=0.5, time_end=4.0, time_plot=0.45)
    with BatchModel(cases) as b:
        b.case_control(**ctl, pixels=True)
        b.run_cases(2000, chunk=8)
        stp, its, k = b.case_status()
        per = b.split(b.state())
        tb, dtb = b.time()
        pix = [b.pixel_read(r) for r in range(b.R)]
    assert k == b.R
    assert len(set(its.tolist())) > 1  # realisations stop at different iterations
    for r, case in enumerate(cases):
        with SPHModel(case) as m:
            m.case_control(**ctl, pixels=True)
            m.run_cases(2000, chunk=8)
            s1, i1, _ = m.case_status()
            st = m.state()
            t, dt = m.time()
            a1, c1 = m.pixel_read(0)
        assert (int(i1[0]), int(s1[0])) == (int(its[r]), int(stp[r]))
        assert (t, dt) == (tb[r], dtb[r])
        for key in ("x", "v", "h", "u", "hsml", "flag"):
            np.testing.assert_array_equal(per[r][key], st[key], err_msg=f"realisation {r} {key}")
        np.testing.assert_array_equal(pix[r][1], c1)
        np.testing.assert_allclose(pix[r][0], a1, rtol=1e-12, atol=0)


def test_case_table_batch_equals_each_problem(tmp_path):
    """MASTER.dat's case table (ic_cases_mat 2) as one batch
    (gfsph_inputs_batch_create) against every problem built by hand (the
    release and const(:) as Input_SW changes them) and run alone."""
    import copy

    from gfsph import Inputs
    from oracle import refio
    case, _ = goldens.case_of("g_voellmy_plane")
    refio.write_inputs(str(tmp_path), case, name="case", topo="topo")
    rows = [[9, 0.2, 3, 0.12, 0.28, 1], [501, 1.0, 2, 0.8, 1.2, 1], [521, 0.0, 2, -3.0, 3.0, 1, 1.0, -1.0]]
    refio.write_master(str(tmp_path), "case", case.params.dt, adapt=2, cases=2, case_rows=rows,
                       stop=(0.05, 2.5, 0.0), dt_driver=0.5, plot_step=1, time_end=1e9)
    with Inputs(str(tmp_path), "case", "topo") as inp:
        tab = inp.cases()["table"]
        with inp.batch(pixels=True) as b:
            b.run_cases(2000, chunk=8)
            stp, its, k = b.case_status()
            per = b.split(b.state())
            tb, _ = b.time()
            pix = [b.pixel_read(r) for r in range(b.R)]
        base = inp.case()
    assert k == len(tab) == 12
    for p in range(len(tab)):
        mu, fac, dx, dy = tab[p, 3], tab[p, 4], tab[p, 5], tab[p, 3 + 3 + 1]
        c = copy.deepcopy(base)
        c.params.const[9] = float(np.float32(mu))
        c.particles.h = c.particles.h * fac
        c.particles.mass = c.particles.mass * fac
        c.particles.x = c.particles.x + np.array([dx, dy])
        c.params.adapt_dt = 2
        with SPHModel(c) as m:
            m.case_control(vstop=0.05, timestop=2.5, dtimestop=0.0, dt=0.5, time_end=1e9, time_plot=0.5,
                           pixels=True)
            m.run_cases(2000, chunk=8)
            _, i1, _ = m.case_status()
            st = m.state()
            t, _ = m.time()
            a1, c1 = m.pixel_read(0)
        assert int(i1[0]) == int(its[p]) and t == tb[p], p
        for key in ("x", "v", "h", "u", "flag"):
            np.testing.assert_array_equal(per[p][key], st[key], err_msg=f"problem {p} {key}")
        np.testing.assert_array_equal(pix[p][1], c1)
        np.testing.assert_allclose(pix[p][0], a1, rtol=1e-12, atol=0)


@pytest.mark.parametrize("tname", ["t2", "t1", "t20", "t401"])
def test_case_table_batch_matches_reference_main(tname, tmp_path):
    """MASTER.dat's case tables as one batch against the reference's own main
    program (tests/golden/g_case_driver_ref.npz, recorded step by step from
    PROGRAM SPH_FEM_Driver_2019): steps per problem, stop flags, time_sph and
    the final states.  With ic_cases_mat 1 the reference starts each problem
    from the dt_sph the previous one ended with (its mode-1 loop does not
    reset it), a sequential dependence the batch does not imitate: every
    realisation starts from MASTER.dat's dt_sph, as the reference's
    ic_cases_mat >= 2 loop does.  There, problem 1 is checked against the
    reference and the others against the oracle's loop started the same
    way."""
    import os
    import sys

    from gfsph import Inputs
    from oracle import refio
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    from make_golden import CASE_TABLES
    from test_oracle_case_driver import base_case, init_problem, problems

    f = goldens.load("g_case_driver_ref")
    rows, opts = CASE_TABLES[tname]
    case = base_case(tname)
    d = tmp_path / "run"
    d.mkdir()
    refio.write_inputs(str(d), case, name="case", topo="topo")
    refio.write_master(str(d), "case", case.params.dt, case_rows=rows, **opts)
    with Inputs(str(d), "case", "topo") as inp:
        with inp.batch(pixels=False) as b:
            b.run_cases(10000, chunk=8)
            stp, its, k = b.case_status()
            per = b.split(b.state())
            tb, _ = b.time()
    nprob = int(f[f"{tname}_nprob"])
    assert k == nprob == len(per)
    od = tmp_path / "oracle"
    od.mkdir()
    cs, _ = problems(od, rows, opts, tname)
    time_plot = opts["plot_step"] * opts["dt_driver"] if opts["cases"] >= 2 else 0.0
    for p in range(nprob):
        key = f"{tname}_p{p + 1}"
        if opts["cases"] == 1 and p > 0:
            st = init_problem(cs[p])
            it, stopped, _, _ = oracle.run_case(cs[p], st, *opts["stop"], opts["dt_driver"], opts["time_end"],
                                                time_plot, max_steps=10000, pixels=False)
            want = {"iters": it, "stop": int(stopped), "time": float(st.dt_time[1]),
                    **{fl: getattr(st, fl) for fl in ("x", "v", "h", "u", "flag")}}
        else:
            want = {"iters": int(f[key + "_iters"]), "stop": int(f[key + "_stop"][-1]),
                    "time": float(f[key + "_time"][-1]), **{fl: f[key + "_" + fl] for fl in ("x", "v", "h", "u", "flag")}}
        assert (int(its[p]), int(stp[p])) == (want["iters"], want["stop"]), (p, its[p], stp[p])
        assert tb[p] == pytest.approx(want["time"], rel=2e-15), p
        for fl in ("x", "v", "h", "u"):
            err = goldens.rel_err(per[p][fl], want[fl])
            dif = float(np.max(np.abs(per[p][fl] - want[fl])))
            # velocities of rounding size (a flow held at rest): 1e-12 m/s absolute
            assert err <= 1e-10 or (fl == "v" and dif <= 1e-12), (p, fl, err, dif)
        np.testing.assert_array_equal(per[p]["flag"], want["flag"])


@pytest.mark.parametrize("mode", [4, 5])
def test_sampled_set_batch_equals_each_problem(tmp_path, mode):
    """A sampled case set (ic_cases_mat 4, or 5 with one geological zone:
    tan(phi) triangular, the initial
    pore pressure's factor -- case variable 401 -- uniform in a histogram)
    over the frictional pore-pressure fixture as one batch, against every
    problem built by hand (const(9) from the table, the pore pressure scaled
    after the initialisation) and run alone under the same case control:
    bit for bit."""
    import copy

    from gfsph import Inputs
    from oracle import refio
    case, _ = goldens.case_of("g_frictional_pwp")
    refio.write_inputs(str(tmp_path), case, name="case", topo="topo")
    rows = [[9, 0.5, 1, 0, 0, 1], [401, 1.0, 1, 0, 0, 1]]
    laws = [(1, [0.42, 0.5, 0.62]), (100, [0.4, 1.0], [1.0])]
    refio.write_master(str(tmp_path), "case", case.params.dt, adapt=0, cases=mode, case_rows=rows, nmc=6, laws=laws,
                       stop=(0.05, 1.53, 0.0), dt_driver=0.5, plot_step=1, time_end=1e9)
    with Inputs(str(tmp_path), "case", "topo") as inp:
        tab = inp.cases()["table"]
        with inp.batch(pixels=False) as b:
            b.run_cases(2000, chunk=8)
            stp, its, k = b.case_status()
            per = b.split(b.state())
            tb, _ = b.time()
        base = inp.case()
    assert k == len(tab) == 6 and len(set(tab[:, 3])) == 6
    for p in range(len(tab)):
        c = copy.deepcopy(base)
        c.params.const[9] = float(np.float32(tab[p, 3]))
        with SPHModel(c) as m:
            m.scale_pwp(tab[p, 4])
            m.case_control(vstop=0.05, timestop=1.53, dtimestop=0.0, dt=0.5, time_end=1e9, time_plot=0.5,
                           pixels=False)
            m.run_cases(2000, chunk=8)
            _, i1, _ = m.case_status()
            st = m.state()
            t, _ = m.time()
        assert int(i1[0]) == int(its[p]) and t == tb[p], p
        for key in ("x", "v", "h", "u", "uaux", "flag"):
            np.testing.assert_array_equal(per[p][key], st[key], err_msg=f"problem {p} {key}")
