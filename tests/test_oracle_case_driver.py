"""The oracle's restatement of the reference's case loop (oracle.run_case:
Driver_Flow_23, "8 SPH_FEM_Driver_2019.f90":1444-1548, and Driver_Flow,
8:1554-1710, with Stop_case_SW, "6 SPH_SW_2019.f90":20114-20177) against the
reference's own main program: oracle/_ref/gfsph_ref_main ran four MASTER.dat
case tables (ic_cases_mat 2 and 1 over the plane fixture's release, the FOSM
sets of ic_cases_mat 20, and a table with case variable 401 -- the initial
pore pressure's factor -- over the frictional pore-pressure fixture) and
recorded every step of every problem (tests/golden/g_case_driver_ref.npz).
For each problem: the number of steps, the stop flag, time_sph (to the
rounding of the adaptive dt_sph it sums) and the final state.

With ic_cases_mat 1 the reference's problem loop (8:62-86) does not reset
dt_sph between problems, unlike the loop of ic_cases_mat >= 2 (8:101-104), so
under adaptive dt each problem starts from the dt_sph its predecessor ended
with; the oracle is given that carried dt_sph here."""
import copy
import os
import sys

import numpy as np
import pytest

import goldens
from gfsph import Inputs
from oracle import oracle, refio

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))


def base_case(tname):
    from make_golden import CASE_BASE

    case, _ = goldens.case_of(CASE_BASE.get(tname, "g_voellmy_plane"))
    return case


def problems(tmp_path, rows, opts, tname="t2"):
    """The cases of every problem of a case table: the release and const(:)
    as Input_SW changes them for the problem (const indices, 501 depth
    factor; 401, the initial pore pressure's factor, kept as pwp_factor for
    init_problem), read back from the table the host reader expands."""
    case = base_case(tname)
    refio.write_inputs(str(tmp_path), case, name="case", topo="topo")
    refio.write_master(str(tmp_path), "case", case.params.dt, case_rows=rows, **opts)
    with Inputs(str(tmp_path), "case", "topo", build=False) as inp:
        t = inp.cases()
    T, nvar = t["table"], t["ncases"]
    out = []
    for p in range(T.shape[0]):
        c = copy.deepcopy(case)
        c.params.adapt_dt = opts["adapt"]
        c.pwp_factor = 1.0
        for v in range(nvar):
            var, val = int(rows[v][0]), float(T[p, 3 + v])
            if var == 501:
                c.particles.h = c.particles.h * val
                c.particles.mass = c.particles.mass * val
            elif var == 401:
                c.pwp_factor = val
            else:
                c.params.const[var] = float(np.float32(val))
        out.append(c)
    return out, t


def init_problem(c):
    """Input_SW for one problem: the oracle's initialisation, then the pore
    pressure (u and Uaux) times the 401 factor (6:1453-1479)."""
    st = oracle.init(c)
    f = getattr(c, "pwp_factor", 1.0)
    if f != 1.0:
        st.u = st.u * f
        st.uaux = st.uaux * f
    return st


@pytest.mark.parametrize("tname", ["t2", "t1", "t20", "t401"])
def test_oracle_case_loop_matches_reference_main(tname, tmp_path):
    from make_golden import CASE_TABLES

    f = goldens.load("g_case_driver_ref")
    rows, opts = CASE_TABLES[tname]
    cs, t = problems(tmp_path, rows, opts, tname)
    assert len(cs) == int(f[f"{tname}_nprob"])
    time_plot = opts["plot_step"] * opts["dt_driver"] if opts["cases"] >= 2 else 0.0
    vstop, timestop, dtimestop = opts["stop"]
    carried = None
    for p, c in enumerate(cs):
        k = f"{tname}_p{p + 1}"
        st = init_problem(c)
        if opts["cases"] == 1 and carried is not None:
            st.dt_time = np.array([carried, 0.0])
        it, stopped, _, _ = oracle.run_case(c, st, vstop, timestop, dtimestop, opts["dt_driver"], opts["time_end"],
                                            time_plot, max_steps=10000, pixels=False)
        carried = float(st.dt_time[0])
        assert it == int(f[k + "_iters"]), (p, it)
        assert stopped == bool(f[k + "_stop"][-1]), p
        # time_sph sums the adaptive dt_sph, which follow the state to rounding
        assert float(st.dt_time[1]) == pytest.approx(float(f[k + "_time"][-1]), rel=2e-15), p
        for fld in ("x", "v", "h", "u"):
            err = goldens.rel_err(getattr(st, fld), f[k + "_" + fld])
            # a flow held at rest (FOSM's tan(phi) = 500 rows) has velocities
            # of rounding size: there an absolute 1e-12 m/s bounds them
            dif = float(np.max(np.abs(getattr(st, fld) - f[k + "_" + fld])))
            assert err <= 1e-10 or (fld == "v" and dif <= 1e-12), (p, fld, err, dif)
        np.testing.assert_array_equal(st.flag, f[k + "_flag"])
