"""The oracle's restatement of the reference's case loop (oracle.run_case:
Driver_Flow_23, "8 SPH_FEM_Driver_2019.f90":1444-1548, and Driver_Flow,
8:1554-1710, with Stop_case_SW, "6 SPH_SW_2019.f90":20114-20177) against the
reference's own main program: oracle/_ref/gfsph_ref_main ran two MASTER.dat
case tables (ic_cases_mat 2 and 1) over the plane fixture's release and
recorded every step of every problem (tests/golden/g_case_driver_ref.npz).
For each problem: the number of steps, the stop flag, time_sph (equal to
the last bit where the steps are cut at plot times) and the final state.

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


def problems(tmp_path, rows, opts):
    """The cases of every problem of a case table: the release and const(:)
    as Input_SW changes them for the problem (const indices, 501 depth
    factor), read back from the table the host reader expands."""
    case, _ = goldens.case_of("g_voellmy_plane")
    refio.write_inputs(str(tmp_path), case, name="case", topo="topo")
    refio.write_master(str(tmp_path), "case", case.params.dt, case_rows=rows, **opts)
    with Inputs(str(tmp_path), "case", "topo", build=False) as inp:
        t = inp.cases()
    T, nvar = t["table"], t["ncases"]
    out = []
    for p in range(T.shape[0]):
        c = copy.deepcopy(case)
        c.params.adapt_dt = opts["adapt"]
        for v in range(nvar):
            var, val = int(rows[v][0]), float(T[p, 3 + v])
            if var == 501:
                c.particles.h = c.particles.h * val
                c.particles.mass = c.particles.mass * val
            else:
                c.params.const[var] = float(np.float32(val))
        out.append(c)
    return out, t


@pytest.mark.parametrize("tname", ["t2", "t1"])
def test_oracle_case_loop_matches_reference_main(tname, tmp_path):
    from make_golden import CASE_TABLES

    f = goldens.load("g_case_driver_ref")
    rows, opts = CASE_TABLES[tname]
    cs, t = problems(tmp_path, rows, opts)
    assert len(cs) == int(f[f"{tname}_nprob"])
    time_plot = opts["plot_step"] * opts["dt_driver"] if opts["cases"] == 2 else 0.0
    vstop, timestop, dtimestop = opts["stop"]
    carried = None
    for p, c in enumerate(cs):
        k = f"{tname}_p{p + 1}"
        st = oracle.init(c)
        if opts["cases"] == 1 and carried is not None:
            st.dt_time = np.array([carried, 0.0])
        it, stopped, _, _ = oracle.run_case(c, st, vstop, timestop, dtimestop, opts["dt_driver"], opts["time_end"],
                                            time_plot, max_steps=10000, pixels=False)
        carried = float(st.dt_time[0])
        assert it == int(f[k + "_iters"]), (p, it)
        assert stopped == bool(f[k + "_stop"][-1]), p
        if opts["cases"] == 2:
            assert float(st.dt_time[1]) == float(f[k + "_time"][-1]), p
        else:
            assert float(st.dt_time[1]) == pytest.approx(float(f[k + "_time"][-1]), rel=2e-15), p
        for fld in ("x", "v", "h", "u"):
            err = goldens.rel_err(getattr(st, fld), f[k + "_" + fld])
            assert err <= 1e-10, (p, fld, err)
        np.testing.assert_array_equal(st.flag, f[k + "_flag"])
