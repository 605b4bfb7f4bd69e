"""Pore pressure on a column of nUaux nodes per particle on the GPU (icpwp =
2: Initialize_Pwp "6 SPH_SW_2019.f90":1743-1800, Get_Pwp_SW 6:10966-11174)
beyond the reference fixtures that test_gpu_parity.py checks
(g_pwp_column_*): RK2 and adaptive dt against the oracle, columns handed
over by the caller (the reference's Init_sph state), a batch against its
realisations run alone, the rejected configurations."""
import copy
import ctypes

import numpy as np
import pytest

import goldens
from gfsph import BatchModel, GfsphError, SPHModel
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-10
TOL_EARLY = 1e-8


def _compare(m, st, tol, what):
    got = m.state()
    for k in ("x", "v", "h", "u", "hsml"):
        err = goldens.rel_err(got[k], getattr(st, k))
        assert err <= tol, f"{what} {k}: {err:.3e}"
    assert np.array_equal(got["flag"], st.flag)
    err = goldens.rel_err(m.pwp_column(), st.uaux)
    assert err <= tol, f"{what} column: {err:.3e}"
    np.testing.assert_array_equal(got["uaux"], m.pwp_column()[:, 0])


@pytest.mark.parametrize("name,integrator,adapt", [("g_pwp_column_cos", 2, 0), ("g_pwp_column_lin", 2, 2),
                                                   ("g_pwp_column_step", 4, 10)])
def test_column_matches_oracle(name, integrator, adapt):
    case, _ = goldens.case_of(name)
    case.params.integrator = integrator
    case.params.adapt_dt = adapt
    st = oracle.init(case)
    with SPHModel(case) as m:
        _compare(m, st, 1e-13, f"{name} init")
        for chunk in (1, 9, 40):
            m.time_integration_sph(chunk)
            oracle.run(case, st, chunk)
            _compare(m, st, TOL_EARLY if chunk < 40 else TOL, f"{name} rk{integrator} adapt{adapt}")


def test_columns_from_the_caller():
    """init_input = 0 (the state Init_sph left): the caller's columns are
    taken as they are, node 1 drives u and the friction."""
    case, _ = goldens.case_of("g_pwp_column_cos")
    st = oracle.init(case)
    rng = np.random.Generator(np.random.Philox(8))
    st.uaux = st.uaux * rng.uniform(0.5, 1.0, st.uaux.shape)
    st.u = st.uaux[:, 0].copy()
    c2 = copy.deepcopy(case)
    c2.particles.h = st.h.copy()
    c2.particles.u = st.u.copy()
    with SPHModel(c2, init_input=False) as m:
        m.set_pwp_column(st.uaux)
        np.testing.assert_array_equal(m.pwp_column(), st.uaux)
        m.time_integration_sph(25)
        oracle.run(case, st, 25)
        _compare(m, st, TOL_EARLY, "caller columns")


def test_batch_equals_single_runs():
    base, _ = goldens.case_of("g_pwp_column_lin")
    cases = []
    for r, (cv, phi) in enumerate(((0.01, 28.0), (0.2, 30.0), (0.05, 26.0))):
        c = copy.deepcopy(base)
        c.params.const[13] = np.pi ** 2 * cv / 4.0
        c.params.const[9] = np.tan(np.radians(phi))
        c.particles.x = c.particles.x + np.array([0.7 * r, 0.2 * r])
        cases.append(c)
    with BatchModel(cases) as b:
        b.time_integration_sph(30)
        per = b.split(b.state())
        cols = np.split(b.pwp_column(), np.cumsum(b.sizes)[:-1])
    for r, c in enumerate(cases):
        with SPHModel(c) as m:
            m.time_integration_sph(30)
            st = m.state()
            for key in ("x", "v", "h", "u", "uaux", "hsml", "flag"):
                np.testing.assert_array_equal(per[r][key], st[key], err_msg=f"realisation {r} {key}")
            np.testing.assert_array_equal(cols[r], m.pwp_column())


def test_rejected_configurations():
    case, _ = goldens.case_of("g_pwp_column_lin")
    for field, value in (("nuaux", 2), ("pwp_bc_type", (3, 1)), ("pwp_bc_val", (-999.0, 0.0)), ("pwp0_type", 5)):
        c = copy.deepcopy(case)
        setattr(c.params, field, value)
        with pytest.raises(GfsphError) as e:
            SPHModel(c)
        assert e.value.code in (-1, -2), field
    plain, _ = goldens.case_of("g_frictional_pwp")
    with SPHModel(plain) as m:
        buf = np.zeros(plain.particles.n)
        with pytest.raises(GfsphError) as e:
            m._check(m.lib.gfsph_get_pwp_column(m.ctx, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        assert e.value.code == -1
