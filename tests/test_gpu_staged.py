"""Pressure re-evaluated in every RK stage on the GPU (ic_K0 = 1: K0
active / passive by the sign of the velocity divergence, "6 SPH_SW_2019.f90":
6736-6760, 6857-6863; ic_cosTheta = 1: g along the terrain normal,
6:6774-6783, 9264-9270), beyond the reference fixtures that
test_gpu_parity.py checks (g_k0, g_costheta_k0, g_k0_pwp): RK2, adaptive dt
and erosion against the oracle; the staged path with K0 = 1 equal, bit for
bit, to the plain step; a batch mixing staged and plain realisations equal
to each run alone."""
import copy

import numpy as np
import pytest

import goldens
from gfsph import AdjointModel, BatchModel, GfsphError, SPHModel
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-10
TOL_EARLY = 1e-8


@pytest.mark.parametrize("name,integrator,adapt", [("g_k0", 2, 0), ("g_costheta_k0", 2, 2), ("g_k0_pwp", 4, 10)])
def test_staged_matches_oracle(name, integrator, adapt):
    case, _ = goldens.case_of(name)
    case.params.integrator = integrator
    case.params.adapt_dt = adapt
    st = oracle.init(case)
    with SPHModel(case) as m:
        for chunk in (1, 9, 40):
            m.time_integration_sph(chunk)
            oracle.run(case, st, chunk)
            got = m.state()
            tol = TOL_EARLY if chunk < 40 else TOL
            for k in ("x", "v", "h", "u", "hsml"):
                err = goldens.rel_err(got[k], getattr(st, k))
                assert err <= tol, f"{name} rk{integrator} adapt{adapt} {k}: {err:.3e}"
            assert np.array_equal(got["flag"], st.flag)


def test_staged_with_erosion_matches_oracle():
    case, _ = goldens.case_of("g_costheta_k0")
    case.params.const.update({21: 1, 22: 0.002})
    st = oracle.init(case)
    with SPHModel(case) as m:
        m.time_integration_sph(30)
        oracle.run(case, st, 30)
        got = m.state()
        for k in ("x", "v", "h", "u"):
            assert goldens.rel_err(got[k], getattr(st, k)) <= TOL_EARLY, k
        assert goldens.rel_err(m.mass(), st.mass) <= TOL_EARLY
        e = m.erosion()
        assert goldens.rel_err(e["eroded"], st.eros[:, 1]) <= 1e-8


def test_unit_k0_equals_plain_step():
    """K0_active = K0_pasive = 1: the per-stage pressures equal the step's,
    so the staged kernels reproduce the plain step exactly."""
    case, _ = goldens.case_of("g_frictional_pwp")
    staged = copy.deepcopy(case)
    staged.params.ic_k0, staged.params.k0_active, staged.params.k0_pasive = 1, 1.0, 1.0
    with SPHModel(case) as a, SPHModel(staged) as b:
        a.time_integration_sph(40)
        b.time_integration_sph(40)
        sa, sb = a.state(), b.state()
        for k in ("x", "v", "h", "u", "uaux", "hsml", "flag"):
            np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)


def test_batch_mixed_equals_single_runs():
    base, _ = goldens.case_of("g_costheta_k0")
    cases = []
    for r, cos in enumerate((1, 0, 1)):
        c = copy.deepcopy(base)
        c.params.const[16] = cos
        c.params.const[9] = float(np.tan(np.radians(12.0 + 2.0 * r)))
        c.particles.x = c.particles.x + np.array([0.4 * r, -0.3 * r])
        cases.append(c)
    with BatchModel(cases) as b:
        b.time_integration_sph(30)
        per = b.split(b.state())
    for r, c in enumerate(cases):
        with SPHModel(c) as m:
            m.time_integration_sph(30)
            st = m.state()
            for key in ("x", "v", "h", "u", "hsml", "flag"):
                np.testing.assert_array_equal(per[r][key], st[key], err_msg=f"realisation {r} {key}")


def test_rejected():
    case, _ = goldens.case_of("g_bingham")
    c = copy.deepcopy(case)
    c.params.const.update({11: -1.0, 16: 1})
    with pytest.raises(GfsphError) as e:
        SPHModel(c)
    assert e.value.code == -2
    c = copy.deepcopy(case)
    c.params.const.update({11: -1.0, 16: 2})
    with pytest.raises(GfsphError) as e:
        SPHModel(c)
    assert e.value.code == -2
    k0, _ = goldens.case_of("g_k0")
    with AdjointModel(k0) as m:
        with pytest.raises(GfsphError) as e:
            m.forward(5, 5)
        assert e.value.code == -2
