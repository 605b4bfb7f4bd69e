"""Erosion on the GPU (const(21) laws 1-3, "6 SPH_SW_2019.f90":8664-8812 and
Get_Erosion_Rate_SW 10289-10404) beyond the reference fixtures, which
test_gpu_parity.py checks (g_erosion_hungr, g_erosion_slope,
g_erosion_stress): RK2 and adaptive dt against the oracle, one particle
(its depth grows too), the terrain state set from outside (restart.top), a
batch against its realisations run alone, and the rejected combinations.

The terrain sums are atomic on the GPU, so the eroded depths agree with the
oracle to rounding (1e-12 field-normwise) rather than bit for bit; they do
not feed back into the particles, whose state in a batch equals the single
runs exactly."""
import copy

import numpy as np
import pytest

import goldens
from gfsph import AdjointModel, BatchModel, GfsphError, SPHModel
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
    assert goldens.rel_err(m.mass(), st.mass) <= tol, what
    e = m.erosion()
    for j, k in enumerate(("z", "eroded", "avail")):
        err = goldens.rel_err(e[k], st.eros[:, j])
        assert err <= max(tol, 1e-12), f"{what} {k}: {err:.3e}"


@pytest.mark.parametrize("name,integrator,adapt", [("g_erosion_hungr", 2, 0), ("g_erosion_slope", 2, 4),
                                                   ("g_erosion_stress", 4, 2), ("g_erosion_hungr", 4, 10),
                                                   ("g_nfric3_erosion", 4, 0), ("g_nfric3_erosion", 2, 0)])
def test_erosion_matches_oracle(name, integrator, adapt):
    """50 steps against the oracle; for the Bingham-erosion law (nfric 3) this
    goes past the reference fixture's 12 steps, into the Bingham regimes the
    oracle's law-equivalence tests pin (test_oracle_laws.py)."""
    case, _ = goldens.case_of(name)
    case.params.integrator = integrator
    case.params.adapt_dt = adapt
    st = oracle.init(case)
    with SPHModel(case) as m:
        for chunk in (1, 9, 40):
            m.time_integration_sph(chunk)
            oracle.run(case, st, chunk)
            _compare(m, st, TOL_EARLY if chunk < 40 else TOL, f"{name} rk{integrator} adapt{adapt}")


def test_single_particle_depth_grows():
    """npoin_s = 1: the eroded depth is added to the particle's depth as well
    (6:8756-8758)."""
    case, _ = goldens.case_of("g_erosion_hungr")
    P = case.particles
    k = P.n // 2
    for a in ("x", "v", "h", "mass", "hsml", "u", "itype"):
        setattr(P, a, getattr(P, a)[k:k + 1].copy())
    P.v[:] = [4.0, 0.5]
    st = oracle.init(case)
    h0 = st.h.copy()
    with SPHModel(case) as m:
        m.time_integration_sph(20)
        oracle.run(case, st, 20)
        _compare(m, st, TOL_EARLY, "one particle")
    assert st.mass[0] > P.mass[0]


def test_erosion_state_from_outside():
    """gfsph_erosion_set (the reference's restart.top): a thin available
    layer stops the terrain update where it runs out."""
    case, _ = goldens.case_of("g_erosion_stress")
    npg = case.dem.nx * case.dem.ny
    rng = np.random.Generator(np.random.Philox(4))
    avail = rng.uniform(0.0, 2e-3, npg)
    eroded = rng.uniform(0.0, 1.0, npg)
    st = oracle.init(case)
    st.eros[:, 1] = eroded
    st.eros[:, 2] = avail
    with SPHModel(case) as m:
        m.set_erosion(0, eroded=eroded, avail=avail)
        m.time_integration_sph(30)
        oracle.run(case, st, 30)
        _compare(m, st, TOL_EARLY, "restart.top")
    assert np.sum(st.eros[:, 2] == 0.0) > 0


def test_batch_equals_single_runs():
    """Realisations with different erosion constants (one without erosion)
    in one batch: particle states and masses exactly as run alone, terrain
    states to rounding."""
    base, _ = goldens.case_of("g_erosion_hungr")
    cases = []
    for r, (k, c48) in enumerate(((0.002, 0.05), (0.004, 0.0), (0.0, 0.0), (0.001, 0.01))):
        c = copy.deepcopy(base)
        c.params.const[22] = k
        c.params.const[48] = c48
        if k == 0.0:
            c.params.const[21] = 0.0
        c.particles.x = c.particles.x + np.array([0.5 * r, -0.3 * r])
        cases.append(c)
    with BatchModel(cases) as b:
        b.time_integration_sph(40)
        per = b.split(b.state())
        mb = np.split(b.mass(), np.cumsum(b.sizes)[:-1])
        eb = [b.erosion(r) for r in range(b.R)]
    for r, c in enumerate(cases):
        with SPHModel(c) as m:
            m.time_integration_sph(40)
            st = m.state()
            for key in ("x", "v", "h", "u", "hsml", "flag"):
                np.testing.assert_array_equal(per[r][key], st[key], err_msg=f"realisation {r} {key}")
            np.testing.assert_array_equal(mb[r], m.mass())
            if c.params.const[21] > 0:
                e = m.erosion()
                for k in ("z", "eroded", "avail"):
                    assert goldens.rel_err(eb[r][k], e[k]) <= 1e-12, (r, k)
            else:
                with pytest.raises(GfsphError):
                    m.erosion()
    # the realisation without erosion keeps its terrain
    np.testing.assert_array_equal(eb[2]["eroded"], 0.0)


def test_rejected_combinations():
    case, _ = goldens.case_of("g_erosion_hungr")
    c = copy.deepcopy(case)
    c.params.const.update({5: 4, 6: 1000.0, 7: 50.0, 8: 1.0})  # Bingham
    with pytest.raises(GfsphError) as e:
        SPHModel(c)
    assert e.value.code == -2
    c = copy.deepcopy(case)
    c.params.const.update({21: 2, 17: 2650.0, 18: 1000.0, 5: 1})  # law 2 needs a frictional law
    with pytest.raises(GfsphError) as e:
        SPHModel(c)
    assert e.value.code == -2
    c = copy.deepcopy(case)
    c.params.const.update({21: 3, 17: 0.0})  # law 3 needs const(17), const(18)
    with pytest.raises(GfsphError) as e:
        SPHModel(c)
    assert e.value.code == -2
    plain, _ = goldens.case_of("g_voellmy_plane")
    with SPHModel(plain) as m:
        with pytest.raises(GfsphError) as e:
            m.erosion()
        assert e.value.code == -1
    with AdjointModel(case) as m:
        with pytest.raises(GfsphError) as e:
            m.forward(10, 5)
        assert e.value.code == -2
