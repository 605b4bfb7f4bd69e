"""Terrain zones on the GPU (ic_topo_props = 10: friction zones, 6:9697-9702
and 10515-10545; erosion zones, 6:10407-10503; Which_Topo_Zone 2:2635-2808)
beyond the reference fixtures that test_gpu_parity.py checks
(g_zones_friction, g_zones_erosion, g_zones_erosion_main): RK2 and adaptive
dt against the oracle, a batch of realisations with their own constants
against each realisation run alone, the terrain file read by the library's
own reader, and the combinations the library refuses."""
import copy

import numpy as np
import pytest

import goldens
from gfsph import AdjointModel, BatchModel, GfsphError, SPHModel
from oracle import oracle, refio

pytestmark = pytest.mark.gpu

TOL = 1e-10
TOL_EARLY = 1e-8


def _compare(m, st, tol, what, eros):
    got = m.state()
    for k in ("x", "v", "h", "u", "hsml"):
        err = goldens.rel_err(got[k], getattr(st, k))
        assert err <= tol, f"{what} {k}: {err:.3e}"
    assert np.array_equal(got["flag"], st.flag)
    if np.ndim(st.uaux) == 2:  # the whole pore-pressure column
        assert goldens.rel_err(m.pwp_column(), st.uaux) <= tol, what
    if eros:
        assert goldens.rel_err(m.mass(), st.mass) <= tol, what
        e = m.erosion()
        for j, k in enumerate(("z", "eroded", "avail")):
            err = goldens.rel_err(e[k], st.eros[:, j])
            assert err <= max(tol, 1e-12), f"{what} {k}: {err:.3e}"


@pytest.mark.parametrize("name,integrator,adapt", [("g_zones_friction", 2, 0), ("g_zones_friction", 4, 10),
                                                   ("g_zones_erosion", 2, 2), ("g_zones_erosion_main", 2, 0),
                                                   ("g_zones_pwp", 2, 0)])
def test_zones_match_oracle(name, integrator, adapt):
    case, _ = goldens.case_of(name)
    case.params.integrator = integrator
    case.params.adapt_dt = adapt
    st = oracle.init(case)
    eros = name in goldens.EROSION_CASES
    with SPHModel(case) as m:
        for chunk in (1, 9, 60):
            m.time_integration_sph(chunk)
            oracle.run(case, st, chunk)
            _compare(m, st, TOL_EARLY if chunk < 60 else TOL, f"{name} rk{integrator} adapt{adapt}", eros)


@pytest.mark.parametrize("name", ["g_zones_friction", "g_zones_erosion", "g_zones_erosion_main", "g_zones_pwp"])
def test_zones_batch_equals_single_runs(name):
    """Three realisations on the zoned terrain with their own tan(phi) and
    releases shifted along y: each equals its run alone bit for bit, its
    masses and erosion state included."""
    base, _ = goldens.case_of(name)
    cases = []
    for r in range(3):
        c = copy.deepcopy(base)
        cst = dict(c.params.const)
        cst[9] = float(cst[9]) * (1.0 + 0.1 * r)
        c.params.const = cst
        c.particles.x = c.particles.x + np.array([0.0, 4.0 * r])
        cases.append(c)
    nsteps = 60
    eros = name in goldens.EROSION_CASES
    with BatchModel(cases) as b:
        b.time_integration_sph(nsteps)
        per = b.split(b.state())
        mass = np.split(b.mass(), np.cumsum(b.sizes)[:-1]) if eros else None
        er = [b.erosion(r) for r in range(3)] if eros else None
    for r, case in enumerate(cases):
        with SPHModel(case) as m:
            m.time_integration_sph(nsteps)
            st = m.state()
            for k in ("x", "v", "h", "u", "hsml"):
                np.testing.assert_array_equal(per[r][k], st[k], err_msg=f"{name} realisation {r} field {k}")
            np.testing.assert_array_equal(per[r]["flag"], st["flag"])
            if eros:
                np.testing.assert_array_equal(mass[r], m.mass())
                e = m.erosion()
                for k in ("z", "eroded", "avail"):  # atomic sums: rounding, not bits
                    assert goldens.rel_err(er[r][k], e[k]) <= 1e-12, (name, r, k)


def test_terrain_file_zones_through_the_reader(tmp_path):
    """g_zones_erosion's files (the terrain's zone block, <terrain>.eros.pts,
    restart.dat, restart.top) read by gfsph_inputs_read and run by
    gfsph_inputs_create: the reference's state after 1, 10 and 100 steps."""
    from gfsph.inputs import Inputs
    case, f = goldens.case_of("g_zones_erosion")
    refio.write_inputs(str(tmp_path), case)
    refio.write_master(str(tmp_path), "case", case.params.dt, 0)
    with Inputs(str(tmp_path), "case", "topo") as inp, inp.model() as m:
        tz = inp.zones()
        np.testing.assert_array_equal(tz.nodes, case.dem.zones.nodes)
        np.testing.assert_array_equal(tz.avail, case.dem.zones.avail)
        done = 0
        for s in (1, 10, 100):
            m.time_integration_sph(s - done)
            done = s
            st = m.state()
            for k in ("x", "v", "h", "u"):
                err = goldens.rel_err(st[k], f[f"s{s}_{k}"])
                assert err <= (TOL if s >= 100 else TOL_EARLY), f"step {s} {k}: {err:.3e}"
            assert goldens.rel_err(m.mass(), f[f"s{s}_mass"]) <= (TOL if s >= 100 else TOL_EARLY)


def test_zone_combinations_refused():
    case, _ = goldens.case_of("g_zones_friction")
    # the adjoint does not cover terrain zones
    with pytest.raises(GfsphError) as e:
        with AdjointModel(case) as m:
            m.forward(5)
    assert e.value.code == -2
    # flux zones (type 4), the drained bed in a pore-pressure zone and
    # structures (geometry 102)
    from gfsph.dem import Zone
    for z, code in ((Zone(4, 0, (4, 0, 10, 0, 10)), -2), (Zone(3, 0, (4, 0, 10, 0, 10, 0, 1, -999.0)), -2),
                    (Zone(1, 0, (102, 0, 0, 0, 0, 0, 0.1)), -2),
                    (Zone(2, 5, (4, 0, 10, 0, 10)), -1)):
        c2 = copy.deepcopy(case)
        c2.dem.zones.zones = list(c2.dem.zones.zones) + [z]
        with pytest.raises(GfsphError) as e:
            SPHModel(c2)
        assert e.value.code == code, (z, e.value)
