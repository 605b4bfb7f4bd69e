"""The oracle's Bingham-erosion law (nfric 3, "6 SPH_SW_2019.f90":9443-9488)
beyond what the reference can pin.

The reference's own run of nfric 3 is pinned by the fixture g_nfric3_erosion
only while the eroded fraction nsoil = (m - m0)/m stays <= 0.01, where the
law is Manning's with n = const(8).  Past that, basal_friction_Ext_SW reads
its yield stress from a local Tauy0 (declared at 6:9377 for law 201) that
hides ExtForces_SW's tauy0 = const(6) and is never set on this branch: the
reference's result is undefined there (its run turns to NaN and faults).
The oracle takes the evident tauy0 = const(6); its three Bingham regimes are
pinned here by equivalence with laws the reference fixtures pin:

* intermediate regime (0.01 <= a_hat <= 50)  ==  nfric 4 (g_bingham) with
  tau_y = const(6) f, K = const(7) f and density 1000 + (rho - 1000) nsoil,
  f = 1 - exp(-const(3) nsoil);
* viscous regime (a_hat > 50)  ==  nfric 2 with const(7) = 24 K f (its
  K/(8 rho) is 3 K f / rho to the bit: scaling by 8 is exact);
* cohesive regime (a_hat < 0.01)  ==  nfric 24 with K = 0;
* nsoil <= 0.01  ==  nfric 1 (Manning) with const(3) = const(8).

Each comparison is one particle on the 30-degree plane, one RK4 step with
the regime held in every stage, compared bit for bit.  The constants are
passed unrounded (the oracle takes const(:) as doubles) so that the
transformed constants equal the products law 3 forms.
"""
import ctypes as C
import math

import numpy as np
import pytest

from gfsph import cases as Cs
from gfsph.dem import plane_runout_dem
from oracle import oracle


def _case(nfric, const):
    dem = plane_runout_dem(nx=81, ny=41, d=5.0, x_break=250.0, blend=50.0)
    parts = Cs.lattice_block(99.0, 99.0, 2.0, 2.0, 1, 1, 2.0, 1.5)
    parts.v[:] = [5.0, 1.0]
    prm = Cs.Params(integrator=4, dt=0.02, h_inf=1e-3, average_velocity=0, const={**const, 5: nfric})
    return Cs.Case("law", prm, dem, parts, 1)


def _run(nfric, cst, mass0_frac=1.0, nsteps=1):
    """One particle, nsteps RK4 steps with const(:) given unrounded."""
    case = _case(nfric, {1: 9.81, 2: 2000.0, 10: 1e-2, 14: 1.0, 15: 1e-2})
    st = oracle.init(case)
    st.mass0 = st.mass * mass0_frac
    p = oracle.make_params(case.params)
    for k, v in cst.items():
        p.cst[k] = v
    p.mass0 = oracle._dp(st.mass0)
    dh = oracle.DemHandle(case.dem)
    rc = oracle.lib().gfo_run(C.byref(p), C.byref(dh.s), 1, oracle._dp(st.x), oracle._dp(st.v), oracle._dp(st.h),
                              oracle._dp(st.mass), oracle._dp(st.hsml), oracle._dp(st.u), oracle._dp(st.uaux),
                              oracle._ip(st.flag), nsteps, oracle._dp(st.tau0), None, None)
    assert rc == 0
    return st


def _soil(st_mass, mass0_frac, c3, tau, K, rho):
    m = float(st_mass[0])
    m0 = m * mass0_frac
    ns = (m - m0) / m
    f = 1.0 - math.exp(-c3 * ns)
    return ns, K * f, tau * f, 1000.0 + (rho - 1000.0) * ns


def _a_hat(st0, K, tau):
    """The regime parameter 6 K |hv| / (tau h^2) at the step start."""
    h = float(st0.h[0])
    hv = h * float(np.hypot(*st0.v[0]))
    return 6.0 * K * hv / (tau * max(h, 1e-2) ** 2)


def _same(a, b):
    for k in ("x", "v", "h", "u"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)


def _base(c3, tau, K, rho=2000.0):
    return {2: rho, 3: c3, 6: tau, 7: K, 8: 0.0}


@pytest.mark.parametrize("frac", [0.8, 0.5])
def test_law3_intermediate_regime_is_bingham(frac):
    c3, tau, K = 5.0, 400.0, 30.0
    st3 = _run(3, _base(c3, tau, K), frac)
    st0 = oracle.init(_case(3, {1: 9.81, 2: 2000.0, 10: 1e-2, 14: 1.0, 15: 1e-2}))
    ns, KB, tB, rB = _soil(st0.mass, frac, c3, tau, K, 2000.0)
    a = _a_hat(st0, KB, tB)
    assert ns > 0.01 and 0.05 < a < 20, (ns, a)
    st4 = _run(4, {2: rB, 6: tB, 7: KB, 8: 1.0})
    _same(st3, st4)
    assert not np.array_equal(st3.v, _run(3, _base(c3, tau, K), 1.0).v)  # the Bingham term acts


def test_law3_viscous_regime_is_laminar():
    c3, tau, K, frac = 3.0, 10.0, 2000.0, 0.7
    st0 = oracle.init(_case(3, {1: 9.81, 2: 2000.0, 10: 1e-2, 14: 1.0, 15: 1e-2}))
    ns, KB, tB, rB = _soil(st0.mass, frac, c3, tau, K, 2000.0)
    assert _a_hat(st0, KB, tB) > 500
    _same(_run(3, _base(c3, tau, K), frac), _run(2, {2: rB, 7: 24.0 * KB}))


def test_law3_cohesive_regime_is_yield_stress():
    c3, tau, K, frac = 4.0, 5000.0, 1e-3, 0.6
    st0 = oracle.init(_case(3, {1: 9.81, 2: 2000.0, 10: 1e-2, 14: 1.0, 15: 1e-2}))
    ns, KB, tB, rB = _soil(st0.mass, frac, c3, tau, K, 2000.0)
    assert _a_hat(st0, KB, tB) < 1e-3
    _same(_run(3, _base(c3, tau, K), frac), _run(24, {2: rB, 6: tB, 7: 0.0, 8: 1.0}))


def test_law3_below_the_soil_threshold_is_manning():
    n = 0.05
    _same(_run(3, {3: 5.0, 6: 400.0, 7: 30.0, 8: n}, 0.995), _run(1, {3: n}))


def test_law3_manning_bounds_the_bingham_term():
    """A Manning term larger than the Bingham one replaces it component by
    component (6:9483-9486)."""
    c3, tau, K, frac = 5.0, 400.0, 30.0, 0.8
    weak = _run(3, _base(c3, tau, K), frac)
    strong = _run(3, {**_base(c3, tau, K), 8: 0.5}, frac)
    manning = _run(1, {3: 0.5})
    _same(strong, manning)
    assert not np.array_equal(weak.v, strong.v)
