"""The oracle's functions that no reference fixture covers, pinned by what
the mathematics fixes.

* RK2 has no counterpart in the reference (its integrators are RK4 and
  variants, "7 SPH_MAIN_2019.f90":242-245): Heun's method on the RK45
  skeleton is exact for constant acceleration and second order for a
  velocity-dependent one.
* The IDW node raster of OutputRes_WIR ("6 SPH_SW_2019.f90":14312-14466) and
  the Monte Carlo pixel accumulators of Get_MC_pixel_Vars_SW (6:20717-20833)
  have closed forms for one particle and for constant fields.
* Central differences (Fdiff_derivs, "2 SPH_FEM_topo_2019.f90":2812-2877)
  and the bilinear lookup (Get_Z_Topo, 2:2395-2491) are exact on planes.
* The kernels (kernel_TOOLS, "5a My_SPH_Tools_2019.f90":2062-2180) integrate
  to one, and their gradients are odd in the separation.
"""
import math

import numpy as np
import pytest

from gfsph import cases as C
from gfsph.dem import DEM, fdiff_derivs
from oracle import oracle


def plane_dem(nx=60, ny=40, d=5.0, a=-0.45, b=0.1, c=200.0):
    xs = np.arange(nx) * d
    ys = np.arange(ny) * d
    X, Y = np.meshgrid(xs, ys)
    return DEM(Z=a * X + b * Y + c, xmin=0.0, ymin=0.0, dx=d, dy=d)


def isolated_particles(points, vel, h=3.0, hsml=1.5):
    n = len(points)
    x = np.array(points, dtype=np.float64)
    v = np.array(vel, dtype=np.float64)
    area = (hsml / 1.5) ** 2
    return C.Particles(x=x, v=v, h=np.full(n, h), mass=np.full(n, area * h), hsml=np.full(n, hsml),
                       u=np.zeros(n), itype=np.ones(n, np.int32))


def run_rk2(case, dt, steps):
    case.params.dt = dt
    st = oracle.init(case)
    oracle.run(case, st, steps)
    return st


def test_rk2_exact_for_constant_acceleration():
    """Particles far apart (no neighbours) moving down the fall line of a
    plane under Coulomb friction (tan(phi) < slope, no turbulence term)
    have a constant acceleration, which Heun's method integrates exactly:
    the positions after time T do not depend on dt, and the velocity gain is
    linear in time."""
    dem = plane_dem(a=-0.7, b=0.0)
    pts = [(40.0 + 20.0 * k, 60.0 + 25.0 * k) for k in range(4)]
    parts = isolated_particles(pts, [(2.0 + k, 0.0) for k in range(4)])
    prm = C.Params(integrator=2, dt=0.05, average_velocity=0, const=C.frictional(20.0))
    case = C.Case("rk2_const", prm, dem, parts, 0)
    a = run_rk2(case, 0.05, 20)
    b = run_rk2(case, 0.025, 40)
    half = run_rk2(case, 0.05, 10)
    assert np.all(a.flag == 0)
    assert np.max(np.abs(a.x - b.x)) <= 1e-11 * np.max(np.abs(a.x))
    gain_T = a.v[:, 0] - parts.v[:, 0]
    gain_half = half.v[:, 0] - parts.v[:, 0]
    assert np.all(gain_T > 1.0)
    np.testing.assert_allclose(gain_T, 2.0 * gain_half, rtol=1e-12)
    np.testing.assert_allclose(a.v[:, 1], 0.0, atol=1e-12)


def test_rk2_is_second_order():
    """With a velocity-dependent (Voellmy) resistance and curved 2-D paths,
    the RK2 error against a converged RK4 solution falls by ~4 when dt is
    halved (dt = 0.02, 0.01, 0.005 s; at larger dt the reversal limiter is
    active and the error is not yet asymptotic)."""
    dem = plane_dem(a=-0.5, b=0.15)
    pts = [(60.0 + 30.0 * (k % 3), 50.0 + 30.0 * (k // 3)) for k in range(6)]
    vel = [(4.0 * math.cos(k), 4.0 * math.sin(k)) for k in range(6)]
    parts = isolated_particles(pts, vel)
    T = 2.0
    ref_case = C.Case("rk4_ref", C.Params(integrator=4, dt=T / 3200, average_velocity=0,
                                          const=C.voellmy(0.2, 200.0)), dem, parts, 0)
    ref = oracle.init(ref_case)
    oracle.run(ref_case, ref, 3200)
    errs = []
    for m in (100, 200, 400):
        case = C.Case("rk2", C.Params(integrator=2, dt=T / m, average_velocity=0, const=C.voellmy(0.2, 200.0)),
                      dem, parts, 0)
        st = run_rk2(case, T / m, m)
        errs.append(np.max(np.abs(st.x - ref.x)))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.3 < r1 < 4.7 and 3.3 < r2 < 4.7, errs


def state_of(x, v, h, u):
    n = len(h)
    st = oracle.State(C.Particles(x=np.array(x, float), v=np.array(v, float), h=np.array(h, float),
                                  mass=np.ones(n), hsml=np.ones(n), u=np.array(u, float),
                                  itype=np.ones(n, np.int32)))
    return st


def test_idw_raster_single_particle_closed_form():
    dem = DEM(Z=np.zeros((10, 12)), xmin=100.0, ymin=50.0, dx=2.0, dy=3.0)
    px, py = 107.3, 61.9
    st = state_of([[px, py]], [[1.5, -0.5]], [2.5], [0.7])
    out, ws = oracle.raster_idw(dem, st)
    distg = math.hypot(dem.dx, dem.dy)
    ix, iy = int((px - dem.xmin) // dem.dx), int((py - dem.ymin) // dem.dy)
    xy = dem.node_xy()
    for g in range(dem.nx * dem.ny):
        gx, gy = g % dem.nx, g // dem.nx
        inside = ix - 1 <= gx <= ix + 2 and iy - 1 <= gy <= iy + 2
        if not inside:
            assert ws[g] == 0.0 and np.all(out[g] == 0.0)
            continue
        d = max(1e-6, math.hypot(px - xy[g, 0], py - xy[g, 1])) / distg
        assert ws[g] == pytest.approx(1.0 / d, rel=1e-14)
        np.testing.assert_allclose(out[g], [2.5, 1.5, -0.5, 0.7 / 2.5], rtol=1e-14)


def test_idw_raster_reproduces_constant_fields():
    dem = DEM(Z=np.zeros((20, 20)), xmin=0.0, ymin=0.0, dx=5.0, dy=5.0)
    rng = np.random.default_rng(3)
    n = 300
    x = rng.uniform(5.0, 90.0, size=(n, 2))
    st = state_of(x, np.tile([[0.8, -2.0]], (n, 1)), np.full(n, 1.7), np.full(n, 0.51))
    out, ws = oracle.raster_idw(dem, st)
    hit = ws >= 1e-6
    assert hit.sum() > 200
    np.testing.assert_allclose(out[hit], np.tile([1.7, 0.8, -2.0, 0.3], (hit.sum(), 1)), rtol=1e-13)
    assert np.all(out[~hit] == 0.0)


def test_mc_pixel_closed_forms():
    dem = DEM(Z=np.zeros((8, 9)), xmin=0.0, ymin=0.0, dx=4.0, dy=5.0)
    px, py = 13.0, 22.0
    st1 = state_of([[px, py]], [[3.0, 4.0]], [2.0], [0.0])
    acc, cnt = oracle.mc_pixel_accum(dem, st1)
    ix, iy = int(px // dem.dx), int(py // dem.dy)
    corners = {iy * dem.nx + ix, iy * dem.nx + ix + 1, (iy + 1) * dem.nx + ix, (iy + 1) * dem.nx + ix + 1}
    for g in range(dem.nx * dem.ny):
        if g in corners:
            assert cnt[g] == 1
            np.testing.assert_array_equal(acc[g], [2.0, 2.0, 5.0, 5.0])
        else:
            assert cnt[g] == 0 and np.all(acc[g] == 0.0)
    st2 = state_of([[px, py]], [[6.0, 8.0]], [1.5], [0.0])
    acc, cnt = oracle.mc_pixel_accum(dem, st2, acc, cnt)
    for g in corners:
        assert cnt[g] == 2
        np.testing.assert_array_equal(acc[g], [3.5, 2.0, 15.0, 10.0])


def test_central_differences_and_lookup_exact_on_a_plane():
    a, b = -0.45, 0.1
    dem = plane_dem(a=a, b=b)
    T = fdiff_derivs(dem.Z, dem.dx, dem.dy).reshape(dem.ny, dem.nx, 9)
    inner = T[1:-1, 1:-1]
    np.testing.assert_allclose(inner[:, :, 1], a, rtol=1e-12)
    np.testing.assert_allclose(inner[:, :, 2], b, rtol=1e-12)
    np.testing.assert_allclose(inner[:, :, 3], 1 + a * a, rtol=1e-12)
    np.testing.assert_allclose(inner[:, :, 4], a * b, rtol=1e-11)
    np.testing.assert_allclose(inner[:, :, 6:9], 0.0, atol=1e-12)
    rng = np.random.default_rng(5)
    xy = np.stack([rng.uniform(dem.dx * 1.2, dem.xmax - dem.dx * 1.2, 50),
                   rng.uniform(dem.dy * 1.2, dem.ymax - dem.dy * 1.2, 50)], axis=1)
    out = oracle.topo_lookup(dem, xy)
    np.testing.assert_allclose(out[:, 1], a, rtol=1e-12)
    np.testing.assert_allclose(out[:, 2], b, rtol=1e-12)


@pytest.mark.parametrize("skf", [1, 2, 3, 4])
def test_kernels_normalised_and_gradient_odd(skf):
    h = 1.3
    support = (2.0 if skf in (1, 4) else 3.0) * h
    nr = 4000
    rs = (np.arange(nr) + 0.5) * support / nr
    ws = np.array([oracle.kernel(skf, r, r, 0.0, h)[0] for r in rs])
    integral = float(np.sum(ws * 2 * np.pi * rs) * support / nr)
    # Gauss (skf 2) is truncated at 3h in the reference: 1 - exp(-9)
    expect = 1.0 - math.exp(-9.0) if skf == 2 else 1.0
    assert integral == pytest.approx(expect, rel=2e-6)
    for r, ang in ((0.4 * h, 0.3), (1.3 * h, 2.0), (1.9 * h, -1.1)):
        dx, dy = r * math.cos(ang), r * math.sin(ang)
        _, gx, gy = oracle.kernel(skf, r, dx, dy, h)
        _, gx2, gy2 = oracle.kernel(skf, r, -dx, -dy, h)
        assert gx == pytest.approx(-gx2, rel=1e-14) and gy == pytest.approx(-gy2, rel=1e-14)
