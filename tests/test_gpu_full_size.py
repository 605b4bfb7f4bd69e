"""The workloads of the measurement at their full sizes, through the C ABI.

* The 10,011-particle plane release (RK4, dt 0.05 s) over 2,000 steps
  against the reference build, with the divergence growth printed step by
  step beside the reference's own build-to-build spread (the same Fortran
  built with fused multiply-adds, a rounding as valid as the first).
* The 200,000-particle fractal release (2048 x 2048 DEM, frictional + pore
  pressure) with the reference's integrator (RK4) against the reference
  build: neighbour sets of steps 1 and 100 (SHA-256 of the sorted pair
  lists), the state after 1, 10 and 100 steps; and with RK2 (no reference
  counterpart) against the oracle.
* The 256-realisation Monte Carlo batch on the 1024 x 1024 DEM: every
  sampled realisation equals its run alone bit for bit.
* The 50,000-particle adjoint workload on the 512 x 512 DEM: gradients
  against central differences and against the FP64 torch reference.

Tolerances are field-normwise (max|a-b| / max|b|, tests/goldens.py).  The
north-star bar is 1e-10 after 100 steps.  Where the reference's own two
builds already differ by more than that on a case (the plane release: its
flow front runs at Courant numbers above 1 and amplifies a rounding by
~1.18x per step), the bar is that spread: no FP64 evaluation can be held
closer to one build of the reference than its other build is.  The factor
2 below covers two independent roundings each as far from the reference as
the FMA build (their distances add at most).
"""
import hashlib

import numpy as np
import pytest

import goldens
from gfsph import AdjointModel, BatchModel, SPHModel
from gfsph import cases as C
from oracle import oracle

pytestmark = pytest.mark.gpu

BAR = 1e-10
SPREAD_FACTOR = 2.0
FIELDS = ("x", "v", "h", "u")


def _sha(ij):
    return hashlib.sha256(np.ascontiguousarray(ij, dtype=np.int32).tobytes()).hexdigest().encode()


def test_gpu_cfg1_divergence_growth():
    """cfg1 over 2,000 steps: GPU vs the reference build at every dump of the
    growth fixture, printed beside the FMA build's distance on the same
    particles.  Up to step 100 the GPU stays within max(1e-10, 2 x spread);
    the whole field at step 100 likewise.  Beyond step 200 the trajectories
    of the two reference builds themselves decorrelate; there the flow's
    integral measures (centroid, mean speed, front and tail, mean depth) are
    compared on the flow's own scale."""
    g = goldens.load("g_cfg1_growth")
    steps = [int(s) for s in g["steps"]]
    idx = g["idx"]
    case = C.cfg_plane_voellmy(nsteps=steps[-1])
    rows = []
    with SPHModel(case) as m:
        done = 0
        for s in steps:
            m.time_integration_sph(s - done)
            done = s
            st = m.state()
            act = st["flag"] == 0
            sp = np.sqrt((st["v"][act] ** 2).sum(axis=1))
            ints = np.array([st["x"][act, 0].mean(), st["x"][act, 1].mean(), sp.mean(), st["x"][act, 0].max(),
                             st["x"][act, 0].min(), st["h"][act].mean(), float(act.sum())])
            errs = {k: goldens.rel_err(st[k][idx], g[f"s{s}_{k}"]) for k in FIELDS}
            if s in (1, 10, 100):
                # per-element relative errors |a - b| / |b| over the sampled
                # particles (elements with b = 0 left out), median and maximum
                pe = []
                for k in ("x", "v", "h"):
                    a, b = np.ravel(st[k][idx]), np.ravel(g[f"s{s}_{k}"])
                    nz = b != 0
                    r = np.abs(a[nz] - b[nz]) / np.abs(b[nz])
                    pe.append(f"{k} median {np.median(r):.1e} max {r.max():.1e}")
                print(f"cfg1 step {s} per element: " + "; ".join(pe))
            fma = {k: goldens.rel_err(g[f"f{s}_{k}"], g[f"s{s}_{k}"]) for k in FIELDS}
            spread = {k: float(g[f"spread{s}_{k}"]) for k in FIELDS}
            full = None
            if s == 100:
                full = {k: goldens.rel_err(st[k], g[f"full100_{k}"]) for k in FIELDS}
            rows.append((s, errs, fma, spread, full, ints, g[f"s{s}_int"], g[f"f{s}_int"]))
    print("\ncfg1 divergence growth (field-normwise; sample of every 23rd particle)")
    print(f"{'step':>5} " + " ".join(f"{'gpu ' + k:>9} {'fma ' + k:>9}" for k in FIELDS) + "  whole-field fma v")
    for s, e, f, sp, full, *_ in rows:
        print(f"{s:5d} " + " ".join(f"{e[k]:9.2e} {f[k]:9.2e}" for k in FIELDS) + f"  {sp['v']:9.2e}"
              + ("" if full is None else "  gpu whole field: " + " ".join(f"{k} {full[k]:.2e}" for k in FIELDS)))
    for s, e, f, sp, full, ints, iref, ifma in rows:
        if s <= 100:
            for k in FIELDS:
                assert e[k] <= max(BAR, SPREAD_FACTOR * f[k], SPREAD_FACTOR * sp[k]), (s, k, e[k], f[k], sp[k])
        if full is not None:
            for k in FIELDS:
                assert full[k] <= max(BAR, SPREAD_FACTOR * sp[k]), (s, k, full[k], sp[k])
        if s >= 200:
            # The two reference builds' trajectories have decorrelated; the
            # flow is compared by integral measures instead (centroid x, y;
            # mean speed; front and tail x; mean depth), each within a
            # tolerance on the flow's own scale (E = front - tail), or the
            # FMA build's distance if larger.  Errors in the physics show up
            # here as metres of runout; rounding shows up as centimetres.
            E = iref[3] - iref[4]
            tol = np.array([2e-3 * E, 2e-3 * E, max(0.05 * abs(iref[2]), 1e-3), 1e-2 * E, 1e-2 * E,
                            1e-2 * abs(iref[5])])
            d_gpu = np.abs(ints[:6] - iref[:6])
            d_fma = np.abs(ifma[:6] - iref[:6])
            print(f"step {s}: extent {E:.1f} m; integral distances gpu {np.array2string(d_gpu, precision=3)} "
                  f"fma {np.array2string(d_fma, precision=3)}")
            assert np.all(d_gpu <= np.maximum(tol, d_fma)), (s, d_gpu, d_fma, tol)
            assert ints[6] == iref[6], s  # the same number of particles left in the domain


def test_gpu_cfg2_full_size_reference():
    """200,000 particles on the 2048 x 2048 fractal DEM, RK4: neighbour set of
    step 1 (from the seeded inputs) bit-exact, of step 100 (from the GPU's
    own state after 99 steps) equal to the reference's; the state after 1,
    10 and 100 steps within the bar (sampled particles and global sums)."""
    import gfsph

    f = goldens.load("g_cfg2_full")
    nsteps = int(f["nsteps"])
    case = C.cfg_fractal_frictional(integrator=4, nsteps=nsteps)
    idx = f["idx"]
    P = case.particles
    ij = gfsph.neighbour_pairs(P.x, P.hsml, np.zeros(P.n, np.int32))
    assert ij.shape[0] == int(f["s1_npairs"]) and _sha(ij) == bytes(f["s1_pairs_sha256"])
    print(f"\ncfg2 step 1: {ij.shape[0]} pairs, SHA-256 equal to the reference's")
    with SPHModel(case) as m:
        done = 0
        for s in [int(v) for v in f["dumps"]]:
            if s == nsteps:
                m.time_integration_sph(s - 1 - done)
                mid = m.state()
                ij = gfsph.neighbour_pairs(mid["x"], mid["hsml"], mid["flag"])
                print(f"cfg2 step {s}: {ij.shape[0]} pairs (reference {int(f[f's{s}_npairs'])})")
                assert ij.shape[0] == int(f[f"s{s}_npairs"]) and _sha(ij) == bytes(f[f"s{s}_pairs_sha256"])
                m.time_integration_sph(1)
            else:
                m.time_integration_sph(s - done)
            done = s
            st = m.state()
            assert int((st["flag"] == 0).sum()) == int(f[f"s{s}_nactive"])
            line = []
            for k in FIELDS:
                err = goldens.rel_err(st[k][idx], f[f"s{s}_{k}"])
                tot = np.sum(st[k], axis=0)
                serr = float(np.max(np.abs(tot - f[f"s{s}_{k}_sum"])) / max(np.max(np.abs(f[f"s{s}_{k}_sum"])), 1e-300))
                spread = float(f[f"spread{s}_{k}"])
                line.append(f"{k} {err:.2e} (sum {serr:.2e}, fma {spread:.2e})")
                bar = BAR if s >= 100 else 1e-8
                assert err <= max(bar, SPREAD_FACTOR * spread), (s, k, err, spread)
                assert serr <= max(bar, SPREAD_FACTOR * spread), (s, k, serr, spread)
            print(f"cfg2 step {s}: " + "; ".join(line))


def test_gpu_cfg2_rk2_matches_oracle():
    """The bench's integrator (RK2, no reference counterpart) at full size
    against the oracle over 10 steps: every particle."""
    case = C.cfg_fractal_frictional(integrator=2, nsteps=10)
    st = oracle.init(case)
    oracle.run(case, st, 10)
    with SPHModel(case) as m:
        m.time_integration_sph(10)
        got = m.state()
    for k in FIELDS:
        err = goldens.rel_err(got[k], getattr(st, k))
        print(f"cfg2 RK2 10 steps {k}: {err:.2e}")
        assert err <= 1e-8, (k, err)
    assert np.array_equal(got["flag"], st.flag)


def test_gpu_cfg3_full_batch_equals_single_runs():
    """All 256 realisations (5k-80k particles each, 7.4M in total) on the
    1024 x 1024 DEM as one batch for 100 steps with pixel statistics every
    20: the smallest, the median and the largest realisation and two
    random ones equal their runs alone bit for bit (state, pixel counts and
    maxima; the pixel sums to the rounding of their atomic order)."""
    cs = C.cfg_mc_hazard(R=256, nsteps=100)
    counts = np.array([c.particles.n for c in cs])
    order = np.argsort(counts)
    rng = np.random.default_rng(7)
    picks = sorted(set([int(order[0]), int(order[len(order) // 2]), int(order[-1])]
                       + [int(v) for v in rng.choice(256, 2, replace=False)]))
    with BatchModel(cs) as b:
        for _ in range(5):
            b.time_integration_sph(20)
            b.pixel_accumulate()
        per = b.split(b.state())
        pix = {r: b.pixel_read(r) for r in picks}
    print(f"\ncfg3: {len(cs)} realisations, {counts.sum()} particles; checked {picks} "
          f"(counts {[int(counts[r]) for r in picks]})")
    for r in picks:
        with SPHModel(cs[r]) as m:
            for _ in range(5):
                m.time_integration_sph(20)
                m.mc_pixel_accumulate()
            st = m.state()
            acc, cnt = m.mc_pixel_read()
        for k in ("x", "v", "h", "u", "hsml", "flag"):
            np.testing.assert_array_equal(per[r][k], st[k], err_msg=f"realisation {r} {k}")
        np.testing.assert_array_equal(pix[r][1], cnt)
        # counts and maxima are exact; the sums are accumulated with hardware
        # FP64 atomics, whose order (so the last bit) differs run to run
        np.testing.assert_array_equal(pix[r][0][:, [1, 3]], acc[:, [1, 3]])
        np.testing.assert_allclose(pix[r][0][:, [0, 2]], acc[:, [0, 2]], rtol=1e-13, atol=0)


def _adjoint_loss(case, target, nsteps):
    with SPHModel(case) as m:
        m.time_integration_sph(nsteps)
        out = m.raster_idw()[0][:, 0]
    return float(((out - target) ** 2).sum())


def _with_const(case, k, value):
    cst = dict(case.params.const)
    cst[k] = value
    prm = C.Params(**{**case.params.__dict__, "const": cst})
    return C.Case(case.name, prm, case.dem, case.particles, case.nsteps)


def test_gpu_cfg4_adjoint_full_size_central_differences():
    """50,000 particles on the 512 x 512 DEM, Voellmy, RK2 dt 0.05 s, 50
    steps with checkpoints every 10: dL/dmu and dL/dxi of the deposit misfit
    against central differences (eps = 1e-6 relative, on the REAL(4)
    constants) within 1e-6 relative, the north-star bar."""
    nsteps = 50
    case = C.cfg_adjoint(nsteps=nsteps)
    with SPHModel(C.cfg_adjoint(nsteps=nsteps, mu=0.17, xi=650.0)) as m:
        m.time_integration_sph(nsteps)
        T = m.raster_idw()[0][:, 0].copy()
    with AdjointModel(case) as a:
        a.forward(nsteps, 10)
        L, lx, lh = a.misfit(T)
        g = a.backward(lam_x=lx, lam_h=lh)
    assert L == pytest.approx(_adjoint_loss(case, T, nsteps), rel=1e-12)
    for k, name in ((9, "mu"), (3, "xi")):
        v = case.params.const[k]
        hi, lo = C.f32(v * (1 + 1e-6)), C.f32(v * (1 - 1e-6))
        fd = (_adjoint_loss(_with_const(case, k, hi), T, nsteps)
              - _adjoint_loss(_with_const(case, k, lo), T, nsteps)) / (hi - lo)
        rel = abs(g[name] - fd) / abs(fd)
        print(f"cfg4 dL/d{name}: adjoint {g[name]:.12e} central difference {fd:.12e} rel {rel:.2e}")
        assert rel <= 1e-6, (name, g[name], fd, rel)


def test_gpu_cfg4_adjoint_full_size_torch_reference():
    """50,000 particles on the 512 x 512 DEM: the adjoint of 4 RK2 steps
    against torch.autograd through the FP64 torch restatement of the step
    (tests/adjoint_reference.py, pinned to the oracle), for L = sum W.x +
    sum U.h: dL/dmu, dL/dxi, dL/dmass of every particle."""
    import torch

    import adjoint_reference as AR

    nsteps = 4
    case = C.cfg_adjoint(nsteps=nsteps)
    n = case.particles.n
    rng = np.random.default_rng(11)
    W, U = rng.normal(size=(n, 2)), rng.normal(size=n)
    c = case.params.cst()
    mu = torch.tensor(c[9], dtype=torch.float64, requires_grad=True)
    xi = torch.tensor(c[3], dtype=torch.float64, requires_grad=True)
    mass = torch.tensor(case.particles.mass, dtype=torch.float64, requires_grad=True)
    x, _, h = AR.run(case, nsteps, mu, xi, mass)
    ((torch.as_tensor(W) * x).sum() + (torch.as_tensor(U) * h).sum()).backward()
    with AdjointModel(case) as m:
        m.forward(nsteps, 2)
        g = m.backward(lam_x=W, lam_h=U)
    print(f"\ncfg4 torch reference: dL/dmu {g['mu']:.10e} vs {mu.grad.item():.10e}, "
          f"dL/dxi {g['xi']:.10e} vs {xi.grad.item():.10e}")
    assert g["mu"] == pytest.approx(mu.grad.item(), rel=1e-8)
    assert g["xi"] == pytest.approx(xi.grad.item(), rel=1e-8)
    gm = mass.grad.numpy()
    assert np.max(np.abs(g["mass"] - gm)) <= 1e-8 * np.max(np.abs(gm))
