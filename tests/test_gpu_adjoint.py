"""Adjoint of the propagation step against central finite differences of
the same forward runs (eps = 1e-6 relative), for the deposit misfit of a
Voellmy release: dL/dmu, dL/dxi and dL/dmass of sampled particles.  mu and
xi are const(9) and const(3), held in REAL(4) by the reference, so the
difference quotients divide by the perturbation actually applied after that
rounding."""
import numpy as np
import pytest

from gfsph import AdjointModel, SPHModel
from gfsph import cases as C
from gfsph.dem import fractal_dem

pytestmark = pytest.mark.gpu

NSTEPS, CKPT = 60, 20


def base_case(mu=0.2, xi=500.0, integrator=2, n=400):
    dem = fractal_dem(65, 2.0, 17, relief=30.0)
    parts = C.elliptical_scar(40.0, 64.0, n, 1.0, 2.0, 5, h_range=(1.0, 3.0))
    prm = C.Params(integrator=integrator, dt=0.02, h_inf=1e-3, const=C.voellmy(mu, xi))
    return C.Case("adj", prm, dem, parts, NSTEPS)


def target_raster():
    case = base_case(0.17, 650.0)
    with SPHModel(case) as m:
        m.time_integration_sph(NSTEPS)
        out, _ = m.raster_idw()
    return out[:, 0].copy()


def loss_of(case, target):
    with AdjointModel(case) as m:
        m.time_integration_sph(NSTEPS)
        L, _, _ = m.misfit(target)
    return L


def with_const(case, k, value):
    c = dict(case.params.const)
    c[k] = value
    prm = C.Params(**{**case.params.__dict__, "const": c})
    return C.Case(case.name, prm, case.dem, case.particles, case.nsteps)


@pytest.mark.parametrize("integrator", [2, 4])
def test_adjoint_matches_central_differences(integrator):
    T = target_raster()
    case = base_case(integrator=integrator)
    with AdjointModel(case) as m:
        m.forward(NSTEPS, CKPT)
        L, lx, lh = m.misfit(T)
        g = m.backward(lam_x=lx, lam_h=lh)
        st = m.state()
    assert np.all(st["flag"] == 0)
    assert L > 0
    # the forward pass with checkpoints is the plain run
    assert loss_of(case, T) == pytest.approx(L, rel=1e-13)  # raster sums use atomics
    eps = 1e-6
    for k, name in ((9, "mu"), (3, "xi")):
        v = case.params.const[k]
        hi, lo = C.f32(v * (1 + eps)), C.f32(v * (1 - eps))
        fd = (loss_of(with_const(case, k, hi), T) - loss_of(with_const(case, k, lo), T)) / (hi - lo)
        assert g[name] == pytest.approx(fd, rel=1e-5), (name, g[name], fd)
    rng = np.random.default_rng(0)
    for i in rng.choice(case.particles.n, size=4, replace=False):
        m0 = case.particles.mass[i]
        P1, P2 = case.particles.copy(), case.particles.copy()
        P1.mass[i] = m0 * (1 + eps)
        P2.mass[i] = m0 * (1 - eps)
        c1 = C.Case(case.name, case.params, case.dem, P1, NSTEPS)
        c2 = C.Case(case.name, case.params, case.dem, P2, NSTEPS)
        fd = (loss_of(c1, T) - loss_of(c2, T)) / (P1.mass[i] - P2.mass[i])
        assert g["mass"][i] == pytest.approx(fd, rel=1e-5, abs=1e-9 * abs(L) / m0), (i, g["mass"][i], fd)


def test_adjoint_rejects_uncovered_laws():
    case = base_case()
    prm = C.Params(integrator=2, dt=0.02, h_inf=1e-3, const=C.voellmy(0.2, 500.0, nfric=7))
    with AdjointModel(C.Case("b", prm, case.dem, case.particles, NSTEPS)) as m:
        with pytest.raises(Exception, match="adjoint"):
            m.forward(10, 5)


def bingham_case(tau_y=1000.0, mu_b=50.0, integrator=2, n=400, seed=5):
    dem = fractal_dem(65, 2.0, 17, relief=30.0)
    parts = C.elliptical_scar(40.0, 64.0, n, 1.0, 2.0, seed, h_range=(1.0, 3.0))
    prm = C.Params(integrator=integrator, dt=0.02, h_inf=1e-3, const=C.bingham(tau_y, mu_b))
    return C.Case("adjb", prm, dem, parts, NSTEPS)


@pytest.mark.parametrize("integrator", [2, 4])
def test_bingham_adjoint_matches_central_differences(integrator):
    """Bingham (nfric 4): its Tau0 is carried from particle to particle, stage
    to stage and step to step (the reference's sticky value), so a particle's
    friction depends on other particles' states; the adjoint routes those
    cotangents back.  dL/dtau_y (const(6)), dL/dmu_B (const(7)) and dL/dmass
    of sampled particles against central differences of the misfit."""
    T = target_raster()
    case = bingham_case(integrator=integrator)
    with AdjointModel(case) as m:
        m.forward(NSTEPS, CKPT)
        L, lx, lh = m.misfit(T)
        g = m.backward(lam_x=lx, lam_h=lh)
        st = m.state()
    assert np.all(st["flag"] == 0)
    assert L > 0
    assert loss_of(case, T) == pytest.approx(L, rel=1e-13)
    eps = 1e-6
    for k, name in ((6, "mu"), (7, "xi")):
        v = case.params.const[k]
        hi, lo = C.f32(v * (1 + eps)), C.f32(v * (1 - eps))
        fd = (loss_of(with_const(case, k, hi), T) - loss_of(with_const(case, k, lo), T)) / (hi - lo)
        print(f"Bingham RK{integrator} dL/dconst({k}): adjoint {g[name]:.10e} central difference {fd:.10e}")
        assert g[name] == pytest.approx(fd, rel=1e-5), (name, g[name], fd)
    rng = np.random.default_rng(0)
    for i in rng.choice(case.particles.n, size=4, replace=False):
        m0 = case.particles.mass[i]
        P1, P2 = case.particles.copy(), case.particles.copy()
        P1.mass[i] = m0 * (1 + eps)
        P2.mass[i] = m0 * (1 - eps)
        c1 = C.Case(case.name, case.params, case.dem, P1, NSTEPS)
        c2 = C.Case(case.name, case.params, case.dem, P2, NSTEPS)
        fd = (loss_of(c1, T) - loss_of(c2, T)) / (P1.mass[i] - P2.mass[i])
        assert g["mass"][i] == pytest.approx(fd, rel=1e-5, abs=1e-9 * abs(L) / m0), (i, g["mass"][i], fd)


@pytest.mark.parametrize("law", ["voellmy", "bingham"])
def test_batched_adjoint_equals_single_runs(law):
    """Three realisations (different parameters and releases) as one batched
    forward + misfit + backward: each realisation's loss and gradients equal
    those of its run alone (to the rounding of the misfit raster's atomic
    sums)."""
    from gfsph import BatchAdjointModel

    T = target_raster()
    if law == "voellmy":
        cases = [base_case(mu, xi) for mu, xi in ((0.2, 500.0), (0.15, 800.0), (0.25, 300.0))]
    else:
        cases = [bingham_case(ty, mb, seed=sd) for ty, mb, sd in ((1000.0, 50.0, 5), (2500.0, 20.0, 6),
                                                                  (600.0, 120.0, 7))]
    single = []
    for c in cases:
        with AdjointModel(c) as m:
            m.forward(NSTEPS, CKPT)
            L, lx, lh = m.misfit(T)
            single.append((L, m.backward(lam_x=lx, lam_h=lh)))
    with BatchAdjointModel(cases) as b:
        b.forward(NSTEPS, CKPT)
        L, lx, lh = b.misfit(T)
        g = b.backward(lam_x=lx, lam_h=lh)
    edges = np.concatenate([[0], np.cumsum([c.particles.n for c in cases])])
    for r, (Ls, gs) in enumerate(single):
        assert L[r] == pytest.approx(Ls, rel=1e-12), r
        assert g["mu"][r] == pytest.approx(gs["mu"], rel=1e-9), (r, g["mu"][r], gs["mu"])
        assert g["xi"][r] == pytest.approx(gs["xi"], rel=1e-9), (r, g["xi"][r], gs["xi"])
        gm = g["mass"][edges[r]:edges[r + 1]]
        np.testing.assert_allclose(gm, gs["mass"], rtol=1e-9, atol=1e-12 * np.abs(gs["mass"]).max())


@pytest.mark.parametrize("integrator", [2, 4])
def test_adjoint_matches_fp64_reference_gradient(integrator):
    """The library's adjoint against torch.autograd through the FP64 torch
    restatement of the step (tests/adjoint_reference.py, pinned to the
    oracle), for L = sum W.x + sum U.h of the final state; and the same
    gradient through the thin torch.autograd.Function over the library."""
    import torch

    import adjoint_reference as AR
    from test_adjoint_reference import small_case

    case = small_case(integrator)
    nsteps, ckpt = 40, 15
    n = case.particles.n
    rng = np.random.default_rng(3)
    W, U = rng.normal(size=(n, 2)), rng.normal(size=n)
    c = case.params.cst()
    mu = torch.tensor(c[9], dtype=torch.float64, requires_grad=True)
    xi = torch.tensor(c[3], dtype=torch.float64, requires_grad=True)
    mass = torch.tensor(case.particles.mass, dtype=torch.float64, requires_grad=True)
    x, _, h = AR.run(case, nsteps, mu, xi, mass)
    ((torch.as_tensor(W) * x).sum() + (torch.as_tensor(U) * h).sum()).backward()
    with AdjointModel(case) as m:
        m.forward(nsteps, ckpt)
        g = m.backward(lam_x=W, lam_h=U)
    assert g["mu"] == pytest.approx(mu.grad.item(), rel=1e-8)
    assert g["xi"] == pytest.approx(xi.grad.item(), rel=1e-8)
    gm = mass.grad.numpy()
    assert np.max(np.abs(g["mass"] - gm)) <= 1e-8 * np.max(np.abs(gm))
    mu2 = torch.tensor(c[9], dtype=torch.float64, requires_grad=True)
    xi2 = torch.tensor(c[3], dtype=torch.float64, requires_grad=True)
    mass2 = torch.tensor(case.particles.mass, dtype=torch.float64, requires_grad=True)
    x2, h2 = AR.GfsphRun.apply(mu2, xi2, mass2, case, nsteps, ckpt)
    ((torch.as_tensor(W) * x2).sum() + (torch.as_tensor(U) * h2).sum()).backward()
    assert mu2.grad.item() == pytest.approx(g["mu"], rel=1e-14)
    np.testing.assert_allclose(mass2.grad.numpy(), g["mass"], rtol=1e-14, atol=0)


@pytest.mark.parametrize("batch", [False, True])
@pytest.mark.parametrize("law", ["voellmy", "bingham"])
def test_adjoint_segments_recompute_the_forward_grids(batch, law):
    """A block sliding ~60 m down a 30 degree plane over 120 steps.  Each
    segment is recomputed from its checkpoint with the grids the forward pass
    built (a single run's box from the restored state, a batch's slot ranges
    saved with the checkpoint) -- not from the state the context last held,
    whose box no longer covers the particles -- so every recomputed step
    equals the forward's bit for bit: the gradient is the same for any
    checkpoint scheme (checkpoints every 40 steps; one segment of 120; one
    segment recorded 15 steps at a time from second-level checkpoints, or 2
    at a time from three levels), and
    equals central differences."""
    from gfsph import BatchAdjointModel
    from gfsph.dem import plane_runout_dem

    dem = plane_runout_dem(nx=61, ny=21, d=5.0, x_break=250.0, blend=50.0)
    parts = C.lattice_block(20.0, 40.0, 20.0, 10.0, 10, 5, 3.0)
    const = C.voellmy(0.1, 1000.0) if law == "voellmy" else C.bingham(800.0, 30.0)
    k_mu = 9 if law == "voellmy" else 6
    prm = C.Params(integrator=2, dt=0.05, h_inf=1e-3, const=const)
    case = C.Case("slide", prm, dem, parts, 120)
    W = np.random.default_rng(2).normal(size=(parts.n, 2))

    def lin_loss(c):
        with SPHModel(c) as m:
            m.time_integration_sph(120)
            x = m.state()["x"]
        return float((W * x).sum()), x

    L0, xe = lin_loss(case)
    assert xe[:, 0].min() > parts.x[:, 0].max()  # the block left its initial box
    grads = {}
    for ckpt, tape in ((40, 0), (120, 0), (120, 15), (120, 2)):
        if batch:
            with BatchAdjointModel([case, case]) as b:
                b.forward(120, ckpt, tape)
                g = b.backward(lam_x=np.concatenate([W, W]))
            grads[ckpt, tape] = float(g["mu"][0])
            assert g["mu"][1] == g["mu"][0]
        else:
            with AdjointModel(case) as m:
                m.forward(120, ckpt, tape)
                grads[ckpt, tape] = m.backward(lam_x=W)["mu"]
    eps = 1e-6
    v = case.params.const[k_mu]
    hi, lo = C.f32(v * (1 + eps)), C.f32(v * (1 - eps))
    fd = (lin_loss(with_const(case, k_mu, hi))[0] - lin_loss(with_const(case, k_mu, lo))[0]) / (hi - lo)
    print(f"{law} slide dL/dconst({k_mu}): {grads}, central difference {fd:.10e}")
    assert grads[40, 0] == grads[120, 0] == grads[120, 15] == grads[120, 2]
    assert grads[40, 0] == pytest.approx(fd, rel=1e-5)
