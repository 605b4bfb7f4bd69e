"""Material parameters, particle releases and the named workloads.

Material behaviour is the reference's ``const(1..110)`` vector, declared
``real(ink)`` i.e. REAL(4) ("4 SPH_SW_Vars_2019.f90":27), so every material
constant is rounded through float32 here exactly as the reference's list
directed read of the .dat file rounds it.  Index meanings
("6 SPH_SW_2019.f90":1151-1169):

  1 g          2 density      3 Voellmy xi / Manning n / mu(I) b
  4 erosion    5 nfric        6 tau_y / r        7 K / mu_B / n
  8 m / visco  9 tan(phi)    10 hfrict0         11 (>=0: no more lines)
 12 tan(phi0) 13 Bfact       14 alpha=h_sat/h   15 comp (minimum h)
 16 ic_cosTheta ...

Particle state follows ``restart.dat`` ("6 SPH_SW_2019.f90":3318-3327):
x, y, vx, vy, h (the SW "rho"), mass (volume), hsml, u, itype.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from .dem import DEM, fractal_dem, plane_runout_dem

__all__ = ["Params", "Particles", "Case", "f32", "voellmy", "frictional", "bingham",
           "lattice_block", "elliptical_scar", "cfg_plane_voellmy", "cfg_fractal_frictional", "cfg_mc_hazard",
           "cfg_adjoint"]


def f32(v: float) -> float:
    """Round through IEEE single precision (REAL(4))."""
    return float(np.float32(v))


@dataclasses.dataclass
class Params:
    """Solver switches; field names follow the reference's variables."""

    integrator: int = 4           # 4: RK45 (SPH_t_Integ_Alg=4); 2: RK2 (Heun)
    pa_sph: int = 2
    nnps: int = 4
    sle: int = 0
    skf: int = 1
    summation_density: int = 1
    average_velocity: int = 1
    nor_density: int = 0
    icpwp: int = 0
    npoin0: int = 0
    dt: float = 0.05
    adapt_dt: int = 0             # ic_adapt_dt_sph: 0 fixed dt; 1, 2, 4, 10 (6:6521-6664)
    h_inf: float = 1.0e-3
    pwp0rel: float = 0.0
    # icpwp = 2: pore pressure on a column of nuaux nodes (Initialize_Pwp,
    # "6 SPH_SW_2019.f90":1743-1800): bottom / top conditions (type 1 value,
    # 2 gradient; values relative to the depth) and the initial profile
    nuaux: int = 0
    pwp_bc_type: tuple = (2, 1)
    pwp_bc_val: tuple = (0.0, 0.0)
    pwp0_type: int = 0
    xirel: float = 0.0
    # K0 (frictional laws, "6 SPH_SW_2019.f90":1373-1390): the pressure times
    # k0_active where the velocity divergence is >= 0, else k0_pasive
    ic_k0: int = 0
    k0_active: float = 1.0
    k0_pasive: float = 1.0
    const: dict = dataclasses.field(default_factory=dict)   # 1-based index -> value

    def cst(self) -> np.ndarray:
        c = np.zeros(111, dtype=np.float64)
        for k, v in self.const.items():
            c[int(k)] = f32(v)
        return c

    @property
    def nfric(self) -> int:
        return int(self.cst()[5] + 0.01)


@dataclasses.dataclass
class Particles:
    x: np.ndarray       # (n, 2)
    v: np.ndarray       # (n, 2)
    h: np.ndarray       # (n,)
    mass: np.ndarray    # (n,)
    hsml: np.ndarray    # (n,)
    u: np.ndarray       # (n,)
    itype: np.ndarray   # (n,) int32

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    def copy(self) -> "Particles":
        return Particles(*(np.array(a, copy=True) for a in dataclasses.astuple(self)))


@dataclasses.dataclass
class Case:
    name: str
    params: Params
    dem: DEM
    particles: Particles
    nsteps: int


def _base_const(rho: float, g: float = 9.81) -> dict:
    return {1: g, 2: rho, 10: 1.0e-2, 14: 1.0, 15: 1.0e-2}


def voellmy(mu: float, xi: float, rho: float = 2000.0, nfric: int = 5) -> dict:
    """Frictional basal law plus Voellmy turbulence (6:9706-9749)."""
    c = _base_const(rho)
    c.update({3: xi, 5: nfric, 9: mu})
    return c


def frictional(phi_deg: float, rho: float = 2000.0, nfric: int = 5, cv: float = 0.0,
               alpha: float = 1.0) -> dict:
    """Pure frictional law; the pore-pressure rate ``Bfact`` is derived from a
    consolidation coefficient ``cv`` as Bfact = pi^2 cv / 4 (Cv = 4 Bfact/pi^2,
    "6 SPH_SW_2019.f90":10926)."""
    c = _base_const(rho)
    c.update({5: nfric, 9: math.tan(math.radians(phi_deg)), 13: math.pi ** 2 * cv / 4.0, 14: alpha})
    return c


def bingham(tau_y: float, mu_b: float, rho: float = 2000.0) -> dict:
    """Bingham law nfric=4 (6:9538-9623); const(8)=1 keeps the unused
    Herschel-Bulkley prefactor finite."""
    c = _base_const(rho)
    c.update({5: 4, 6: tau_y, 7: mu_b, 8: 1.0})
    return c


def lattice_block(x0: float, y0: float, lx: float, ly: float, nx: int, ny: int, h0: float,
                  hsml_fac: float = 1.5) -> Particles:
    """Regular nx x ny lattice filling [x0,x0+lx] x [y0,y0+ly] (cell-centred),
    thickness h0, volume per particle dx*dy*h0, hsml = hsml_fac*sqrt(dx*dy)."""
    dx, dy = lx / nx, ly / ny
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny))
    x = np.stack([x0 + (ix.ravel() + 0.5) * dx, y0 + (iy.ravel() + 0.5) * dy], axis=1)
    n = x.shape[0]
    h = np.full(n, h0)
    return Particles(x=x, v=np.zeros((n, 2)), h=h, mass=dx * dy * h, hsml=np.full(n, hsml_fac * math.sqrt(dx * dy)),
                     u=np.zeros(n), itype=np.ones(n, dtype=np.int32))


def elliptical_scar(cx: float, cy: float, n: int, spacing: float, aspect: float, seed: int,
                    h_range: tuple[float, float] = (2.0, 8.0), hsml_fac: float = 1.5) -> Particles:
    """``n`` lattice particles of spacing ``spacing`` filling an ellipse of axis
    ratio ``aspect`` (a/b) centred at (cx, cy), taken in order of normalised
    radius; thickness ~ U(h_range) from a Philox stream seeded with ``seed``."""
    b = math.sqrt(n * spacing * spacing / (math.pi * aspect)) * 1.08
    a = aspect * b
    nxh, nyh = int(a / spacing) + 2, int(b / spacing) + 2
    ix, iy = np.meshgrid(np.arange(-nxh, nxh + 1), np.arange(-nyh, nyh + 1))
    px, py = ix.ravel() * spacing, iy.ravel() * spacing
    rr = (px / a) ** 2 + (py / b) ** 2
    order = np.lexsort((px, py, rr))[:n]
    order = np.sort(order)
    x = np.stack([cx + px[order], cy + py[order]], axis=1)
    rng = np.random.Generator(np.random.Philox(seed))
    h = rng.uniform(h_range[0], h_range[1], size=n)
    vol = spacing * spacing * h
    return Particles(x=x, v=np.zeros((n, 2)), h=h, mass=vol, hsml=np.full(n, hsml_fac * spacing),
                     u=np.zeros(n), itype=np.ones(n, dtype=np.int32))


def cfg_plane_voellmy(nx_p: int = 141, ny_p: int = 71, nsteps: int = 2000, dem_cells=(200, 100)) -> Case:
    """Plain FP64 CPU-runnable case: 200x100-cell DEM at 5 m, planar 30 deg
    slope for x < 600 m blending into a flat runout; a 100 x 50 m block, 5 m
    thick, on a regular lattice (141 x 71 = 10,011 particles), at rest;
    Voellmy mu = 0.2, xi = 500 m/s2; cubic spline, hsml = 1.5 spacing; RK4,
    dt = 0.05 s."""
    dem = plane_runout_dem(nx=dem_cells[0] + 1, ny=dem_cells[1] + 1, d=5.0)
    parts = lattice_block(200.0, 225.0, 100.0, 50.0, nx_p, ny_p, 5.0, 1.5)
    prm = Params(integrator=4, dt=0.05, h_inf=1.0e-3, const=voellmy(0.2, 500.0))
    return Case("plane_voellmy", prm, dem, parts, nsteps)


def cfg_fractal_frictional(n: int = 200_000, dem_n: int = 2048, d: float = 2.0, nsteps: int = 5000,
                           integrator: int = 2, seed_dem: int = 1002, seed_h: int = 2002,
                           spacing: float = 1.0) -> Case:
    """Plain FP64 GPU drop-in at scale: fractal DEM 2048 x 2048 at 2 m (H 0.8,
    400 m relief, carved 25-35 deg valley); elliptical release of 200,000
    particles, thickness ~ U(2, 8) m; frictional phi = 28 deg with pore
    pressure (initial ratio 0.6, consolidation coefficient 0.01 m2/s); RK2,
    dt = 0.02 s."""
    dem = fractal_dem(dem_n, d, seed_dem)
    L = (dem_n - 1) * d
    cx = 0.14 * L
    yc = 0.5 * L
    # place the scar on the valley axis near its head
    ix = int(round(cx / d))
    col = dem.Z[:, ix]
    iy = int(np.argmin(np.abs(np.arange(dem_n) * d - yc) + 1e-9 * col))
    parts = elliptical_scar(cx, iy * d, n, spacing, 2.0, seed_h)
    prm = Params(integrator=integrator, dt=0.02, h_inf=1.0e-3, icpwp=3, pwp0rel=0.6,
                 const=frictional(28.0, cv=0.01))
    return Case("fractal_frictional", prm, dem, parts, nsteps)


def cfg_mc_hazard(R: int = 256, dem_n: int = 1024, d: float = 5.0, n_range=(5_000, 80_000), nsteps: int = 3000,
                  seed: int = 3003, spacing: float = 2.0, jitter: float = 50.0, integrator: int = 2,
                  dt: float = 0.05, dem: DEM | None = None) -> list:
    """Monte Carlo hazard batch: fractal DEM 1024 x 1024 at 5 m (H 0.8, 400 m
    relief, carved valley); R realisations with Voellmy mu ~ U(0.08, 0.35),
    xi ~ LogNormal(ln 400, 0.5) m/s2, a release of N ~ logU(5k, 80k)
    particles (elliptical scar, spacing 2 m, hsml 3 m, thickness ~ U(2, 8) m)
    on the valley axis, its centre jittered by U(-50, 50) m in x and y; RK2,
    dt = 0.05 s.  All draws come from one Philox stream seeded with ``seed``
    (the reference's sampler, "8 SPH_FEM_Driver_2019.f90":1282-1369, is not
    reproducible and has no log-uniform law)."""
    dem = fractal_dem(dem_n, d, seed) if dem is None else dem
    L = (dem.nx - 1) * dem.dx
    cx0 = 0.14 * L
    ix = int(round(cx0 / dem.dx))
    col = dem.Z[:, ix]
    iy = int(np.argmin(np.abs(np.arange(dem.ny) * dem.dy - 0.5 * L) + 1e-9 * col))
    rng = np.random.Generator(np.random.Philox(seed))
    mus = rng.uniform(0.08, 0.35, size=R)
    xis = np.exp(rng.normal(math.log(400.0), 0.5, size=R))
    ns = np.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1]), size=R)).astype(int)
    jx = rng.uniform(-jitter, jitter, size=R)
    jy = rng.uniform(-jitter, jitter, size=R)
    out = []
    for r in range(R):
        parts = elliptical_scar(cx0 + jx[r], iy * dem.dy + jy[r], int(ns[r]), spacing, 2.0, seed * 1000 + r)
        prm = Params(integrator=integrator, dt=dt, h_inf=1.0e-3, const=voellmy(float(mus[r]), float(xis[r])))
        out.append(Case(f"mc{r}", prm, dem, parts, nsteps))
    return out


def cfg_adjoint(n: int = 50_000, dem_n: int = 512, d: float = 5.0, nsteps: int = 1000, mu: float = 0.2,
                xi: float = 500.0, seed: int = 4004, spacing: float = 2.0, dt: float = 0.05) -> Case:
    """Adjoint back-analysis: fractal DEM 512 x 512 at 5 m (H 0.8, 400 m
    relief, carved valley); elliptical release of 50,000 particles (spacing
    2 m, hsml 3 m, thickness ~ U(2, 8) m) on the valley axis at x = 0.25 L;
    Voellmy with
    (mu, xi) = (0.2, 500) as the starting guess (the target deposit comes from
    mu = 0.17, xi = 650); RK2, dt = 0.05 s, 1,000 steps."""
    dem = fractal_dem(dem_n, d, seed)
    L = (dem.nx - 1) * dem.dx
    cx = 0.25 * L
    ix = int(round(cx / dem.dx))
    col = dem.Z[:, ix]
    iy = int(np.argmin(np.abs(np.arange(dem.ny) * dem.dy - 0.5 * L) + 1e-9 * col))
    parts = elliptical_scar(cx, iy * dem.dy, n, spacing, 2.0, seed + 1)
    prm = Params(integrator=2, dt=dt, h_inf=1.0e-3, const=voellmy(mu, xi))
    return Case("adjoint", prm, dem, parts, nsteps)


def cfg_calibration(R: int = 64, n: int = 1_000_000, dem_n: int = 4096, d: float = 1.0, nsteps: int = 2000,
                    seed: int = 1005, draw_seed: int = 2005, spacing: float = 1.0, dt: float = 0.02,
                    dem: DEM | None = None) -> list:
    """Batched adjoint calibration: fractal DEM 4096 x 4096 at 1 m (H 0.8,
    400 m relief, carved valley); one elliptical release of 1,000,000
    particles (spacing 1 m, hsml 1.5 m, thickness ~ U(2, 8) m) on the valley
    axis at x = 0.2 L; R parameter draws of the Bingham law, tau_y ~ U(500,
    5000) Pa and mu_B ~ U(10, 200) Pa s (one Philox stream seeded with
    ``draw_seed``); RK2, dt = 0.02 s, 2,000 steps.  Every draw runs the same
    release: only the rheology differs."""
    dem = fractal_dem(dem_n, d, seed) if dem is None else dem
    L = (dem.nx - 1) * dem.dx
    cx = 0.2 * L
    ix = int(round(cx / dem.dx))
    col = dem.Z[:, ix]
    iy = int(np.argmin(np.abs(np.arange(dem.ny) * dem.dy - 0.5 * L) + 1e-9 * col))
    parts = elliptical_scar(cx, iy * dem.dy, n, spacing, 2.0, seed + 1)
    rng = np.random.Generator(np.random.Philox(draw_seed))
    taus = rng.uniform(500.0, 5000.0, size=R)
    mus = rng.uniform(10.0, 200.0, size=R)
    return [Case(f"cal{r}", Params(integrator=2, dt=dt, h_inf=1.0e-3, const=bingham(float(taus[r]), float(mus[r]))),
                 dem, parts, nsteps) for r in range(R)]
