"""Digital elevation model container and synthetic terrain generators.

The DEM is a regular node grid (npoigx x npoigy) stored row-major in x, node
``ipoig = iy*nx + ix`` (0-based), as ``SPH_FEM_topo_2019`` numbers it
("2 SPH_FEM_topo_2019.f90":2437-2442).  Nodal fields follow the reference's
``topol(1:9, ipoig)`` layout ("2 SPH_FEM_topo_2019.f90":69-77):
Z, Zx, Zy, 1+Zx^2, Zx*Zy, 1+Zy^2, Zxx/rt, Zxy/rt, Zyy/rt.

Slopes are computed by central finite differences, the reference's
``Fdiff_derivs`` used for explicit grids with ictop >= 100
("2 SPH_FEM_topo_2019.f90":2812-2877): boundary nodes and nodes outside the
model keep zero derivatives.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

__all__ = ["DEM", "Zone", "TerrainZones", "fdiff_derivs", "plane_runout_dem", "fractal_dem"]


def fdiff_derivs(Z: np.ndarray, dx: float, dy: float, in_model: np.ndarray | None = None) -> np.ndarray:
    """Return topol[ny*nx, 9] from Z[ny, nx] ("2 SPH_FEM_topo_2019.f90":2812-2877)."""
    ny, nx = Z.shape
    T = np.zeros((ny, nx, 9), dtype=np.float64)
    T[:, :, 0] = Z
    if nx < 3 or ny < 3:
        return T.reshape(ny * nx, 9)
    del2x, del2y = 2.0 * dx, 2.0 * dy
    delx2, dely2, delxy = dx * dx, dy * dy, dx * dy
    zc = Z[1:-1, 1:-1]
    zx = (Z[1:-1, 2:] - Z[1:-1, :-2]) / del2x
    zy = (Z[2:, 1:-1] - Z[:-2, 1:-1]) / del2y
    rt = (1.0 + zx * zx + zy * zy) ** 0.5
    zxx = (Z[1:-1, :-2] - 2.0 * zc + Z[1:-1, 2:]) / delx2
    zyy = (Z[:-2, 1:-1] - 2.0 * zc + Z[2:, 1:-1]) / dely2
    zxy = (Z[2:, 2:] - Z[:-2, 2:] - Z[2:, :-2] + Z[:-2, :-2]) / (4.0 * delxy)
    inner = T[1:-1, 1:-1]
    inner[:, :, 1] = zx
    inner[:, :, 2] = zy
    inner[:, :, 3] = 1.0 + zx * zx
    inner[:, :, 4] = zx * zy
    inner[:, :, 5] = 1.0 + zy * zy
    inner[:, :, 6] = zxx / rt
    inner[:, :, 7] = zxy / rt
    inner[:, :, 8] = zyy / rt
    if in_model is not None:
        mask = np.asarray(in_model).reshape(ny, nx) == 0
        T[:, :, 1:][mask] = 0.0
    return T.reshape(ny * nx, 9)


@dataclasses.dataclass
class Zone:
    """One terrain zone of a terrain file's ic_topo_props = 10 block
    ("2 SPH_FEM_topo_2019.f90":516-576): basal_type (1 friction, 2
    erosion), basal_law (the zone's erosion law 0..3) and basal_props(1:16)
    (1: the geometry -- 1, 2, 3 a range of Z, x, y; 4 a rectangle; 5 a turned
    rectangle; 101 erodible points; 2-6 its bounds; 7.. the law's
    constants)."""

    type: int
    law: int
    props: tuple

    def __post_init__(self):
        p = [float(v) for v in self.props] + [0.0] * 16
        self.props = tuple(p[:16])


@dataclasses.dataclass
class TerrainZones:
    """The terrain zones of a DEM.  ``nodes`` [ny*nx] int32: bit z set when
    the node is in zone z's node list (geometry 101); ``avail`` [ny*nx] the
    terrain's topol(15) (erodible depths of geometry-101 zones, 0
    elsewhere); ``eros_pts``: the points file those come from
    ({"pts": [n, 3] x, y, depth, "type": type_Eros, "dx_s": spacing}), as the
    reference reads it from <terrain>.eros.pts."""

    zones: list
    nodes: np.ndarray | None = None
    avail: np.ndarray | None = None
    eros_pts: dict | None = None


@dataclasses.dataclass
class DEM:
    """Regular DEM grid. ``Z`` has shape (ny, nx); coordinates are exact
    multiples of the spacing so that the reference's Xmaxg-Xming spacing
    recovers ``dx`` bit for bit."""

    Z: np.ndarray
    xmin: float
    ymin: float
    dx: float
    dy: float
    in_model: np.ndarray | None = None
    zones: TerrainZones | None = None

    @property
    def nx(self) -> int:
        return int(self.Z.shape[1])

    @property
    def ny(self) -> int:
        return int(self.Z.shape[0])

    @property
    def xmax(self) -> float:
        return self.xmin + (self.nx - 1) * self.dx

    @property
    def ymax(self) -> float:
        return self.ymin + (self.ny - 1) * self.dy

    def topol(self) -> np.ndarray:
        return fdiff_derivs(np.ascontiguousarray(self.Z, dtype=np.float64), self.dx, self.dy, self.in_model)

    def node_xy(self) -> np.ndarray:
        xs = self.xmin + np.arange(self.nx) * self.dx
        ys = self.ymin + np.arange(self.ny) * self.dy
        X, Y = np.meshgrid(xs, ys)
        return np.stack([X.ravel(), Y.ravel()], axis=1)


def plane_runout_dem(nx: int = 201, ny: int = 101, d: float = 5.0, slope_deg: float = 30.0,
                     x_break: float = 600.0, blend: float = 100.0) -> DEM:
    """Planar slope of ``slope_deg`` for x < x_break - blend/2, a C1 transition
    whose gradient ramps linearly to zero over ``blend`` metres, then a flat
    runout plain (the 200x100-cell, 5 m test terrain)."""
    t = math.tan(math.radians(slope_deg))
    x = np.arange(nx) * d
    x1, x2 = x_break - blend / 2.0, x_break + blend / 2.0
    # dZ/dx = -s(x); integrate piecewise in closed form
    z = np.empty_like(x)
    for k, xv in enumerate(x):
        if xv <= x1:
            z[k] = -t * xv
        elif xv <= x2:
            u = xv - x1
            z[k] = -t * x1 - t * (u - u * u / (2.0 * blend))
        else:
            z[k] = -t * x1 - t * (blend / 2.0)
    z = z - z.min()
    Z = np.repeat(z[None, :], ny, axis=0)
    return DEM(Z=Z.astype(np.float64), xmin=0.0, ymin=0.0, dx=d, dy=d)


def _diamond_square(k: int, H: float, rng: np.random.Generator) -> np.ndarray:
    n = 2 ** k + 1
    a = np.zeros((n, n), dtype=np.float64)
    a[0, 0], a[0, -1], a[-1, 0], a[-1, -1] = rng.standard_normal(4)
    step = n - 1
    scale = 1.0
    while step > 1:
        half = step // 2
        # diamond step: centres of squares
        c = (a[0:-1:step, 0:-1:step] + a[0:-1:step, step::step] + a[step::step, 0:-1:step]
             + a[step::step, step::step]) * 0.25
        a[half:-1:step, half:-1:step] = c + scale * rng.standard_normal(c.shape)
        # square step: edge midpoints
        for (r0, c0) in ((0, half), (half, 0)):
            idx_r, idx_c = np.meshgrid(np.arange(r0, n, step), np.arange(c0, n, step), indexing="ij")
            acc = np.zeros(idx_r.shape)
            cnt = np.zeros(idx_r.shape)
            for (drr, dcc) in ((-half, 0), (half, 0), (0, -half), (0, half)):
                rr, cc = idx_r + drr, idx_c + dcc
                ok = (rr >= 0) & (rr < n) & (cc >= 0) & (cc < n)
                acc[ok] += a[rr[ok], cc[ok]]
                cnt[ok] += 1
            a[idx_r, idx_c] = acc / cnt + scale * rng.standard_normal(idx_r.shape)
        step = half
        scale *= 2.0 ** (-H)
    return a


def fractal_dem(n: int, d: float, seed: int, H: float = 0.8, relief: float = 400.0,
                valley_deg: tuple[float, float] = (25.0, 35.0), valley_len_frac: float = 0.45,
                valley_width: float | None = None) -> DEM:
    """Diamond-square fractal terrain (Hurst exponent ``H``, ``relief`` metres
    peak-to-peak) on an n x n grid of spacing ``d`` with a channelled valley
    carved along +x: starting near x = 0.1 L at the terrain top, its bed drops
    at a slope varying smoothly between ``valley_deg`` over ``valley_len_frac``
    of the domain, then flattens into a gentle (2 deg) runout fan.  The RNG is
    numpy's Philox counter-based generator seeded with ``seed``."""
    rng = np.random.Generator(np.random.Philox(seed))
    k = max(1, int(math.ceil(math.log2(n - 1))))
    a = _diamond_square(k, H, rng)[:n, :n]
    a = (a - a.min()) / max(a.max() - a.min(), 1e-12) * relief
    L = (n - 1) * d
    x = np.arange(n) * d
    y = np.arange(n) * d
    Xg, Yg = np.meshgrid(x, y)
    # valley axis meanders gently in y
    phase = rng.uniform(0, 2 * math.pi)
    yc = 0.5 * L + 0.08 * L * np.sin(2 * math.pi * x / L + phase)
    th0, th1 = (math.radians(v) for v in valley_deg)
    x0 = 0.1 * L
    x1 = x0 + valley_len_frac * L
    s = np.clip((x - x0) / max(x1 - x0, 1e-9), 0.0, 1.0)
    slope = np.where(x < x1, np.tan(th0 + (th1 - th0) * 0.5 * (1 - np.cos(math.pi * s))), math.tan(math.radians(2.0)))
    slope = np.where(x < x0, slope[np.argmax(x >= x0)], slope)
    bed = np.cumsum(slope) * d
    bed = bed[-1] - bed  # decreasing along +x
    bed = bed - bed.min()
    bed = bed / max(bed.max(), 1e-12) * (relief + 0.6 * relief)
    w = valley_width if valley_width is not None else 0.12 * L
    dist = np.abs(Yg - yc[None, :])
    channel = bed[None, :] + (relief * 0.5) * (dist / w) ** 2
    Z = np.minimum(a + bed[None, :] * 0.6, channel)
    Z = Z - Z.min()
    return DEM(Z=np.ascontiguousarray(Z, dtype=np.float64), xmin=0.0, ymin=0.0, dx=d, dy=d)
