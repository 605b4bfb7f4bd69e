"""FP64 reference gradient of the propagation step: the same step written
with torch operations on the CPU, differentiated by torch.autograd.

Test infrastructure only.  It restates, for the laws the adjoint covers
(nfric = 5 frictional / Voellmy, pa_sph = 2, skf = 1, icpwp = 0, sle = 0),
the step of "7 SPH_MAIN_2019.f90":746-1005 as the oracle does: each step's
neighbour set comes from the oracle (the predicate is not differentiable,
and the forward pass holds it fixed), the sums follow sum_density_Tools
("5a My_SPH_Tools_2019.f90":1763-1859), IntForces_SW (6:6892-6941), the
stages ExtForces_SW / basal_friction_Ext_SW with the limiter (6:8532-8657,
6:9355-9901) and av_vel (7:3042-3099).  ``GfsphRun`` is the thin
torch.autograd.Function over the GPU forward and its C-ABI vector-Jacobian
product, used by the gradient tests only.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from gfsph import AdjointModel
from gfsph import cases as C
from oracle import oracle

F64 = torch.float64


def kernel(dx, dy, mh):
    """Cubic spline W and grad W (kernel_TOOLS, skf = 1) for pair vectors."""
    r = torch.sqrt(dx * dx + dy * dy)
    q = r / mh
    factor = 15.0 / (7.0 * math.pi * mh * mh)
    w1 = factor * (2.0 / 3.0 - q * q + 0.5 * q ** 3)
    a1 = factor * (-2.0 + 1.5 * q) / (mh * mh)
    t = 2.0 - q
    w2 = factor / 6.0 * t ** 3
    a2 = -factor * 0.5 * t * t / mh / torch.where(r > 0, r, torch.ones_like(r))
    in1, in2 = q <= 1.0, q <= 2.0
    zero = torch.zeros_like(r)
    w = torch.where(in1, w1, torch.where(in2, w2, zero))
    a = torch.where(in1, a1, torch.where(in2, a2, zero))
    return w, a * dx, a * dy


def slopes(dem_t, xx, yy):
    """Bilinear nodal slopes of Get_Z_Topo (2:2395-2491); dem_t = (nx, ny,
    xmin, ymin, dx, dy, xmax, ymax, Zx[nodes], Zy[nodes])."""
    nx, ny, xmin, ymin, ddx, ddy, xmax, ymax, ZX, ZY = dem_t
    xv, yv = xx.detach().numpy(), yy.detach().numpy()
    ipx = ((xv - xmin) / ddx + 1).astype(np.int64)
    ipy = ((yv - ymin) / ddy + 1).astype(np.int64)
    ipx = np.where((ipx == nx) & (np.abs(xmax - xv) <= 1e-3), ipx - 1, ipx)
    ipy = np.where((ipy == ny) & (np.abs(ymax - yv) <= 1e-3), ipy - 1, ipy)
    ipx = np.clip(ipx, 1, nx - 1)
    ipy = np.clip(ipy, 1, ny - 1)
    n1 = (ipy - 1) * nx + (ipx - 1)
    n2, n4 = n1 + 1, n1 + nx
    n3 = n4 + 1
    xref = torch.as_tensor(xmin + (ipx - 1) * ddx, dtype=F64)
    yref = torch.as_tensor(ymin + (ipy - 1) * ddy, dtype=F64)
    xi = (xx - xref) / ddx
    eta = (yy - yref) / ddy
    sh = [(1 - xi) * (1 - eta), xi * (1 - eta), xi * eta, (1 - xi) * eta]
    nodes = [n1, n2, n3, n4]
    zx = sum(s * ZX[nd] for s, nd in zip(sh, nodes))
    zy = sum(s * ZY[nd] for s, nd in zip(sh, nodes))
    return zx, zy


def ssqrt(v):
    """sqrt with a zero (not NaN) derivative at 0."""
    pos = v > 0
    return torch.where(pos, torch.sqrt(torch.where(pos, v, torch.ones_like(v))), torch.zeros_like(v))


def ext_force(P, dem_t, xx, yy, vx, vy, rf, inx, iny, mu, xi):
    """ExtForces_SW for nfric = 5 without curvature or pore pressure."""
    g, dt, comp = P["g"], P["dt"], P["comp"]
    small = rf < comp
    h = torch.where(small, torch.full_like(rf, comp), rf)
    hu = torch.where(small, torch.zeros_like(rf), h * vx)
    hv = torch.where(small, torch.zeros_like(rf), h * vy)
    hinv = torch.where(small, torch.zeros_like(rf), 1.0 / h)
    hvmod = ssqrt(hu * hu + hv * hv)
    vmod = hvmod * hinv
    big = hvmod >= 1e-5
    hvmodinv = torch.where(big, 1.0 / torch.where(big, hvmod, torch.ones_like(hvmod)), torch.zeros_like(hvmod))
    zx, zy = slopes(dem_t, xx, yy)
    exx = -g * zx
    exy = -g * zy
    gtanfi = torch.clamp(g * mu * 1.0, min=0.0) * torch.ones_like(rf)
    frictx = gtanfi * hu * hvmodinv * h
    fricty = gtanfi * hv * hvmodinv * h
    if float(xi.detach() if torch.is_tensor(xi) else xi) > 0.0:
        vmod = ssqrt(vx * vx + vy * vy)
        frictx = frictx + g * vmod * vx / xi
        fricty = fricty + g * vmod * vy / xi
    dcx, dcy = -frictx * hinv, -fricty * hinv
    ax, ay = exx + inx, exy + iny
    moving = vmod > 1e-5
    # moving: reversal limiter
    vsx, vsy = vx + dt * (ax + dcx), vy + dt * (ay + dcy)
    vdot = vsx * vx + vsy * vy
    vm_safe = torch.where(moving, vmod, torch.ones_like(vmod))
    tx, ty = vx / vm_safe, vy / vm_safe
    accelT = ax * tx + ay * ty
    deccelT = dcx * tx + dcy * ty
    rev = moving & (vdot < 0) & (accelT * deccelT < 0)
    fx_rev, fy_rev = -(-vx / dt - ax) * h, -(-vy / dt - ay) * h
    # static
    acc = ssqrt(ax * ax + ay * ay)
    acc_safe = torch.where(acc > 1e-5, acc, torch.ones_like(acc))
    fsx, fsy = gtanfi * ax / acc_safe, gtanfi * ay / acc_safe
    fricT = ssqrt(fsx * fsx + fsy * fsy)
    stop = acc <= fricT
    fsx = torch.where(stop, ax, fsx)
    fsy = torch.where(stop, ay, fsy)
    fsx = torch.where(acc > 1e-5, fsx, torch.zeros_like(fsx)) * h
    fsy = torch.where(acc > 1e-5, fsy, torch.zeros_like(fsy)) * h
    frictx = torch.where(moving, torch.where(rev, fx_rev, frictx), fsx)
    fricty = torch.where(moving, torch.where(rev, fy_rev, fricty), fsy)
    return exx - frictx / h, exy - fricty / h


def step(P, dem_t, x, v, m, hs):
    """One step; returns (x, v, h) with the step-start neighbour set fixed."""
    n = x.shape[0]
    ij, _, _ = oracle.pairs(x.detach().numpy(), hs.detach().numpy(), skf=1)
    i = torch.as_tensor(np.concatenate([ij[:, 0], ij[:, 1]]), dtype=torch.long)
    j = torch.as_tensor(np.concatenate([ij[:, 1], ij[:, 0]]), dtype=torch.long)
    dx = x[i, 0] - x[j, 0]
    dy = x[i, 1] - x[j, 1]
    mh = (hs[i] + hs[j]) / 2.0
    w, gx, gy = kernel(dx, dy, mh)
    w0 = 15.0 / (7.0 * math.pi * hs * hs) * (2.0 / 3.0)
    S = w0 * m + torch.zeros(n, dtype=F64).index_add(0, i, m[j] * w)
    rf = torch.clamp(S, min=P["h_inf"])
    Pp = P["g"] * 0.5 * rf * rf / (rf * rf)
    c = -(Pp[i] + Pp[j])
    ainx = torch.zeros(n, dtype=F64).index_add(0, i, m[j] * (c * gx))
    ainy = torch.zeros(n, dtype=F64).index_add(0, i, m[j] * (c * gy))
    x0x, x0y, v0x, v0y = x[:, 0], x[:, 1], v[:, 0], v[:, 1]
    xx, xy, vx, vy = x0x, x0y, v0x, v0y
    ex = ey = torch.zeros(n, dtype=F64)
    Dx = Dy = Dvx = Dvy = torch.zeros(n, dtype=F64)
    for f1, f2 in zip(P["f1"], P["f2"]):
        xx, xy = x0x + f1 * P["dt"] * vx, x0y + f1 * P["dt"] * vy
        vx, vy = v0x + f1 * P["dt"] * (ex + ainx), v0y + f1 * P["dt"] * (ey + ainy)
        ex, ey = ext_force(P, dem_t, xx, xy, vx, vy, rf, ainx, ainy, P["mu"], P["xi"])
        Dvx, Dvy = Dvx + f2 * (ex + ainx), Dvy + f2 * (ey + ainy)
        Dx, Dy = Dx + f2 * vx, Dy + f2 * vy
    xn = torch.stack([x0x + P["fin"] * Dx, x0y + P["fin"] * Dy], 1)
    vp = torch.stack([v0x + P["fin"] * Dvx, v0y + P["fin"] * Dvy], 1)
    if P["av_vel"]:
        cij = (2 * m[j]) * w / (S[i] + S[j])
        axx = torch.zeros(n, dtype=F64).index_add(0, i, -cij * (vp[i, 0] - vp[j, 0]))
        ayy = torch.zeros(n, dtype=F64).index_add(0, i, -cij * (vp[i, 1] - vp[j, 1]))
        vp = torch.stack([vp[:, 0] + 0.3 * axx, vp[:, 1] + 0.3 * ayy], 1)
    return xn, vp, S


def params_of(case, mu, xi):
    prm = case.params
    c = prm.cst()
    if prm.integrator == 4:
        f1, f2, fin = (0.0, 0.5, 0.5, 1.0), (1.0, 2.0, 2.0, 1.0), prm.dt / 6.0
    else:
        f1, f2, fin = (0.0, 1.0), (1.0, 1.0), prm.dt / 2.0
    return {"g": c[1], "dt": prm.dt, "comp": c[15], "h_inf": prm.h_inf, "mu": mu, "xi": xi, "f1": f1, "f2": f2,
            "fin": fin, "av_vel": bool(prm.average_velocity)}


def dem_tensors(dem):
    T = dem.topol()
    return (dem.nx, dem.ny, dem.xmin, dem.ymin, dem.dx, dem.dy, dem.xmax, dem.ymax,
            torch.as_tensor(T[:, 1], dtype=F64), torch.as_tensor(T[:, 2], dtype=F64))


def run(case, nsteps, mu, xi, mass, x0=None, v0=None):
    """Differentiable run: returns final (x, v, h)."""
    P = params_of(case, mu, xi)
    dem_t = dem_tensors(case.dem)
    x = torch.as_tensor(case.particles.x, dtype=F64) if x0 is None else x0
    v = torch.as_tensor(case.particles.v, dtype=F64) if v0 is None else v0
    hs = torch.as_tensor(case.particles.hsml, dtype=F64)
    h = torch.as_tensor(case.particles.h, dtype=F64)
    for _ in range(nsteps):
        x, v, h = step(P, dem_t, x, v, mass, hs)
    return x, v, h


class GfsphRun(torch.autograd.Function):
    """Final (x, h) of ``nsteps`` GPU steps as a function of (mu, xi, mass);
    backward is the library's adjoint."""

    @staticmethod
    def forward(ctx, mu, xi, mass, case, nsteps, ckpt):
        c = dict(case.params.const)
        c[9], c[3] = float(mu), float(xi)
        prm = C.Params(**{**case.params.__dict__, "const": c})
        P = case.particles.copy()
        P.mass = mass.detach().numpy().copy()
        m = AdjointModel(C.Case(case.name, prm, case.dem, P, nsteps))
        m.forward(nsteps, ckpt)
        st = m.state()
        ctx.model = m
        return torch.as_tensor(st["x"], dtype=F64), torch.as_tensor(st["h"], dtype=F64)

    @staticmethod
    def backward(ctx, gx, gh):
        m = ctx.model
        g = m.backward(lam_x=gx.detach().numpy(), lam_h=gh.detach().numpy())
        m.close()
        return (torch.tensor(g["mu"], dtype=F64), torch.tensor(g["xi"], dtype=F64),
                torch.as_tensor(g["mass"], dtype=F64), None, None, None)
