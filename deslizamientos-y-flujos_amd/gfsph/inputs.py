"""The reference's input files through the C ABI (gfsph_inputs_*).

``Inputs(dir, name)`` reads what Init_sph reads ("7 SPH_MAIN_2019.f90":74):
<name>.MASTER.dat (Driver_Input run control, optional), <name>.dat (Input_SW),
<topo>.top (Topo_main; scattered DTMs gridded on the GPU as Get_DTM does,
slopes and curvatures projected on the GPU as Get_Topol does) and the release
(<name>.pts gridded on the GPU as Read_Source_plus_New does, or restart.dat).
``model()`` creates the propagation context on it; ``case()`` rebuilds the
same inputs as a ``cases.Case`` (for the oracle and the batch API).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import cases as Cs
from .api import BatchModel, CGrid, CParams, CRun, GfsphError, SPHModel, _dp, _ip, load_library
from .dem import DEM, TerrainZones, Zone

__all__ = ["Inputs"]


class FileDEM(DEM):
    """A DEM whose grid and nodal terrain terms come from the input files:
    the grid extent is the one read (Xmaxg, Ymaxg), and ``topol`` holds
    Get_Topol's nodal terms when they were projected."""

    def __init__(self, Z, xmin, ymin, dx, dy, in_model, xmax, ymax, topol9):
        super().__init__(Z=Z, xmin=xmin, ymin=ymin, dx=dx, dy=dy, in_model=in_model)
        self._xmax, self._ymax, self.topol9 = xmax, ymax, topol9

    @property
    def xmax(self) -> float:
        return self._xmax

    def topol(self) -> np.ndarray:
        return self.topol9 if self.topol9 is not None else super().topol()

    @property
    def ymax(self) -> float:
        return self._ymax


class Inputs:
    def __init__(self, directory: str, name: str, topo: str | None = None, build: bool = True):
        self.lib = load_library()
        h = C.c_void_p()
        f = self.lib.gfsph_inputs_read if build else self.lib.gfsph_inputs_parse
        rc = f(directory.encode(), name.encode(), topo.encode() if topo else None, C.byref(h))
        self.h = h
        if rc != 0:
            msg = self.lib.gfsph_inputs_error(h).decode() if h.value else ""
            self.close()
            raise GfsphError(rc, msg)
        self.built = build
        self.params, self.run, self.grid = CParams(), CRun(), CGrid()
        n = C.c_int32()
        self._check(self.lib.gfsph_inputs_info(h, C.byref(self.params), C.byref(self.run), C.byref(self.grid),
                                               C.byref(n)))
        self.n = n.value

    def _check(self, rc):
        if rc != 0:
            raise GfsphError(rc, self.lib.gfsph_inputs_error(self.h).decode())

    def dem(self) -> dict:
        g = self.grid
        npg = g.nx * g.ny
        Z = np.zeros(npg)
        im = np.zeros(npg, np.int32)
        top = np.zeros((npg, 9)) if not g.fdiff else None
        self._check(self.lib.gfsph_inputs_dem(self.h, _dp(Z), _ip(im), _dp(top) if top is not None else None))
        return {"Z": Z.reshape(g.ny, g.nx), "in_model": im.reshape(g.ny, g.nx), "topol": top}

    def zones(self):
        """The terrain zones read (ic_topo_props = 10) as TerrainZones, or
        None: the zones, their node-list bits and topol(15) after
        Get_Eros_Pts."""
        from .api import CZone
        nz = C.c_int32()
        self._check(self.lib.gfsph_inputs_zones(self.h, C.byref(nz), None, None, None))
        if nz.value == 0:
            return None
        g = self.grid
        zs = (CZone * nz.value)()
        nodes = np.zeros(g.nx * g.ny, np.int32)
        avail = np.zeros(g.nx * g.ny)
        self._check(self.lib.gfsph_inputs_zones(self.h, None, C.cast(zs, C.POINTER(CZone)), _ip(nodes), _dp(avail)))
        return TerrainZones([Zone(z.type, z.law, tuple(z.props)) for z in zs], nodes=nodes, avail=avail)

    def particles(self) -> dict:
        n = self.n
        out = {"x": np.zeros((n, 2)), "v": np.zeros((n, 2)), "h": np.zeros(n), "mass": np.zeros(n),
               "hsml": np.zeros(n), "u": np.zeros(n), "itype": np.zeros(n, np.int32)}
        self._check(self.lib.gfsph_inputs_particles(self.h, *(_dp(out[k]) for k in ("x", "v", "h", "mass", "hsml",
                                                                                     "u")), _ip(out["itype"])))
        return out

    def case(self, nsteps: int = 0) -> Cs.Case:
        """The inputs as a Case: Params (const(:) as read), the terrain and the release."""
        p = self.params
        const = {k: float(p.cst[k]) for k in range(1, 111) if p.cst[k] != 0.0}
        prm = Cs.Params(integrator=p.integrator, pa_sph=p.pa_sph, nnps=p.nnps, sle=p.sle, skf=p.skf,
                        summation_density=p.summation_density, average_velocity=p.average_velocity,
                        nor_density=p.nor_density, icpwp=p.icpwp, npoin0=p.npoin0, dt=p.dt, h_inf=p.h_inf,
                        pwp0rel=p.pwp0rel, adapt_dt=self.run.adapt_dt if self.run.present else 0, const=const,
                        nuaux=p.nuaux, pwp_bc_type=(p.pwp_bc_type[0], p.pwp_bc_type[1]),
                        pwp_bc_val=(p.pwp_bc_val[0], p.pwp_bc_val[1]), pwp0_type=p.pwp0_type, xirel=p.xirel,
                        ic_k0=p.ic_k0, k0_active=p.k0_active, k0_pasive=p.k0_pasive)
        d = self.dem()
        g = self.grid
        dem = FileDEM(d["Z"], g.xmin, g.ymin, g.dx, g.dy, d["in_model"], g.xmax, g.ymax, d["topol"])
        dem.zones = self.zones()
        q = self.particles()
        parts = Cs.Particles(x=q["x"], v=q["v"], h=q["h"], mass=q["mass"], hsml=q["hsml"], u=q["u"],
                             itype=q["itype"])
        return Cs.Case("inputs", prm, dem, parts, nsteps)

    def model(self) -> SPHModel:
        """gfsph_inputs_create: the propagation context on these inputs
        (Input_SW's initialisation, ic_adapt_dt_sph and restart time applied)."""
        ctx = C.c_void_p()
        self._check(self.lib.gfsph_inputs_create(self.h, C.byref(ctx)))
        return SPHModel.adopt(self.lib, ctx, self.case())

    def cases(self) -> dict:
        """MASTER.dat's case table: mode (ic_cases_mat), OC_mat
        [nproblems, 3 + ncases + 2] and the stop thresholds."""
        mode, nc, npb = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self.lib.gfsph_inputs_cases(self.h, C.byref(mode), C.byref(nc), C.byref(npb), None, None))
        table = np.zeros((npb.value, 3 + nc.value + 2))
        stop = np.zeros(3)
        self._check(self.lib.gfsph_inputs_cases(self.h, None, None, None, _dp(table), _dp(stop)))
        return {"mode": mode.value, "ncases": nc.value, "table": table, "vstop": stop[0], "timestop": stop[1],
                "dtimestop": stop[2]}

    def resample(self, seed: int) -> None:
        """Draws a sampled case set (ic_cases_mat 3, 4, 31, 32) again from
        another seed (the sets are drawn from seed 1 when read)."""
        self._check(self.lib.gfsph_inputs_resample(self.h, C.c_uint64(int(seed))))

    def batch(self, pixels: bool = False) -> BatchModel:
        """gfsph_inputs_batch_create: every problem of the case table as one
        realisation, under the reference's per-problem loop (case control)."""
        ctx = C.c_void_p()
        self._check(self.lib.gfsph_inputs_batch_create(self.h, int(pixels), C.byref(ctx)))
        base = self.case()
        npb = self.cases()["table"].shape[0]
        return BatchModel.adopt(self.lib, ctx, [base] * npb)

    def close(self) -> None:
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.gfsph_inputs_destroy(self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
