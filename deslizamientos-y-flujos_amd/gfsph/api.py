"""Host-side mirror of the reference's SPH_MAIN_2019 interface over the C ABI.

The reference drives one simulation through module procedures with shared
module state ("7 SPH_MAIN_2019.f90":47-61): ``Init_sph`` reads inputs and
allocates, ``time_integration_sph`` advances one step, ``Out_plot_sph`` and
``Manage_pixel_Hazard_SPH`` produce rasters.  ``SPHModel`` keeps the same
verbs over one ``gfsph_ctx`` (include/gfsph.h); all computation runs in the
HIP library, and a missing library is an error, never a CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._build import LIB

__all__ = ["GfsphError", "SPHModel", "BatchModel", "AdjointModel", "neighbour_pairs", "load_library",
           "write_netcdf", "EXPORTED_SYMBOLS"]

EXPORTED_SYMBOLS = (
    "gfsph_create", "gfsph_step", "gfsph_sync", "gfsph_get_state", "gfsph_neighbour_pairs",
    "gfsph_raster_idw", "gfsph_mc_pixel_accumulate", "gfsph_mc_pixel_read", "gfsph_profile_step",
    "gfsph_counts", "gfsph_last_error", "gfsph_destroy", "gfsph_batch_create", "gfsph_batch_pixel_accumulate",
    "gfsph_batch_pixel_read", "gfsph_batch_hazard", "gfsph_autotune", "gfsph_adjoint_forward", "gfsph_adjoint_set_tape",
    "gfsph_adjoint_backward", "gfsph_raster_misfit", "gfsph_set_adaptive_dt", "gfsph_time", "gfsph_set_time",
    "gfsph_inputs_parse", "gfsph_inputs_read", "gfsph_inputs_info", "gfsph_inputs_dem", "gfsph_inputs_particles",
    "gfsph_inputs_create", "gfsph_inputs_error", "gfsph_inputs_destroy", "gfsph_write_netcdf",
    "gfsph_batch_cases", "gfsph_batch_status", "gfsph_inputs_cases", "gfsph_inputs_batch_create",
    "gfsph_get_mass", "gfsph_erosion_read", "gfsph_erosion_set", "gfsph_get_pwp_column", "gfsph_set_pwp_column",
    "gfsph_neighbour_pairs_noalloc", "gfsph_set_mass0", "gfsph_batch_scale_pwp", "gfsph_inputs_resample",
    "gfsph_terrain_eros_pts", "gfsph_inputs_zones",
)

STATUS = {0: "ok", -1: "invalid argument", -2: "unsupported option", -3: "HIP runtime failure",
          -4: "non-finite state", -5: "capacity exceeded"}


class GfsphError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gfsph status {code} ({STATUS.get(code, '?')}): {msg}")
        self.code = code


class CParams(C.Structure):
    _fields_ = [("integrator", C.c_int32), ("pa_sph", C.c_int32), ("nnps", C.c_int32), ("sle", C.c_int32),
                ("skf", C.c_int32), ("summation_density", C.c_int32), ("average_velocity", C.c_int32),
                ("nor_density", C.c_int32), ("icpwp", C.c_int32), ("npoin0", C.c_int32),
                ("init_input", C.c_int32), ("check_finite", C.c_int32),
                ("dt", C.c_double), ("h_inf", C.c_double), ("pwp0rel", C.c_double), ("cst", C.c_double * 111),
                ("nuaux", C.c_int32), ("pwp_bc_type", C.c_int32 * 2), ("pwp0_type", C.c_int32),
                ("pwp_bc_val", C.c_double * 2), ("xirel", C.c_double),
                ("ic_k0", C.c_int32), ("pad_k0", C.c_int32), ("k0_active", C.c_double), ("k0_pasive", C.c_double)]


class CZone(C.Structure):
    _fields_ = [("type", C.c_int32), ("law", C.c_int32), ("props", C.c_double * 16)]


class CDem(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("xmin", C.c_double), ("ymin", C.c_double),
                ("dx", C.c_double), ("dy", C.c_double), ("xmax", C.c_double), ("ymax", C.c_double),
                ("Z", C.POINTER(C.c_double)), ("in_model", C.POINTER(C.c_int32)),
                ("topol", C.POINTER(C.c_double)), ("nzones", C.c_int32), ("pad_z", C.c_int32),
                ("zones", C.POINTER(CZone)), ("zone_nodes", C.POINTER(C.c_int32)),
                ("eros_avail", C.POINTER(C.c_double))]


class CRun(C.Structure):
    _fields_ = [("present", C.c_int32), ("t_integ_alg", C.c_int32), ("adapt_dt", C.c_int32),
                ("maxtimestep", C.c_int32), ("print_step", C.c_int32), ("save_step", C.c_int32),
                ("plot_step", C.c_int32), ("pad_", C.c_int32), ("dt", C.c_double), ("time_end", C.c_double),
                ("dt_sph", C.c_double), ("time_restart", C.c_double)]


class CGrid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("xmin", C.c_double), ("ymin", C.c_double),
                ("dx", C.c_double), ("dy", C.c_double), ("xmax", C.c_double), ("ymax", C.c_double),
                ("xshift", C.c_double), ("yshift", C.c_double), ("fdiff", C.c_int32), ("pad_", C.c_int32)]


class CCaseCtl(C.Structure):
    _fields_ = [("vstop", C.c_double), ("timestop", C.c_double), ("dtimestop", C.c_double), ("dt", C.c_double),
                ("time_end", C.c_double), ("time_plot", C.c_double), ("pixels", C.c_int32), ("pad_", C.c_int32)]


_lib = None


def load_library(path: str | None = None):
    """Loads libgfsph.so (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(p)
    dp, ip, vp = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.c_void_p
    L.gfsph_create.argtypes = [C.POINTER(CParams), C.POINTER(CDem), C.c_int32, dp, dp, dp, dp, dp, dp, ip,
                               C.POINTER(vp)]
    L.gfsph_step.argtypes = [vp, C.c_int32]
    L.gfsph_sync.argtypes = [vp]
    L.gfsph_get_state.argtypes = [vp, dp, dp, dp, dp, dp, dp, ip]
    L.gfsph_neighbour_pairs.argtypes = [C.c_int32, C.c_int32, dp, dp, ip, C.c_int64, C.POINTER(C.c_int64), ip]
    # device pointers and a hipStream_t, passed as addresses
    L.gfsph_neighbour_pairs_noalloc.argtypes = [C.c_int32, C.c_int32, vp, vp, vp, C.c_int64, vp, vp, vp]
    L.gfsph_raster_idw.argtypes = [vp, dp, dp]
    L.gfsph_mc_pixel_accumulate.argtypes = [vp]
    L.gfsph_mc_pixel_read.argtypes = [vp, dp, ip, C.c_int32]
    L.gfsph_profile_step.argtypes = [vp, C.c_int32, C.c_int32, ip, C.POINTER(C.c_char_p), dp]
    L.gfsph_counts.argtypes = [vp, ip, ip]
    L.gfsph_last_error.argtypes = [vp]
    L.gfsph_last_error.restype = C.c_char_p
    L.gfsph_destroy.argtypes = [vp]
    L.gfsph_destroy.restype = None
    L.gfsph_batch_create.argtypes = [C.POINTER(CParams), C.POINTER(CDem), C.c_int32, ip, dp, dp, dp, dp, dp, dp, dp,
                                     C.POINTER(vp)]
    L.gfsph_batch_pixel_accumulate.argtypes = [vp]
    L.gfsph_batch_pixel_read.argtypes = [vp, C.c_int32, dp, ip]
    L.gfsph_batch_hazard.argtypes = [vp, C.c_double, dp, ip, dp, dp]
    L.gfsph_autotune.argtypes = [vp, C.c_int32, ip]
    L.gfsph_adjoint_forward.argtypes = [vp, C.c_int32, C.c_int32]
    L.gfsph_adjoint_set_tape.argtypes = [vp, C.c_int32]
    L.gfsph_adjoint_backward.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp, dp]
    L.gfsph_raster_misfit.argtypes = [vp, dp, dp, dp, dp]
    L.gfsph_set_adaptive_dt.argtypes = [vp, C.c_int32]
    L.gfsph_time.argtypes = [vp, dp, dp]
    L.gfsph_set_time.argtypes = [vp, C.c_double]
    L.gfsph_batch_cases.argtypes = [vp, C.POINTER(CCaseCtl)]
    L.gfsph_inputs_cases.argtypes = [vp, ip, ip, ip, dp, dp]
    L.gfsph_inputs_batch_create.argtypes = [vp, C.c_int32, C.POINTER(vp)]
    L.gfsph_batch_status.argtypes = [vp, ip, ip, ip]
    L.gfsph_get_mass.argtypes = [vp, dp]
    L.gfsph_set_mass0.argtypes = [vp, dp]
    L.gfsph_batch_scale_pwp.argtypes = [vp, dp]
    L.gfsph_inputs_resample.argtypes = [vp, C.c_uint64]
    L.gfsph_terrain_eros_pts.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, dp, dp,
                                         C.c_int32, dp, C.c_int32, C.c_double, dp, dp, ip, dp]
    L.gfsph_inputs_zones.argtypes = [vp, ip, C.POINTER(CZone), ip, dp]
    L.gfsph_get_pwp_column.argtypes = [vp, dp]
    L.gfsph_set_pwp_column.argtypes = [vp, dp]
    L.gfsph_erosion_read.argtypes = [vp, C.c_int32, dp, dp, dp]
    L.gfsph_erosion_set.argtypes = [vp, C.c_int32, dp, dp, dp]
    for f in (L.gfsph_inputs_parse, L.gfsph_inputs_read):
        f.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(vp)]
    L.gfsph_inputs_info.argtypes = [vp, C.POINTER(CParams), C.POINTER(CRun), C.POINTER(CGrid), ip]
    L.gfsph_inputs_dem.argtypes = [vp, dp, ip, dp]
    L.gfsph_inputs_particles.argtypes = [vp, dp, dp, dp, dp, dp, dp, ip]
    L.gfsph_inputs_create.argtypes = [vp, C.POINTER(vp)]
    L.gfsph_inputs_error.argtypes = [vp]
    L.gfsph_inputs_error.restype = C.c_char_p
    L.gfsph_inputs_destroy.argtypes = [vp]
    L.gfsph_inputs_destroy.restype = None
    cpp = C.POINTER(C.c_char_p)
    L.gfsph_write_netcdf.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_int32, cpp, cpp, cpp, C.POINTER(dp), C.c_double, C.c_char_p]
    if path is None:
        _lib = L
    return L


def write_netcdf(path: str, dem, rasters: dict, title: str = "", fill: float = float("nan")) -> None:
    """Rasters on the terrain grid of ``dem`` (xmin, ymin, dx, dy, nx, ny) to
    a NetCDF classic file.  ``rasters``: name -> (array [ny, nx], units,
    long_name)."""
    L = load_library()
    names = list(rasters)
    arrs = [np.ascontiguousarray(np.asarray(rasters[k][0], dtype=np.float64).reshape(dem.ny, dem.nx)) for k in names]
    cs = lambda xs: (C.c_char_p * len(xs))(*[x.encode() for x in xs])  # noqa: E731
    dps = (C.POINTER(C.c_double) * len(arrs))(*[_dp(a) for a in arrs])
    rc = L.gfsph_write_netcdf(path.encode(), dem.nx, dem.ny, dem.xmin, dem.ymin, dem.dx, dem.dy, len(names),
                              cs(names), cs([rasters[k][1] for k in names]), cs([rasters[k][2] for k in names]), dps,
                              fill, title.encode())
    if rc != 0:
        raise GfsphError(rc, f"writing {path}: invalid names or sizes, or the file cannot be written")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def make_cparams(prm, init_input: bool = True, check_finite: bool = False) -> CParams:
    p = CParams()
    p.integrator, p.pa_sph, p.nnps, p.sle, p.skf = prm.integrator, prm.pa_sph, prm.nnps, prm.sle, prm.skf
    p.summation_density, p.average_velocity, p.nor_density = (prm.summation_density, prm.average_velocity,
                                                              prm.nor_density)
    p.icpwp, p.npoin0 = prm.icpwp, prm.npoin0
    p.init_input, p.check_finite = int(init_input), int(check_finite)
    p.dt, p.h_inf, p.pwp0rel = prm.dt, prm.h_inf, prm.pwp0rel
    c = prm.cst()
    for k in range(111):
        p.cst[k] = c[k]
    p.nuaux = int(getattr(prm, "nuaux", 0))
    bt, bv = getattr(prm, "pwp_bc_type", (2, 1)), getattr(prm, "pwp_bc_val", (0.0, 0.0))
    p.pwp_bc_type[0], p.pwp_bc_type[1] = int(bt[0]), int(bt[1])
    p.pwp_bc_val[0], p.pwp_bc_val[1] = float(bv[0]), float(bv[1])
    p.pwp0_type = int(getattr(prm, "pwp0_type", 0))
    p.xirel = float(getattr(prm, "xirel", 0.0))
    p.ic_k0 = int(getattr(prm, "ic_k0", 0))
    p.k0_active = float(getattr(prm, "k0_active", 1.0))
    p.k0_pasive = float(getattr(prm, "k0_pasive", 1.0))
    return p


class _DemArgs:
    """C view of a DEM; keeps the numpy buffers alive."""

    def __init__(self, dem):
        self._Z = np.ascontiguousarray(dem.Z, dtype=np.float64)
        self._im = None if dem.in_model is None else np.ascontiguousarray(dem.in_model, dtype=np.int32)
        d = CDem()
        d.nx, d.ny = dem.nx, dem.ny
        d.xmin, d.ymin, d.dx, d.dy, d.xmax, d.ymax = dem.xmin, dem.ymin, dem.dx, dem.dy, dem.xmax, dem.ymax
        d.Z = _dp(self._Z)
        d.in_model = _ip(self._im) if self._im is not None else C.POINTER(C.c_int32)()
        # nodal terrain terms projected from the input files (Get_Topol); else
        # finite differences on the device
        t9 = getattr(dem, "topol9", None)
        self._t9 = None if t9 is None else np.ascontiguousarray(t9, dtype=np.float64)
        d.topol = _dp(self._t9) if self._t9 is not None else C.POINTER(C.c_double)()
        # terrain zones (gfsph.dem.TerrainZones): friction and erosion zones
        tz = getattr(dem, "zones", None)
        self._zones = self._zn = self._ea = None
        if tz is not None and len(tz.zones):
            self._zones = (CZone * len(tz.zones))()
            for k, z in enumerate(tz.zones):
                self._zones[k].type = int(z.type)
                self._zones[k].law = int(z.law)
                for j, v in enumerate(z.props):
                    self._zones[k].props[j] = float(v)
            d.nzones = len(tz.zones)
            d.zones = C.cast(self._zones, C.POINTER(CZone))
            if tz.nodes is not None:
                self._zn = np.ascontiguousarray(tz.nodes, dtype=np.int32)
                d.zone_nodes = _ip(self._zn)
            if tz.avail is not None:
                self._ea = np.ascontiguousarray(tz.avail, dtype=np.float64)
                d.eros_avail = _dp(self._ea)
        self.c = d


def terrain_eros_pts(dem, pts, type_eros: int, dx_s: float):
    """Get_Eros_Pts ("2 SPH_FEM_topo_2019.f90":909-1110) on ``dem``'s grid
    (host code of the library): returns (Z after type_eros, topol(15), node
    list as a bool array, basal_props(2:5)).  pts [n, 3]: x, y, depth."""
    L = load_library()
    Z = np.ascontiguousarray(dem.Z, dtype=np.float64).ravel().copy()
    avail = np.zeros_like(Z)
    listed = np.zeros(Z.shape[0], np.int32)
    box = np.zeros(4)
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    rc = L.gfsph_terrain_eros_pts(dem.nx, dem.ny, dem.xmin, dem.ymin, dem.dx, dem.dy, None, None, pts.shape[0],
                                  _dp(pts), int(type_eros), float(dx_s), _dp(Z), _dp(avail), _ip(listed), _dp(box))
    if rc != 0:
        raise GfsphError(rc, "gfsph_terrain_eros_pts")
    return Z.reshape(dem.Z.shape), avail, listed.astype(bool), box


class SPHModel:
    """One propagation run on the GPU (Init_sph + time_integration_sph)."""

    def __init__(self, case, init_input: bool = True, check_finite: bool = False):
        self.lib = load_library()
        self.case = case
        P = case.particles
        self.n = P.n
        dem = case.dem
        self._dem = _DemArgs(dem)
        d = self._dem.c
        p = make_cparams(case.params, init_input, check_finite)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (P.x, P.v, P.h, P.mass, P.hsml, P.u)]
        it = np.ascontiguousarray(P.itype, dtype=np.int32)
        ctx = C.c_void_p()
        rc = self.lib.gfsph_create(C.byref(p), C.byref(d), self.n, *(_dp(a) for a in arrs), _ip(it), C.byref(ctx))
        self.ctx = ctx
        self.R = 1
        if rc != 0:
            msg = self.lib.gfsph_last_error(ctx).decode() if ctx.value else ""
            self.close()
            raise GfsphError(rc, msg)
        self.nsteps_done = 0
        self._apply_time_control(case.params)

    @classmethod
    def adopt(cls, lib, ctx, case):
        """A model around a context created elsewhere (gfsph_inputs_create)."""
        m = cls.__new__(cls)
        m.lib, m.case, m.ctx, m.R = lib, case, ctx, 1
        m.n = case.particles.n
        m._dem = None
        m.nsteps_done = 0
        return m

    def _apply_time_control(self, params) -> None:
        mode = int(getattr(params, "adapt_dt", 0))
        if mode:
            try:
                self.set_adaptive_dt(mode)
            except GfsphError:
                self.close()
                raise

    def _check(self, rc):
        if rc != 0:
            raise GfsphError(rc, self.lib.gfsph_last_error(self.ctx).decode())

    # --- reference verbs ---------------------------------------------
    def time_integration_sph(self, nsteps: int = 1) -> None:
        self._check(self.lib.gfsph_step(self.ctx, int(nsteps)))
        self.nsteps_done += int(nsteps)

    def sync(self) -> None:
        self._check(self.lib.gfsph_sync(self.ctx))

    def state(self) -> dict:
        n = self.n
        out = {"x": np.zeros((n, 2)), "v": np.zeros((n, 2)), "h": np.zeros(n), "u": np.zeros(n),
               "uaux": np.zeros(n), "hsml": np.zeros(n), "flag": np.zeros(n, np.int32)}
        self._check(self.lib.gfsph_get_state(self.ctx, _dp(out["x"]), _dp(out["v"]), _dp(out["h"]), _dp(out["u"]),
                                             _dp(out["uaux"]), _dp(out["hsml"]), _ip(out["flag"])))
        return out

    def pwp_column(self) -> np.ndarray:
        """icpwp = 2: the pore-pressure columns [n][nuaux], node 0 at the bed."""
        col = np.zeros((self.n, int(self.case.params.nuaux)))
        self._check(self.lib.gfsph_get_pwp_column(self.ctx, _dp(col)))
        return col

    def set_pwp_column(self, col) -> None:
        c = np.ascontiguousarray(col, dtype=np.float64)
        if c.shape != (self.n, int(self.case.params.nuaux)):
            raise ValueError("columns must be [n][nuaux]")
        self._check(self.lib.gfsph_set_pwp_column(self.ctx, _dp(c)))

    def mass(self) -> np.ndarray:
        """Particle masses (volumes) in the caller's order; erosion changes them."""
        m = np.zeros(self.n)
        self._check(self.lib.gfsph_get_mass(self.ctx, _dp(m)))
        return m

    def scale_pwp(self, factor) -> None:
        """Initial pore pressure (u and the Uaux column) of every realisation
        times its factor: case variable 401 of Input_SW."""
        f = np.ascontiguousarray(np.broadcast_to(np.asarray(factor, dtype=np.float64), (self.R,)))
        self._check(self.lib.gfsph_batch_scale_pwp(self.ctx, _dp(f)))

    def set_mass0(self, mass0) -> None:
        """The masses before any erosion (mass0), read by the Bingham-erosion
        law nfric = 3; create takes the masses it is given."""
        m = np.ascontiguousarray(mass0, dtype=np.float64)
        if m.shape != (self.n,):
            raise ValueError("mass0 must hold one value per particle")
        self._check(self.lib.gfsph_set_mass0(self.ctx, _dp(m)))

    def erosion(self, r: int = 0) -> dict:
        """The terrain's erosion state of realisation r: current Z
        (topol(1,:)), eroded depth (topol(14,:)), available depth
        (topol(15,:)), each [ny*nx]."""
        npg = self.case.dem.nx * self.case.dem.ny
        out = {k: np.zeros(npg) for k in ("z", "eroded", "avail")}
        self._check(self.lib.gfsph_erosion_read(self.ctx, int(r), _dp(out["z"]), _dp(out["eroded"]),
                                                _dp(out["avail"])))
        return out

    def set_erosion(self, r: int = 0, z=None, eroded=None, avail=None) -> None:
        """Replaces parts of realisation r's erosion state (None: keep), as
        the reference's restart.top does."""
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (z, eroded, avail)]
        npg = self.case.dem.nx * self.case.dem.ny
        if any(a is not None and a.size != npg for a in arrs):
            raise ValueError("erosion arrays must have nx*ny values")
        ptr = [C.POINTER(C.c_double)() if a is None else _dp(a) for a in arrs]
        self._check(self.lib.gfsph_erosion_set(self.ctx, int(r), *ptr))

    def raster_idw(self):
        npg = self.case.dem.nx * self.case.dem.ny
        out = np.zeros((npg, 4))
        ws = np.zeros(npg)
        self._check(self.lib.gfsph_raster_idw(self.ctx, _dp(out), _dp(ws)))
        return out, ws

    def mc_pixel_accumulate(self) -> None:
        self._check(self.lib.gfsph_mc_pixel_accumulate(self.ctx))

    def mc_pixel_read(self, reset: bool = False):
        npg = self.case.dem.nx * self.case.dem.ny
        acc = np.zeros((npg, 4))
        cnt = np.zeros(npg, np.int32)
        self._check(self.lib.gfsph_mc_pixel_read(self.ctx, _dp(acc), _ip(cnt), int(reset)))
        return acc, cnt

    def profile_step(self, nrep: int = 10) -> dict:
        cap = 16
        nk = C.c_int32()
        names = (C.c_char_p * cap)()
        ms = np.zeros(cap)
        self._check(self.lib.gfsph_profile_step(self.ctx, int(nrep), cap, C.byref(nk), names, _dp(ms)))
        self.nsteps_done += int(nrep)
        return {names[k].decode(): float(ms[k]) for k in range(nk.value)}

    def autotune(self, reps: int = 3) -> tuple:
        """Work-group sizes of the density, force and XSPH gathers chosen by
        timing on the current state (needs one completed step)."""
        ch = np.zeros(3, np.int32)
        self._check(self.lib.gfsph_autotune(self.ctx, int(reps), _ip(ch)))
        return tuple(int(v) for v in ch)

    def set_adaptive_dt(self, mode: int) -> None:
        """ic_adapt_dt_sph: 0 fixed dt_sph, 1, 2, 4 or 10 (Adaptive_dt_sph_SW)."""
        self._check(self.lib.gfsph_set_adaptive_dt(self.ctx, int(mode)))

    def time(self):
        """(time_sph, dt_sph of the last step); arrays over the realisations
        of a batch."""
        t = np.zeros(self.R)
        dt = np.zeros(self.R)
        self._check(self.lib.gfsph_time(self.ctx, _dp(t), _dp(dt)))
        if self.R == 1:
            return float(t[0]), float(dt[0])
        return t, dt

    def pixel_read(self, r: int = 0):
        """Per-realisation pixel accumulators (batch accumulation or case
        control plot times) of realisation r."""
        npg = self.case.dem.nx * self.case.dem.ny
        acc = np.zeros((npg, 4))
        cnt = np.zeros(npg, np.int32)
        self._check(self.lib.gfsph_batch_pixel_read(self.ctx, int(r), _dp(acc), _ip(cnt)))
        return acc, cnt

    # --- case control (Driver_Flow_23 + Stop_case_SW per realisation) ---
    def case_control(self, vstop: float, timestop: float, dtimestop: float, dt: float, time_end: float,
                     time_plot: float, pixels: bool = True) -> None:
        ctl = CCaseCtl(vstop, timestop, dtimestop, dt, time_end, time_plot, int(pixels), 0)
        self._check(self.lib.gfsph_batch_cases(self.ctx, C.byref(ctl)))

    def case_status(self):
        """(stopped[R], itimestep[R], number stopped)"""
        st = np.zeros(self.R, np.int32)
        it = np.zeros(self.R, np.int32)
        k = C.c_int32()
        self._check(self.lib.gfsph_batch_status(self.ctx, _ip(st), _ip(it), C.byref(k)))
        return st, it, k.value

    def run_cases(self, max_steps: int, chunk: int = 64) -> int:
        """Steps until every realisation has stopped or max_steps loop
        iterations have run; returns the iterations run."""
        done = 0
        while done < max_steps:
            k = min(chunk, max_steps - done)
            self.time_integration_sph(k)
            done += k
            if self.case_status()[2] == self.R:
                break
        return done

    def counts(self):
        n, na = C.c_int32(), C.c_int32()
        self._check(self.lib.gfsph_counts(self.ctx, C.byref(n), C.byref(na)))
        return n.value, na.value

    def close(self) -> None:
        if getattr(self, "ctx", None) is not None and self.ctx.value:
            self.lib.gfsph_destroy(self.ctx)
        self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class BatchModel(SPHModel):
    """Many independent realisations on one DEM advanced together (the
    reference's Monte Carlo loop run in one launch sequence).  ``cases``
    share the DEM and the solver switches; each brings its particles and its
    const(:) vector.  ``state()`` returns all particles stacked in case
    order; ``split(state)`` cuts it per realisation."""

    def __init__(self, cases, init_input: bool = True, check_finite: bool = False):
        self.lib = load_library()
        self.cases = list(cases)
        self.case = self.cases[0]
        self.R = len(self.cases)
        self.sizes = np.array([c.particles.n for c in self.cases], dtype=np.int32)
        self.n = int(self.sizes.sum())
        self._dem = _DemArgs(self.case.dem)
        p = make_cparams(self.case.params, init_input, check_finite)
        cst = np.ascontiguousarray(np.stack([c.params.cst() for c in self.cases]), dtype=np.float64)
        arrs = [np.ascontiguousarray(np.concatenate([getattr(c.particles, k) for c in self.cases]), dtype=np.float64)
                for k in ("x", "v", "h", "mass", "hsml", "u")]
        ctx = C.c_void_p()
        rc = self.lib.gfsph_batch_create(C.byref(p), C.byref(self._dem.c), self.R, _ip(self.sizes), _dp(cst),
                                         *(_dp(a) for a in arrs), C.byref(ctx))
        self.ctx = ctx
        if rc != 0:
            msg = self.lib.gfsph_last_error(ctx).decode() if ctx.value else ""
            self.close()
            raise GfsphError(rc, msg)
        self.nsteps_done = 0
        self._apply_time_control(self.case.params)

    @classmethod
    def adopt(cls, lib, ctx, cases):
        """A batch around a context created elsewhere (gfsph_inputs_batch_create)."""
        b = cls.__new__(cls)
        b.lib, b.ctx, b.cases = lib, ctx, list(cases)
        b.case = b.cases[0]
        b.R = len(b.cases)
        b.sizes = np.array([c.particles.n for c in b.cases], dtype=np.int32)
        b.n = int(b.sizes.sum())
        b._dem = None
        b.nsteps_done = 0
        return b

    def split(self, st: dict) -> list:
        edges = np.concatenate([[0], np.cumsum(self.sizes)])
        return [{k: v[edges[r]:edges[r + 1]] for k, v in st.items()} for r in range(self.R)]

    def pixel_accumulate(self) -> None:
        self._check(self.lib.gfsph_batch_pixel_accumulate(self.ctx))


    def hazard(self, hthr: float = 0.5) -> dict:
        npg = self.case.dem.nx * self.case.dem.ny
        stats = np.zeros((npg, 4, 2))
        nprob = np.zeros(npg, np.int32)
        maxs = np.zeros((npg, 2))
        pexc = np.zeros(npg)
        self._check(self.lib.gfsph_batch_hazard(self.ctx, float(hthr), _dp(stats), _ip(nprob), _dp(maxs), _dp(pexc)))
        return {"mean": stats[:, :, 0], "std": stats[:, :, 1], "nprob": nprob, "max_h": maxs[:, 0],
                "max_v": maxs[:, 1], "p_exceed": pexc}

    def write_hazard(self, path: str, hthr: float = 0.5, title: str = "") -> dict:
        """hazard(hthr) as a NetCDF classic file: mean and standard deviation
        over the realisations of the average depth and speed at each node,
        maxima, the number of realisations reaching the node and P(max h >
        hthr)."""
        hz = self.hazard(hthr)
        r = {"mean_h": (hz["mean"][:, 0], "m", "mean over realisations of the time-averaged depth"),
             "std_h": (hz["std"][:, 0], "m", "standard deviation over realisations of the time-averaged depth"),
             "mean_v": (hz["mean"][:, 2], "m s-1", "mean over realisations of the time-averaged speed"),
             "std_v": (hz["std"][:, 2], "m s-1", "standard deviation over realisations of the time-averaged speed"),
             "max_h": (hz["max_h"], "m", "maximum depth over realisations and plot times"),
             "max_v": (hz["max_v"], "m s-1", "maximum speed over realisations and plot times"),
             "n_reached": (hz["nprob"].astype(np.float64), "1", "number of realisations reaching the node"),
             "p_exceed": (hz["p_exceed"], "1", f"probability that the maximum depth exceeds {hthr} m")}
        write_netcdf(path, self.case.dem, r, title or f"Monte Carlo hazard, {self.R} realisations")
        return hz


class _Adjoint:
    """The reverse mode of a context: ``forward(nsteps, ckpt)`` advances and
    keeps checkpoints, ``misfit(target)`` gives the deposit misfit of the
    current state (per realisation) and its cotangents, ``backward(...)`` the
    gradients with respect to the law's two parameters — (mu, xi) =
    (const(9), const(3)) for nfric 5, (tau_y, mu_B) = (const(6), const(7)) for
    Bingham, nfric 4 — the particle masses (volumes) and the initial
    positions and velocities."""

    def forward(self, nsteps: int, ckpt_every: int = 50, tape_steps: int = 0) -> None:
        """Advance nsteps with checkpoints every ckpt_every steps; the backward
        pass records at most tape_steps steps at once (0: a whole segment)."""
        self._check(self.lib.gfsph_adjoint_set_tape(self.ctx, int(tape_steps)))
        self._check(self.lib.gfsph_adjoint_forward(self.ctx, int(nsteps), int(ckpt_every)))
        self.nsteps_done += int(nsteps)

    def _nreal(self) -> int:
        return int(getattr(self, "R", 1))

    def misfit(self, target):
        t = np.ascontiguousarray(target, dtype=np.float64).ravel()
        lx = np.zeros((self.n, 2))
        lh = np.zeros(self.n)
        L = np.zeros(self._nreal())
        self._check(self.lib.gfsph_raster_misfit(self.ctx, _dp(t), _dp(L), _dp(lx), _dp(lh)))
        return (float(L[0]) if self._nreal() == 1 else L), lx, lh

    def backward(self, lam_x=None, lam_v=None, lam_h=None) -> dict:
        def arg(a, shape):
            return None if a is None else np.ascontiguousarray(np.broadcast_to(a, shape), dtype=np.float64)

        lx, lv, lh = arg(lam_x, (self.n, 2)), arg(lam_v, (self.n, 2)), arg(lam_h, (self.n,))
        R = self._nreal()
        gmu, gxi = np.zeros(R), np.zeros(R)
        gm = np.zeros(self.n)
        gx = np.zeros((self.n, 2))
        gv = np.zeros((self.n, 2))
        nul = C.POINTER(C.c_double)()
        self._check(self.lib.gfsph_adjoint_backward(self.ctx, _dp(lx) if lx is not None else nul,
                                                    _dp(lv) if lv is not None else nul,
                                                    _dp(lh) if lh is not None else nul, _dp(gmu), _dp(gxi),
                                                    _dp(gm), _dp(gx), _dp(gv)))
        if R == 1:
            return {"mu": float(gmu[0]), "xi": float(gxi[0]), "mass": gm, "x0": gx, "v0": gv}
        return {"mu": gmu, "xi": gxi, "mass": gm, "x0": gx, "v0": gv}


class AdjointModel(_Adjoint, SPHModel):
    """One run with its reverse mode (see _Adjoint)."""


class BatchAdjointModel(_Adjoint, BatchModel):
    """A batch of realisations with their reverse mode: every realisation's
    misfit against one target raster and its own parameter gradients, from
    one batched forward and one batched backward sweep."""


def neighbour_pairs(x, hsml, flag=None, skf: int = 1) -> np.ndarray:
    """Neighbour set (i<j, 0-based, sorted) computed on the GPU."""
    L = load_library()
    x = np.ascontiguousarray(x, dtype=np.float64)
    hsml = np.ascontiguousarray(hsml, dtype=np.float64)
    n = x.shape[0]
    fl = np.zeros(n, np.int32) if flag is None else np.ascontiguousarray(flag, dtype=np.int32)
    cap = max(16, 40 * n)
    for _ in range(3):
        ij = np.zeros((cap, 2), np.int32)
        npairs = C.c_int64()
        rc = L.gfsph_neighbour_pairs(skf, n, _dp(x), _dp(hsml), _ip(fl), cap, C.byref(npairs), _ip(ij))
        if rc == -5:
            cap = int(npairs.value) + 16
            continue
        if rc != 0:
            raise GfsphError(rc, "gfsph_neighbour_pairs")
        ij = ij[: npairs.value]
        order = np.lexsort((ij[:, 1], ij[:, 0]))
        return ij[order]
    raise GfsphError(-5, "pair capacity")
