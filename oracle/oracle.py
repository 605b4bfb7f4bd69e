"""ctypes binding of the oracle library (liboracle_gfsph.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to check the GPU's results or to time the CPU
baseline.  The product path
(the HIP library) never calls it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_gfsph.so")


class GfoParams(C.Structure):
    _fields_ = [("integrator", C.c_int32), ("pa_sph", C.c_int32), ("skf", C.c_int32), ("sle", C.c_int32),
                ("summation_density", C.c_int32), ("average_velocity", C.c_int32), ("nor_density", C.c_int32),
                ("icpwp", C.c_int32), ("npoin0", C.c_int32), ("rev_order", C.c_int32),
                ("adapt_dt", C.c_int32), ("pad2_", C.c_int32), ("dt", C.c_double), ("h_inf", C.c_double), ("pwp0rel", C.c_double),
                ("cst", C.c_double * 111), ("nuaux", C.c_int32), ("pwp_bc_type", C.c_int32 * 2),
                ("pwp0_type", C.c_int32), ("pwp_bc_val", C.c_double * 2), ("xirel", C.c_double),
                ("ic_k0", C.c_int32), ("pad3_", C.c_int32), ("k0_active", C.c_double), ("k0_pasive", C.c_double),
                ("mass0", C.POINTER(C.c_double))]


class GfoZone(C.Structure):
    _fields_ = [("type", C.c_int32), ("law", C.c_int32), ("props", C.c_double * 16)]


class GfoDem(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("xmax_set", C.c_int32), ("pad_", C.c_int32),
                ("xmin", C.c_double), ("ymin", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
                ("xmax", C.c_double), ("ymax", C.c_double),
                ("topol", C.POINTER(C.c_double)), ("in_model", C.POINTER(C.c_int32)),
                ("coorg", C.POINTER(C.c_double)), ("nzones", C.c_int32), ("pad2_", C.c_int32),
                ("zones", C.POINTER(GfoZone)), ("zone_nodes", C.POINTER(C.c_int32))]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "liboracle_gfsph.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.gfo_kernel.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, dp, dp, dp]
        L.gfo_pairs.argtypes = [C.c_int, C.c_int, dp, dp, ip, C.c_int, C.c_int64, ip, dp, dp]
        L.gfo_pairs.restype = C.c_int64
        L.gfo_topo_lookup.argtypes = [C.POINTER(GfoDem), C.c_int, dp, dp]
        L.gfo_which_zone.argtypes = [C.POINTER(GfoDem), dp, C.c_int, dp, C.c_int, ip]
        L.gfo_fdiff_derivs.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, ip, dp]
        L.gfo_init.argtypes = [C.POINTER(GfoParams), C.POINTER(GfoDem), C.c_int, dp, dp, dp, dp, dp, dp, ip]
        L.gfo_run.argtypes = [C.POINTER(GfoParams), C.POINTER(GfoDem), C.c_int, dp, dp, dp, dp, dp, dp, dp, ip,
                              C.c_int, dp, dp, dp]
        L.gfo_raster_idw.argtypes = [C.POINTER(GfoDem), C.c_int, dp, dp, dp, dp, ip, dp, dp]
        L.gfo_mc_pixel_accum.argtypes = [C.POINTER(GfoDem), C.c_int, dp, dp, dp, ip, dp, ip]
        L.gfo_mc_hazard.argtypes = [C.c_int64, C.c_int, dp, ip, C.c_double, dp, ip, dp, dp]
        L.gfo_adaptive_dt.argtypes = [C.POINTER(GfoParams), C.POINTER(GfoDem), C.c_int, dp, dp, dp, dp, ip, C.c_double]
        L.gfo_adaptive_dt.restype = C.c_double
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def make_params(prm) -> GfoParams:
    p = GfoParams()
    p.integrator, p.pa_sph, p.skf, p.sle = prm.integrator, prm.pa_sph, prm.skf, prm.sle
    p.summation_density, p.average_velocity, p.nor_density = prm.summation_density, prm.average_velocity, prm.nor_density
    p.icpwp, p.npoin0 = prm.icpwp, prm.npoin0
    p.adapt_dt = getattr(prm, "adapt_dt", 0)
    p.dt, p.h_inf, p.pwp0rel = prm.dt, prm.h_inf, prm.pwp0rel
    c = prm.cst()
    for k in range(111):
        p.cst[k] = c[k]
    p.nuaux = int(getattr(prm, "nuaux", 0)) if prm.icpwp == 2 else 0
    bt, bv = getattr(prm, "pwp_bc_type", (2, 1)), getattr(prm, "pwp_bc_val", (0.0, 0.0))
    p.pwp_bc_type[0], p.pwp_bc_type[1] = int(bt[0]), int(bt[1])
    p.pwp_bc_val[0], p.pwp_bc_val[1] = float(bv[0]), float(bv[1])
    p.pwp0_type = int(getattr(prm, "pwp0_type", 0))
    p.xirel = float(getattr(prm, "xirel", 0.0))
    p.ic_k0 = int(getattr(prm, "ic_k0", 0))
    p.k0_active = float(getattr(prm, "k0_active", 1.0))
    p.k0_pasive = float(getattr(prm, "k0_pasive", 1.0))
    return p


def ncol(prm) -> int:
    """Nodes of the pore-pressure column per particle (1: a single value)."""
    return int(prm.nuaux) if prm.icpwp == 2 and int(getattr(prm, "nuaux", 0)) > 1 else 1


class DemHandle:
    """Keeps the numpy buffers alive while the C struct points at them."""

    def __init__(self, dem, topol=None):
        self.topol = np.ascontiguousarray(dem.topol() if topol is None else topol, dtype=np.float64)
        self.in_model = None if dem.in_model is None else np.ascontiguousarray(dem.in_model, dtype=np.int32)
        d = GfoDem()
        d.nx, d.ny, d.xmax_set = dem.nx, dem.ny, 1
        d.xmin, d.ymin, d.dx, d.dy = dem.xmin, dem.ymin, dem.dx, dem.dy
        d.xmax, d.ymax = dem.xmax, dem.ymax
        d.topol = _dp(self.topol)
        d.in_model = _ip(self.in_model) if self.in_model is not None else C.POINTER(C.c_int32)()
        d.coorg = C.POINTER(C.c_double)()
        tz = getattr(dem, "zones", None)
        self._zones = self._zn = None
        if tz is not None and len(tz.zones):
            self._zones = (GfoZone * len(tz.zones))()
            for k, z in enumerate(tz.zones):
                self._zones[k].type, self._zones[k].law = int(z.type), int(z.law)
                for j, v in enumerate(z.props):
                    self._zones[k].props[j] = float(v)
            d.nzones = len(tz.zones)
            d.zones = C.cast(self._zones, C.POINTER(GfoZone))
            if tz.nodes is not None:
                self._zn = np.ascontiguousarray(tz.nodes, dtype=np.int32)
                d.zone_nodes = _ip(self._zn)
        self.s = d


def kernel(skf, r, dx, dy, h):
    w, a, b = C.c_double(), C.c_double(), C.c_double()
    lib().gfo_kernel(skf, r, dx, dy, h, C.byref(w), C.byref(a), C.byref(b))
    return w.value, a.value, b.value


def pairs(x, hsml, flag=None, skf=1, brute=False):
    """Neighbour set: returns (ij[np,2] int32 with i<j, w[np], dwdx[np,2])."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    hsml = np.ascontiguousarray(hsml, dtype=np.float64)
    n = x.shape[0]
    flag = np.zeros(n, np.int32) if flag is None else np.ascontiguousarray(flag, dtype=np.int32)
    L = lib()
    npairs = L.gfo_pairs(n, skf, _dp(x), _dp(hsml), _ip(flag), int(brute), 0, None, None, None)
    ij = np.zeros((max(npairs, 1), 2), np.int32)
    w = np.zeros(max(npairs, 1))
    dw = np.zeros((max(npairs, 1), 2))
    L.gfo_pairs(n, skf, _dp(x), _dp(hsml), _ip(flag), int(brute), npairs, _ip(ij), _dp(w), _dp(dw))
    return ij[:npairs], w[:npairs], dw[:npairs]


def topo_lookup(dem, xy):
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    out = np.zeros((xy.shape[0], 9))
    h = DemHandle(dem)
    lib().gfo_topo_lookup(C.byref(h.s), xy.shape[0], _dp(xy), _dp(out))
    return out


class State:
    def copy(self):
        c = State.__new__(State)
        for k, v in self.__dict__.items():
            setattr(c, k, None if v is None else v.copy())
        return c

    def __init__(self, parts, n_uaux=True):
        self.x = np.ascontiguousarray(parts.x, dtype=np.float64).copy()
        self.v = np.ascontiguousarray(parts.v, dtype=np.float64).copy()
        self.h = np.ascontiguousarray(parts.h, dtype=np.float64).copy()
        self.mass = np.ascontiguousarray(parts.mass, dtype=np.float64).copy()
        self.mass0 = self.mass.copy()  # masses before any erosion (mass0, 5:41)
        self.hsml = np.ascontiguousarray(parts.hsml, dtype=np.float64).copy()
        self.u = np.ascontiguousarray(parts.u, dtype=np.float64).copy()
        self.uaux = np.zeros_like(self.h)
        self.flag = np.zeros(self.h.shape[0], np.int32)
        self.tau0 = np.zeros(1)  # Bingham Tau0 carried between calls
        self.dt_time = None      # [dt_sph, time_sph] carried between calls
        self.eros = None         # [npoig, 3] topol(1), topol(14), topol(15) when erosion is on


def erosion_on(prm, dem=None) -> bool:
    """ic_erosion (6:1180-1184, 6:1288-1290): const(21) > 1e-9, or the older
    form const(4) > 1e-4, or an erosion zone of the terrain."""
    c = prm.cst()
    tz = getattr(dem, "zones", None)
    zones = tz is not None and any(int(z.type) == 2 for z in tz.zones)
    return bool(c[21] > 1e-9 or c[4] > 1e-4 or zones)


def erosion_grid(case):
    """Initial erosion state of the terrain: topol(1) = Z, topol(14) = 0 and
    the available depth topol(15) = const(48) if given, else 999 with a
    surface law (6:1294-1301); with erosion zones alone the terrain's
    topol(15) with values below 1e-6 set to 0 (6:1302-1307)."""
    c = case.params.cst()
    g = np.zeros((case.dem.nx * case.dem.ny, 3))
    g[:, 0] = np.asarray(case.dem.Z, dtype=np.float64).ravel()
    c21 = 1.0 if c[4] > 1e-4 else c[21]
    tz = getattr(case.dem, "zones", None)
    if c21 > 1e-9 or tz is None:
        g[:, 2] = c[48] if c[48] > 1e-6 else 999.0
    elif tz.avail is not None:
        a = np.asarray(tz.avail, dtype=np.float64).ravel()
        g[:, 2] = np.where(a < 1e-6, 0.0, a)
    return g


def which_zone(dem, xy, btype, zc=None):
    """Which_Topo_Zone (2:2635-2808) at the points xy [n, 2]: zone index or -1."""
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    out = np.zeros(xy.shape[0], np.int32)
    dh = DemHandle(dem)
    lib().gfo_which_zone(C.byref(dh.s), None if zc is None else _dp(np.ascontiguousarray(zc, dtype=np.float64)),
                         xy.shape[0], _dp(xy), int(btype), _ip(out))
    return out


def init(case, state=None):
    """Input_SW initialisation (re-summed h, pwp) -> State."""
    st = State(case.particles) if state is None else state
    if ncol(case.params) > 1 and st.uaux.ndim == 1:
        st.uaux = np.zeros((st.h.shape[0], ncol(case.params)))
    p = make_params(case.params)
    dh = DemHandle(case.dem)
    lib().gfo_init(C.byref(p), C.byref(dh.s), st.h.shape[0], _dp(st.x), _dp(st.h), _dp(st.mass), _dp(st.hsml),
                   _dp(st.u), _dp(st.uaux), _ip(st.flag))
    if erosion_on(case.params, case.dem) and st.eros is None:
        st.eros = erosion_grid(case)
    return st


def run(case, st, nsteps, rev_order=False):
    """Advances ``st`` by ``nsteps``.  rev_order=True sums every pair loop in
    reverse order: the spread between the two orders measures how strongly a
    case amplifies rounding (its conditioning)."""
    p = make_params(case.params)
    p.rev_order = int(rev_order)
    p.mass0 = _dp(st.mass0)
    if st.dt_time is None:
        st.dt_time = np.array([case.params.dt, 0.0])
    dh = DemHandle(case.dem)
    if erosion_on(case.params, case.dem) and st.eros is None:
        st.eros = erosion_grid(case)
    rc = lib().gfo_run(C.byref(p), C.byref(dh.s), st.h.shape[0], _dp(st.x), _dp(st.v), _dp(st.h), _dp(st.mass),
                       _dp(st.hsml), _dp(st.u), _dp(st.uaux), _ip(st.flag), nsteps, _dp(st.tau0),
                       _dp(st.dt_time), None if st.eros is None else _dp(st.eros))
    if rc != 0:
        raise RuntimeError(f"gfo_run failed: {rc}")
    return st


def raster_idw(dem, st):
    np_ = dem.nx * dem.ny
    out = np.zeros((np_, 4))
    ws = np.zeros(np_)
    dh = DemHandle(dem)
    lib().gfo_raster_idw(C.byref(dh.s), st.h.shape[0], _dp(st.x), _dp(st.v), _dp(st.h), _dp(st.u), _ip(st.flag),
                         _dp(out), _dp(ws))
    return out, ws


def mc_pixel_accum(dem, st, acc=None, cnt=None):
    np_ = dem.nx * dem.ny
    acc = np.zeros((np_, 4)) if acc is None else acc
    cnt = np.zeros(np_, np.int32) if cnt is None else cnt
    dh = DemHandle(dem)
    lib().gfo_mc_pixel_accum(C.byref(dh.s), st.h.shape[0], _dp(st.x), _dp(st.v), _dp(st.h), _ip(st.flag),
                             _dp(acc), _ip(cnt))
    return acc, cnt


def mc_hazard(acc, cnt, hthr=0.5):
    """Norm_MC_pixel_Vars_SW ("6 SPH_SW_2019.f90":20838-20908) over the
    per-problem pixel accumulators acc[R][npoig][4], cnt[R][npoig]: returns
    {"pr": [npoig][4][4] MC_Pr rows (sum, sum of squares, mean, standard
    deviation) x variables (average h, max h, average |v|, max |v|),
    "nprob", "maxs" [npoig][2], "pexc"}."""
    acc = np.ascontiguousarray(acc, dtype=np.float64)
    cnt = np.ascontiguousarray(cnt, dtype=np.int32)
    R, npg = cnt.shape
    pr = np.zeros((npg, 4, 4))
    nprob = np.zeros(npg, np.int32)
    maxs = np.zeros((npg, 2))
    pexc = np.zeros(npg)
    lib().gfo_mc_hazard(npg, R, _dp(acc), _ip(cnt), float(hthr), _dp(pr), _ip(nprob), _dp(maxs), _dp(pexc))
    return {"pr": pr, "nprob": nprob, "maxs": maxs, "pexc": pexc}


def run_case(case, st, vstop, timestop, dtimestop, dt_driver, time_end, time_plot, max_steps, pixels=True):
    """The step loop the reference runs for one case of a case table or
    Monte Carlo set (Driver_Flow_23, "8 SPH_FEM_Driver_2019.f90":1444-1548)
    with Stop_case_SW ("6 SPH_SW_2019.f90":20114-20177): adaptive dt_sph,
    the step cut to end on the next plot time and restored after it, pixel
    accumulation at plot times (Get_MC_pixel_Vars_SW), the stop conditions.
    Advances ``st``; returns (iterations, stopped, pixel acc, pixel cnt).
    (The reference's Driver_Flow_23 lives in the file of its main program,
    which the reference build here replaces with its harness.)"""
    import copy
    prm = make_params(case.params)
    dh = DemHandle(case.dem)
    n = st.h.shape[0]
    adapt = int(getattr(case.params, "adapt_dt", 0))
    if st.dt_time is None:
        st.dt_time = np.array([case.params.dt, 0.0])
    fixed = copy.deepcopy(case)
    fixed.params.adapt_dt = 0
    npg = case.dem.nx * case.dem.ny
    acc, cnt = np.zeros((npg, 4)), np.zeros(npg, np.int32)
    tplot, tglob = 0.0, float(st.dt_time[1])
    dt_sph = float(st.dt_time[0])
    it = 0
    stopped = False
    while it < max_steps and not stopped:
        it += 1
        if adapt:
            dt_sph = lib().gfo_adaptive_dt(C.byref(prm), C.byref(dh.s), n, _dp(st.x), _dp(st.v), _dp(st.h),
                                           _dp(st.hsml), _ip(st.flag), dt_sph)
        dt_keep = dt_sph
        if time_plot > 0.0:
            tpr = tplot + dt_sph
            if tpr >= time_plot:
                dt_sph = int(tpr / time_plot) * time_plot - tplot
            tplot = tplot + dt_sph
        st.dt_time[0] = dt_sph
        run(fixed, st, 1)
        tglob = tglob + dt_sph
        if time_plot > 0.0 and tplot >= time_plot:
            if pixels:
                mc_pixel_accum(case.dem, st, acc, cnt)
            dt_sph = dt_keep
            tplot = 0.0
        act = st.flag == 0
        v = float(np.sum(np.abs(st.v[act, 0]))) / n
        if it > 1:
            if v < vstop and it > 10:
                stopped = True
            if int(np.sum(st.flag != 0)) > n // 5:
                stopped = True
            if st.dt_time[1] > timestop:
                stopped = True
            if dt_driver < dtimestop:
                stopped = True
        if tglob >= time_end:
            stopped = True
    st.dt_time[0] = dt_sph
    return it, stopped, acc, cnt
