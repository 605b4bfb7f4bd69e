"""Loads the golden fixtures written by tests/golden/make_golden.py."""
from __future__ import annotations

import glob
import json
import os

import numpy as np

from gfsph import cases as C
from gfsph.dem import DEM, TerrainZones, Zone

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

STATE_CASES = ["g_voellmy_plane", "g_frictional_pwp", "g_bingham", "g_curvature7", "g_mui8", "g_icpwp1_sle1",
               "g_manning1", "g_perzyna9", "g_hb25", "g_bagnold11", "g_kernel_skf2", "g_kernel_skf3",
               "g_kernel_skf4", "g_domain_exit", "g_adaptive_dt2", "g_adaptive_dt7", "g_adaptive_dt10", "g_erosion_hungr",
               "g_erosion_slope", "g_erosion_stress", "g_pwp_column_lin", "g_pwp_column_cos", "g_pwp_column_step",
               "g_k0", "g_costheta_k0", "g_k0_pwp", "g_nfric22", "g_hb23_m1", "g_hb23_m2", "g_hb23_m3",
               "g_nfric3_erosion", "g_sle2", "g_sle3", "g_zones_friction", "g_zones_erosion",
               "g_zones_erosion_main", "g_zones_pwp"]
STAGED_CASES = ["g_k0", "g_costheta_k0", "g_k0_pwp"]
PWP_COLUMN_CASES = ["g_pwp_column_lin", "g_pwp_column_cos", "g_pwp_column_step", "g_zones_pwp"]
EROSION_CASES = ["g_erosion_hungr", "g_erosion_slope", "g_erosion_stress", "g_nfric3_erosion", "g_zones_erosion",
                 "g_zones_erosion_main"]
ZONE_CASES = ["g_zones_friction", "g_zones_erosion", "g_zones_erosion_main", "g_zones_pwp"]
ALL_CASES = STATE_CASES + ["g_random_neighbours"]


def load(name: str) -> dict:
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def case_of(name: str):
    """Rebuilds the Case (inputs) of a fixture and returns (case, fixture)."""
    f = load(name)
    pj = json.loads(bytes(f["params_json"]).decode())
    cst = f["const"]
    const = {k: float(cst[k]) for k in range(1, 111) if cst[k] != 0.0}
    for k in ("pwp_bc_type", "pwp_bc_val"):
        if k in pj:
            pj[k] = tuple(pj[k])
    prm = C.Params(**pj, const=const)
    xmin, ymin, dx, dy = (float(v) for v in f["dem_geom"])
    dem = DEM(Z=f["dem_Z"], xmin=xmin, ymin=ymin, dx=dx, dy=dy)
    if "zone_type" in f:
        zs = [Zone(int(t), int(l), tuple(p)) for t, l, p in zip(f["zone_type"], f["zone_law"], f["zone_props"])]
        ep = None
        if "eros_pts" in f:
            ep = {"pts": f["eros_pts"], "type": int(f["eros_type_dxs"][0]), "dx_s": float(f["eros_type_dxs"][1])}
        dem.zones = TerrainZones(zs, nodes=f.get("zone_nodes"), avail=f.get("zone_avail"), eros_pts=ep)
    n = f["p_x"].shape[0]
    P = C.Particles(x=f["p_x"], v=f["p_v"], h=f["p_h"], mass=f["p_mass"], hsml=f["p_hsml"], u=f["p_u"],
                    itype=np.ones(n, np.int32))
    return C.Case(name, prm, dem, P, int(f["nsteps"])), f


def rel_err(a, b):
    """Field-normwise relative error max|a-b| / max|b| (max|b| floored at
    1e-300).  Normwise, because a per-particle relative bound is not
    attainable by any reordering of the sums: particles near the yield
    threshold of the frictional law amplify a one-ulp change of the pressure
    sum (reversing the reference's own summation order moves such a
    particle's velocity by ~1e-10 relative to the field)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if not a.size:
        return 0.0
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))
