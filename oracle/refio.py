"""Drive the reference Fortran build (oracle/_ref/gfsph_ref) on a case.

Writes the reference's own positional ASCII inputs -- <name>.dat
("6 SPH_SW_2019.f90":812-1570), an explicit-grid <topo>.top with ictop=100
("2 SPH_FEM_topo_2019.f90":208-252, finite-difference slopes) and an exact
restart.dat particle set (icunk_s=200, "6 SPH_SW_2019.f90":3270-3340, 17
significant digits) -- runs the harness and reads its binary dumps.

TEST INFRASTRUCTURE ONLY (golden-fixture generation and the optional
reference CPU baseline).
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_EXE = os.path.join(HERE, "_ref", "gfsph_ref")
# the same harness with time_integration_sph replaced by the product's
# ISO_C_BINDING drop-in (SPH_MAIN_GFSPH -> libgfsph.so)
REF_GPU_EXE = os.path.join(HERE, "_ref", "gfsph_ref_gpu")


# stdout of the last harness run (the harness prints its step-loop time)
LAST_STDOUT = ""


def available() -> bool:
    return os.path.exists(REF_EXE)


def _r(v: float) -> str:
    return repr(float(v))


def write_inputs(d: str, case, name: str = "case", topo: str = "topo", release=None) -> None:
    """The reference's input files for ``case``: <name>.dat, <topo>.top and
    restart.dat (the release as a restart, icunk_s 200), or with ``release``
    = ("rect", (npoinx, npoiny), (x0, y0, x1, y1, b, d0, facthsml)) or
    ("ellipse", (xp0, yp0, a, b, theta0, hp0), (drr, facthsml)) the
    release generated by the reference itself (icunk_s 11, 12: Read_rect_SW
    6:4426-4497, Read_ellipsoidal_SW 6:3818-3899)."""
    prm, dem, P = case.params, case.dem, case.particles
    c = prm.cst()
    # const(16:30), (31:45), (46:60) follow on lines of their own when const(11),
    # const(30), const(45) are negative (Input_SW 6:1186-1204)
    if any(c[k] != 0.0 for k in range(61, 111)):
        raise ValueError("only const(1:60) are written by this helper")
    nblk = 0
    for b, (lo, flag) in enumerate(((16, 11), (31, 30), (46, 45))):
        if any(c[k] != 0.0 for k in range(lo, lo + 15)):
            nblk = b + 1
    for b, flag in enumerate((11, 30, 45)):
        if b < nblk and not c[flag] < 0:
            raise ValueError(f"const({flag}) must be negative to pass const({16 + 15 * b}:...)")
    L = []
    L += ["1", "synthetic case written for the reference oracle"]
    L += ["ic_SW_Alg", "0"]
    L += ["nhist", "0"]
    L += ["ndimn", "2"]
    L += ["ic_soil ic_water ic_vps ic_abs", "1 0 0 0"]
    if release is None:
        L += ["Soil unkno  h_inf_SW", f"200 {_r(prm.h_inf)}"]
        L += ["restart name", name]
    elif release[0] == "rect":
        L += ["Soil unkno  h_inf_SW", f"11 {_r(prm.h_inf)}"]
        L += ["npoinx npoiny", " ".join(str(int(v)) for v in release[1])]
        L += ["x0 y0 x1 y1 b d0 facthsml", " ".join(_r(v) for v in release[2])]
    elif release[0] == "rect_ellipsoid":
        L += ["Soil unkno  h_inf_SW", f"13 {_r(prm.h_inf)}"]
        L += ["npoinx npoiny", " ".join(str(int(v)) for v in release[1])]
        L += ["x0 y0 x1 y1 b d0 facthsml", " ".join(_r(v) for v in release[2])]
    elif release[0] == "cylinder":  # (xp0, yp0, rp0, hp0, icdens), (rp1, hp1) or None, (nlinep, facthsml)
        c5 = release[1]
        L += ["Soil unkno  h_inf_SW", f"1 {_r(prm.h_inf)}"]
        L += ["xp0 yp0 rp0 hp0 icdens", " ".join(_r(v) for v in c5[:4]) + f" {int(c5[4])}"]
        if int(c5[4]) == 2:
            L += ["rp1 hp1", " ".join(_r(v) for v in release[2])]
        L += ["nlinep facthsml", f"{int(release[3][0])} {_r(release[3][1])}"]
    elif release[0] == "ellipse":
        L += ["Soil unkno  h_inf_SW", f"12 {_r(prm.h_inf)}"]
        L += ["xp0 yp0 a b theta0 hp0", " ".join(_r(v) for v in release[1])]
        L += ["drr facthsml", " ".join(_r(v) for v in release[2])]
    else:
        raise ValueError(f"release form {release[0]}")
    L += ["pa_sph nnps sle skf", f"{prm.pa_sph} {prm.nnps} {prm.sle} {prm.skf}"]
    tf = lambda b: "T" if b else "F"
    L += ["sum_den av_vel virt_part nor_dens",
          f"{tf(prm.summation_density)} {tf(prm.average_velocity)} F {tf(prm.nor_density)}"]
    L += ["const 1..15", " ".join(_r(c[k]) for k in range(1, 16))]
    for b in range(nblk):
        lo = 16 + 15 * b
        L += [f"const {lo}..{lo + 14}", " ".join(_r(c[k]) for k in range(lo, lo + 15))]
    nfric = int(c[5] + 0.01)
    if (5 <= nfric <= 10) or nfric in (26, 27):
        k0 = int(getattr(prm, "ic_k0", 0))
        L += ["K0 activated?", str(k0)]
        if k0 == 1:  # 6:1383-1389
            L += ["K0_active K0_pasive", f"{_r(prm.k0_active)} {_r(prm.k0_pasive)}"]
    elif int(getattr(prm, "ic_k0", 0)):
        raise ValueError("the reference reads ic_K0 only for frictional laws")
    L += ["icpwp", str(prm.icpwp)]
    if prm.icpwp in (1, 3):
        L += ["pwp0rel", _r(prm.pwp0rel)]
    elif prm.icpwp == 2:  # Initialize_Pwp (6:1743-1768)
        bt, bv = prm.pwp_bc_type, prm.pwp_bc_val
        L += ["nUaux", str(int(prm.nuaux))]
        L += ["BcT1 BcV1 BcT2 BcV2", f"{int(bt[0])} {_r(bv[0])} {int(bt[1])} {_r(bv[1])}"]
        L += ["pwp0_type pwp0rel", f"{int(prm.pwp0_type)} {_r(prm.pwp0rel)}"]
        if int(prm.pwp0_type) == 4:
            L += ["xirel", _r(prm.xirel)]
    elif prm.icpwp != 0:
        raise ValueError("icpwp must be 0, 1, 2 or 3")
    L += ["coarse mesh saving utility?", "0"]
    L += ["control points?", "0"]
    L += ["GID filter", " ".join(str(int(m)) for m in getattr(case, "gid_mask", (1, 1, 1, 1, 0, 1, 0, 0, 0, 1, 0, 0)))]
    L += ["T_change_to_W", "1.e+12"]
    with open(os.path.join(d, name + ".dat"), "w") as f:
        f.write("\n".join(L) + "\n")

    nx, ny = dem.nx, dem.ny
    xs = dem.xmin + np.arange(nx) * dem.dx
    ys = dem.ymin + np.arange(ny) * dem.dy
    Z = np.asarray(dem.Z, dtype=np.float64)
    with open(os.path.join(d, topo + ".top"), "w") as f:
        f.write("explicit grid, finite-difference slopes\n100\nnpoigx npoigy\n")
        f.write(f"{nx} {ny}\nipoig x y z\n")
        k = 1
        rows = []
        for iy in range(ny):
            for ix in range(nx):
                rows.append(f"{k} {_r(xs[ix])} {_r(ys[iy])} {_r(Z[iy, ix])}")
                k += 1
        f.write("\n".join(rows) + "\n")
        tz = getattr(dem, "zones", None)
        if tz is None or not tz.zones:
            f.write("topo props (zones) 0 = none\n0\n")
        else:  # ic_topo_props = 10 (Topo_main 2:516-576)
            f.write(f"topo props: basal zones\n10\nntopo_zones\n{len(tz.zones)}\n")
            for z in tz.zones:
                f.write(f"basal type\n{int(z.type)}\nbasal law\n{int(z.law)}\n")
                f.write("props 1..6\n" + " ".join(_r(v) for v in z.props[:6]) + "\n")
                f.write("props 7..16\n" + " ".join(_r(v) for v in z.props[6:16]) + "\n")
    tz = getattr(dem, "zones", None)
    if tz is not None and tz.eros_pts is not None:  # Get_Eros_Pts (2:943-978)
        ep = tz.eros_pts
        pts = np.asarray(ep["pts"], dtype=np.float64)
        with open(os.path.join(d, topo + ".eros.pts"), "w") as f:
            f.write("erodible points written for the reference oracle\n")
            f.write(f"{pts.shape[0]} {int(ep['type'])} {_r(ep['dx_s'])}\nx y depth\n")
            f.write("\n".join(f"{_r(a)} {_r(b)} {_r(cc)}" for a, b, cc in pts) + "\n")
    eros_zones = tz is not None and any(int(z.type) == 2 for z in tz.zones)

    if c[21] > 1e-9 or c[4] > 1e-4 or eros_zones:
        # erosion with restart.dat input reads the terrain's erosion state from
        # restart.top (Get_Topo_restart, "2 SPH_FEM_topo_2019.f90":2880-2921):
        # Z, eroded depth 0, available depth (reset by Input_SW right after)
        with open(os.path.join(d, "restart.top"), "w") as f:
            f.write("terrain erosion state written for the reference oracle\n")
            f.write(f"{nx * ny} 0.0\nipoig topol(1) topol(14) topol(15)\n")
            avail = np.full(nx * ny, c[48] if c[48] > 1e-6 else 999.0)
            if not (c[21] > 1e-9 or c[4] > 1e-4):  # zones alone: the terrain's topol(15)
                avail = np.zeros(nx * ny) if tz.avail is None else np.asarray(tz.avail, dtype=np.float64).ravel()
            rows = []
            k = 1
            for iy in range(ny):
                for ix in range(nx):
                    rows.append(f"{k} {_r(Z[iy, ix])} 0.0 {_r(avail[k - 1])}")
                    k += 1
            f.write("\n".join(rows) + "\n")

    with open(os.path.join(d, "restart.dat"), "w") as f:
        f.write("restart written for the reference oracle\n")
        f.write(f"{P.n} 0.0\n")
        f.write("i itype x y vx vy h mass hsml u\n")
        rows = []
        for i in range(P.n):
            rows.append(" ".join([str(i + 1), str(int(P.itype[i])), _r(P.x[i, 0]), _r(P.x[i, 1]), _r(P.v[i, 0]),
                                  _r(P.v[i, 1]), _r(P.h[i]), _r(P.mass[i]), _r(P.hsml[i]), _r(P.u[i])]))
        f.write("\n".join(rows) + "\n")


def _law_lines(law) -> list:
    """One sampled variable: (law, params) for a continuous law (8:737-745) or
    (100, nodes, counts) for a histogram (8:1391-1406)."""
    if law[0] == 100:
        _, xh, ph = law
        return ["probab_law", "100", "histogram np_h", str(len(xh)), "nodes", " ".join(_r(v) for v in xh),
                "counts", " ".join(_r(v) for v in ph)]
    prm = list(law[1]) + [0.0] * (10 - len(law[1]))
    return ["probab_law", str(int(law[0])), "params", " ".join(_r(v) for v in prm)]


def write_master(d: str, name: str, dt: float, adapt: int = 0, cases: int = 0, case_rows=(), stop=(0.0, 1e9, 0.0),
                 lists=(), dt_driver: float | None = None, time_end: float = 1.e9, plot_step: int = 1000000,
                 nmc: int = 0, laws=(), fosm=None, records=()) -> None:
    """<name>.MASTER.dat in Driver_Input's layout ("8 SPH_FEM_Driver_2019.f90":
    229-357, 475-1236): SPH only, RK4, dt_sph and ic_adapt_dt_sph, optionally a
    case table: ic_cases_mat 1 or 2 (rows of Ccases_mat, statistic lists as
    (values, probabilities) pairs, the stop thresholds), 3 (as 2, then nmc
    rows drawn), 4 / 31 (rows, stop, nmc, one law per variable), 20 / 21
    (rows, stop, fosm = (means, sigmas)), 32 (stop, ncases, records (variable,
    law, params) of which the first with a law > 0 are the variables, nmc)."""
    dtd = dt if dt_driver is None else dt_driver
    L = ["1", "run control written for the reference's input path",
         "if_sph if_gfl ic_cases_mat", f"1 0 {int(cases)}",
         "SPH_problem_type SPH_t_integ_Alg", "1 4",
         "sph problem name", name,
         "dt time_end maxtimesteps", f"{_r(dtd)} {_r(time_end)} 1000000",
         "print_step save_step plot_step", f"1000000 1000000 {int(plot_step)}",
         "dt_sph ic_adapt", f"{_r(dt)} {int(adapt)}",
         "Ntime curves max pts", "0 0"]
    if cases == 32:
        nvar = sum(1 for r in records if r[1] > 0)
        L += ["vstop timestop dtimestop", " ".join(_r(v) for v in stop)]
        L += ["ncases_mat mx_cases ninfo_mat", f"{nvar} {len(records)} 12"]
        L += ["n_MCarlo_problems n_GZones", f"{int(nmc)} 1"]
        L += ["var law params"]
        for var, law, prm in records:
            prm = list(prm) + [0.0] * (10 - len(prm))
            L += [f"{int(var)} {int(law)} " + " ".join(_r(v) for v in prm)]
    elif cases:
        ninfo = max(len(r) for r in case_rows)
        L += ["ncases_mat ninfo_mat", f"{len(case_rows)} {ninfo}", "index val0 nval min max law ..."]
        L += [" ".join(_r(v) for v in list(r) + [0.0] * (ninfo - len(r))) for r in case_rows]
        if lists:
            L += ["statistic lists"]
            for vals, probs in lists:
                L += [" ".join(_r(v) for v in vals), " ".join(_r(v) for v in probs)]
        L += ["vstop timestop dtimestop", " ".join(_r(v) for v in stop)]
        if cases == 3:
            L += ["n_MCarlo_problems", str(int(nmc))]
        elif cases in (4, 5, 31):
            L += ["n_MCarlo_problems" + (" n_GZones" if cases in (5, 31) else ""),
                  f"{int(nmc)}" + (" 1" if cases in (5, 31) else "")]
            for law in laws:
                L += _law_lines(law)
        elif cases in (20, 21):
            mu, sg = fosm
            L += ["means sigmas", " ".join(_r(v) for v in mu), " ".join(_r(v) for v in sg)]
    L += ["win", "0"]
    with open(os.path.join(d, name + ".MASTER.dat"), "w") as f:
        f.write("\n".join(L) + "\n")


def read_dump(path: str) -> dict:
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw, dtype=np.int32, count=6)
    magic, ver, n, ndimn, has_uaux, npairs = (int(v) for v in hdr)
    assert magic == 1195725648 and ver == 1 and ndimn == 2
    off = 24
    out = {}

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a.copy()

    out["time"] = float(take(np.float64, 1)[0])
    out["x"] = take(np.float64, 2 * n).reshape(n, 2)
    out["v"] = take(np.float64, 2 * n).reshape(n, 2)
    out["h"] = take(np.float64, n)
    out["mass"] = take(np.float64, n)
    out["hsml"] = take(np.float64, n)
    out["u"] = take(np.float64, n)
    out["uaux"] = take(np.float64, n) if has_uaux else None
    out["flag"] = take(np.int32, n)
    out["x0"] = take(np.float64, 2 * n).reshape(n, 2)
    if npairs > 0:
        pi = take(np.int32, npairs)
        pj = take(np.int32, npairs)
        out["pair_w"] = take(np.float64, npairs)
        out["pair_dwdx"] = take(np.float64, 2 * npairs).reshape(npairs, 2)
        out["pairs"] = np.stack([pi - 1, pj - 1], axis=1)
    else:
        out["pairs"] = np.zeros((0, 2), np.int32)
        out["pair_w"] = np.zeros(0)
        out["pair_dwdx"] = np.zeros((0, 2))
    return out


def read_eros(path: str) -> np.ndarray:
    """eros_<step>.bin written by the harness when erosion is on:
    [npoig][3] = topol(1), topol(14), topol(15)."""
    raw = open(path, "rb").read()
    return np.frombuffer(raw, dtype=np.float64).reshape(-1, 3).copy()


def read_topo(path: str) -> dict:
    """topo.bin written by the harness after Init_sph (the terrain grid,
    topol(1:9,:) and poig_in_model)."""
    raw = open(path, "rb").read()
    magic, ver, nx, ny = (int(v) for v in np.frombuffer(raw, dtype=np.int32, count=4))
    assert magic == 1195725648 and ver == 2
    g = np.frombuffer(raw, dtype=np.float64, count=8, offset=16)
    npg = nx * ny
    top = np.frombuffer(raw, dtype=np.float64, count=9 * npg, offset=80).reshape(npg, 9).copy()
    im = np.frombuffer(raw, dtype=np.int32, count=npg, offset=80 + 72 * npg).copy()
    return {"nx": nx, "ny": ny, "xmin": g[0], "xmax": g[1], "ymin": g[2], "ymax": g[3], "dx": g[4], "dy": g[5],
            "xshift": g[6], "yshift": g[7], "topol": top, "in_model": im}


def run_reference_files(src_dir: str, name: str, files, nsteps: int, dt: float, adapt: int, dump_steps,
                        topo: str | None = None, timeout: float = 3600.0) -> tuple:
    """Run the reference harness on input files the reference reads itself
    (<name>.dat, <topo>.top, <name>.pts or restart.dat) copied from src_dir;
    returns ({step: dump}, topo dict)."""
    if not os.path.exists(REF_EXE):
        raise FileNotFoundError(REF_EXE)
    tmp = tempfile.mkdtemp(prefix="gfsph_ref_")
    try:
        for f in files:
            with open(os.path.join(src_dir, f), "rb") as a, open(os.path.join(tmp, f), "wb") as b:
                b.write(a.read())
        steps = sorted(set(int(s) for s in dump_steps))
        with open(os.path.join(tmp, "harness.ctl"), "w") as f:
            f.write(f"{name}\n{nsteps} {_r(dt)} {int(adapt)}\n")
            f.write(f"{len(steps)} " + " ".join(str(s) for s in steps) + "\n")
        r = subprocess.run([REF_EXE], input=(topo or name) + "\n" + "\n" * 8, cwd=tmp, capture_output=True,
                           text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"reference failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        out = {s: read_dump(os.path.join(tmp, f"dump_{s}.bin")) for s in steps}
        return out, read_topo(os.path.join(tmp, "topo.bin"))
    finally:
        subprocess.run(["rm", "-rf", tmp])


def read_zones(path: str, npts: int) -> dict:
    """zones.bin of the harness: topol(1) and topol(15) as Topo_main left
    them, and Which_Topo_Zone's answer (1-based zone, -999 none) for basal
    types 1 and 2 at each sample point."""
    raw = open(path, "rb").read()
    magic, npg, n = (int(v) for v in np.frombuffer(raw, dtype=np.int32, count=3))
    assert magic == 1195725648 and n == npts
    t = np.frombuffer(raw, dtype=np.float64, count=2 * npg, offset=12).reshape(npg, 2).copy()
    z = np.frombuffer(raw, dtype=np.int32, count=2 * n, offset=12 + 16 * npg).reshape(n, 2).copy()
    return {"z": t[:, 0], "avail": t[:, 1], "zone1": z[:, 0], "zone2": z[:, 1]}


def run_reference_zones(case, points, timeout: float = 600.0) -> dict:
    """The reference's Topo_main on ``case``'s terrain (its zones and
    erodible points), then Which_Topo_Zone at ``points`` [n, 2] (harness
    zones.ctl mode; the run goes on through Init_sph, 0 steps)."""
    if not os.path.exists(REF_EXE):
        raise FileNotFoundError(REF_EXE)
    tmp = tempfile.mkdtemp(prefix="gfsph_ref_")
    try:
        write_inputs(tmp, case)
        pts = np.asarray(points, dtype=np.float64)
        with open(os.path.join(tmp, "zones.ctl"), "w") as f:
            f.write(f"{pts.shape[0]}\n" + "\n".join(f"{_r(a)} {_r(b)}" for a, b in pts) + "\n")
        with open(os.path.join(tmp, "harness.ctl"), "w") as f:
            f.write(f"case\n0 {_r(case.params.dt)} 0\n0\n")
        r = subprocess.run([REF_EXE], input="topo\n" + "\n" * 8, cwd=tmp, capture_output=True, text=True,
                           timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"reference failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        return read_zones(os.path.join(tmp, "zones.bin"), pts.shape[0])
    finally:
        subprocess.run(["rm", "-rf", tmp])


def generated_top(ictop: int, extent, nx: int, ny: int, params) -> str:
    """A terrain file of a generated grid (Topo_main 2:262-444): ictop (1-7,
    -7; +100 for finite-difference slopes), Xming Xmaxg Yming Ymaxg, npoigx
    npoigy and the formula's parameters (ictop 4: the x's, then the Z's)."""
    base = abs(ictop) % 100 if ictop > 0 else ictop
    L = ["generated terrain written for the reference oracle", str(int(ictop)), "Xming Xmaxg Yming Ymaxg",
         " ".join(_r(v) for v in extent), "npoigx npoigy", f"{int(nx)} {int(ny)}", "terrain parameters"]
    if base == 4:
        n = len(params) // 2
        L += [str(n), "x and Z of the points", " ".join(_r(v) for v in params[:n]), " ".join(_r(v) for v in params[n:])]
    else:
        L += [" ".join(_r(v) for v in params)]
    L += ["topo props (zones) 0 = none", "0"]
    return "\n".join(L) + "\n"


def run_reference_topo(case, top_text: str, timeout: float = 600.0) -> dict:
    """The reference's Init_sph on ``case`` with its terrain file replaced by
    ``top_text``; returns the terrain it built (topo.bin)."""
    if not os.path.exists(REF_EXE):
        raise FileNotFoundError(REF_EXE)
    tmp = tempfile.mkdtemp(prefix="gfsph_ref_")
    try:
        write_inputs(tmp, case)
        with open(os.path.join(tmp, "topo.top"), "w") as f:
            f.write(top_text)
        with open(os.path.join(tmp, "harness.ctl"), "w") as f:
            f.write(f"case\n0 {_r(case.params.dt)} 0\n0\n")
        r = subprocess.run([REF_EXE], input="topo\n" + "\n" * 8, cwd=tmp, capture_output=True, text=True,
                           timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"reference failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        return read_topo(os.path.join(tmp, "topo.bin"))
    finally:
        subprocess.run(["rm", "-rf", tmp])


def run_reference(case, nsteps: int, dump_steps, workdir: str | None = None, timeout: float = 3600.0,
                  keep: bool = False, exe: str = REF_EXE, gid_final: bool = False, env: dict | None = None,
                  release=None) -> dict:
    """Run the reference on ``case`` for ``nsteps`` RK4 steps; return
    {step: dump}.  Raises if the reference build is absent."""
    if not os.path.exists(exe):
        raise FileNotFoundError(exe)
    if case.params.integrator != 4:
        raise ValueError("the reference integrates with RK4 only")
    tmp = workdir or tempfile.mkdtemp(prefix="gfsph_ref_")
    write_inputs(tmp, case, release=release)
    steps = sorted(set(int(s) for s in dump_steps))
    with open(os.path.join(tmp, "harness.ctl"), "w") as f:
        f.write("case\n")
        f.write(f"{nsteps} {_r(case.params.dt)} {int(getattr(case.params, 'adapt_dt', 0))}\n")
        f.write(f"{len(steps)} " + " ".join(str(s) for s in steps) + "\n")
    if gid_final:  # the reference's GiD writer also runs after the last step
        open(os.path.join(tmp, "gid_final.ctl"), "w").close()
    r = subprocess.run([exe], input="topo\n" + "\n" * 8, cwd=tmp, capture_output=True, text=True,
                       timeout=timeout, env=env)
    global LAST_STDOUT
    LAST_STDOUT = r.stdout
    if r.returncode != 0:
        raise RuntimeError(f"reference failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
    out = {s: read_dump(os.path.join(tmp, f"dump_{s}.bin")) for s in steps}
    for s in steps:
        pe = os.path.join(tmp, f"eros_{s}.bin")
        out[s]["eros"] = read_eros(pe) if os.path.exists(pe) else None
        pc = os.path.join(tmp, f"uaux_{s}.bin")
        if os.path.exists(pc):  # the whole pore-pressure column Uaux(1:nUaux, :)
            out[s]["ucol"] = np.frombuffer(open(pc, "rb").read(), dtype=np.float64).reshape(
                out[s]["h"].shape[0], -1).copy()
    if not keep and workdir is None:
        subprocess.run(["rm", "-rf", tmp])
    return out


def read_mc_pixel(path: str) -> dict:
    """mc_pixel.bin of the harness's Monte Carlo pixel mode: for every DEM
    node whose MC_Pr the reference allocated, the node (0-based), the number
    of problems that reached it (nr_problem_pix) and MC_Pr(1:4, 1:4) as
    [row][variable] (rows: sum over problems, sum of squares, mean, standard
    deviation; variables: average h, max h, average |v|, max |v|)."""
    raw = open(path, "rb").read()
    magic, npg, nst = (int(v) for v in np.frombuffer(raw, dtype=np.int32, count=3))
    assert magic == 1195725648
    node = np.zeros(nst, np.int64)
    nprob = np.zeros(nst, np.int32)
    pr = np.zeros((nst, 4, 4))
    off = 12
    for k in range(nst):
        a, b = np.frombuffer(raw, dtype=np.int32, count=2, offset=off)
        off += 8
        node[k], nprob[k] = a - 1, b
        # MC_Pr(1:4, 1:4) column-major: element (row r, var q) at r + 4 q
        pr[k] = np.frombuffer(raw, dtype=np.float64, count=16, offset=off).reshape(4, 4).T
        off += 128
    return {"npoig": npg, "node": node, "nprob": nprob, "mc_pr": pr}


def run_reference_pixels(cases, nsteps: int, plot_every: int, timeout: float = 3600.0) -> dict:
    """The reference's Monte Carlo pixel statistics over ``cases`` (one
    problem each: same terrain and release, its own const(:)), run one after
    another by the harness: Get_MC_pixel_Vars_SW every ``plot_every`` steps
    and Norm_MC_pixel_Vars_SW(1) after each problem's ``nsteps`` steps
    ("6 SPH_SW_2019.f90":20717-20908).  Returns read_mc_pixel's dict."""
    if not os.path.exists(REF_EXE):
        raise FileNotFoundError(REF_EXE)
    tmp = tempfile.mkdtemp(prefix="gfsph_refpix_")
    try:
        names = []
        for k, c in enumerate(cases):
            if c.params.integrator != 4:
                raise ValueError("the reference integrates with RK4 only")
            nm = f"p{k + 1}"
            write_inputs(tmp, c, name=nm)
            names.append(nm)
        dt = cases[0].params.dt
        with open(os.path.join(tmp, "harness.ctl"), "w") as f:
            f.write(f"{names[0]}\n{nsteps} {_r(dt)} {int(getattr(cases[0].params, 'adapt_dt', 0))}\n0\n")
            f.write(f"{len(names)} {int(plot_every)}\n" + " ".join(names) + "\n")
        r = subprocess.run([REF_EXE], input="topo\n" + "\n" * 8, cwd=tmp, capture_output=True, text=True,
                           timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"reference failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        return read_mc_pixel(os.path.join(tmp, "mc_pixel.bin"))
    finally:
        subprocess.run(["rm", "-rf", tmp])


REF_MAIN_EXE = os.path.join(HERE, "_ref", "gfsph_ref_main")


def read_case_trace(path: str) -> list:
    """case_trace.bin of the reference's main program (oracle/ref_main_probe.f90):
    one record per Driver_Flow_23 step, in order."""
    raw = open(path, "rb").read()
    out = []
    off = 0
    while off < len(raw):
        magic, prob, it, stop, n = (int(v) for v in np.frombuffer(raw, dtype=np.int32, count=5, offset=off))
        assert magic == 1195725648
        off += 20
        t = np.frombuffer(raw, dtype=np.float64, count=4, offset=off)
        off += 32
        rec = {"problem": prob, "itimestep": it, "stop": stop, "time_sph": float(t[0]), "dt_sph": float(t[1]),
               "time": float(t[2]), "t_plot_reset": float(t[3])}
        for k, cnt in (("x", 2 * n), ("v", 2 * n), ("h", n), ("u", n)):
            a = np.frombuffer(raw, dtype=np.float64, count=cnt, offset=off)
            off += a.nbytes
            rec[k] = a.reshape(n, 2).copy() if cnt == 2 * n else a.copy()
        rec["flag"] = np.frombuffer(raw, dtype=np.int32, count=n, offset=off).copy()
        off += 4 * n
        out.append(rec)
    return out


def run_reference_main(case, name: str = "case", timeout: float = 3600.0, **master) -> list:
    """Runs the reference's own main program (PROGRAM SPH_FEM_Driver_2019,
    built as oracle/_ref/gfsph_ref_main) on ``case``'s input files and a
    <name>.MASTER.dat written by write_master(**master); returns the per-step
    records of its case loop (read_case_trace)."""
    if not os.path.exists(REF_MAIN_EXE):
        raise FileNotFoundError(REF_MAIN_EXE)
    tmp = tempfile.mkdtemp(prefix="gfsph_refmain_")
    try:
        write_inputs(tmp, case, name=name, topo="topo")
        write_master(tmp, name, case.params.dt, **master)
        # the main program opens <name>.master.dat (8:217); the reference was
        # written on a file system that ignores case
        os.rename(os.path.join(tmp, name + ".MASTER.dat"), os.path.join(tmp, name + ".master.dat"))
        r = subprocess.run([REF_MAIN_EXE], input=f"{name}\ntopo\n" + "\n" * 16, cwd=tmp, capture_output=True,
                           text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"reference main failed ({r.returncode}):\n{r.stdout[-3000:]}\n{r.stderr[-2000:]}")
        return read_case_trace(os.path.join(tmp, "case_trace.bin"))
    finally:
        subprocess.run(["rm", "-rf", tmp])
