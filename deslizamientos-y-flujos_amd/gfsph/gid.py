"""GiD post-process text output of a single-phase run: the ``.post.msh``
mesh and ``.post.res`` result blocks the reference writes for GiD.

Host-side formatting of arrays the library already produced (the DEM, the
particle state, the IDW raster of ``SPHModel.raster_idw``).  Layout follows
the reference's writers:

* mesh: ``OutputMesh_WIR`` ("6 SPH_SW_2019.f90":12603-12800).  Three node
  sets share one coordinate block -- the DEM nodes 1..npoig (Q1), the same
  nodes again as npoig+1..2*npoig (Q2), the particles as 2*npoig+i -- then
  the Q1 and Q2 quadrilaterals over the DEM cells whose four nodes are in
  the model, and one point element per particle.
* results: ``OutputRes_WIR`` ("6 SPH_SW_2019.f90":14500-14917), blocks of
  the GiD mask (``Gid_Mask_SW``, 6:1536-1559): 1 height soil, 2
  displacement, 3 velocity, 4 pore pressure (icpwp 1-3), 5 eroded depth,
  available depth and the terrain under and over the flow (erosion), 6 Z
  topo (first output only: the two-phase ``hs_plus_Z_topo`` form is not
  built), 7 with the value 999 (the mass-based height "hs",
  6:14661-14713), 8 water height (0 without water particles), 9 the rise
  of the flow surface since the run's first output (eta), 10 joint height,
  11 the IDW sum of the smoothing lengths (1) or the Froude number (-1);
  12 writes nothing in two dimensions, as there.  A mask entry of 1 writes
  the block, 9 (-9 for 11) writes it without the nodes no particle reaches
  (6:14279-14281); block 4 = -1 (total pressure), 7 = 1 (Peclet / law 202)
  and other values raise ``ValueError``.

Numbers are written as shortest round-trip decimals; GiD reads the files
as whitespace-separated tokens, so the values (not the spacing of the
reference's list-directed output) are what the two have in common.
"""

from __future__ import annotations

import numpy as np

BUILT_BLOCKS = (1, 2, 3, 4, 5, 6, 8, 9, 10, 11, 12)


def _f(v: float) -> str:
    return repr(float(v))


def dem_nodes(dem):
    """Coordinates (x, y, z) of the DEM nodes, node ipoig = iy*nx + ix + 1."""
    ix = np.arange(dem.nx)
    iy = np.arange(dem.ny)
    X = np.broadcast_to(dem.xmin + ix * dem.dx, (dem.ny, dem.nx))
    Y = np.broadcast_to((dem.ymin + iy * dem.dy)[:, None], (dem.ny, dem.nx))
    return X.ravel(), Y.ravel(), np.asarray(dem.Z, dtype=np.float64).ravel()


def z_at(dem, x: np.ndarray) -> np.ndarray:
    """Terrain height under each position by the bilinear element
    interpolation of Get_Z_Topo ("2 SPH_FEM_topo_2019.f90":2395-2491):
    element of the node below-left, shape functions (1-xi)(1-eta), xi(1-eta),
    xi eta, (1-xi)eta on its nodes n1, n2, n3, n4 (counter-clockwise)."""
    x = np.asarray(x, dtype=np.float64).reshape(-1, 2)
    Z = np.asarray(dem.Z, dtype=np.float64).ravel()
    nx, ny = dem.nx, dem.ny
    xx, yy = x[:, 0], x[:, 1]
    ipx = ((xx - dem.xmin) / dem.dx + 1).astype(np.int64)
    ipy = ((yy - dem.ymin) / dem.dy + 1).astype(np.int64)
    ipx = np.where((ipx == nx) & (np.abs(dem.xmax - xx) <= 1.0e-3), ipx - 1, ipx)
    ipy = np.where((ipy == ny) & (np.abs(dem.ymax - yy) <= 1.0e-3), ipy - 1, ipy)
    ipx = np.clip(ipx, 1, nx - 1)
    ipy = np.clip(ipy, 1, ny - 1)
    n1 = (ipy - 1) * nx + (ipx - 1)
    n2, n4 = n1 + 1, n1 + nx
    n3 = n4 + 1
    xi = (xx - (dem.xmin + (ipx - 1) * dem.dx)) / dem.dx
    eta = (yy - (dem.ymin + (ipy - 1) * dem.dy)) / dem.dy
    sh1, sh2, sh3, sh4 = (1.0 - xi) * (1.0 - eta), xi * (1.0 - eta), xi * eta, (1.0 - xi) * eta
    return sh1 * Z[n1] + sh2 * Z[n2] + sh3 * Z[n3] + sh4 * Z[n4]


def cell_nodes(dem, x: np.ndarray) -> np.ndarray:
    """The four nodes n1, n2, n3, n4 (0-based, counter-clockwise from the one
    below-left) of each position's DEM cell, as Get_Z_Topo and the result
    writers find it: cell index (x - xmin)/dx + 1 truncated, moved one cell
    back on the last row or column within 1e-3 of the grid's edge."""
    x = np.asarray(x, dtype=np.float64).reshape(-1, 2)
    nx, ny = dem.nx, dem.ny
    xx, yy = x[:, 0], x[:, 1]
    ipx = ((xx - dem.xmin) / dem.dx + 1).astype(np.int64)
    ipy = ((yy - dem.ymin) / dem.dy + 1).astype(np.int64)
    ipx = np.where((ipx == nx) & (np.abs(dem.xmax - xx) <= 1.0e-3), ipx - 1, ipx)
    ipy = np.where((ipy == ny) & (np.abs(dem.ymax - yy) <= 1.0e-3), ipy - 1, ipy)
    ipx = np.clip(ipx, 1, nx - 1)
    ipy = np.clip(ipy, 1, ny - 1)
    n1 = (ipy - 1) * nx + (ipx - 1)
    return np.stack([n1, n1 + 1, n1 + 1 + nx, n1 + nx], axis=1)


def hs_nodes(dem, x: np.ndarray, mass: np.ndarray, live: np.ndarray) -> np.ndarray:
    """The mass-based height of the "hs" block (6:14661-14713): every particle
    in the domain adds its mass to the four nodes of its cell, in particle
    order; the sums are divided by 4 dx dy."""
    nodes = cell_nodes(dem, x)[live]
    m = np.asarray(mass, dtype=np.float64)[live]
    hg = np.zeros(dem.nx * dem.ny)
    np.add.at(hg, nodes.ravel(), np.repeat(m, 4))  # per particle n1..n4, particles in order
    return hg / (4.0 * (dem.dx * dem.dy))


def quads(dem) -> np.ndarray:
    """Q1 quadrilaterals [keleg][4] (1-based DEM nodes) over the cells whose
    four nodes are in the model, in the reference's row-by-row order."""
    nx, ny = dem.nx, dem.ny
    inm = (np.ones(nx * ny, np.int32) if dem.in_model is None
           else np.asarray(dem.in_model, np.int32).ravel())
    iy, ix = np.meshgrid(np.arange(ny - 1), np.arange(nx - 1), indexing="ij")
    p = (iy * nx + ix).ravel()  # 0-based node of the lower-left corner
    keep = (inm[p] * inm[p + 1] * inm[p + nx + 1] * inm[p + nx]) == 1
    p = p[keep] + 1
    return np.stack([p, p + 1, p + nx + 1, p + nx], axis=1)


def write_post_msh(path: str, dem, x: np.ndarray, itype: np.ndarray | None = None) -> np.ndarray:
    """Writes the ``.post.msh`` mesh for particles at x [n][2] and returns
    their terrain heights (the reference keeps them as Z00 for the
    displacement block)."""
    x = np.asarray(x, dtype=np.float64).reshape(-1, 2)
    n = x.shape[0]
    npoig = dem.nx * dem.ny
    itype = np.ones(n, np.int32) if itype is None else np.asarray(itype)
    X, Y, Z = dem_nodes(dem)
    zp = z_at(dem, x)
    q = quads(dem)
    L = [" MESH SPH_quads   dimension   3    ElemType Quadrilateral  Nnode 4", "  coordinates"]
    for base in (0, npoig):
        L += [f" {base + k + 1} {_f(X[k])} {_f(Y[k])} {_f(Z[k])}" for k in range(npoig)]
    L += [f" {2 * npoig + i + 1} {_f(x[i, 0])} {_f(x[i, 1])} {_f(zp[i])}" for i in range(n)]
    L.append("  End coordinates")
    for name, off in (("SPH_Q1", 0), ("SPH_Q2", npoig)):
        L += [f" MESH {name}   dimension   3    ElemType Quadrilateral  Nnode 4", "  Elements"]
        L += [f" {e + 1} {q[e, 0] + off} {q[e, 1] + off} {q[e, 2] + off} {q[e, 3] + off}     1"
              for e in range(q.shape[0])]
        L.append("  End elements")
    L += [" MESH SPH_points dimension 3   ElemType Point  Nnode 1", "  Elements"]
    L += [f" {i + 1} {2 * npoig + i + 1} {abs(int(itype[i]))}" for i in range(n)]
    L.append("  End elements")
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")
    return zp


def _block(title: str, time: float, rows) -> list[str]:
    return [f' Result {title}  {_f(time)}  Vector OnNodes "where"', "  Values", *rows, "  End Values"]


def hsml_sum(dem, x: np.ndarray, hsml: np.ndarray, live: np.ndarray) -> np.ndarray:
    """The smoothing-length sum of the raster (auxv(13), 6:14381-14383): every
    particle in the domain adds hsml/d to the nodes of its 4 x 4 stencil, d
    its distance to the node over the cell diagonal (at least 1e-6); the sum
    is written as it is (the reference divides it by the weights only for
    hrelSat, 6:14444-14446)."""
    x = np.asarray(x, dtype=np.float64).reshape(-1, 2)
    hs = np.asarray(hsml, dtype=np.float64)
    nx, ny = dem.nx, dem.ny
    distg = (dem.dx * dem.dx + dem.dy * dem.dy) ** 0.5
    out = np.zeros(nx * ny)
    for i in np.nonzero(live)[0]:
        xx, yy = x[i]
        ipx = int((xx - dem.xmin) / dem.dx + 1)
        ipy = int((yy - dem.ymin) / dem.dy + 1)
        for inx in range(max(1, ipx - 1), min(nx, ipx + 2) + 1):
            for iny in range(max(1, ipy - 1), min(ny, ipy + 2) + 1):
                gx = dem.xmin + (inx - 1) * dem.dx
                gy = dem.ymin + (iny - 1) * dem.dy
                a = (xx - gx) * (xx - gx)
                b = (yy - gy) * (yy - gy)
                d = max(1.0e-6, (a + b) ** 0.5) / distg
                out[(iny - 1) * nx + inx - 1] += hs[i] / d
    return out


def post_res_blocks(dem, time: float, raster: np.ndarray, x: np.ndarray, v: np.ndarray, x00: np.ndarray,
                    z00: np.ndarray, out_domain: np.ndarray | None = None, mask=(1, 1, 1, 1, 0, 1, 0, 0, 0, 1, 0, 0),
                    first: bool = True, icpwp: int = 0, mass: np.ndarray | None = None,
                    wsum: np.ndarray | None = None, hsml: np.ndarray | None = None, erosion: dict | None = None,
                    basal_erosion: bool = False, history: dict | None = None, cst=None) -> list[str]:
    """Lines of one output time: raster [npoig][4] = (h, vx, vy, pwp) of the
    IDW raster, x/v [n][2] the particles, x00/z00 where they started
    (write_post_msh's return), out_domain [n] nonzero for particles out of
    the domain, mask the 12 GiD mask entries, first for the first output of
    the run (the Z topo block), icpwp the pore-pressure model, mass [n] the
    particle masses (needed by the "hs" block, mask entry 7 = 999).  Further
    blocks need: wsum [npoig] the raster's weight sums (mask value 9),
    hsml [n] (block 11 = 1), erosion = {"z", "eroded", "avail"} the
    terrain's topol(1), (14), (15) (block 5; block 9 reads the current Z
    from it when given), basal_erosion for terrain erosion zones (block 5),
    history a dict kept across the run's outputs (block 9 stores the first
    output's surface there), cst const(:) (block 11 = -1)."""
    mask = [int(m) for m in mask] + [0] * (12 - len(mask))
    skip = [False] * 12
    for k, m in enumerate(mask, 1):
        if m in (9, -9):  # the block, without the nodes no particle reaches (6:14279-14281)
            skip[k - 1] = True
            mask[k - 1] = 1 if m == 9 else -1
            m = mask[k - 1]
        ok = m in (0, 1) or (k == 7 and m == 999) or (k == 11 and m == -1)
        if not ok:
            raise ValueError(f"GiD mask entry {k} = {m}: not built")
        if m == 1 and k not in BUILT_BLOCKS:
            raise ValueError(f"GiD result block {k} is not built (blocks {BUILT_BLOCKS})")
    raster = np.asarray(raster, dtype=np.float64).reshape(-1, 4)
    x = np.asarray(x, dtype=np.float64).reshape(-1, 2)
    v = np.asarray(v, dtype=np.float64).reshape(-1, 2)
    npoig = dem.nx * dem.ny
    n = x.shape[0]
    live = np.ones(n, bool) if out_domain is None else (np.asarray(out_domain) == 0)
    _, _, Ztop = dem_nodes(dem)
    if any(skip) and wsum is None:
        raise ValueError("GiD mask value 9 needs the raster's weight sums (wsum)")
    empty = np.zeros(npoig, bool) if wsum is None else (np.asarray(wsum, dtype=np.float64).ravel() <= 1.0e-6)

    def nodes(k):  # node indices block k writes
        return [j for j in range(npoig) if not (skip[k - 1] and empty[j])]

    L: list[str] = []
    if mask[0]:
        L += _block('"height soil" "Height soil"', time,
                    [f" {k + 1} 0. 0. {_f(raster[k, 0])}" for k in range(npoig)])
    if mask[1]:
        zp = z_at(dem, x)
        d = x - np.asarray(x00, dtype=np.float64).reshape(-1, 2)
        dz = zp - np.asarray(z00, dtype=np.float64)
        L += _block('"dis" "Disp"', time, [f" {i + 2 * npoig + 1} {_f(d[i, 0])} {_f(d[i, 1])} {_f(dz[i])}"
                                            for i in range(n) if live[i]])
    if mask[2]:
        rows = [f" {k + 1} {_f(raster[k, 1])} {_f(raster[k, 2])} 0.0" for k in nodes(3)]
        rows += [f" {i + 2 * npoig + 1} {_f(v[i, 0])} {_f(v[i, 1])} 0.0" for i in range(n) if live[i]]
        L += _block('"vel" "veloc"', time, rows)
    if mask[3] and icpwp in (1, 2, 3):
        L += _block('"Pwp" "pwpress"', time, [f" {k + 1} 0.0 0.0 {_f(raster[k, 3])}" for k in nodes(4)])
    zcur = Ztop if erosion is None else np.asarray(erosion["z"], dtype=np.float64).ravel()
    if mask[4]:  # 6:14591-14636
        if erosion is None:
            raise ValueError("GiD result block 5 needs the erosion state")
        er = np.asarray(erosion["eroded"], dtype=np.float64).ravel()
        av = np.asarray(erosion["avail"], dtype=np.float64).ravel()
        L += _block('"erodedTopo" "-erodedTopo-"', time, [f" {k + 1} 0. 0. {_f(er[k])}" for k in range(npoig)])
        L += _block('"AvailErosD" "ErosDepth"', time, [f" {k + 1} 0. 0. {_f(av[k])}" for k in range(npoig)])
        under = zcur - er if basal_erosion else zcur
        over = zcur + raster[:, 0] - er if basal_erosion else zcur + raster[:, 0]
        rows = [f" {k + 1} 0. 0. {_f(under[k])}" for k in range(npoig)]
        rows += [f" {npoig + k + 1} 0. 0. {_f(over[k])}" for k in range(npoig)]
        L += _block('"hs over topo" "-hs over topo-"', time, rows)
    if mask[5] and first:
        L += _block('" Z topo " "   Z   "', time, [f" {k + 1} 0. 0. {_f(Ztop[k])}" for k in range(npoig)])
    if mask[6] == 999:
        if mass is None:
            raise ValueError("GiD mask entry 7 = 999 (the hs block) needs the particle masses")
        hg = hs_nodes(dem, x, mass, live)
        L += _block('"hs" "hs"', time, [f" {k + 1} 0. 0. {_f(hg[k])}" for k in range(npoig) if hg[k] > 0.005])
    if mask[7]:  # water height: no water particles here, 0 (6:14749-14758)
        L += _block('"height water" "Height water"', time, [f" {k + 1} 0. 0. 0.0" for k in nodes(8)])
    if mask[8]:  # the flow surface's rise since the first output (6:14500-14503, 14762-14784)
        if history is None:
            raise ValueError("GiD result block 9 needs a history dict kept across the run's outputs")
        surf = raster[:, 0] + zcur
        if "etag" not in history:
            history["etag"] = surf.copy()
        eta = surf - history["etag"]
        L += _block('"eta" " sobreelev"', time, [f" {k + 1} 0. 0. {_f(eta[k])}" for k in nodes(9)])
    if mask[9]:
        rows = [f" {k + 1} 0. 0. {_f(raster[k, 0])}" for k in nodes(10)]
        rows += [f" {npoig + k + 1} 0. 0. {_f(raster[k, 0])}" for k in nodes(10)]
        L += _block('"jointh" "JOINTh"', time, rows)
    if mask[10] == 1:  # 6:14810-14843 (no hrelSat)
        if hsml is None:
            raise ValueError("GiD result block 11 needs the smoothing lengths")
        hg = hsml_sum(dem, x, hsml, live)
        L += _block('"hsml" "-HSML-"', time, [f" {k + 1} 0. 0. {_f(hg[k])}" for k in nodes(11)])
    elif mask[10] == -1:  # Froude number (6:14845-14869), g = const(1) (ic_cosTheta = 0)
        if cst is None:
            raise ValueError("GiD result block 11 = -1 needs const(:)")
        c = np.asarray(cst, dtype=np.float64)
        rows = []
        for k in nodes(11):
            vvv = (raster[k, 1] * raster[k, 1] + raster[k, 2] * raster[k, 2]) ** 0.5
            hhh = raster[k, 0]
            fr = 0.0
            if hhh > c[10]:
                fr = vvv / ((hhh * c[1]) ** 0.5)
                if c[5] == 8:
                    if hhh < c[6]:
                        hhh = c[6]
                    fr = fr * (c[6] / hhh)
            rows.append(f" {k + 1} 0. 0. {_f(fr)}")
        L += _block('"F or I" "-F or I -"', time, rows)
    return L


def write_post_res(path: str, *args, append: bool = False, **kw) -> None:
    """Writes (or appends) one output time of the ``.post.res`` file; the
    arguments are those of post_res_blocks.  A new file starts with the
    GiD header line."""
    L = post_res_blocks(*args, **kw)
    with open(path, "a" if append else "w") as f:
        if not append:
            f.write(" GiD Post Results File 1.0\n")
        f.write("\n".join(L) + "\n")


def read_tokens(path: str) -> dict:
    """Parses a GiD mesh or results file into {section: float array of rows}
    (coordinates, each mesh's elements, each result block as "result <name>
    #<k>" for its k-th output time, k from 0, with the time under "time
    <name> #<k>"); used to compare files value by value."""
    out: dict = {}
    cur, rows, mesh = None, [], None
    seen: dict = {}
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s:
                continue
            low = s.lower()
            if low.startswith("mesh "):
                mesh = s.split()[1]
            elif low == "coordinates":
                cur, rows = "coordinates", []
            elif low == "elements":
                cur, rows = f"elements {mesh}", []
            elif low.startswith("result "):
                parts = s.split('"')
                k = seen.get(parts[1], 0)
                seen[parts[1]] = k + 1
                cur, rows = f"result {parts[1]} #{k}", []
                out[f"time {parts[1]} #{k}"] = np.array([float(parts[4].split()[0])])
            elif low == "values":
                pass
            elif low in ("end coordinates", "end elements", "end values"):
                out[cur] = np.array(rows, dtype=np.float64)
                cur = None
            elif cur is not None:
                if '"' in s:  # the rest of a header line the writer wrapped
                    continue
                rows.append([float(t) for t in s.split()])
    return out
