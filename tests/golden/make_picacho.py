"""Generates the Picacho end-to-end fixture from the reference Fortran build.

The reference ships one complete example, the Picacho landslide
(EJEMPLOS/Picacho Landslide/Input files: Picacho.MASTER.dat, Picacho.dat,
Picacho.top -- a scattered 10 m DTM, ictop 10 -- and Picacho.pts -- 2402
release points gridded at 10 m, icunk 7).  This script copies those data
files, gzipped, to tests/golden/picacho/ and runs the reference harness
(oracle/_ref/gfsph_ref: the reference's own Init_sph on those files, then
Driver_Flow steps with ic_adapt_dt_sph = 1 and dt_sph = 0.1 as the MASTER
file sets) to record:

  * the terrain Init_sph built (grid geometry; SHA-256 of Z, topol(2:9,:) and
    poig_in_model, and their values at 4096 sampled nodes), and
  * the particle state (x, v, h, mass, hsml, u, Uaux, flags, time_sph) after
    Init_sph and after 1, 10 and 100 steps

into tests/golden/picacho.npz.  Data only: the example's input files and the
reference's outputs.

    python tests/golden/make_picacho.py
"""
from __future__ import annotations

import gzip
import hashlib
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from oracle import refio  # noqa: E402

SRC = "/root/reference/EJEMPLOS/Picacho Landslide/Input files"
FILES = ("Picacho.MASTER.dat", "Picacho.dat", "Picacho.top", "Picacho.pts")
OUT_DIR = os.path.join(HERE, "picacho")
DUMPS = (0, 1, 10, 100)


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main() -> None:
    os.makedirs(OUT_DIR, exist_ok=True)
    for f in FILES:
        with open(os.path.join(SRC, f), "rb") as a, gzip.open(os.path.join(OUT_DIR, f + ".gz"), "wb", 9) as b:
            shutil.copyfileobj(a, b)
    out, topo = refio.run_reference_files(SRC, "Picacho", FILES[1:], max(DUMPS), 0.1, 1, DUMPS)
    top = topo["topol"]
    npg = top.shape[0]
    rng = np.random.default_rng(7)
    sample = np.sort(rng.choice(npg, size=min(4096, npg), replace=False)).astype(np.int64)
    f = {
        "grid": np.array([topo["xmin"], topo["xmax"], topo["ymin"], topo["ymax"], topo["dx"], topo["dy"],
                          topo["xshift"], topo["yshift"]]),
        "grid_n": np.array([topo["nx"], topo["ny"]], np.int64),
        "sha_Z": sha(top[:, 0]), "sha_topol": sha(top[:, 1:9]), "sha_in_model": sha(topo["in_model"]),
        "sample": sample, "sample_topol": top[sample], "sample_in_model": topo["in_model"][sample],
        "dumps": np.array(DUMPS, np.int64),
    }
    for s, d in out.items():
        for k in ("x", "v", "h", "mass", "hsml", "u", "flag"):
            f[f"s{s}_{k}"] = d[k]
        if d["uaux"] is not None:
            f[f"s{s}_uaux"] = d["uaux"]
        f[f"s{s}_time"] = np.array(d["time"])
    np.savez_compressed(os.path.join(HERE, "picacho.npz"), **f)
    print("picacho: grid", topo["nx"], "x", topo["ny"], "particles", out[0]["x"].shape[0],
          "time at step", max(DUMPS), out[max(DUMPS)]["time"])


if __name__ == "__main__":
    main()
