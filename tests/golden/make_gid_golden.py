"""Writes tests/golden/gid_frictional_pwp.npz: the GiD mesh and result blocks
the reference's own writers (OutputMesh_WIR at Init_sph, OutputRes_SW after
the last step) produce for the g_frictional_pwp fixture case, parsed into
arrays.  Needs the reference build oracle/_ref/gfsph_ref (oracle/build_ref.sh);
run from the repository root:  python tests/golden/make_gid_golden.py"""
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "deslizamientos-y-flujos_amd"), os.path.join(ROOT, "tests"),
                os.path.join(ROOT, "oracle")]
import goldens  # noqa: E402
import refio  # noqa: E402
from gfsph import gid  # noqa: E402

NSTEPS = 100


def main():
    case, _ = goldens.case_of("g_frictional_pwp")
    tmp = tempfile.mkdtemp(prefix="gfsph_gid_")
    refio.run_reference(case, NSTEPS, [NSTEPS], workdir=tmp, gid_final=True)
    out = {}
    for kind in ("msh", "res"):
        for k, v in gid.read_tokens(os.path.join(tmp, f"case.post.{kind}")).items():
            out[f"{kind}: {k}"] = v
    out["nsteps"] = np.array(NSTEPS)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "gid_frictional_pwp.npz"), **out)
    print(sorted(out))
    # the mass-based height block ("hs", GiD mask entry 7 = 999)
    case.gid_mask = (1, 0, 0, 0, 0, 0, 999, 0, 0, 0, 0, 0)
    tmp = tempfile.mkdtemp(prefix="gfsph_gid_")
    refio.run_reference(case, NSTEPS, [NSTEPS], workdir=tmp, gid_final=True)
    out = {f"res: {k}": v for k, v in gid.read_tokens(os.path.join(tmp, "case.post.res")).items()}
    out["nsteps"] = np.array(NSTEPS)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "gid_hs_frictional_pwp.npz"), **out)
    print(sorted(out))
    # erosion, water height, eta, smoothing lengths / Froude number and the
    # mask value 9 (nodes no particle reaches left out), on an erosion-zone
    # run (terrain erosion zones) and a surface-law erosion run
    for fname, name, mask in (("gid_zones_erosion", "g_zones_erosion", (1, 0, 9, 0, 1, 0, 0, 9, 1, 9, 1, 0)),
                              ("gid_erosion_hungr", "g_erosion_hungr", (0, 0, 0, 0, 1, 0, 0, 1, 9, 0, -9, 1))):
        case, _ = goldens.case_of(name)
        case.gid_mask = mask
        tmp = tempfile.mkdtemp(prefix="gfsph_gid_")
        refio.run_reference(case, NSTEPS, [NSTEPS], workdir=tmp, gid_final=True)
        out = {f"res: {k}": v for k, v in gid.read_tokens(os.path.join(tmp, "case.post.res")).items()}
        out["nsteps"] = np.array(NSTEPS)
        out["mask"] = np.array(mask)
        np.savez_compressed(os.path.join(ROOT, "tests", "golden", fname + ".npz"), **out)
        print(sorted(out))


if __name__ == "__main__":
    main()
