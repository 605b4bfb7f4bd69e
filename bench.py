"""Throughput of the Geoflow-SPH propagation step on one MI355X.

Workload (the 200k-particle drop-in configuration): fractal DEM 2048 x 2048
at 2 m (diamond-square, H 0.8, 400 m relief, carved 25-35 deg valley);
elliptical release of 200,000 particles, thickness ~ U(2, 8) m; frictional
phi = 28 deg with pore pressure (initial ratio 0.6, Cv 0.01 m2/s); RK2,
dt = 0.02 s; FP64.  Synthetic data generated from fixed seeds.

A "step" is one time_integration_sph call over all particles (neighbour
build, density, pressure force, RK stages with rheology and pore pressure,
XSPH).  Inputs are resident in HBM before the timed region.

Prints one JSON line (see the metric contract in the repository docs).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
for _p in (REPO, os.path.join(REPO, "deslizamientos-y-flujos_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

PEAK_HBM_GBS = 8000.0

# Algorithmic HBM bytes per particle per launch of each step kernel (FP64,
# compulsory own-state traffic; neighbour gathers assumed served on chip).
#   k_density  : read pos 16, hsml 8, mass 8, flag 4; write S 8, P 8       = 52
#   k_force_rk : read pos 16, vel 16, h 8, S 8, P 8, hsml 8, mass 8, u 8,
#                uaux 8, flag 4; write pos 16, vel 16, h 8, hsml 8, u 8,
#                uaux 8, flag 4                                            = 164
#   k_xsph     : read pos 16, hsml 8, mass 8, h_new 8, vel_new 16, flags 8,
#                id 4; write vel 16                                        = 84
B_ALG = {"k_density": 52.0, "k_force_rk": 164.0, "k_xsph": 84.0}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=200_000)
    ap.add_argument("--integrator", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-steps", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding of realisations is a design note, not a built path)", file=sys.stderr)
        return 2

    from gfsph import SPHModel, cases

    t0 = time.time()
    case = cases.cfg_fractal_frictional(n=args.n, integrator=args.integrator)
    print(f"# workload generated in {time.time() - t0:.1f}s: n={case.particles.n} dem={case.dem.nx}x{case.dem.ny}",
          file=sys.stderr, flush=True)
    m = SPHModel(case)
    m.time_integration_sph(args.warmup)
    m.sync()
    t1 = time.perf_counter()
    m.time_integration_sph(args.steps)
    m.sync()
    t2 = time.perf_counter()
    elapsed = t2 - t1
    n = case.particles.n
    value = n * args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps

    # per-kernel device time with HIP events on the library's stream
    prof = m.profile_step(nrep=20)
    kern = {k: v for k, v in prof.items() if k in B_ALG}
    dom = max(kern, key=kern.get)
    ms = kern[dom]
    achieved = B_ALG[dom] * n / (ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None,
                "kernel_ms": {k: round(v, 4) for k, v in prof.items()}}
    st = m.state()
    n_active = int((st["flag"] == 0).sum())
    m.close()

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(case, args.cpu_sample_steps)

    out = {
        "metric": "particle-steps/s at 200k",
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (fractal DEM seed 1002, release seed 2002)",
        "config": {"workload": "fractal DEM 2048x2048 @2m, 200k-particle elliptical release, frictional 28deg + "
                               "pore pressure, RK2 dt=0.02 s", "particles": n, "active_after": n_active,
                   "integrator": "RK2" if args.integrator == 2 else "RK4", "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def cpu_baseline(case, nsteps):
    """The oracle (serial C restatement, 1 thread) on a bounded sample of the
    same workload: all particles, ``nsteps`` steps."""
    try:
        from oracle import oracle
    except Exception as e:  # the oracle ships with the repository
        return {"error": str(e)}
    st = oracle.init(case)
    t = time.perf_counter()
    oracle.run(case, st, nsteps)
    dt = time.perf_counter() - t
    return {"value": round(case.particles.n * nsteps / dt, 1), "unit": "particle-steps/s", "cores": 1,
            "kind": "port", "sample": f"{case.particles.n} particles x {nsteps} steps (oracle, serial C FP64)",
            "cpu": cpu_model()}


if __name__ == "__main__":
    sys.exit(main())
