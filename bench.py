"""Throughput of the Geoflow-SPH propagation step on one MI355X.

Workload (the 200k-particle drop-in configuration): fractal DEM 2048 x 2048
at 2 m (diamond-square, H 0.8, 400 m relief, carved 25-35 deg valley);
elliptical release of 200,000 particles, thickness ~ U(2, 8) m; frictional
phi = 28 deg with pore pressure (initial ratio 0.6, Cv 0.01 m2/s); RK2,
dt = 0.02 s, 5,000 steps; FP64.  Synthetic data generated from fixed seeds.

A "step" is one time_integration_sph call over all particles (neighbour
build, density, pressure force, RK stages with rheology and pore pressure,
XSPH).  The cost of a step changes as the slide evolves (the flow converges
in the valley and neighbour counts grow), so the timed window is taken from
the developed flow: a fresh context whose inputs are resident in HBM is
advanced --develop steps (default: half the configured run) untimed, then
--warmup steps, then exactly --steps steps are timed.  The whole configured
run (5,000 steps from t = 0) is timed as well and reported under
"full_run".  Per-kernel device times ("kernel_ms", the roofline) are taken
with HIP events over the next --steps steps of the same context.

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

# Algorithmic HBM bytes per particle-step of each pass (FP64, compulsory
# own-state traffic per dependency pass; neighbour gathers are assumed served
# on chip):
#   neighbour_build  pass A  read x, hsml; write key, id; radix passes     64
#   k_density        pass B  read x, m, hsml, id; write h                  44
#   k_pforce         pass C  read x, m, hsml, h; write a                   56
#   k_rk             pass D  read x, v, h, h_prev, U, a, flag;
#                            write x, v, U, h_prev, flag                 128
#   k_xsph           pass E  read x^n, v, m, h, hsml; write v              72
# The pair geometry and h are frozen over the RK stages, so the internal
# force is gathered once per step and k_rk runs every stage in registers;
# x^n needs no extra copy (pass D', 16 B) because the step writes its outputs
# to a second buffer.  Total 364 B per particle-step with XSPH.
B_ALG = {"neighbour_build": 64.0, "k_density": 44.0, "k_pforce": 56.0, "k_rk": 128.0, "k_xsph": 72.0}

# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on
# this workload (200k particles, the steps of the developed-flow window after
# 1,000 untimed steps), each scaled by what the counter reads for a device
# copy of known size (fetch x2.0, write x1.0)
MEASURED_TRAFFIC_200K = {"k_density": 39.10e6, "k_pforce": 33.11e6, "k_xsph": 38.64e6, "k_rk": 49.97e6}

# What bounds the gathers instead of HBM, per CU and cycle from rocprofv3 --pmc
# on the 256-realisation MC batch (7.37M particles; profiles/gather_counters_mc256.md):
# VALU issue (one wave64 instruction per CU per cycle at most), the load-data
# return unit TD, L1 tag lookups (one per cycle at most).  Static figures of
# that profile, not measured by this run.
GATHER_LIMITS_MC = {
    "source": "profiles/gather_counters_mc256.md", "workload": "MC batch, 256 realisations, 7,366,052 particles",
    "k_density": {"valu_busy": 0.73, "td_busy": 0.96, "l1_lookups_per_clk_per_cu": 0.77},
    "k_pforce": {"valu_busy": 0.87, "td_busy": 0.94, "l1_lookups_per_clk_per_cu": 0.86},
    "k_xsph": {"valu_busy": 0.70, "td_busy": 0.96, "l1_lookups_per_clk_per_cu": 0.94},
}


def cpu_info() -> dict:
    """The host CPU as the OS reports it: the full model string, the vendor,
    family/model/stepping, the logical CPUs online (nproc) and those this
    process may run on."""
    info = {"model": platform.processor() or "unknown"}
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and "model_name" not in info:
                info["model"] = v
                info["model_name"] = v
            elif k in ("vendor_id", "cpu family", "model", "stepping", "cpu MHz") and k.replace(" ", "_") not in info:
                info[k.replace(" ", "_")] = v
    except OSError:
        pass
    info.pop("model_name", None)
    info["nproc"] = os.cpu_count()
    try:
        info["nproc_affinity"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    return info


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--run-steps", type=int, default=5000, help="steps of the configured run")
    ap.add_argument("--develop", type=int, default=-1,
                    help="untimed steps before the timed window (default: half of --run-steps)")
    ap.add_argument("--no-full-run", action="store_true", help="skip timing the whole configured run")
    ap.add_argument("--profile-steps", type=int, default=0, help="steps of the event-timed run (0: --steps)")
    ap.add_argument("--traffic", type=float, default=None,
                    help="measured HBM bytes per launch of the dominant kernel (from rocprofv3 --pmc)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a plain device copy of known size (counter calibration) and exit")
    ap.add_argument("--n", type=int, default=200_000)
    ap.add_argument("--integrator", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-steps", type=int, default=40)
    ap.add_argument("--no-mc", action="store_true", help="skip the batched Monte Carlo measurement")
    ap.add_argument("--mc-real", type=int, default=256)
    ap.add_argument("--mc-steps", type=int, default=3000)
    ap.add_argument("--mc-seq-sample", type=int, default=0,
                    help="realisations timed one by one for the sequential cost (0: all of them)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint measurement")
    ap.add_argument("--no-cal", action="store_true", help="skip the batched calibration (Bingham forward + adjoint)")
    ap.add_argument("--cal-draws", type=int, default=64)
    ap.add_argument("--cal-particles", type=int, default=1_000_000)
    ap.add_argument("--cal-steps", type=int, default=20,
                    help="steps of the bounded calibration sample (the configured run is 2,000)")
    ap.add_argument("--cal-ckpt", type=int, default=5)
    ap.add_argument("--cal-seq-sample", type=int, default=0,
                    help="sequential draws timed (0: all; else a sample, the total extrapolated)")
    ap.add_argument("--adj-steps", type=int, default=1000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding of realisations is a design note, not a built path)", file=sys.stderr)
        return 2

    from gfsph import SPHModel, cases

    if args.calibrate_copy:
        return calibrate_copy()

    t0 = time.time()
    case = cases.cfg_fractal_frictional(n=args.n, integrator=args.integrator, nsteps=args.run_steps)
    print(f"# workload generated in {time.time() - t0:.1f}s: n={case.particles.n} dem={case.dem.nx}x{case.dem.ny}",
          file=sys.stderr, flush=True)
    n = case.particles.n
    develop = args.run_steps // 2 if args.develop < 0 else args.develop

    # the whole configured run from t = 0 on a fresh context
    full = None
    if not args.no_full_run:
        with SPHModel(case) as w:  # first-use costs (graph capture, code load) off the clock
            w.time_integration_sph(16)
            w.sync()
        with SPHModel(case) as mf:
            mf.sync()
            t1 = time.perf_counter()
            mf.time_integration_sph(args.run_steps)
            mf.sync()
            tf = time.perf_counter() - t1
        full = {"steps": args.run_steps, "s": round(tf, 4), "ms_per_step": round(1e3 * tf / args.run_steps, 5),
                "particle_steps_per_s": round(n * args.run_steps / tf, 1)}
        print(f"# full run: {args.run_steps} steps in {tf:.3f}s", file=sys.stderr, flush=True)

    # the timed window: develop, warm up, then exactly --steps steps
    m = SPHModel(case)
    m.time_integration_sph(develop + args.warmup)
    m.sync()
    t1 = time.perf_counter()
    m.time_integration_sph(args.steps)
    m.sync()
    t2 = time.perf_counter()
    elapsed = t2 - t1
    value = n * args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    print(f"# timed steps {develop + args.warmup + 1}..{develop + args.warmup + args.steps} in {elapsed:.4f}s",
          file=sys.stderr, flush=True)
    window_state = None if args.no_cpu_baseline else (m.state(), m.mass())

    # per-kernel device time: the next --steps steps of the same context with
    # HIP events around every kernel group on the library's stream
    prof_steps = args.steps if args.profile_steps <= 0 else args.profile_steps
    prof = m.profile_step(nrep=prof_steps)
    st = m.state()
    n_active = int((st["flag"] == 0).sum())
    m.close()
    kern = {k: v for k, v in prof.items() if k in B_ALG and k != "neighbour_build"}
    dom = max(kern, key=kern.get)
    ms = kern[dom]
    achieved = B_ALG[dom] * n / (ms * 1e-3) / 1e9
    step_bytes = sum(B_ALG.values()) * n
    traffic = args.traffic
    if traffic is None and n == 200_000:
        traffic = MEASURED_TRAFFIC_200K.get(dom)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                "algorithmic_bytes_per_launch": B_ALG[dom] * n,
                "step_frac": round(step_bytes / (ms_per_step * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                "window": f"steps {develop + args.warmup + args.steps + 1}..{develop + args.warmup + args.steps + prof_steps}",
                "kernel_ms": {k: round(v, 4) for k, v in prof.items()},
                "gather_limits": GATHER_LIMITS_MC}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(case, window_state, develop + args.warmup, args.cpu_sample_steps)
    mc = None if args.no_mc else mc_hazard(args.mc_real, args.mc_steps, args.mc_seq_sample)
    adj = None if args.no_adjoint else adjoint_cost(args.adj_steps)
    cal = None if args.no_cal else calibration(args.cal_draws, args.cal_particles, args.cal_steps, args.cal_ckpt,
                                               args.cal_seq_sample)

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
        "config": {"workload": f"fractal DEM 2048x2048 @2m, 200k-particle elliptical release, frictional 28deg + "
                               f"pore pressure, RK2 dt=0.02 s, {args.run_steps} steps",
                   "particles": n, "active_after": n_active,
                   "timed_window": f"steps {develop + args.warmup + 1}..{develop + args.warmup + args.steps}",
                   "integrator": "RK2" if args.integrator == 2 else "RK4", "parallelism": "single GPU"},
        "full_run": full,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "mc": mc,
        "adjoint": adj,
        "calibration": cal,
    }
    print(json.dumps(out), flush=True)
    return 0


def calibration(R, n, nsteps, ckpt, nseq):
    """Batched adjoint calibration (fractal DEM 4096 x 4096 at 1 m, one
    release of n particles, R Bingham draws tau_y ~ U(500, 5000) Pa, mu_B ~
    U(10, 200) Pa s, RK2 dt 0.02 s): one batched forward + deposit misfit +
    backward over all draws, from context creation to the R gradients on the
    host, against the same for each draw run alone through the single-run
    adjoint.  The misfit target is the deposit of a forward run at (tau_y,
    mu_B) = (2000 Pa, 60 Pa s).  A bounded sample of ``nsteps`` steps of the
    configured 2,000 (the JSON says so); the checkpoints every ``ckpt``
    steps and one segment's records are what fit the GPU's memory for R x n
    particles at this horizon."""
    from gfsph import AdjointModel, BatchAdjointModel, SPHModel, cases

    t0 = time.perf_counter()
    cs = cases.cfg_calibration(R=R, n=n, nsteps=nsteps)
    print(f"# calibration: {R} draws x {n} particles, {nsteps} steps, workload {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    tprm = cases.Params(integrator=2, dt=cs[0].params.dt, h_inf=1.0e-3, const=cases.bingham(2000.0, 60.0))
    with SPHModel(cases.Case("target", tprm, cs[0].dem, cs[0].particles, nsteps)) as m:
        m.time_integration_sph(nsteps)
        T = m.raster_idw()[0][:, 0].copy()
    t0 = time.perf_counter()
    with BatchAdjointModel(cs) as b:
        b.forward(nsteps, ckpt)
        L, lx, lh = b.misfit(T)
        g = b.backward(lam_x=lx, lam_h=lh)
    t_batch = time.perf_counter() - t0
    print(f"# calibration batch {t_batch:.2f}s", file=sys.stderr, flush=True)
    picks = list(range(R)) if nseq <= 0 or nseq >= R else list(range(nseq))
    ts, check = [], []
    for r in picks:
        t1 = time.perf_counter()
        with AdjointModel(cs[r]) as a:
            a.forward(nsteps, ckpt)
            Lr, lxr, lhr = a.misfit(T)
            gr = a.backward(lam_x=lxr, lam_h=lhr)
        ts.append(time.perf_counter() - t1)
        if r < 3:
            check.append({"draw": r, "loss_rel": abs(L[r] - Lr) / abs(Lr),
                          "grad_tau_y_rel": abs(g["mu"][r] - gr["mu"]) / max(abs(gr["mu"]), 1e-300),
                          "grad_mu_B_rel": abs(g["xi"][r] - gr["xi"]) / max(abs(gr["xi"]), 1e-300)})
        if len(ts) % 16 == 0:
            print(f"# calibration sequential: {len(ts)} of {len(picks)} draws, {sum(ts):.1f}s", file=sys.stderr,
                  flush=True)
    measured = len(picks) == R
    t_seq = float(np.sum(ts)) if measured else float(np.mean(ts)) * R
    print(f"# calibration sequential {'measured' if measured else 'extrapolated'} {t_seq:.2f}s", file=sys.stderr,
          flush=True)
    return {"metric": "gradients/hr", "value": round(R * 3600.0 / t_batch, 1), "unit": "gradients/hr",
            "draws": R, "particles_per_draw": n, "particles": R * n, "steps": nsteps,
            "steps_configured": 2000, "sample": f"{nsteps} of the configured 2,000 steps (bounded sample)",
            "ckpt_every": ckpt, "batch_s": round(t_batch, 3), "sequential_s": round(t_seq, 3),
            "sequential_kind": "measured" if measured else "extrapolated", "sequential_runs": len(ts),
            "speedup_vs_sequential": round(t_seq / t_batch, 2), "batch_vs_single_check": check,
            "grad_tau_y_first": [float(v) for v in g["mu"][:3]], "grad_mu_B_first": [float(v) for v in g["xi"][:3]]}


def calibrate_copy() -> int:
    """A torch device-to-device copy of 2^25 doubles (256 MiB read + 256 MiB
    written): the FETCH_SIZE / WRITE_SIZE counters of this kernel give the
    scale between counter readings and bytes."""
    import torch

    a = torch.arange(1 << 25, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": a.numel() * 8, "copy_bytes_written": b.numel() * 8}))
    return 0


def mc_hazard(R, nsteps, nseq, plot_every=20):
    """Batched Monte Carlo hazard workload (fractal DEM 1024 x 1024 at 5 m,
    R realisations of 5k-80k particles, Voellmy mu ~ U(0.08, 0.35), xi ~
    LogNormal(ln 400, 0.5), RK2 dt 0.05 s, pixel statistics every
    ``plot_every`` steps, then the hazard rasters): the whole batch is timed
    from context creation to the rasters on the host; the sequential cost is
    the time of all R realisations run one by one through the plain drop-in
    (same steps and pixel accumulation), each from context creation to its
    pixel accumulators on the host (with 0 < ``nseq`` < R only a sample is
    run and the total is extrapolated by a least-squares line of run time
    against particle count; the JSON says which)."""
    from gfsph import BatchModel, SPHModel, cases

    cs = cases.cfg_mc_hazard(R=R, nsteps=nsteps)
    n = sum(c.particles.n for c in cs)
    print(f"# MC: {R} realisations, {n} particles, {nsteps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    b = BatchModel(cs)
    # the work-group sizes and lanes per particle are tuned once the flows
    # have developed (the first steps, from rest, favour other shapes)
    tune_at = min(100, nsteps)
    blocks = None
    done = 0
    while done < nsteps:
        if done == tune_at and blocks is None:
            blocks = b.autotune(reps=2)
        k = min(plot_every - done % plot_every, nsteps - done)
        if done < tune_at:
            k = min(k, tune_at - done)
        b.time_integration_sph(k)
        done += k
        if done % plot_every == 0:
            b.pixel_accumulate()
    hz = b.hazard(0.5)
    t_batch = time.perf_counter() - t0
    b.close()
    print(f"# MC batch {t_batch:.2f}s (blocks {blocks})", file=sys.stderr, flush=True)
    if nseq <= 0 or nseq >= R:
        picks = list(range(R))
    else:
        picks = sorted(np.random.default_rng(1).choice(R, size=nseq, replace=False).tolist())
    ts, ns = [], []
    for r in picks:
        ns.append(cs[r].particles.n)
        t1 = time.perf_counter()
        m = SPHModel(cs[r])
        for _ in range(nsteps // plot_every):
            m.time_integration_sph(plot_every)
            m.mc_pixel_accumulate()
        m.time_integration_sph(nsteps % plot_every)
        m.mc_pixel_read()
        m.close()
        ts.append(time.perf_counter() - t1)
        if len(ts) % 32 == 0:
            print(f"# MC sequential: {len(ts)} of {len(picks)} runs, {sum(ts):.1f}s", file=sys.stderr, flush=True)
    measured = len(picks) == R
    if measured:
        t_seq = float(np.sum(ts))
    elif len(ts) >= 3 and len(set(ns)) > 1:
        b1, b0 = np.polyfit(np.asarray(ns, float), np.asarray(ts), 1)
        t_seq = float(R * b0 + b1 * n)
    else:
        t_seq = float(np.mean(ts)) * R
    print(f"# MC sequential {'measured' if measured else 'extrapolated'} {t_seq:.2f}s from {len(ts)} runs",
          file=sys.stderr, flush=True)
    return {"metric": "MC realisations/hr", "value": round(R * 3600.0 / t_batch, 1), "unit": "realisations/hr",
            "realisations": R, "steps": nsteps, "particles": n, "batch_s": round(t_batch, 3),
            "particle_steps_per_s": round(n * nsteps / t_batch, 1),
            "sequential_s": round(t_seq, 3), "sequential_kind": "measured" if measured else "extrapolated",
            "sequential_runs": len(ts), "sequential_run_s_min_max": [round(min(ts), 4), round(max(ts), 4)],
            "speedup_vs_sequential": round(t_seq / t_batch, 2), "autotuned_blocks": list(blocks),
            "max_h": round(float(hz["max_h"].max()), 3), "exceed_nodes": int((hz["p_exceed"] > 0).sum())}


def adjoint_cost(nsteps, ckpt=50, fd_steps=50):
    """Adjoint back-analysis workload (fractal DEM 512 x 512 at 5 m, 50,000
    particles, Voellmy, RK2 dt 0.05 s): the time of the forward pass with
    checkpoints + deposit misfit + reverse sweep over the time of one plain
    forward pass of the same steps; then dL/dmu and dL/dxi against central
    differences (eps = 1e-6, relative, applied to the REAL(4) constants) over
    ``fd_steps`` steps, a horizon on which the difference quotient has
    converged (over long horizons the misfit of the windowed IDW raster
    jumps when a particle crosses a DEM cell, so no difference quotient
    converges there)."""
    from gfsph import AdjointModel, SPHModel, cases

    def target(n):
        with SPHModel(cases.cfg_adjoint(nsteps=n, mu=0.17, xi=650.0)) as m:
            m.time_integration_sph(n)
            return m.raster_idw()[0][:, 0].copy()

    case = cases.cfg_adjoint(nsteps=nsteps)
    T = target(nsteps)
    with SPHModel(case) as w:
        w.time_integration_sph(16)
        w.sync()
    m = SPHModel(case)
    m.sync()
    t0 = time.perf_counter()
    m.time_integration_sph(nsteps)
    m.sync()
    t_fwd = time.perf_counter() - t0
    m.close()
    a = AdjointModel(case)
    a.sync()
    t0 = time.perf_counter()
    a.forward(nsteps, ckpt)
    L, lx, lh = a.misfit(T)
    g = a.backward(lam_x=lx, lam_h=lh)
    t_adj = time.perf_counter() - t0
    a.close()
    # difference check on the short horizon
    Ts = target(fd_steps)
    short = cases.cfg_adjoint(nsteps=fd_steps)
    with AdjointModel(short) as a2:
        a2.forward(fd_steps, ckpt)
        _, lx2, lh2 = a2.misfit(Ts)
        g2 = a2.backward(lam_x=lx2, lam_h=lh2)

    def loss(c):
        with SPHModel(c) as mm:
            mm.time_integration_sph(fd_steps)
            out = mm.raster_idw()[0][:, 0]
        return float(((out - Ts) ** 2).sum())

    fd = {}
    for k, name in ((9, "mu"), (3, "xi")):
        v = short.params.const[k]
        hi, lo = cases.f32(v * (1 + 1e-6)), cases.f32(v * (1 - 1e-6))
        vals = []
        for val in (hi, lo):
            cst = dict(short.params.const)
            cst[k] = val
            prm = cases.Params(**{**short.params.__dict__, "const": cst})
            vals.append(loss(cases.Case(short.name, prm, short.dem, short.particles, fd_steps)))
        q = (vals[0] - vals[1]) / (hi - lo)
        fd[name] = {"adjoint": g2[name], "fd": q, "rel": abs(g2[name] - q) / abs(q)}
    print(f"# adjoint: forward {t_fwd:.3f}s, forward+adjoint {t_adj:.3f}s", file=sys.stderr, flush=True)
    return {"metric": "adjoint/forward cost ratio", "value": round(t_adj / t_fwd, 3), "unit": "x",
            "higher_is_better": False, "steps": nsteps, "ckpt_every": ckpt, "particles": case.particles.n,
            "forward_s": round(t_fwd, 4), "forward_plus_adjoint_s": round(t_adj, 4), "loss": L,
            "grad_mu": g["mu"], "grad_xi": g["xi"], "fd_check_steps": fd_steps, "fd_check": fd}


def cpu_baseline(case, window_state, start_step, nsteps):
    """The oracle (serial C restatement of the reference's step, 1 thread) on
    a bounded sample of the same workload: all particles, ``nsteps`` steps
    from the state the GPU run reached at the start of the timed window (so
    both time the same developed flow)."""
    try:
        from oracle import oracle
    except Exception as e:  # the oracle ships with the repository
        return {"error": str(e)}
    st = oracle.init(case)
    if window_state is not None:
        g, mass = window_state
        st.x, st.v, st.h, st.u, st.hsml = (np.ascontiguousarray(g[k], dtype=np.float64).copy()
                                           for k in ("x", "v", "h", "u", "hsml"))
        st.uaux = np.ascontiguousarray(g["uaux"], dtype=np.float64).copy()
        st.flag = np.ascontiguousarray(g["flag"], dtype=np.int32).copy()
        st.mass = np.ascontiguousarray(mass, dtype=np.float64).copy()
    t = time.perf_counter()
    oracle.run(case, st, nsteps)
    dt = time.perf_counter() - t
    where = f"steps {start_step + 1}..{start_step + nsteps}" if window_state is not None else f"steps 1..{nsteps}"
    return {"value": round(case.particles.n * nsteps / dt, 1), "unit": "particle-steps/s", "cores": 1,
            "kind": "port", "sample": f"{case.particles.n} particles x {nsteps} steps ({where}, the GPU run's "
                                      f"state at step {start_step} as input; oracle, serial C FP64)",
            "seconds": round(dt, 3), "cpu": cpu_info()}


if __name__ == "__main__":
    sys.exit(main())
